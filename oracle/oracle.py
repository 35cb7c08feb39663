"""TEST INFRASTRUCTURE ONLY: ctypes binding of the CPU oracle (mk_oracle.cpp).

Imported by tests/, ``__graft_entry__.smoke()`` and the cpu_baseline leg of
bench.py, always as the checker. The product never imports this module.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from monkeygl_amd._capi import MK_NUM_LABELS, PlaneParams, RenderParams

LIB = Path(__file__).resolve().parent / "_build" / "libmk_oracle.so"
_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} is not built; run `python -m monkeygl_amd.build`")
        L = C.CDLL(str(LIB))
        vp, i = C.c_void_p, C.c_int
        L.mko_set_threads.restype = i
        L.mko_set_threads.argtypes = [i]
        L.mko_render.restype = i
        L.mko_render.argtypes = [C.POINTER(RenderParams), vp, i, i, i, vp, vp, vp, i, i, vp, vp]
        L.mko_render_rows.restype = i
        L.mko_render_rows.argtypes = [C.POINTER(RenderParams), vp, i, i, i, vp, vp, vp, i, i, i, i, vp, vp]
        L.mko_render_plane.restype = i
        L.mko_render_plane.argtypes = [C.POINTER(PlaneParams), vp, i, i, i, i, i, vp, vp]
        L.mko_render_cpr.restype = i
        L.mko_render_cpr.argtypes = [i, i, vp, vp, i, i, vp, i, i, i, i, i, vp, vp]
        L.mko_marching_cubes.restype = C.c_uint64
        L.mko_marching_cubes.argtypes = [vp, i, i, i, vp, C.c_float, vp, C.c_uint64]
        L.mko_fit_render.restype = i
        L.mko_fit_render.argtypes = [C.POINTER(RenderParams), vp, i, i, i, vp, i, C.c_double, vp]
        L.mko_fit_render_rows.restype = i
        L.mko_fit_render_rows.argtypes = [C.POINTER(RenderParams), vp, i, i, i, vp, i, C.c_double, i, i, vp]
        L.mko_textout.restype = i
        L.mko_textout.argtypes = [C.c_char_p, i, i, i, i, vp, vp, vp, vp, vp, vp, i, i]
        L.mko_mc_table.restype = None
        L.mko_mc_table.argtypes = [vp, vp]
        L.mko_tex3_linear.restype = C.c_double
        L.mko_tex3_linear.argtypes = [vp, i, i, i, C.c_double, C.c_double, C.c_double]
        L.mko_tex1_linear.restype = i
        L.mko_tex1_linear.argtypes = [vp, i, C.c_double, vp]
        _lib = L
    return _lib


def build_native(out_dir) -> Path:
    """Compiles mk_oracle.cpp for the host this runs on (-O3 -march=native
    -ffp-contract=off -fopenmp) into out_dir, for timing the CPU baseline. The
    library built with the package targets x86-64-v3, since it is built on one
    machine and run on another. Raises when the compiler fails."""
    import subprocess

    out = Path(out_dir) / "libmk_oracle_native.so"
    src = Path(__file__).resolve().parent / "mk_oracle.cpp"
    inc = Path(__file__).resolve().parent.parent / "include"
    subprocess.run(["g++", "-O3", "-march=native", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared", "-std=c++17",
                    "-I", str(inc), str(src), "-o", str(out)], check=True, capture_output=True)
    return out


def use_library(path=None) -> None:
    """Switches the binding to another build of mk_oracle.cpp (None: the one
    built with the package); takes effect at the next call."""
    global LIB, _lib
    LIB = Path(path) if path else Path(__file__).resolve().parent / "_build" / "libmk_oracle.so"
    _lib = None


def set_threads(n: int) -> int:
    return lib().mko_set_threads(n)


def _luts(luts: dict):
    """luts: {label: float32 array (n,4)} -> (pointer array, length array, keepalive)."""
    keep = []
    ptrs = (C.c_void_p * MK_NUM_LABELS)()
    lens = (C.c_int * MK_NUM_LABELS)()
    for l, a in luts.items():
        arr = np.ascontiguousarray(a, dtype=np.float32).reshape(-1, 4)
        keep.append(arr)
        ptrs[l] = arr.ctypes.data
        lens[l] = arr.shape[0]
    return ptrs, lens, keep


def render(p: RenderParams, vol_zyx: np.ndarray, luts: dict, mask_zyx=None, precision: str = "f64",
           weights8: bool = False, touched: bool = False):
    v = np.ascontiguousarray(vol_zyx, dtype=np.int16)
    nz, ny, nx = v.shape
    m = None if mask_zyx is None else np.ascontiguousarray(mask_zyx, dtype=np.uint8)
    ptrs, lens, keep = _luts(luts)
    out = np.zeros((p.height, p.width, 3), np.uint8)
    t = np.zeros(v.shape, np.uint8) if touched else None
    rc = lib().mko_render(C.byref(p), v.ctypes.data, nx, ny, nz, None if m is None else m.ctypes.data,
                          C.cast(ptrs, C.c_void_p), C.cast(lens, C.c_void_p), 0 if precision == "f64" else 1,
                          int(weights8), out.ctypes.data, None if t is None else t.ctypes.data)
    assert rc == 0
    return (out, t) if touched else out


def render_rows(p: RenderParams, vol_zyx: np.ndarray, luts: dict, y0: int, y1: int, mask_zyx=None,
                precision: str = "f64", weights8: bool = False):
    """Rows [y0, y1) of the frame p describes (same rays as the full frame)."""
    v = np.ascontiguousarray(vol_zyx, dtype=np.int16)
    nz, ny, nx = v.shape
    m = None if mask_zyx is None else np.ascontiguousarray(mask_zyx, dtype=np.uint8)
    ptrs, lens, keep = _luts(luts)
    out = np.zeros((y1 - y0, p.width, 3), np.uint8)
    rc = lib().mko_render_rows(C.byref(p), v.ctypes.data, nx, ny, nz, None if m is None else m.ctypes.data,
                               C.cast(ptrs, C.c_void_p), C.cast(lens, C.c_void_p), 0 if precision == "f64" else 1,
                               int(weights8), y0, y1, out.ctypes.data, None)
    assert rc == 0
    return out


def render_plane(p: PlaneParams, vol_zyx: np.ndarray, precision: str = "f64", weights8: bool = False):
    v = np.ascontiguousarray(vol_zyx, dtype=np.int16)
    nz, ny, nx = v.shape
    out = np.zeros((p.height, p.width), np.int16)
    rc = lib().mko_render_plane(C.byref(p), v.ctypes.data, nx, ny, nz, 0 if precision == "f64" else 1,
                                int(weights8), out.ctypes.data, None)
    assert rc == 0
    return out


def render_cpr(width: int, points, dirs, invert_z: bool, vol_zyx, sampling: int = 0, precision: str = "f64",
               weights8: bool = False):
    v = np.ascontiguousarray(vol_zyx, dtype=np.int16)
    nz, ny, nx = v.shape
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    drs = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
    out = np.zeros((pts.shape[0], width), np.int16)
    rc = lib().mko_render_cpr(width, pts.shape[0], pts.ctypes.data, drs.ctypes.data, int(invert_z), sampling,
                              v.ctypes.data, nx, ny, nz, 0 if precision == "f64" else 1, int(weights8),
                              out.ctypes.data, None)
    assert rc == 0
    return out


def tex3_linear(vol_zyx, fx: float, fy: float, fz: float) -> float:
    v = np.ascontiguousarray(vol_zyx, dtype=np.int16)
    nz, ny, nx = v.shape
    return lib().mko_tex3_linear(v.ctypes.data, nx, ny, nz, fx, fy, fz)


def tex1_linear(lut: np.ndarray, s: float) -> np.ndarray:
    a = np.ascontiguousarray(lut, dtype=np.float32).reshape(-1, 4)
    out = np.zeros(4, np.float64)
    lib().mko_tex1_linear(a.ctypes.data, a.shape[0], s, out.ctypes.data)
    return out


def marching_cubes(vol_zyx, box, iso: float) -> np.ndarray:
    """Triangles (n, 3, 3) float32, same contract as mk_marching_cubes."""
    v = np.ascontiguousarray(vol_zyx, dtype=np.int16)
    nz, ny, nx = v.shape
    b = np.ascontiguousarray(box, dtype=np.int32)
    n = lib().mko_marching_cubes(v.ctypes.data, nx, ny, nz, b.ctypes.data, iso, None, 0)
    out = np.empty((n, 3, 3), np.float32)
    if n:
        lib().mko_marching_cubes(v.ctypes.data, nx, ny, nz, b.ctypes.data, iso, out.ctypes.data, n)
    return out


def mc_table():
    tri = np.zeros((256, 16), np.int8)
    cnt = np.zeros(256, np.uint8)
    lib().mko_mc_table(tri.ctypes.data, cnt.ctypes.data)
    return tri, cnt


def fit_render(p: RenderParams, vol_zyx, lut, scale: float) -> np.ndarray:
    """Fitting-rule image (height, width, 3) in float64 (FP64 restatement of mk_tf_forward)."""
    v = np.ascontiguousarray(vol_zyx, dtype=np.int16)
    nz, ny, nx = v.shape
    l = np.ascontiguousarray(lut, dtype=np.float64).reshape(-1, 4)
    out = np.zeros((p.height, p.width, 3), np.float64)
    rc = lib().mko_fit_render(C.byref(p), v.ctypes.data, nx, ny, nz, l.ctypes.data, l.shape[0], float(scale),
                              out.ctypes.data)
    assert rc == 0
    return out


def fit_render_rows(p: RenderParams, vol_zyx, lut, scale: float, y0: int, y1: int) -> np.ndarray:
    """Rows [y0, y1) of fit_render's image, (y1 - y0, width, 3) float64."""
    v = np.ascontiguousarray(vol_zyx, dtype=np.int16)
    nz, ny, nx = v.shape
    l = np.ascontiguousarray(lut, dtype=np.float64).reshape(-1, 4)
    out = np.zeros((y1 - y0, p.width, 3), np.float64)
    rc = lib().mko_fit_render_rows(C.byref(p), v.ctypes.data, nx, ny, nz, l.ctypes.data, l.shape[0], float(scale),
                                   int(y0), int(y1), out.ctypes.data)
    assert rc == 0
    return out


def glyph_table(glyphs: dict):
    """{(font, byte): (h, w) uint8 mask} -> (widths, heights, offsets, atlas)
    arrays over font * 256 + byte; -1 where there is no glyph."""
    gw = np.full(3 * 256, -1, np.int32)
    gh = np.full(3 * 256, -1, np.int32)
    go = np.zeros(3 * 256, np.int64)
    parts, off = [], 0
    for (font, code), m in sorted(glyphs.items()):
        m = np.ascontiguousarray(m, np.uint8)
        k = font * 256 + code
        gh[k], gw[k] = m.shape
        go[k] = off
        parts.append(m.reshape(-1))
        off += m.size
    atlas = np.concatenate(parts) if parts else np.zeros(1, np.uint8)
    return gw, gh, go, np.ascontiguousarray(atlas)


def textout(img: np.ndarray, text: bytes, x: int, y: int, font: int, fmt: int, color, table) -> bool:
    """AnnotationUtils::Textout2Image on an (h, w, 3) uint8 image in place."""
    gw, gh, go, atlas = table
    c = np.ascontiguousarray(color, np.float32).reshape(3)
    assert img.dtype == np.uint8 and img.flags["C_CONTIGUOUS"]
    h, w = img.shape[:2]
    return bool(lib().mko_textout(bytes(text), int(x), int(y), int(font), int(fmt), c.ctypes.data, gw.ctypes.data,
                                  gh.ctypes.data, go.ctypes.data, atlas.ctypes.data, img.ctypes.data, w, h))
