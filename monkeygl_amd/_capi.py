"""ctypes binding of the C ABI in ``include/monkeygl_hip.h``.

The structs mirror the header field for field. ``load()`` raises if the HIP
library has not been built: there is no CPU fallback anywhere in the product.
"""
from __future__ import annotations

import ctypes as C
import importlib.util
import sys
import weakref
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libmonkeygl_hip.so"

MK_MAX_OBJECTS = 64
MK_NUM_LABELS = MK_MAX_OBJECTS + 1
MK_MAX_DIM = 131072

MK_OK, MK_ERR_ARG, MK_ERR_STATE, MK_ERR_HIP, MK_ERR_NOMEM, MK_ERR_RANGE = range(6)
MK_RENDER_VR, MK_RENDER_MIP, MK_RENDER_SURFACE = range(3)
MK_MPR_AVERAGE, MK_MPR_MIP, MK_MPR_MINIP = range(3)
MK_SAMPLE_LINEAR, MK_SAMPLE_NEAREST = range(2)
MK_ITEM_RENDER, MK_ITEM_PLANE, MK_ITEM_CPR = range(3)
MK_MASK_ADD, MK_MASK_UPDATE = range(2)

# exported symbols, in header order
SYMBOLS = [
    "mk_device_count", "mk_context_create", "mk_context_destroy", "mk_last_error",
    "mk_context_stream", "mk_synchronize", "mk_device_alloc", "mk_device_free", "mk_copy_to_host",
    "mk_copy_to_device", "mk_copy_volume_data", "mk_copy_volume_data_device",
    "mk_copy_mask_data", "mk_copy_mask_data_device", "mk_clear_mask", "mk_merge_object_mask",
    "mk_merge_object_mask_device", "mk_get_mask_data", "mk_origin_slice_async", "mk_set_transfer_func",
    "mk_render", "mk_render_async", "mk_set_vr_trip_budget", "mk_render_footprint", "mk_render_plane", "mk_render_plane_async",
    "mk_render_cpr", "mk_render_cpr_async", "mk_marching_cubes", "mk_marching_cubes_async", "mk_marching_cubes_noalloc",
    "mk_tf_forward", "mk_tf_backward", "mk_set_orientation_mesh", "mk_orientation_workspace_bytes",
    "mk_merge_orientation_box_async", "mk_draw_polyline_async", "mk_set_glyph", "mk_draw_text_async", "mk_png_max_bytes", "mk_png_workspace_bytes", "mk_png_encode_async",
    "mk_base64_max_bytes", "mk_base64_encode_async", "mk_batcher_create", "mk_batcher_destroy", "mk_batcher_last_error",
    "mk_batch_submit", "mk_batch_query", "mk_batch_wait",
]


class AlphaWWWL(C.Structure):
    _fields_ = [("alpha", C.c_float), ("ww", C.c_float), ("wl", C.c_float)]


class VOI(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("left", "right", "posterior", "anterior", "head", "foot")]


class RenderParams(C.Structure):
    _fields_ = [
        ("width", C.c_int), ("height", C.c_int), ("render_type", C.c_int), ("invert_z", C.c_int),
        ("max_len_spacing", C.c_float * 3), ("spacing", C.c_float * 3), ("spacing_voxel", C.c_float * 3),
        ("x_translate", C.c_float), ("y_translate", C.c_float), ("scale", C.c_float),
        ("transform", C.c_float * 9), ("voi", VOI), ("background", C.c_float * 4),
        ("alpha_wwwl", AlphaWWWL * MK_NUM_LABELS), ("step_scale", C.c_float), ("exit_alpha", C.c_float),
    ]


class PlaneParams(C.Structure):
    _fields_ = [
        ("width", C.c_int), ("height", C.c_int), ("mpr_type", C.c_int), ("sampling", C.c_int),
        ("invert_z", C.c_int), ("spacing", C.c_float * 3), ("dir_h", C.c_float * 3), ("dir_v", C.c_float * 3),
        ("dir_n", C.c_float * 3), ("pt_left_top", C.c_float * 3), ("pixel_spacing", C.c_float),
        ("half_num", C.c_float),
    ]


class BatchItem(C.Structure):
    _fields_ = [
        ("kind", C.c_int), ("context", C.c_int), ("out_offset", C.c_uint64), ("render", RenderParams),
        ("plane", PlaneParams), ("cpr_width", C.c_int), ("cpr_height", C.c_int), ("cpr_invert_z", C.c_int),
        ("cpr_sampling", C.c_int), ("cpr_points_offset", C.c_uint64), ("cpr_dirs_offset", C.c_uint64),
    ]


class MkError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"mk status {status}: {msg}")
        self.status = status


_lib = None


def load() -> C.CDLL:
    """Loads the HIP library; raises when it is missing (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is not built; run `python -m monkeygl_amd.build`")
    # One HIP runtime per process. torch ships its own copy of the runtime with
    # the same SONAME (libamdhip64.so.7) as ROCm's, and asks for it as
    # libamdhip64.so: if this library bound ROCm's copy, a later `import torch`
    # would load a second runtime that cannot see the GPU. So when torch is
    # installed and not yet imported, its copy is loaded first, by path,
    # without importing torch: this library binds to it by SONAME, and torch,
    # when imported, finds the same file already loaded. Without torch the
    # library binds ROCm's runtime.
    if "torch" not in sys.modules:
        try:
            spec = importlib.util.find_spec("torch")
        except (ImportError, ValueError):
            spec = None
        if spec is not None and spec.origin:
            rt = Path(spec.origin).resolve().parent / "lib" / "libamdhip64.so"
            if rt.exists():
                C.CDLL(str(rt), mode=C.RTLD_GLOBAL)
    lib = C.CDLL(str(LIB_PATH))
    vp, i, st = C.c_void_p, C.c_int, C.c_int
    sig = {
        "mk_device_count": (i, []),
        "mk_context_create": (st, [i, C.POINTER(vp)]),
        "mk_context_destroy": (None, [vp]),
        "mk_last_error": (C.c_char_p, [vp]),
        "mk_context_stream": (vp, [vp]),
        "mk_synchronize": (st, [vp]),
        "mk_device_alloc": (st, [vp, C.c_uint64, C.POINTER(vp)]),
        "mk_device_free": (st, [vp, vp]),
        "mk_copy_to_host": (st, [vp, vp, vp, C.c_uint64]),
        "mk_copy_to_device": (st, [vp, vp, vp, C.c_uint64]),
        "mk_copy_volume_data": (st, [vp, vp, i, i, i]),
        "mk_copy_volume_data_device": (st, [vp, vp, i, i, i]),
        "mk_copy_mask_data": (st, [vp, vp, i, i, i]),
        "mk_copy_mask_data_device": (st, [vp, vp, i, i, i]),
        "mk_clear_mask": (st, [vp]),
        "mk_set_orientation_mesh": (st, [vp, vp, i]),
        "mk_orientation_workspace_bytes": (C.c_uint64, [vp]),
        "mk_merge_orientation_box_async": (st, [vp, vp, vp, i, i, vp, C.c_uint64, vp]),
        "mk_draw_polyline_async": (st, [vp, vp, i, vp, vp, i, i, vp]),
        "mk_set_glyph": (st, [vp, i, i, i, i, vp]),
        "mk_draw_text_async": (st, [vp, C.c_char_p, i, i, i, i, vp, vp, i, i, vp]),
        "mk_merge_object_mask": (st, [vp, vp, i, i, i, i, i]),
        "mk_merge_object_mask_device": (st, [vp, vp, i, i, i, i, i]),
        "mk_get_mask_data": (st, [vp, vp]),
        "mk_origin_slice_async": (st, [vp, i, vp, vp]),
        "mk_set_transfer_func": (st, [vp, i, vp, i]),
        "mk_render": (st, [vp, C.POINTER(RenderParams), vp]),
        "mk_render_async": (st, [vp, C.POINTER(RenderParams), vp, vp]),
        "mk_render_footprint": (st, [vp, C.POINTER(RenderParams), vp, vp, vp, vp]),
        "mk_set_vr_trip_budget": (st, [vp, C.c_int]),
        "mk_render_plane": (st, [vp, C.POINTER(PlaneParams), vp]),
        "mk_render_plane_async": (st, [vp, C.POINTER(PlaneParams), vp, vp]),
        "mk_render_cpr": (st, [vp, i, i, vp, vp, i, i, vp]),
        "mk_render_cpr_async": (st, [vp, i, i, vp, vp, i, i, vp, vp]),
        "mk_marching_cubes": (st, [vp, vp, C.c_float, vp, C.c_uint64, C.POINTER(C.c_uint64)]),
        "mk_marching_cubes_async": (st, [vp, vp, C.c_float, vp, C.c_uint64, vp, vp]),
        "mk_marching_cubes_noalloc": (st, [vp, vp, C.c_float, vp, C.c_uint64, vp, vp]),
        "mk_tf_forward": (st, [vp, C.POINTER(RenderParams), vp, i, C.c_float, vp, vp, vp]),
        "mk_tf_backward": (st, [vp, C.POINTER(RenderParams), vp, i, C.c_float, vp, vp, vp, vp, vp]),
        "mk_png_max_bytes": (C.c_uint64, [i, i, i]),
        "mk_png_workspace_bytes": (C.c_uint64, [i, i, i]),
        "mk_png_encode_async": (st, [vp, i, i, i, vp, C.c_uint64, vp, C.c_uint64, vp, vp]),
        "mk_base64_max_bytes": (C.c_uint64, [C.c_uint64]),
        "mk_base64_encode_async": (st, [vp, vp, C.c_uint64, vp, C.c_uint64, vp, vp]),
        "mk_batcher_create": (st, [C.POINTER(vp), i, i, i, i, C.POINTER(vp)]),
        "mk_batcher_destroy": (None, [vp]),
        "mk_batcher_last_error": (C.c_char_p, [vp]),
        "mk_batch_submit": (st, [vp, i, C.POINTER(BatchItem), i, vp, C.c_uint64, vp]),
        "mk_batch_query": (i, [vp, i]),
        "mk_batch_wait": (st, [vp, i]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a) -> int:
    if a is None:
        return 0
    if isinstance(a, int):
        return a
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return int(a.data_ptr())  # DeviceBuffer, or a device array that exposes data_ptr()


def _nbytes(a) -> int:
    if isinstance(a, DeviceBuffer):
        return a.nbytes
    return int(a.numel() * a.element_size())


class DeviceBuffer:
    """Device memory of a context's device (mk_device_alloc), for callers that
    use no device-array library. Copies run on the context's stream and
    return when done, after the work queued before them there. The context
    frees its live buffers before it is destroyed."""

    def __init__(self, ctx: "Context", nbytes: int):
        if nbytes <= 0:
            raise ValueError("a device buffer needs at least one byte")
        self.ctx = ctx
        self.nbytes = int(nbytes)
        h = C.c_void_p()
        ctx._check(ctx.lib.mk_device_alloc(ctx.h, self.nbytes, C.byref(h)))
        self.ptr = int(h.value)
        ctx._buffers.add(self)
        self._keep = None   # device or host objects a queued call still reads

    def data_ptr(self) -> int:
        return self.ptr

    def numel(self) -> int:
        return self.nbytes

    def element_size(self) -> int:
        return 1

    def to_host(self, dtype=np.uint8, count: int | None = None, offset: int = 0) -> np.ndarray:
        dt = np.dtype(dtype)
        n = (self.nbytes - offset) // dt.itemsize if count is None else int(count)
        out = np.empty(n, dt)
        if n:
            if offset + n * dt.itemsize > self.nbytes:
                raise ValueError("read past the end of the device buffer")
            self.ctx._check(self.ctx.lib.mk_copy_to_host(self.ctx.h, out.ctypes.data, self.ptr + offset,
                                                         n * dt.itemsize))
        return out

    def from_host(self, arr: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(arr)
        if a.nbytes > self.nbytes:
            raise ValueError("host array larger than the device buffer")
        self.ctx._check(self.ctx.lib.mk_copy_to_device(self.ctx.h, self.ptr, a.ctypes.data, a.nbytes))
        return self

    @classmethod
    def of(cls, ctx: "Context", arr: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(arr)
        return cls(ctx, max(a.nbytes, 1)).from_host(a)

    def close(self) -> None:
        if self.ptr and self.ctx.h:
            self.ctx.lib.mk_device_free(self.ctx.h, self.ptr)
            self.ctx._buffers.discard(self)
        self.ptr = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """One render context: resident volume, mask and LUTs on one device."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        st = self.lib.mk_context_create(device, C.byref(h))
        if st != MK_OK:
            raise MkError(st, self.lib.mk_last_error(None).decode())
        self.h = h
        self.shape = None
        self._buffers = weakref.WeakSet()

    def close(self) -> None:
        if getattr(self, "h", None):
            for b in list(self._buffers):
                b.close()
            self.lib.mk_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st: int) -> None:
        if st != MK_OK:
            raise MkError(st, self.lib.mk_last_error(self.h).decode())

    @property
    def stream(self) -> int:
        return self.lib.mk_context_stream(self.h) or 0

    def synchronize(self) -> None:
        self._check(self.lib.mk_synchronize(self.h))

    def set_volume(self, vol_zyx: np.ndarray) -> None:
        """vol_zyx: int16 array indexed [z, y, x] (x fastest in memory)."""
        v = np.ascontiguousarray(vol_zyx, dtype=np.int16)
        nz, ny, nx = v.shape
        self._check(self.lib.mk_copy_volume_data(self.h, _ptr(v), nx, ny, nz))
        self.shape = (nz, ny, nx)

    def set_volume_device(self, t, shape_zyx) -> None:
        nz, ny, nx = shape_zyx
        self._check(self.lib.mk_copy_volume_data_device(self.h, _ptr(t), nx, ny, nz))
        self.shape = (nz, ny, nx)

    def set_mask(self, mask_zyx: np.ndarray) -> None:
        m = np.ascontiguousarray(mask_zyx, dtype=np.uint8)
        nz, ny, nx = m.shape
        self._check(self.lib.mk_copy_mask_data(self.h, _ptr(m), nx, ny, nz))

    def clear_mask(self) -> None:
        self._check(self.lib.mk_clear_mask(self.h))

    def merge_mask(self, obj_zyx: np.ndarray, label: int, mode: int = 0) -> None:
        """mk_merge_object_mask: obj_zyx nonzero = inside (uint8, the volume's
        shape); MK_MASK_ADD or MK_MASK_UPDATE."""
        m = np.ascontiguousarray(obj_zyx, dtype=np.uint8)
        nz, ny, nx = m.shape
        self._check(self.lib.mk_merge_object_mask(self.h, _ptr(m), nx, ny, nz, int(label), int(mode)))

    def merge_mask_device(self, t, shape_zyx, label: int, mode: int = 0) -> None:
        nz, ny, nx = shape_zyx
        self._check(self.lib.mk_merge_object_mask_device(self.h, _ptr(t), nx, ny, nz, int(label), int(mode)))

    def get_mask(self) -> np.ndarray:
        """The label mask as a (z, y, x) uint8 array."""
        out = np.empty(self.shape, np.uint8)
        self._check(self.lib.mk_get_mask_data(self.h, _ptr(out)))
        return out

    def set_orientation_mesh(self, stl: bytes | None) -> int:
        """A binary STL file's bytes (or None to remove the figure); returns
        the facet count."""
        if not stl:
            self._check(self.lib.mk_set_orientation_mesh(self.h, None, 0))
            return 0
        n = int.from_bytes(stl[80:84], "little")
        if n <= 0 or len(stl) < 84 + 50 * n:
            raise ValueError("not a binary STL")
        body = np.frombuffer(stl, np.uint8, 50 * n, 84)
        self._check(self.lib.mk_set_orientation_mesh(self.h, _ptr(np.ascontiguousarray(body)), n))
        return n

    def orientation_workspace_bytes(self) -> int:
        return int(self.lib.mk_orientation_workspace_bytes(self.h))

    def merge_orientation_box_async(self, rotate9, dev_rgb, width: int, height: int, dev_workspace,
                                    stream: int = 0) -> None:
        m = np.ascontiguousarray(rotate9, dtype=np.float32).reshape(9)
        ws = _nbytes(dev_workspace)
        self._check(self.lib.mk_merge_orientation_box_async(self.h, _ptr(m), _ptr(dev_rgb), width, height,
                                                            _ptr(dev_workspace), ws, stream))

    def draw_polyline_async(self, dev_xy, n: int, color3, dev_rgb, width: int, height: int, stream: int = 0) -> None:
        c = np.ascontiguousarray(color3, dtype=np.float32).reshape(3)
        self._check(self.lib.mk_draw_polyline_async(self.h, _ptr(dev_xy), int(n), _ptr(c), _ptr(dev_rgb), width, height,
                                                    stream))

    def set_glyph(self, font_size: int, code: int, mask) -> None:
        """mk_set_glyph: mask is a (height, width) uint8 array, or None to remove the glyph."""
        if mask is None:
            self._check(self.lib.mk_set_glyph(self.h, int(font_size), int(code), 0, 0, None))
            return
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        self._check(self.lib.mk_set_glyph(self.h, int(font_size), int(code), m.shape[1], m.shape[0], _ptr(m)))

    def draw_text_async(self, text: bytes, x: int, y: int, font_size: int, fmt: int, color3, dev_rgb, width: int,
                        height: int, stream: int = 0) -> None:
        c = np.ascontiguousarray(color3, dtype=np.float32).reshape(3)
        self._check(self.lib.mk_draw_text_async(self.h, bytes(text), int(x), int(y), int(font_size), int(fmt),
                                                _ptr(c), _ptr(dev_rgb), int(width), int(height), stream))

    def origin_slice_async(self, z: int, dev_out, stream: int = 0) -> None:
        self._check(self.lib.mk_origin_slice_async(self.h, int(z), _ptr(dev_out), stream))

    def set_lut(self, label: int, lut: np.ndarray) -> None:
        a = np.ascontiguousarray(lut, dtype=np.float32).reshape(-1, 4)
        self._check(self.lib.mk_set_transfer_func(self.h, label, _ptr(a), a.shape[0]))

    def render(self, p: RenderParams) -> np.ndarray:
        out = np.empty((p.height, p.width, 3), np.uint8)
        self._check(self.lib.mk_render(self.h, C.byref(p), _ptr(out)))
        return out

    def render_async(self, p: RenderParams, dev_out, stream: int = 0) -> None:
        self._check(self.lib.mk_render_async(self.h, C.byref(p), _ptr(dev_out), stream))

    def set_vr_trip_budget(self, trips: int) -> None:
        """Loop trips of the first VR pass (0: one pass, -1: chosen per frame); the frame is the same for every value."""
        self._check(self.lib.mk_set_vr_trip_budget(self.h, int(trips)))

    def render_footprint(self, p: RenderParams, dev_out, dev_touch_vol, dev_touch_mask=None, stream: int = 0) -> None:
        self._check(self.lib.mk_render_footprint(self.h, C.byref(p), _ptr(dev_out), _ptr(dev_touch_vol),
                                                 _ptr(dev_touch_mask), stream))

    def render_plane(self, p: PlaneParams) -> np.ndarray:
        out = np.empty((p.height, p.width), np.int16)
        self._check(self.lib.mk_render_plane(self.h, C.byref(p), _ptr(out)))
        return out

    def render_plane_async(self, p: PlaneParams, dev_out, stream: int = 0) -> None:
        self._check(self.lib.mk_render_plane_async(self.h, C.byref(p), _ptr(dev_out), stream))

    def render_cpr(self, width: int, points: np.ndarray, dirs: np.ndarray, invert_z: bool,
                   sampling: int = MK_SAMPLE_LINEAR) -> np.ndarray:
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        drs = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
        h = pts.shape[0]
        out = np.empty((h, width), np.int16)
        self._check(self.lib.mk_render_cpr(self.h, width, h, _ptr(pts), _ptr(drs), int(invert_z), sampling, _ptr(out)))
        return out

    def render_cpr_async(self, width: int, height: int, dev_points, dev_dirs, invert_z: bool, sampling: int,
                         dev_out, stream: int = 0) -> None:
        self._check(self.lib.mk_render_cpr_async(self.h, width, height, _ptr(dev_points), _ptr(dev_dirs),
                                                 int(invert_z), sampling, _ptr(dev_out), stream))

    def marching_cubes(self, box, iso: float, capacity: int | None = None) -> np.ndarray:
        """Triangles (n, 3, 3) float32 in voxel units; box = (x0, y0, z0, x1, y1, z1)."""
        b = np.ascontiguousarray(box, dtype=np.int32)
        n = C.c_uint64(0)
        self._check(self.lib.mk_marching_cubes(self.h, _ptr(b), iso, None, 0, C.byref(n)))
        cap = n.value if capacity is None else min(capacity, n.value)
        out = np.empty((cap, 3, 3), np.float32)
        if cap:
            self._check(self.lib.mk_marching_cubes(self.h, _ptr(b), iso, _ptr(out), cap, C.byref(n)))
        return out

    def marching_cubes_async(self, box, iso: float, dev_out, capacity: int, dev_count, stream: int = 0) -> None:
        b = np.ascontiguousarray(box, dtype=np.int32)
        self._check(self.lib.mk_marching_cubes_async(self.h, _ptr(b), iso, _ptr(dev_out), capacity, _ptr(dev_count),
                                                     stream))

    def marching_cubes_noalloc(self, box, iso: float, dev_out, capacity: int, dev_count, stream: int = 0) -> None:
        """mk_marching_cubes_noalloc: as marching_cubes_async, with the GPU's fixed scratch (nothing allocated)."""
        b = np.ascontiguousarray(box, dtype=np.int32)
        self._check(self.lib.mk_marching_cubes_noalloc(self.h, _ptr(b), iso, _ptr(dev_out), capacity,
                                                       _ptr(dev_count), stream))

    def tf_forward(self, p: RenderParams, dev_lut, n: int, scale: float, dev_rgb, stream: int = 0,
                   dev_rgb64=None) -> None:
        """dev_rgb64 (optional): the colours in float64 as well, for tf_backward."""
        self._check(self.lib.mk_tf_forward(self.h, C.byref(p), _ptr(dev_lut), n, scale, _ptr(dev_rgb),
                                           _ptr(dev_rgb64), stream))

    def tf_backward(self, p: RenderParams, dev_lut, n: int, scale: float, dev_grad_rgb, dev_grad_lut, dev_grad_scale,
                    stream: int = 0, dev_rgb64=None) -> None:
        """dev_rgb64 (optional): tf_forward's float64 colours of this view; without them each ray is marched twice."""
        self._check(self.lib.mk_tf_backward(self.h, C.byref(p), _ptr(dev_lut), n, scale, _ptr(dev_grad_rgb),
                                            _ptr(dev_rgb64), _ptr(dev_grad_lut), _ptr(dev_grad_scale), stream))


class Batcher:
    """Batched launches over several contexts of one device (``mk_batcher``)."""

    def __init__(self, contexts: list, max_items: int, max_tiles: int, nslots: int = 2):
        self.lib = load()
        self.contexts = list(contexts)   # keeps the contexts alive
        arr = (C.c_void_p * len(self.contexts))(*[c.h for c in self.contexts])
        h = C.c_void_p()
        st = self.lib.mk_batcher_create(arr, len(self.contexts), max_items, max_tiles, nslots, C.byref(h))
        if st != MK_OK:
            raise MkError(st, self.lib.mk_batcher_last_error(None).decode())
        self.h = h
        self.nslots = nslots
        self.max_items = max_items
        self.max_tiles = max_tiles

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.mk_batcher_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st: int) -> None:
        if st != MK_OK:
            raise MkError(st, self.lib.mk_batcher_last_error(self.h).decode())

    def submit(self, slot: int, items, dev_arena, arena_bytes: int, stream: int = 0) -> None:
        arr = (BatchItem * len(items))(*items)
        self._check(self.lib.mk_batch_submit(self.h, slot, arr, len(items), _ptr(dev_arena), arena_bytes, stream))

    def query(self, slot: int) -> bool:
        r = self.lib.mk_batch_query(self.h, slot)
        if r < 0:
            raise MkError(MK_ERR_HIP, self.lib.mk_batcher_last_error(self.h).decode())
        return r == 1

    def wait(self, slot: int) -> None:
        self._check(self.lib.mk_batch_wait(self.h, slot))
