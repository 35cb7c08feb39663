"""Seeded synthetic workloads of the render path (no datasets are available).

Volumes are int16 arrays indexed [z, y, x] (x fastest in memory, the layout
the reference's SetVolumeArray hands to the device) with the reference's
default orientation (direction z = (0, 0, -1), so VolumeInfo::Need2InvertZ
is false).

* ``blob_volume``    -- Gaussian blobs on a -1024 background plus integer
                        uniform noise, clamped to [-1024, 3071].
* ``body_volume``    -- blob "chambers" plus a network of cubic-spline vessels
                        (radius in mm, constant HU per vessel).
* ``ellipsoid_mask`` -- uint8 label mask of seeded ellipsoids.
* ``tf_points``      -- seeded transfer-function control points.
* ``centerline``     -- seeded cubic-spline centreline for CPR.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass

import numpy as np

BACKGROUND = -1024
HU_MIN, HU_MAX = -1024, 3071


@dataclass
class VolumeSpec:
    name: str
    dims: tuple          # (nx, ny, nz)
    spacing: tuple       # mm (x, y, z)
    n_blobs: int
    sigma: tuple         # voxel units (lo, hi)
    amplitude: tuple
    n_vessels: int = 0
    seed: int = 0


SMALL = VolumeSpec("blobs128", (128, 128, 128), (1.0, 1.0, 1.0), 8, (4, 20), (200, 2000))
CARDIAC = VolumeSpec("cardiac512x512x361", (512, 512, 361), (0.351, 0.351, 0.3), 12, (8, 40), (800, 1400), 40)
BODY = VolumeSpec("body512x512x1559", (512, 512, 1559), (0.7422, 0.7422, 1.0), 24, (8, 40), (800, 1400), 40)
TF_FIT = VolumeSpec("blobs256", (256, 256, 256), (1.0, 1.0, 1.0), 16, (4, 20), (200, 2000))


def _add_blob(vol: np.ndarray, c, sigma: float, amp: float) -> None:
    nz, ny, nx = vol.shape
    r = int(np.ceil(3.5 * sigma))
    lo = [max(0, int(np.floor(c[i])) - r) for i in range(3)]
    hi = [min((nx, ny, nz)[i], int(np.floor(c[i])) + r + 1) for i in range(3)]
    if any(lo[i] >= hi[i] for i in range(3)):
        return
    x = (np.arange(lo[0], hi[0], dtype=np.float32) - np.float32(c[0])) ** 2
    y = (np.arange(lo[1], hi[1], dtype=np.float32) - np.float32(c[1])) ** 2
    z = (np.arange(lo[2], hi[2], dtype=np.float32) - np.float32(c[2])) ** 2
    inv = np.float32(-0.5 / (sigma * sigma))
    g = np.exp(z[:, None, None] * inv) * np.exp(y[None, :, None] * inv) * np.exp(x[None, None, :] * inv)
    vol[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]] += np.float32(amp) * g


def _spline(ctrl: np.ndarray, n: int) -> np.ndarray:
    """Catmull-Rom spline through the control points, n samples."""
    P = np.vstack([ctrl[:1], ctrl, ctrl[-1:]])
    segs = len(ctrl) - 1
    t = np.linspace(0, segs, n, endpoint=False)
    i = np.minimum(t.astype(int), segs - 1)
    u = (t - i)[:, None]
    p0, p1, p2, p3 = P[i], P[i + 1], P[i + 2], P[i + 3]
    out = 0.5 * ((2 * p1) + (-p0 + p2) * u + (2 * p0 - 5 * p1 + 4 * p2 - p3) * u ** 2 + (-p0 + 3 * p1 - 3 * p2 + p3) * u ** 3)
    return np.vstack([out, ctrl[-1:]])


def _add_vessel(vol: np.ndarray, pts_vox: np.ndarray, radius_mm: float, hu: float, spacing) -> None:
    nz, ny, nx = vol.shape
    sp = np.asarray(spacing, np.float32)
    rv = np.ceil(radius_mm / sp).astype(int)
    for p in pts_vox:
        c = np.floor(p).astype(int)
        lo = np.maximum(c - rv, 0)
        hi = np.minimum(c + rv + 2, (nx, ny, nz))
        if np.any(lo >= hi):
            continue
        x = ((np.arange(lo[0], hi[0]) - p[0]) * sp[0]) ** 2
        y = ((np.arange(lo[1], hi[1]) - p[1]) * sp[1]) ** 2
        z = ((np.arange(lo[2], hi[2]) - p[2]) * sp[2]) ** 2
        inside = (z[:, None, None] + y[None, :, None] + x[None, None, :]) <= radius_mm * radius_mm
        sub = vol[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]
        np.maximum(sub, np.where(inside, np.float32(hu), np.float32(-np.inf)), out=sub)


def make_volume(spec: VolumeSpec, seed: int | None = None) -> np.ndarray:
    rng = np.random.default_rng(spec.seed if seed is None else seed)
    nx, ny, nz = spec.dims
    vol = np.full((nz, ny, nx), np.float32(BACKGROUND), np.float32)
    margin = np.array([nx, ny, nz], float) / 8.0
    for _ in range(spec.n_blobs):
        c = rng.uniform(margin, np.array([nx, ny, nz]) - margin)
        sigma = rng.uniform(*spec.sigma)
        amp = rng.uniform(*spec.amplitude)
        _add_blob(vol, c, sigma, amp)
    for _ in range(spec.n_vessels):
        k = rng.integers(5, 9)
        ctrl = rng.uniform(margin, np.array([nx, ny, nz]) - margin, size=(k, 3))
        # sample densely enough that consecutive points are < 0.5 voxel apart
        length = np.sum(np.linalg.norm(np.diff(ctrl, axis=0), axis=1))
        pts = _spline(ctrl, int(length * 2.5) + 8)
        _add_vessel(vol, pts, rng.uniform(1.0, 4.0), rng.uniform(300.0, 600.0), spec.spacing)
    noise = rng.integers(-20, 21, size=vol.shape, dtype=np.int16)
    out = np.rint(vol)
    np.clip(out, HU_MIN, HU_MAX, out=out)
    out = out.astype(np.int16)
    out += noise
    np.clip(out, HU_MIN, HU_MAX, out=out)
    return out


def ellipsoid_mask(dims, n_labels: int = 4, seed: int = 1) -> np.ndarray:
    """Labels 1..n_labels as seeded axis-aligned ellipsoids (later labels win)."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = dims
    m = np.zeros((nz, ny, nx), np.uint8)
    for lab in range(1, n_labels + 1):
        c = rng.uniform(0.25, 0.75, 3) * np.array([nx, ny, nz])
        r = rng.uniform(0.08, 0.22, 3) * np.array([nx, ny, nz])
        lo = np.maximum(np.floor(c - r).astype(int), 0)
        hi = np.minimum(np.ceil(c + r).astype(int) + 1, [nx, ny, nz])
        x = ((np.arange(lo[0], hi[0]) - c[0]) / r[0]) ** 2
        y = ((np.arange(lo[1], hi[1]) - c[1]) / r[1]) ** 2
        z = ((np.arange(lo[2], hi[2]) - c[2]) / r[2]) ** 2
        inside = (z[:, None, None] + y[None, :, None] + x[None, None, :]) <= 1.0
        sub = m[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]
        sub[inside] = lab
    return m


def tf_points(seed: int = 0, n_points: int = 4) -> dict:
    """Seeded control points on the reference's [0, 99] TF domain: the first
    is transparent, the rest partly opaque."""
    rng = np.random.default_rng(seed)
    pos = np.sort(rng.choice(np.arange(15, 99), size=n_points, replace=False))
    pts = {}
    for i, p in enumerate(pos):
        r, g, b = rng.uniform(0.2, 1.0, 3)
        a = 0.0 if i == 0 else float(rng.uniform(0.05, 0.8))
        pts[int(p)] = (float(r), float(g), float(b), a)
    return pts


def centerline(dims, n: int = 200, seed: int = 2) -> np.ndarray:
    """Seeded cubic-spline centreline of n points (voxel units) through the
    middle of the volume, roughly along z."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = dims
    k = 6
    z = np.linspace(0.1 * nz, 0.9 * nz, k)
    x = nx / 2 + rng.uniform(-0.15, 0.15, k) * nx
    y = ny / 2 + rng.uniform(-0.15, 0.15, k) * ny
    ctrl = np.stack([x, y, z], axis=1)
    pts = _spline(ctrl, n - 1)
    return pts[:n]


# ---------------------------------------------------------------------------
# Transfer-function fitting: the 256^3 blob volume (TF_FIT), 8 views, a
# 256-entry RGBA LUT and an opacity scale, MSE to images of a seeded
# ground-truth LUT

FIT_LUT_ENTRIES = 256
FIT_SCALE_TRUE, FIT_SCALE_START = 1.2, 1.5


def fit_views(size: int = 512, n: int = 8):
    """Render arguments of the fitting views (label 0 window 2200 / 0). The
    angles are generic: at multiples of 90 degrees whole pixel columns run
    exactly on a volume face, where float sample positions and an FP64 check's
    positions fall on opposite sides of it."""
    from . import params as P

    out = []
    for k in range(n):
        info = P.RenderInfo()
        info.set_vr_size(size, size)
        info.rotate(45.0 * k + 15.0, 20.0 * (k % 3) - 17.0)
        out.append(P.render_params(TF_FIT.dims, TF_FIT.spacing, info, {0: P.LabelState(1.0, 2200.0, 0.0)},
                                   width=size, height=size))
    return out


def fit_lut(seed: int, n: int = FIT_LUT_ENTRIES) -> np.ndarray:
    """Seeded (n, 4) float32 LUT: colours in [0.1, 1), alphas in [0, 0.08).
    Seed 1 is the ground truth, seed 2 the start of the fit."""
    rng = np.random.default_rng(seed)
    a = np.empty((n, 4), np.float32)
    a[:, :3] = rng.uniform(0.1, 1.0, (n, 3))
    a[:, 3] = rng.uniform(0.0, 0.08, n)
    return a


# ---------------------------------------------------------------------------
# Render serving: three resident volumes and a Poisson request stream

SERVE_MIX = (("vr", 0.60), ("mip", 0.20), ("mpr", 0.15), ("cpr", 0.05))
SERVE_SIZES = (256, 512, 768, 1024)


@dataclass
class ServedVolume:
    spec: VolumeSpec
    vol: np.ndarray
    mask: np.ndarray | None
    luts: dict            # label -> (n, 4) float32
    labels: dict          # label -> params.LabelState
    line: np.ndarray      # CPR centreline (voxel units)


def serving_scene(specs=None, seed: int = 0) -> list:
    """The served volumes: the 128^3 blob set, the cardiac-sized set and the
    body-sized set with a 4-label mask, each with its own TFs and a seeded
    centreline."""
    from . import params as P

    specs = specs if specs is not None else [(SMALL, 0), (CARDIAC, 0), (BODY, 4)]
    out = []
    for k, (spec, n_labels) in enumerate(specs):
        vol = make_volume(spec)
        mask = ellipsoid_mask(spec.dims, n_labels, seed=1 + k) if n_labels else None
        luts, labels = {}, {}
        for l in range(n_labels + 1):
            luts[l] = P.transfer_function(tf_points(100 + 10 * k + l))
            labels[l] = P.LabelState(alpha=(0.4 if l == 0 and n_labels else 1.0), ww=1200.0 + 200.0 * l,
                                     wl=-200.0 + 60.0 * l)
        out.append(ServedVolume(spec, vol, mask, luts, labels, centerline(spec.dims, 200, seed=2 + k)))
    return out


def serving_requests(scene: list, n: int, qps: float, seed: int = 0, sizes=SERVE_SIZES, mix=SERVE_MIX):
    """n requests with Poisson arrivals at ``qps`` (t_arrival in seconds from
    0), volumes uniform over the scene, kinds by ``mix``, square viewports
    uniform over ``sizes``, random camera rotations, MPR planes of random
    type / slab mode / sampling around a random crosshair, and stretched or
    straightened CPRs at a random rotation."""
    from . import cpr as G
    from . import params as P
    from ._capi import MK_MPR_AVERAGE, MK_MPR_MINIP, MK_MPR_MIP, MK_RENDER_MIP, MK_RENDER_VR, MK_SAMPLE_LINEAR, \
        MK_SAMPLE_NEAREST
    from .serve import KIND_CPR, KIND_PLANE, KIND_RENDER, Request

    rng = np.random.default_rng(seed)
    kinds = [k for k, _ in mix]
    probs = np.array([p for _, p in mix], float)
    probs /= probs.sum()
    t = 0.0
    reqs = []
    cprs = {}
    for i in range(n):
        t += rng.exponential(1.0 / qps) if qps > 0 else 0.0
        vi = int(rng.integers(len(scene)))
        sv = scene[vi]
        dims, sp = sv.spec.dims, sv.spec.spacing
        kind = kinds[int(rng.choice(len(kinds), p=probs))]
        size = int(sizes[int(rng.integers(len(sizes)))])
        if kind in ("vr", "mip"):
            info = P.RenderInfo()
            info.set_vr_size(size, size)
            info.rotate(float(rng.uniform(0.0, 360.0)), float(rng.uniform(-60.0, 60.0)))
            rt = MK_RENDER_VR if kind == "vr" else MK_RENDER_MIP
            p = P.render_params(dims, sp, info, sv.labels, render_type=rt, width=size, height=size)
            reqs.append(Request(KIND_RENDER, vi, render=p, t_arrival=t, tag=(i, kind)))
        elif kind == "mpr":
            plane = int(rng.integers(3))
            mode = (MK_MPR_AVERAGE, MK_MPR_MIP, MK_MPR_MINIP)[int(rng.integers(3))]
            samp = MK_SAMPLE_NEAREST if rng.random() < 0.5 else MK_SAMPLE_LINEAR
            ps = min(sp)
            thick = ps if rng.random() < 0.5 else float(rng.uniform(2.0, 10.0))
            cross = [float(rng.uniform(0.3, 0.7)) * dims[a] * sp[a] for a in range(3)]
            p = P.plane_params(dims, sp, plane, mpr_type=mode, sampling=samp, thickness=thick, crosshair=cross,
                               size=(size, size))
            reqs.append(Request(KIND_PLANE, vi, plane=p, t_arrival=t, tag=(i, kind)))
        else:
            which = G.PLANE_STRETCHED_CPR if rng.random() < 0.5 else G.PLANE_STRAIGHTENED_CPR
            key = (vi, which)
            if key not in cprs:
                info = G.CPRInfo()
                info.set_spacing(sp)
                info.set_cpr_line_voxel(sv.line)
                cprs[key] = info
            info = cprs[key]
            info.rotate_cpr(float(rng.uniform(0.0, 360.0)), which)
            pts, dirs, w, h = info.get_cpr_info(which)
            samp = MK_SAMPLE_NEAREST if rng.random() < 0.5 else MK_SAMPLE_LINEAR
            reqs.append(Request(KIND_CPR, vi, cpr_width=w, cpr_points=pts, cpr_dirs=dirs, cpr_sampling=samp,
                                t_arrival=t, tag=(i, "cpr")))
    return reqs


# ---- orientation figures (binary STL) ------------------------------------------
# The reference ships its figure as resources/human.stl (34,660 facets) and
# loads it at start-up; these synthetic figures stand in for it: a body of UV
# ellipsoids whose pole rings hold zero-area facets, plus two large facets.

def uv_ellipsoid(center, radii, nu: int = 18, nv: int = 12) -> np.ndarray:
    """(n, 3, 3) facets of a UV ellipsoid; the pole rows are degenerate
    triangles (two vertices at the pole), as UV spheres have."""
    c, r = np.asarray(center, np.float64), np.asarray(radii, np.float64)
    th = np.linspace(0, np.pi, nv + 1)
    ph = np.linspace(0, 2 * np.pi, nu + 1)
    P = np.stack([np.outer(np.sin(th), np.cos(ph)), np.outer(np.sin(th), np.sin(ph)),
                  np.outer(np.cos(th), np.ones_like(ph))], -1) * r + c
    out = []
    for i in range(nv):
        for j in range(nu):
            a, b, cc, d = P[i, j], P[i + 1, j], P[i + 1, j + 1], P[i, j + 1]
            out.append((a, b, cc))
            out.append((a, cc, d))
    return np.asarray(out, np.float32)


def figure(seed: int = 0, fine: bool = False) -> np.ndarray:
    """A seeded body of UV ellipsoids (head, torso, arms, legs), 2,722 facets;
    fine=True tessellates it into ~35,000, the size of the reference's figure."""
    k = 3.6 if fine else 1.0
    rng = np.random.default_rng(seed)

    def part(c, r, nu, nv):
        return uv_ellipsoid(c, r, int(round(nu * k)), int(round(nv * k)))

    parts = [part((0, 0, 160), (12, 11, 14), 20, 14),          # head
             part((0, 0, 100), (24, 14, 45), 28, 18),          # torso
             part((-34, 0, 105), (6, 6, 38), 12, 10),          # arms
             part((34, 0, 105), (6, 6, 38), 12, 10),
             part((-12, 0, 30), (8, 8, 45), 14, 12),           # legs
             part((12, 0, 30), (8, 8, 45), 14, 12)]
    f = np.concatenate(parts)
    f += rng.normal(0, 0.05, f.shape).astype(np.float32)   # break exact symmetries
    big = np.array([[[-40, 20, -10], [40, 20, -10], [0, 20, 60]],
                    [[-40, -25, 190], [40, -25, 190], [0, -25, -15]]], np.float32)
    return np.concatenate([f, big])


def stl_bytes(facets: np.ndarray) -> bytes:
    f = np.asarray(facets, np.float32).reshape(-1, 3, 3)
    out = bytearray(b"synthetic figure".ljust(80, b" "))
    out += struct.pack("<I", f.shape[0])
    n = np.cross(f[:, 1] - f[:, 0], f[:, 2] - f[:, 0])
    for k in range(f.shape[0]):
        out += np.asarray(n[k], np.float32).tobytes() + f[k].tobytes() + b"\0\0"
    return bytes(out)
