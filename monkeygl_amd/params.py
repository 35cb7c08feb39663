"""Kernel-argument producers of the render path, restated in Python.

These compute what the reference's host classes hand to its kernels, with the
reference's float32/float64 arithmetic kept operation for operation:

* ``transfer_function``   -- TransferFunction::GetTransferFunction
                             (core/TransferFunction.cpp:38-140)
* ``RenderInfo``          -- RenderInfo + Methods::ComputeTransformMatrix /
                             MatrixMul (core/RenderInfo.cpp:28-134,
                             core/Methods.cpp:52-147)
* ``init_common``         -- DataManager::InitCommon (core/DataManager.h:201-216)
* ``need_to_invert_z``    -- VolumeInfo::Need2InvertZ (core/VolumeInfo.cpp:142-148)
* ``plane_params``        -- MPRInfo::ResetPlaneInfos / GetPlaneInitSize and
                             MPRImageLayer::GetGrayscaleData
                             (core/MPRInfo.cpp:37-156, core/MPRImageLayer.cpp:84-201)

The golden vectors in tests/golden (produced by compiling the reference's own
TransferFunction.cpp, Methods.cpp and RenderInfo.cpp) pin these functions.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from ._capi import (MK_MPR_AVERAGE, MK_NUM_LABELS, MK_RENDER_VR, MK_SAMPLE_LINEAR, PlaneParams,
                    RenderParams)

F32 = np.float32
PI = 3.141592654  # core/Defines.h:33


# ---------------------------------------------------------------------------
# transfer function


def _lerp_segments(points: dict, lo: int, hi: int, out: np.ndarray, channels) -> None:
    items = sorted(points.items())
    pos_prev, val_prev = items[0]
    for pos_cur, val_cur in items[1:]:
        n_part = pos_cur - pos_prev + 1
        for j in range(pos_prev, pos_cur + 1):
            idx = j - lo
            for c in channels:
                a = F32(val_prev[c]) if isinstance(val_prev, tuple) else F32(val_prev)
                b = F32(val_cur[c]) if isinstance(val_cur, tuple) else F32(val_cur)
                # prev + (cur - prev) * (j - posPrev) / nLenPart, float arithmetic
                out[idx, c] = F32(a + F32(F32(F32(b - a) * F32(j - pos_prev)) / F32(n_part)))
        pos_prev, val_prev = pos_cur, val_cur


def transfer_function(rgba_points: dict, alpha_points: dict | None = None, n: int = 100):
    """Control points -> (n, 4) float32 LUT, or None when fewer than 2 points
    lie in [0, n-1]. rgba_points: {pos: (r, g, b, a)}; alpha_points: {pos: a}."""
    lo, hi = 0, n - 1
    pts = {int(k): tuple(float(x) for x in v) for k, v in rgba_points.items() if lo <= int(k) <= hi}
    if len(pts) < 2:
        return None
    if min(pts) != lo:
        pts[lo] = pts[min(pts)]
    if max(pts) != hi:
        pts[hi] = pts[max(pts)]
    out = np.zeros((n, 4), np.float32)
    _lerp_segments(pts, lo, hi, out, (0, 1, 2, 3))
    if alpha_points and len(alpha_points) > 1:
        ap = {int(k): float(v) for k, v in alpha_points.items() if lo <= int(k) <= hi}
        if ap:
            if min(ap) != lo:
                ap[lo] = 0.0
            if max(ap) != hi:
                ap[hi] = ap[max(ap)]
            if len(ap) > 1:
                _lerp_segments(ap, lo, hi, out, (3,))
    return out


# ---------------------------------------------------------------------------
# camera


def _identity() -> np.ndarray:
    return np.eye(3, dtype=np.float32).reshape(9)


def matrix_mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Methods::MatrixMul for 3x3 float matrices, same accumulation order."""
    out = np.zeros(9, np.float32)
    for i in range(3):
        for j in range(3):
            acc = F32(a[i * 3] * b[j])
            for k in range(1, 3):
                acc = F32(acc + F32(a[i * 3 + k] * b[k * 3 + j]))
            out[i * 3 + j] = acc
    return out


def _cs(deg: float) -> tuple:
    # cos((deg)/180.0f*PI): float division, then double multiply and cos
    r = float(F32(F32(deg) / F32(180.0))) * PI
    return F32(math.cos(r)), F32(math.sin(r))


def compute_transform(rot, rot_t, m, m_t, fx_rotate: float, fz_rotate: float, scale: float) -> None:
    """Methods::ComputeTransformMatrix, updating the four float[9] in place."""
    cx, sx = _cs(fx_rotate)
    cz, sz = _cs(fz_rotate)
    S = np.zeros(9, np.float32)
    S[0] = S[4] = S[8] = F32(F32(1.0) / F32(scale))
    Rx = _identity(); Rx[4] = cx; Rx[5] = -sx; Rx[7] = sx; Rx[8] = cx
    Rz = _identity(); Rz[0] = cz; Rz[1] = sz; Rz[3] = -sz; Rz[4] = cz
    m[:] = matrix_mul(m, S)
    m[:] = matrix_mul(m, Rz)
    m[:] = matrix_mul(m, Rx)
    temp = matrix_mul(Rx, Rz)
    rot[:] = matrix_mul(temp, rot)
    RxT = _identity(); RxT[4] = cx; RxT[5] = sx; RxT[7] = -sx; RxT[8] = cx
    RzT = _identity(); RzT[0] = cz; RzT[1] = -sz; RzT[3] = sz; RzT[4] = cz
    ST = np.zeros(9, np.float32)
    ST[0] = ST[4] = ST[8] = F32(scale)
    temp = matrix_mul(RxT, RzT)
    temp = matrix_mul(temp, ST)
    m_t[:] = matrix_mul(temp, m_t)
    rot_t[:] = matrix_mul(rot_t, RzT)
    rot_t[:] = matrix_mul(rot_t, RxT)


class RenderInfo:
    """core/RenderInfo.cpp: VR size, pan, zoom and the accumulated rotation."""

    def __init__(self):
        self.width, self.height = 512, 512
        self._init()

    def _init(self):
        self.x_translate = F32(0.0)
        self.y_translate = F32(0.0)
        self.total_scale = F32(1.0)
        self.rot = _identity()
        self.rot_t = _identity()
        self.m = _identity()
        self.m_t = _identity()

    def set_vr_size(self, w: int, h: int) -> bool:
        delta = float((w * h) // (768 * 768))  # integer division, then float
        if delta > 1:
            self.width, self.height = int(w / delta), int(h / delta)
        else:
            self.width, self.height = w, h
        if self.width % 2:
            self.width += 1
        return True

    def _ct(self, fx, fz, s):
        compute_transform(self.rot, self.rot_t, self.m, self.m_t, fx, fz, s)

    def rotate(self, fx_rotate: float, fy_rotate: float) -> None:
        # the x angle drives Rz and the y angle Rx (RenderInfo.cpp:77-80)
        self._ct(fy_rotate, fx_rotate, 1.0)

    def zoom(self, ratio: float) -> float:
        self.total_scale = F32(self.total_scale * F32(ratio))
        self._ct(0.0, 0.0, ratio)
        return float(self.total_scale)

    def pan(self, dx: float, dy: float) -> None:
        self.x_translate = F32(self.x_translate + F32(dx))
        self.y_translate = F32(self.y_translate + F32(dy))

    def _preset(self, fx, fz):
        # Init() resets the total scale to 1 before it is used (RenderInfo.cpp:82-123)
        self._init()
        self._ct(fx, fz, float(self.total_scale))

    def anterior(self):
        self._preset(0.0, 0.0)

    def posterior(self):
        self._preset(0.0, 180.0)

    def left(self):
        self._preset(0.0, -90.0)

    def right(self):
        self._preset(0.0, 90.0)

    def head(self):
        self._preset(90.0, 180.0)

    def foot(self):
        self._preset(-90.0, 0.0)


# ---------------------------------------------------------------------------
# volume geometry


def init_common(spacing, dims):
    """DataManager::InitCommon: (spacing, spacing_voxel, max_len_spacing), float32."""
    sx, sy, sz = (F32(s) for s in spacing)
    w, h, d = dims
    sv = [F32(F32(1.0) / F32(w)), F32(F32(1.0) / F32(h)), F32(F32(1.0) / F32(d))]
    max_len = max(F32(w) * sx, max(F32(h) * sy, F32(d) * sz))
    mls = [F32(F32(max_len) / F32(F32(w) * sx)), F32(F32(max_len) / F32(F32(h) * sy)),
           F32(F32(max_len) / F32(F32(d) * sz))]
    return [sx, sy, sz], sv, mls


def need_to_invert_z(dir_x=(1.0, 0.0, 0.0), dir_y=(0.0, 1.0, 0.0), dir_z=(0.0, 0.0, -1.0)) -> bool:
    n = np.cross(np.asarray(dir_x, float), np.asarray(dir_y, float))
    return float(np.dot(n, np.asarray(dir_z, float))) > 0


@dataclass
class LabelState:
    alpha: float = 1.0
    ww: float = 400.0
    wl: float = 40.0


def render_params(dims, spacing, info: RenderInfo, labels: dict | None = None, render_type: int = MK_RENDER_VR,
                  invert_z: bool = False, background=(0.0, 0.0, 0.0, 1.0), step_scale: float = 1.0,
                  exit_alpha: float = 0.995, width: int | None = None, height: int | None = None) -> RenderParams:
    """Arguments of cu_render as VRImageLayer::GetRGBData assembles them
    (core/VRImageLayer.cpp:64-97). labels: {label: LabelState}; labels not
    given keep AlphaAndWWWL() = (0, 0, 0), as the reference's table does."""
    p = RenderParams()
    p.width = width if width is not None else info.width
    p.height = height if height is not None else info.height
    p.render_type = render_type
    p.invert_z = int(invert_z)
    sp, sv, mls = init_common(spacing, dims)
    for i in range(3):
        p.spacing[i] = sp[i]
        p.spacing_voxel[i] = sv[i]
        p.max_len_spacing[i] = mls[i]
    p.x_translate = info.x_translate
    p.y_translate = info.y_translate
    p.scale = info.total_scale
    for i in range(9):
        p.transform[i] = info.m[i]
    w, h, d = dims
    p.voi.left, p.voi.right = 0, w - 1
    p.voi.posterior, p.voi.anterior = 0, h - 1
    p.voi.head, p.voi.foot = 0, d - 1
    for i in range(4):
        p.background[i] = background[i]
    labels = labels if labels is not None else {0: LabelState()}
    for l, st in labels.items():
        p.alpha_wwwl[l].alpha = st.alpha
        p.alpha_wwwl[l].ww = st.ww if st.ww > 0 else 1.0
        p.alpha_wwwl[l].wl = st.wl
    p.step_scale = step_scale
    p.exit_alpha = exit_alpha
    return p


# ---------------------------------------------------------------------------
# MPR planes

PLANE_AXIAL, PLANE_SAGITTAL, PLANE_CORONAL = 0, 1, 2
PLANE_AXIAL_OBLIQUE, PLANE_SAGITTAL_OBLIQUE, PLANE_CORONAL_OBLIQUE = 3, 4, 5

_PLANE_DIRS = {
    PLANE_AXIAL: ((1, 0, 0), (0, 1, 0)),
    PLANE_AXIAL_OBLIQUE: ((1, 0, 0), (0, 1, 0)),
    PLANE_SAGITTAL: ((0, 1, 0), (0, 0, -1)),
    PLANE_SAGITTAL_OBLIQUE: ((0, 1, 0), (0, 0, -1)),
    PLANE_CORONAL: ((1, 0, 0), (0, 0, -1)),
    PLANE_CORONAL_OBLIQUE: ((1, 0, 0), (0, 0, -1)),
}


def plane_init_size(dims, spacing, plane: int):
    """MPRInfo::GetPlaneInitSize -> (width, height, number)."""
    ms = min(spacing)
    w, h, d = dims
    sx, sy, sz = spacing
    if plane in (PLANE_AXIAL, PLANE_AXIAL_OBLIQUE):
        return int(w * sx / ms), int(h * sy / ms), int(d * sz / ms)
    if plane in (PLANE_SAGITTAL, PLANE_SAGITTAL_OBLIQUE):
        return int(h * sy / ms), int(d * sz / ms), int(w * sx / ms)
    return int(w * sx / ms), int(d * sz / ms), int(h * sy / ms)


def plane_params(dims, spacing, plane: int, invert_z: bool = False, mpr_type: int = MK_MPR_AVERAGE,
                 sampling: int = MK_SAMPLE_LINEAR, thickness: float | None = None, crosshair=None,
                 size=None, dir_h=None, dir_v=None) -> PlaneParams:
    """cu_renderPlane_* arguments as MPRImageLayer::GetGrayscaleData builds
    them for the default plane frames of MPRInfo::ResetPlaneInfos. ``size``
    overrides the plane size (the reference's is GetPlaneInitSize)."""
    dh, dv = _PLANE_DIRS[plane]
    dh = np.asarray(dir_h if dir_h is not None else dh, float)
    dv = np.asarray(dir_v if dir_v is not None else dv, float)
    dn = np.cross(dh, dv)
    ps = float(min(spacing))
    w, h, _ = plane_init_size(dims, spacing, plane)
    if size is not None:
        w, h = size
    if w % 2:
        w += 1
    if crosshair is None:
        crosshair = [0.5 * dims[i] * spacing[i] for i in range(3)]
    center = np.asarray(crosshair, float)  # center projected on its own plane
    lt = center - dh * (0.5 * w * ps)
    lt = lt - dv * (0.5 * h * ps)
    thick = ps if thickness is None else float(thickness)
    n_slice = int(thick / ps)
    n_slice = 1 if n_slice < 1 else n_slice
    half = F32(F32(1.0) * F32(n_slice - 1) / F32(2))
    p = PlaneParams()
    p.width, p.height = w, h
    p.mpr_type = mpr_type
    p.sampling = sampling
    p.invert_z = int(invert_z)
    for i in range(3):
        p.spacing[i] = F32(spacing[i])
        p.dir_h[i] = F32(dh[i])
        p.dir_v[i] = F32(dv[i])
        p.dir_n[i] = F32(dn[i])
        p.pt_left_top[i] = F32(lt[i])
    p.pixel_spacing = F32(ps)
    p.half_num = half
    return p
