"""MPR plane state: the frames, sizes and crosshair mk_render_plane samples.

Restates core/MPRInfo.cpp (plane frames 96-156, plane index 304-326, rotate
matrix 328-345, Browse 347-354, PanCrossHair 355-362, RotateCrossHair
364-427, UpdatePlaneSize 429-495, crosshair/image mapping 497-536, 2D/3D
directions 538-621, GetCenterPointPlane 623-625), core/PlaneInfo.h:37-45,
core/DataManager.cpp:509-525 (GetVertexes) and the kernel-argument assembly
of MPRImageLayer::GetGrayscaleData (core/MPRImageLayer.cpp:84-201).
Arithmetic follows the reference's types: points and directions in double
(every Direction3d construction normalises), crosshair angles and Browse
deltas in float, sizes truncated to int, kernel arguments rounded to float.
"""
from __future__ import annotations

import math

import numpy as np

from ._capi import MK_MPR_AVERAGE, MK_SAMPLE_LINEAR, PlaneParams
from .cpr import PI, cross, length_vector_in_line

PLANE_AXIAL, PLANE_SAGITTAL, PLANE_CORONAL = 0, 1, 2
PLANE_AXIAL_OBLIQUE, PLANE_SAGITTAL_OBLIQUE, PLANE_CORONAL_OBLIQUE = 3, 4, 5
PLANE_VR = 6
MPR_PLANES = range(PLANE_AXIAL, PLANE_CORONAL_OBLIQUE + 1)

_F32 = np.float32


def _f(x) -> float:
    return float(_F32(x))


def _dir(x, y, z):
    """Direction3d(x, y, z): normalised (NaN for a zero vector)."""
    n = math.sqrt(x * x + y * y + z * z)
    if n == 0.0:
        return (math.nan, math.nan, math.nan)
    return (x / n, y / n, z / n)


def _half_down(n: int) -> int:
    """(n - 1) / 2 in C integer arithmetic (truncation toward zero)."""
    return int((n - 1) / 2)


def _add(a, b):
    return (a[0] + b[0], a[1] + b[1], a[2] + b[2])


def _sub(a, b):
    return (a[0] - b[0], a[1] - b[1], a[2] - b[2])


def _mul(d, s):
    return (d[0] * s, d[1] * s, d[2] * s)


def distance_point_to_plane(pt, n, pp) -> float:
    """Methods::Distance_Point2Plane(pt, dirNorm, ptPlane) (core/Methods.cpp:588-592)."""
    v = _sub(pt, pp)
    return v[0] * n[0] + v[1] * n[1] + v[2] * n[2]


def distance_point_to_plane_hv(pt, h, v, pp) -> float:
    """The (dirH, dirV) overload: normal = dirH x dirV (core/Methods.cpp:582-586)."""
    return distance_point_to_plane(pt, cross(h, v), pp)


def project_point_to_plane(pt, n, pp):
    d = distance_point_to_plane(pt, n, pp)
    return (pt[0] - d * n[0], pt[1] - d * n[1], pt[2] - d * n[2])


def project_point_to_plane_hv(pt, h, v, pp):
    return project_point_to_plane(pt, cross(h, v), pp)


def transfer_point(m, pt):
    """Methods::GetTransferPoint (core/Methods.cpp:667-679): m (3x3 double) times pt."""
    r = []
    for i in range(3):
        s = 0.0
        for j in range(3):
            s += m[i][j] * pt[j]
        r.append(s)
    return tuple(r)


def plane_init_size(dims, spacing, plane: int):
    """MPRInfo::GetPlaneInitSize (core/MPRInfo.cpp:37-94) -> (width, height,
    number), or None for undefined input."""
    if min(dims) <= 0 or min(spacing) <= 0:
        return None
    ms = spacing[0] if spacing[0] < spacing[1] else spacing[1]
    ms = ms if ms < spacing[2] else spacing[2]
    w, h, d = dims
    sx, sy, sz = spacing
    if plane in (PLANE_AXIAL, PLANE_AXIAL_OBLIQUE):
        return int(w * sx / ms), int(h * sy / ms), int(d * sz / ms)
    if plane in (PLANE_SAGITTAL, PLANE_SAGITTAL_OBLIQUE):
        return int(h * sy / ms), int(d * sz / ms), int(w * sx / ms)
    if plane in (PLANE_CORONAL, PLANE_CORONAL_OBLIQUE):
        return int(w * sx / ms), int(d * sz / ms), int(h * sy / ms)
    if plane == PLANE_VR:
        return 512, 512, 1
    return None


_CROSS = {
    PLANE_AXIAL: (PLANE_SAGITTAL, PLANE_CORONAL),
    PLANE_SAGITTAL: (PLANE_AXIAL, PLANE_CORONAL),
    PLANE_CORONAL: (PLANE_AXIAL, PLANE_SAGITTAL),
    PLANE_AXIAL_OBLIQUE: (PLANE_SAGITTAL_OBLIQUE, PLANE_CORONAL_OBLIQUE),
    PLANE_SAGITTAL_OBLIQUE: (PLANE_AXIAL_OBLIQUE, PLANE_CORONAL_OBLIQUE),
    PLANE_CORONAL_OBLIQUE: (PLANE_AXIAL_OBLIQUE, PLANE_SAGITTAL_OBLIQUE),
}


class PlaneInfo:
    """core/PlaneInfo.h:30-57 (constructor core/PlaneInfo.cpp:27-31)."""

    def __init__(self, plane: int):
        self.plane = plane
        self.crosshair_angle = 0.0          # float
        self.dir_h = (1.0, 0.0, 0.0)
        self.dir_v = (0.0, 1.0, 0.0)
        self.width = self.height = self.number = 0
        self.pixel_spacing = 1.0
        self.thickness = 1.0
        self.mpr_type = MK_MPR_AVERAGE

    def norm(self):
        return cross(self.dir_h, self.dir_v)

    def left_top(self, center):
        lt = _sub(center, _mul(self.dir_h, 0.5 * self.width * self.pixel_spacing))
        return _sub(lt, _mul(self.dir_v, 0.5 * self.height * self.pixel_spacing))


class MPRInfo:
    """The reference's MPR state for one volume (dims in voxels, spacing in mm)."""

    def __init__(self, dims, spacing):
        self.dims = tuple(int(d) for d in dims)
        self.spacing = tuple(float(s) for s in spacing)
        self.planes: dict = {}
        self.crosshair = (0.0, 0.0, 0.0)
        self.center = (0.0, 0.0, 0.0)
        self.reset_plane_infos()

    @property
    def min_spacing(self) -> float:
        # VolumeInfo::GetMinSpacing (core/VolumeInfo.h:91-96)
        s = self.spacing
        m = s[0] if s[0] < s[1] else s[1]
        return m if m < s[2] else s[2]

    def vertexes(self):
        """DataManager::GetVertexes (core/DataManager.cpp:509-525)."""
        xl, yl, zl = (self.dims[i] * self.spacing[i] for i in range(3))
        return [(0.0, 0.0, 0.0), (xl, 0.0, 0.0), (0.0, yl, 0.0), (xl, yl, 0.0), (0.0, 0.0, zl), (xl, 0.0, zl),
                (0.0, yl, zl), (xl, yl, zl)]

    def reset_plane_infos(self) -> bool:
        """MPRInfo::ResetPlaneInfos (core/MPRInfo.cpp:96-156)."""
        d, s = self.dims, self.spacing
        self.crosshair = (0.5 * d[0] * s[0], 0.5 * d[1] * s[1], 0.5 * d[2] * s[2])
        self.planes = {}
        frames = {PLANE_AXIAL: ((1, 0, 0), (0, 1, 0)), PLANE_AXIAL_OBLIQUE: ((1, 0, 0), (0, 1, 0)),
                  PLANE_SAGITTAL: ((0, 1, 0), (0, 0, -1)), PLANE_SAGITTAL_OBLIQUE: ((0, 1, 0), (0, 0, -1)),
                  PLANE_CORONAL: ((1, 0, 0), (0, 0, -1)), PLANE_CORONAL_OBLIQUE: ((1, 0, 0), (0, 0, -1))}
        for pt in MPR_PLANES:
            info = PlaneInfo(pt)
            info.width, info.height, info.number = plane_init_size(d, s, pt)
            h, v = frames[pt]
            info.dir_h = _dir(*map(float, h))
            info.dir_v = _dir(*map(float, v))
            info.pixel_spacing = self.min_spacing
            info.thickness = self.min_spacing
            self.planes[pt] = info
            self.center = self.crosshair
        return True

    # -- per-plane settings ---------------------------------------------------
    def set_mpr_type(self, t: int) -> None:
        for info in self.planes.values():
            info.mpr_type = t

    def update_thickness(self, val: float) -> None:
        for info in self.planes.values():
            info.thickness = float(val)

    def set_thickness(self, val: float, plane: int) -> None:
        if plane in self.planes:
            self.planes[plane].thickness = float(val)

    def get_thickness(self, plane: int) -> float:
        return self.planes[plane].thickness if plane in self.planes else -1.0

    def get_pixel_spacing(self, plane: int) -> float:
        return self.planes[plane].pixel_spacing if plane in self.planes else 1.0

    def get_plane_size(self, plane: int):
        """DataManager::GetPlaneSize -> MPRInfo::GetPlaneSize (MPR planes only)."""
        if plane not in self.planes:
            return None
        return self.planes[plane].width, self.planes[plane].height

    def get_plane_number(self, plane: int):
        return self.planes[plane].number if plane in self.planes else None

    def get_plane_index(self, plane: int):
        """MPRInfo::GetPlaneIndex (core/MPRInfo.cpp:304-314)."""
        if plane not in self.planes:
            return None
        info = self.planes[plane]
        dist = distance_point_to_plane_hv(self.crosshair, info.dir_h, info.dir_v, self.center)
        return int(dist / self.min_spacing) + _half_down(info.number)

    def set_plane_index(self, index: int, plane: int) -> None:
        """MPRInfo::SetPlaneIndex (core/MPRInfo.cpp:316-326)."""
        if plane not in self.planes:
            return
        info = self.planes[plane]
        n = info.norm()
        dist = (index - _half_down(info.number)) * self.min_spacing
        proj = project_point_to_plane_hv(self.crosshair, info.dir_h, info.dir_v, self.center)
        self.crosshair = _add(proj, _mul(n, dist))

    def get_plane_rotate_matrix(self, plane: int):
        """MPRInfo::GetPlaneRotateMatrix (core/MPRInfo.cpp:328-345): float[9] rows dirH, dirV, dirH x dirV."""
        if plane not in self.planes:
            return None
        info = self.planes[plane]
        z = cross(info.dir_h, info.dir_v)
        return np.array(list(info.dir_h) + list(info.dir_v) + list(z), np.float32)

    # -- interaction ----------------------------------------------------------
    def browse(self, delta: float, plane: int) -> None:
        """MPRInfo::Browse (core/MPRInfo.cpp:347-354): move the crosshair along the plane normal."""
        if plane not in self.planes:
            return
        self.crosshair = _add(self.crosshair, _mul(self.planes[plane].norm(), _f(delta)))

    def pan_crosshair(self, fx: float, fy: float, plane: int) -> None:
        """MPRInfo::PanCrossHair (core/MPRInfo.cpp:355-362)."""
        p = self.transfer_image_to_voxel(_f(fx), _f(fy), plane)
        if p is not None:
            self.crosshair = p

    def rotate_crosshair(self, angle: float, plane: int) -> None:
        """MPRInfo::RotateCrossHair (core/MPRInfo.cpp:364-427): turn the two
        crossing planes about this plane's normal through the crosshair, to the
        absolute crosshair angle ``angle`` (degrees)."""
        info = self.planes[plane]
        a = info.norm()
        f_angle = _F32(angle) - _F32(info.crosshair_angle)       # float - float
        info.crosshair_angle = _f(angle)
        f_angle = _f(float(f_angle / _F32(180)) * PI)             # (float / 180) * double -> float
        x, y, z = a
        xx, yy, zz = x * x, y * y, z * z
        s, c = math.sin(f_angle), math.cos(f_angle)
        ct = 1.0 - c
        m = [[xx + (1 - xx) * c, x * y * ct - z * s, x * z * ct + y * s],
             [x * y * ct + z * s, yy + (1 - yy) * c, y * z * ct - x * s],
             [x * z * ct - y * s, y * z * ct + x * s, zz + (1 - zz) * c]]
        for cp in _CROSS[plane]:
            ci = self.planes[cp]
            h, v = ci.dir_h, ci.dir_v
            center = self.center_point_plane(ci.norm())
            fw = ci.pixel_spacing * ci.width
            fh = ci.pixel_spacing * ci.height
            lt = _sub(_sub(center, _mul(h, fw / 2)), _mul(v, fh / 2))
            rt = _sub(_add(center, _mul(h, fw / 2)), _mul(v, fh / 2))
            lb = _add(_sub(center, _mul(h, fw / 2)), _mul(v, fh / 2))
            lt = _add(transfer_point(m, _sub(lt, self.crosshair)), self.crosshair)
            rt = _add(transfer_point(m, _sub(rt, self.crosshair)), self.crosshair)
            lb = _add(transfer_point(m, _sub(lb, self.crosshair)), self.crosshair)
            ci.dir_h = _dir(rt[0] - lt[0], rt[1] - lt[1], rt[2] - lt[2])
            ci.dir_v = _dir(lb[0] - lt[0], lb[1] - lt[1], lb[2] - lt[2])
            self.update_plane_size(cp)

    def update_plane_size(self, plane: int) -> None:
        """MPRInfo::UpdatePlaneSize (core/MPRInfo.cpp:429-495): the plane's
        extent over the projections of the volume's corners."""
        if plane not in self.planes:
            return
        info = self.planes[plane]
        sp = self.min_spacing
        xn = xp = yn = yp = zn = zp = 0.0
        for vtx in self.vertexes():
            proj = project_point_to_plane_hv(vtx, info.dir_h, info.dir_v, self.crosshair)
            x = length_vector_in_line(proj, info.dir_h, self.crosshair)
            if x >= 0:
                xp = xp if xp > x else x
            else:
                xn = xn if xn < x else x
            y = length_vector_in_line(proj, info.dir_v, self.crosshair)
            if y >= 0:
                yp = yp if yp > y else y
            else:
                yn = yn if yn < y else y
            zz = distance_point_to_plane_hv(vtx, info.dir_h, info.dir_v, self.crosshair)
            if zz >= 0:
                zp = zp if zp > zz else zz
            else:
                zn = zn if zn < zz else zz
        info.width = int((xp - xn) / sp)
        info.height = int((yp - yn) / sp)
        info.number = int((zp - zn) / sp)

    # -- mappings -------------------------------------------------------------
    def center_point_plane(self, n):
        """MPRInfo::GetCenterPointPlane (core/MPRInfo.cpp:623-625)."""
        return project_point_to_plane(self.center, n, self.crosshair)

    def get_crosshair_point(self, plane: int):
        """MPRInfo::GetCrossHairPoint (core/MPRInfo.cpp:497-511): image coordinates of the crosshair."""
        if plane not in self.planes:
            return None
        info = self.planes[plane]
        vec = _sub(self.crosshair, self.center_point_plane(info.norm()))
        lh = vec[0] * info.dir_h[0] + vec[1] * info.dir_h[1] + vec[2] * info.dir_h[2]
        lv = vec[0] * info.dir_v[0] + vec[1] * info.dir_v[1] + vec[2] * info.dir_v[2]
        return (info.width - 1) / 2.0 + lh / info.pixel_spacing, (info.height - 1) / 2.0 + lv / info.pixel_spacing

    def transfer_image_to_voxel(self, xi: float, yi: float, plane: int):
        """MPRInfo::TransferImage2Voxel (core/MPRInfo.cpp:513-536); the point is in mm."""
        if plane not in self.planes:
            return None
        info = self.planes[plane]
        lt = info.left_top(self.center_point_plane(info.norm()))
        p = _add(lt, _mul(info.dir_h, xi * info.pixel_spacing))
        return _add(p, _mul(info.dir_v, yi * info.pixel_spacing))

    def get_direction(self, plane: int):
        """MPRInfo::GetDirection (core/MPRInfo.cpp:538-611): in-image 2D
        directions of the two crossing planes' normals."""
        group = (PLANE_AXIAL, PLANE_SAGITTAL, PLANE_CORONAL) if plane in (PLANE_AXIAL, PLANE_SAGITTAL,
                                                                           PLANE_CORONAL) else \
            (PLANE_AXIAL_OBLIQUE, PLANE_SAGITTAL_OBLIQUE, PLANE_CORONAL_OBLIQUE)
        if plane not in _CROSS or not all(g in self.planes for g in group):
            return None
        if plane in (PLANE_AXIAL, PLANE_SAGITTAL, PLANE_CORONAL):
            return (1.0, 0.0), (0.0, 1.0)
        hs, vs = self.planes[plane].dir_h, self.planes[plane].dir_v
        if plane == PLANE_AXIAL_OBLIQUE:
            a, b = PLANE_SAGITTAL_OBLIQUE, PLANE_CORONAL_OBLIQUE
        elif plane == PLANE_SAGITTAL_OBLIQUE:
            a, b = PLANE_CORONAL_OBLIQUE, PLANE_AXIAL_OBLIQUE
        else:
            a, b = PLANE_SAGITTAL_OBLIQUE, PLANE_AXIAL_OBLIQUE
        d3h, d3v = self.planes[a].norm(), self.planes[b].norm()

        def d2(x, y):
            n = math.sqrt(x * x + y * y)
            return (math.nan, math.nan) if n == 0.0 else (x / n, y / n)

        lh = d3h[0] * hs[0] + d3h[1] * hs[1] + d3h[2] * hs[2]
        lv = d3h[0] * vs[0] + d3h[1] * vs[1] + d3h[2] * vs[2]
        dh = d2(lh, lv)
        lh = d3v[0] * hs[0] + d3v[1] * hs[1] + d3v[2] * hs[2]
        lv = d3v[0] * vs[0] + d3v[1] * vs[1] + d3v[2] * vs[2]
        return dh, d2(lh, lv)

    def get_direction_3d(self, plane: int):
        if plane not in self.planes:
            return None
        return self.planes[plane].dir_h, self.planes[plane].dir_v

    # -- kernel arguments -----------------------------------------------------
    def plane_params(self, plane: int, invert_z: bool = False, sampling: int = MK_SAMPLE_LINEAR) -> PlaneParams | None:
        """MPRImageLayer::GetGrayscaleData (core/MPRImageLayer.cpp:84-201):
        size padded to even, left-top corner from the plane's centre point,
        slab half count from thickness / pixel spacing, all rounded to float."""
        if plane not in self.planes:
            return None
        info = self.planes[plane]
        w, h = info.width, info.height
        if w % 2:
            w += 1
        n = info.norm()
        ps = info.pixel_spacing
        center = self.center_point_plane(n)
        lt = _sub(center, _mul(info.dir_h, 0.5 * w * ps))
        lt = _sub(lt, _mul(info.dir_v, 0.5 * h * ps))
        n_slice = int(info.thickness / ps)
        n_slice = 1 if n_slice < 1 else n_slice
        half = _f(_f(1.0 * (n_slice - 1)) / 2)
        p = PlaneParams()
        p.width, p.height = w, h
        p.mpr_type = info.mpr_type
        p.sampling = sampling
        p.invert_z = int(invert_z)
        for i in range(3):
            p.spacing[i] = _f(self.spacing[i])
            p.dir_h[i] = _f(info.dir_h[i])
            p.dir_v[i] = _f(info.dir_v[i])
            p.dir_n[i] = _f(n[i])
            p.pt_left_top[i] = _f(lt[i])
        p.pixel_spacing = _f(ps)
        p.half_num = half
        return p


# -- batch series (core/BatchInfo.h, DataManager::GetBatchDirection3D) ------------

def batch_direction_3d(dir_h, dir_v, angle: float):
    """DataManager::GetBatchDirection3D (core/DataManager.cpp:613-644) for a
    plane frame (dir_h, dir_v): dir_h turned by ``angle`` degrees about the
    plane normal a = dir_h x dir_v (the axis-angle rotation matrix, applied by
    GetTransferPoint), and a itself as the vertical direction."""
    a = cross(dir_h, dir_v)
    t = angle / 180 * PI
    x, y, z = a
    xx, yy, zz = x * x, y * y, z * z
    s, c = math.sin(t), math.cos(t)
    ct = 1.0 - c
    m = [[xx + (1 - xx) * c, x * y * ct - z * s, x * z * ct + y * s],
         [x * y * ct + z * s, yy + (1 - yy) * c, y * z * ct - x * s],
         [x * z * ct - y * s, y * z * ct + x * s, zz + (1 - zz) * c]]
    return _dir(*transfer_point(m, dir_h)), a


class BatchInfo:
    """core/BatchInfo.h:29-54: a series of ``num`` slab images of one frame.
    Points and lengths are in mm, like the MPR state."""

    def __init__(self, dir_h=(1.0, 0.0, 0.0), dir_v=(0.0, 1.0, 0.0), center=(0.0, 0.0, 0.0), length_h=0.0,
                 length_v=0.0, pixel_spacing=0.0, slice_thickness=0.0, slice_distance=0.0, num=0,
                 mpr_type=MK_MPR_AVERAGE):
        self.dir_h, self.dir_v, self.center = tuple(dir_h), tuple(dir_v), tuple(center)
        self.length_h, self.length_v = float(length_h), float(length_v)
        self.pixel_spacing = float(pixel_spacing)
        self.slice_thickness, self.slice_distance = float(slice_thickness), float(slice_distance)
        self.num = int(num)
        self.mpr_type = int(mpr_type)

    def width(self) -> int:
        return int(self.length_h / self.pixel_spacing)

    def height(self) -> int:
        return int(self.length_v / self.pixel_spacing)


def batch_plane_params(info: BatchInfo, spacing, invert_z: bool = False,
                       sampling: int = MK_SAMPLE_LINEAR) -> list | None:
    """Kernel arguments of every image of a batch series, or None when the
    series is empty or degenerate.

    The reference declares the series (core/BatchInfo.h) and the entry point
    (Render::GetBatchData, core/Render.cpp:117-120) but leaves it unimplemented.
    The series here: ``num`` parallel slabs spanned by (dir_h, dir_v), image k
    centred at center + (k - (num - 1) / 2) * slice_distance * n with
    n = dir_h x dir_v, each width() x height() pixels of pixel_spacing and a
    slab of slice_thickness combined by mpr_type, with the per-image arithmetic
    of the MPR planes (left-top corner from the centre, slab half count from
    thickness / pixel spacing, everything rounded to float)."""
    if info.num <= 0 or info.pixel_spacing <= 0:
        return None
    w, h = info.width(), info.height()
    if w <= 0 or h <= 0:
        return None
    dh, dv = _dir(*info.dir_h), _dir(*info.dir_v)
    n = cross(dh, dv)
    if any(math.isnan(c) for c in dh + dv + n):
        return None
    ps = info.pixel_spacing
    n_slice = int(info.slice_thickness / ps)
    n_slice = 1 if n_slice < 1 else n_slice
    half = _f(_f(1.0 * (n_slice - 1)) / 2)
    out = []
    for k in range(info.num):
        center = _add(info.center, _mul(n, (k - (info.num - 1) / 2.0) * info.slice_distance))
        lt = _sub(center, _mul(dh, 0.5 * w * ps))
        lt = _sub(lt, _mul(dv, 0.5 * h * ps))
        p = PlaneParams()
        p.width, p.height = w, h
        p.mpr_type = info.mpr_type
        p.sampling = sampling
        p.invert_z = int(invert_z)
        for i in range(3):
            p.spacing[i] = _f(spacing[i])
            p.dir_h[i] = _f(dh[i])
            p.dir_v[i] = _f(dv[i])
            p.dir_n[i] = _f(n[i])
            p.pt_left_top[i] = _f(lt[i])
        p.pixel_spacing = _f(ps)
        p.half_num = half
        out.append(p)
    return out
