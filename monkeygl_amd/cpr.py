"""CPR geometry: the per-row points and directions mk_render_cpr samples.

Restates the host double-precision geometry of core/CPRInfo.cpp:45-238 with
its helpers (core/Methods.cpp:563-633 distance, projection and rotation;
core/Direction.h:55-154 normalising constructors, cross, negative;
core/Point.h:89-103 distances; core/VolumeInfo.cpp:317-323 Patient2Voxel).
Arithmetic follows the reference's types: coordinates in double, rotation
angles handed to RotatePoint3D as float (cosf / sinf), radii read back
through the float getters (core/CPRInfo.h:51-56), widths truncated to int.

Quirks kept as the reference has them:
* a straightened CPR leaves its last row unset: point (0, 0, 0), direction
  Direction3d() = (1, 0, 0) (core/Point.h:32-34, core/Direction.h:32-37);
* a direction built from a zero vector normalises to NaN, and the
  ``Length() < 0.99`` fallback never fires on NaN;
* a centreline point on the start-end axis can give sqrt of a slightly
  negative number, NaN, which the stretched radius maximum skips.
"""
from __future__ import annotations

import math

import numpy as np

PI = 3.141592654   # core/Defines.h:33

PLANE_STRETCHED_CPR = "stretched"
PLANE_STRAIGHTENED_CPR = "straightened"

_F32 = np.float32


def _normalize(v):
    # Direction::Normalize (core/Direction.h:144-154): divide by the length,
    # NaN for a zero vector, as IEEE division gives
    x, y, z = v
    n = math.sqrt(x * x + y * y + z * z)
    if n == 0.0:
        return (math.nan, math.nan, math.nan)
    return (x / n, y / n, z / n)


def direction(a, b):
    """Direction3d(ptStart, ptEnd): normalised b - a."""
    return _normalize((b[0] - a[0], b[1] - a[1], b[2] - a[2]))


def cross(a, b):
    return _normalize((a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]))


def dot(a, b) -> float:
    v = 0.0
    v += a[0] * b[0]
    v += a[1] * b[1]
    v += a[2] * b[2]
    return v


def length(a) -> float:
    return math.sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2])


def distance(a, b) -> float:
    return math.sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]))


def patient_distance(a, b, sp) -> float:
    ds = (a[0] - b[0]) * (a[0] - b[0]) * sp[0] * sp[0] + (a[1] - b[1]) * (a[1] - b[1]) * sp[1] * sp[1]
    ds += (a[2] - b[2]) * (a[2] - b[2]) * sp[2] * sp[2]
    return math.sqrt(ds)


def length_vector_in_line(pt, d, pl) -> float:
    """Methods::Length_VectorInLine (core/Methods.cpp:570-574)."""
    return (pt[0] - pl[0]) * d[0] + (pt[1] - pl[1]) * d[1] + (pt[2] - pl[2]) * d[2]


def distance_point_to_line(pt, d, pl) -> float:
    """Methods::Distance_Point2Line (core/Methods.cpp:563-568)."""
    lv = distance(pt, pl)
    lp = length_vector_in_line(pt, d, pl)
    q = lv * lv - lp * lp
    return math.sqrt(q) if q >= 0.0 else math.nan   # C sqrt of a rounding-negative value: NaN


def projection_point_to_plane(pt, n, pp):
    """Methods::Projection_Point2Plane (core/Methods.cpp:588-604)."""
    dist = (pt[0] - pp[0]) * n[0] + (pt[1] - pp[1]) * n[1] + (pt[2] - pp[2]) * n[2]
    return (pt[0] - dist * n[0], pt[1] - dist * n[1], pt[2] - dist * n[2])


def _cosf_sinf(theta: float):
    t = float(_F32(theta))
    # correctly rounded float cos / sin of the float angle
    return float(_F32(math.cos(t))), float(_F32(math.sin(t)))


def rotate_point(pt, axis, center, theta: float):
    """Methods::RotatePoint3D (core/Methods.cpp:606-633): rotation of pt
    about the axis through center by theta (radians, taken as float)."""
    c, s = _cosf_sinf(theta)
    dx, dy, dz = axis
    x0, y0, z0 = pt[0] - center[0], pt[1] - center[1], pt[2] - center[2]
    k = float(_F32(1.0 - c))   # int 1 minus a float is a float subtraction
    x = x0 * (dx * dx * k + c) + y0 * (dx * dy * k - dz * s) + z0 * (dx * dz * k + dy * s)
    y = x0 * (dy * dx * k + dz * s) + y0 * (dy * dy * k + c) + z0 * (dy * dz * k - dx * s)
    z = x0 * (dz * dx * k - dy * s) + y0 * (dy * dz * k + dx * s) + z0 * (dz * dz * k + c)
    return (x + center[0], y + center[1], z + center[2])


def first_direction_projection(pt, n):
    """CPRInfo::FirstDirectionProjection (core/CPRInfo.cpp:187-208)."""
    if abs(n[1]) > 0.3:
        q = (pt[0], pt[1] - n[2] / n[1], pt[2] + 1)
    elif abs(n[0]) > 0.3:
        q = (pt[0] - n[2] / n[0], pt[1], pt[2] + 1)
    else:
        q = (pt[0] + 1, pt[1], pt[2] - n[0] / n[2])
    return direction(pt, q)


def patient_to_voxel(pt, origin, dir_x, dir_y, dir_z, spacing):
    """VolumeInfo::Patient2Voxel (core/VolumeInfo.cpp:317-323)."""
    return (length_vector_in_line(pt, dir_x, origin) / spacing[0],
            length_vector_in_line(pt, dir_y, origin) / spacing[1],
            length_vector_in_line(pt, dir_z, origin) / spacing[2])


class CPRInfo:
    """Host CPR state (core/CPRInfo.cpp:30-40 defaults)."""

    def __init__(self):
        self.dir_stretched = (0.0, 0.0, 1.0)
        self.angle_stretched = 0.0
        self.angle_straightened = 0.0
        self.min_spacing = 1.0
        self.spacing = (1.0, 1.0, 1.0)
        self.stretched_radius = 0.0
        self.straightened_radius = 50.0
        self.origin_stretched = (0.0, 0.0, 0.0)
        self.origin_straightened = (0.0, 0.0, 0.0)
        self.line: list = []

    # float getters (core/CPRInfo.h:51-56)
    def _r_stretched(self) -> float:
        return float(_F32(self.stretched_radius))

    def _r_straightened(self) -> float:
        return float(_F32(self.straightened_radius))

    def set_spacing(self, spacing) -> None:
        self.spacing = tuple(float(s) for s in spacing)
        m = self.spacing[0] if self.spacing[0] < self.spacing[1] else self.spacing[1]
        self.min_spacing = m if m < self.spacing[2] else self.spacing[2]

    def rotate_cpr(self, angle: float, plane: str) -> bool:
        if plane == PLANE_STRETCHED_CPR:
            self.angle_stretched += float(_F32(angle))
            return True
        if plane == PLANE_STRAIGHTENED_CPR:
            self.angle_straightened += float(_F32(angle))
            return True
        return False

    def set_cpr_line_voxel(self, pts) -> bool:
        self.line = [tuple(float(c) for c in p) for p in np.asarray(pts, np.float64).reshape(-1, 3)]
        return self._update()

    def set_cpr_line_patient(self, pts, origin, dir_x, dir_y, dir_z) -> bool:
        """SetCPRLinePatient with the volume's Patient2Voxel frame."""
        self.line = [patient_to_voxel(tuple(map(float, p)), origin, dir_x, dir_y, dir_z, self.spacing)
                     for p in np.asarray(pts, np.float64).reshape(-1, 3)]
        return self._update()

    def _update(self) -> bool:
        """CPRInfo::UpdateCPRInfo (core/CPRInfo.cpp:210-238)."""
        if len(self.line) < 2:
            return False
        start, end = self.line[0], self.line[-1]
        self.dir_stretched = direction(start, end)
        r = 0.0
        for p in self.line[1:]:
            d = distance_point_to_line(p, self.dir_stretched, start)
            if d > r:
                r = d
        self.stretched_radius = r + 100.0
        d0 = first_direction_projection(start, self.dir_stretched)
        rs, rt = self._r_stretched(), self._r_straightened()
        self.origin_stretched = (start[0] + d0[0] * rs, start[1] + d0[1] * rs, start[2] + d0[2] * rs)
        self.origin_straightened = (start[0] + d0[0] * rt, start[1] + d0[1] * rt, start[2] + d0[2] * rt)
        self.angle_stretched = 0.0
        self.angle_straightened = 0.0
        return True

    def get_cpr_info(self, plane: str):
        """(points (h, 3), dirs (h, 3), width, height) in voxel units, or None
        (CPRInfo::GetCPRInfo, core/CPRInfo.cpp:68-79)."""
        if plane == PLANE_STRETCHED_CPR:
            return self._stretched()
        if plane == PLANE_STRAIGHTENED_CPR:
            return self._straightened()
        return None

    def _stretched(self):
        """CPRInfo::GetCPRInfoStretched (core/CPRInfo.cpp:81-115)."""
        if len(self.line) < 2:
            return None
        width = int(self._r_stretched() * 2.0 + 1)
        if width % 2:
            width += 1
        start = self.line[0]
        rot = rotate_point(self.origin_stretched, self.dir_stretched, start, self.angle_stretched * PI / 180.0)
        d = direction(rot, start)
        pts = [projection_point_to_plane(start, d, rot)]
        prev = pts[0]
        for p in self.line[1:]:
            cur = projection_point_to_plane(p, d, rot)
            if patient_distance(cur, prev, self.spacing) >= self.min_spacing:
                prev = cur
                pts.append(prev)
        points = np.array(pts, np.float64)
        dirs = np.tile(np.array(d, np.float64), (len(pts), 1))
        return points, dirs, width, len(pts)

    def _straightened(self):
        """CPRInfo::GetCPRInfoStraightened (core/CPRInfo.cpp:117-160)."""
        if len(self.line) < 2:
            return None
        width = int(self._r_straightened() * 2.0 + 1)
        if width % 2:
            width += 1
        n = len(self.line)
        points = np.zeros((n, 3), np.float64)
        dirs = np.zeros((n, 3), np.float64)
        dirs[:, 0] = 1.0   # rows the loop never sets keep Direction3d() = (1, 0, 0)
        start = self.line[0]
        path = direction(start, self.line[1])
        rot = rotate_point(self.origin_straightened, path, start, self.angle_straightened * PI / 180.0)
        d = direction(rot, start)
        points[0] = rot
        dirs[0] = d
        r = distance(rot, start)
        for i in range(1, n - 1):
            pre, cur, nxt = self.line[i - 1], self.line[i], self.line[i + 1]
            d1 = tuple(dirs[i - 1])
            d2 = direction(pre, cur)
            d3 = direction(cur, nxt)
            dt = cross(d1, d2)
            di = cross(d3, dt)
            if length(di) < float(_F32(0.99)):
                di = d1
            if dot(di, d1) < 0:
                di = _normalize((-di[0], -di[1], -di[2]))
            dirs[i] = di
            points[i] = (cur[0] - di[0] * r, cur[1] - di[1] * r, cur[2] - di[2] * r)
        return points, dirs, width, n
