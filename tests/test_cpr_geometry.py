"""CPR geometry (monkeygl_amd/cpr.py) against the reference's own helpers and
closed forms.

The helpers are pinned by golden vectors written by the reference's
Methods.cpp / Direction.h / Point.h compiled unmodified
(tests/golden/make_golden.py). CPRInfo.cpp itself does not compile here (it
includes DataManager.h -> CuDataManager.h -> cuda_runtime.h), so the stretched
and straightened producers are pinned by closed forms on straight centrelines
and by invariants on curved ones."""
import json
import math
from pathlib import Path

import numpy as np
import pytest

from monkeygl_amd import cpr as G
from monkeygl_amd import workloads as W

GOLD = json.loads((Path(__file__).resolve().parent / "golden" / "cpr_geometry.json").read_text())


def _h(xs):
    return tuple(float.fromhex(x) for x in xs)


def test_direction_matches_reference_bits():
    for c in GOLD["direction"]:
        assert G.direction(_h(c["a"]), _h(c["b"])) == _h(c["out"])


def test_cross_matches_reference_bits():
    for c in GOLD["cross"]:
        a = G.direction(_h(c["a0"]), _h(c["a1"]))
        b = G.direction(_h(c["b0"]), _h(c["b1"]))
        assert G.cross(a, b) == _h(c["out"])


def test_projection_matches_reference_bits():
    for c in GOLD["project"]:
        n = G.direction(_h(c["n0"]), _h(c["n1"]))
        assert G.projection_point_to_plane(_h(c["pt"]), n, _h(c["pp"])) == _h(c["out"])


def test_distance_to_line_matches_reference_bits():
    for c in GOLD["dist_line"]:
        d = G.direction(_h(c["d0"]), _h(c["d1"]))
        assert G.distance_point_to_line(_h(c["pt"]), d, _h(c["pl"])) == _h(c["out"])[0]


def test_patient_distance_matches_reference_bits():
    for c in GOLD["patient_dist"]:
        assert G.patient_distance(_h(c["a"]), _h(c["b"]), _h(c["sp"])) == _h(c["out"])[0]


def test_rotate_point_matches_reference():
    """cosf/sinf of the float angle: the restatement rounds the double cos/sin
    to float (correctly rounded); the C library's cosf may differ by one float
    ulp, which moves a point by at most |r| * 2^-24 (relative 6e-8)."""
    exact = 0
    for c in GOLD["rotate"]:
        ax = G.direction(_h(c["ax0"]), _h(c["ax1"]))
        got = G.rotate_point(_h(c["pt"]), ax, _h(c["center"]), float.fromhex(c["theta"]))
        ref = _h(c["out"])
        r = math.dist(_h(c["pt"]), _h(c["center"]))
        assert all(abs(g - e) <= 2 * r * 2.0 ** -24 + 1e-12 for g, e in zip(got, ref)), (got, ref)
        exact += got == ref
    assert exact >= len(GOLD["rotate"]) * 7 // 8


def _straight_line(n=50, x=100.0, y=120.0, z0=10.0, dz=1.0):
    return np.array([(x, y, z0 + dz * i) for i in range(n)], np.float64)


def test_stretched_straight_line_closed_form():
    line = _straight_line()
    info = G.CPRInfo()
    info.set_spacing((0.5, 0.5, 0.8))
    assert info.set_cpr_line_voxel(line)
    pts, dirs, w, h = info.get_cpr_info(G.PLANE_STRETCHED_CPR)
    # radius 0 + 100 -> width 201 -> padded to even
    assert w == 202
    # axis (0, 0, 1): FirstDirectionProjection takes its third branch, (1, 0, 0)
    assert np.array_equal(dirs, np.tile([-1.0, 0.0, 0.0], (h, 1)))
    # projected onto the plane x = x0 + 100; points kept every >= 0.5 mm
    # (spacing z 0.8 mm per 1-voxel step: every point is kept)
    assert h == len(line)
    assert np.allclose(pts, line + [100.0, 0.0, 0.0], atol=1e-12)
    # a quarter turn about the line swings the plane to y
    info.rotate_cpr(90.0, G.PLANE_STRETCHED_CPR)
    pts, dirs, w, h = info.get_cpr_info(G.PLANE_STRETCHED_CPR)
    assert np.allclose(dirs[0], [0.0, -1.0, 0.0], atol=1e-6)
    assert np.allclose(pts[:, 1], line[0, 1] + 100.0, atol=1e-4)


def test_stretched_keeps_min_spacing():
    line = _straight_line(n=40, dz=0.25)
    info = G.CPRInfo()
    info.set_spacing((1.0, 1.0, 1.0))
    info.set_cpr_line_voxel(line)
    pts, _, _, h = info.get_cpr_info(G.PLANE_STRETCHED_CPR)
    # 0.25 mm steps, min spacing 1 mm: every fourth point
    assert h == 10
    assert np.allclose(np.diff(pts[:, 2]), 1.0)


def test_straightened_straight_line_closed_form():
    line = _straight_line(n=30)
    info = G.CPRInfo()
    info.set_spacing((1.0, 1.0, 1.0))
    info.set_cpr_line_voxel(line)
    pts, dirs, w, h = info.get_cpr_info(G.PLANE_STRAIGHTENED_CPR)
    assert (w, h) == (102, 30)   # radius 50 -> 101 -> even
    # offset direction from the rotated origin to the start: (-1, 0, 0);
    # along a straight path every row keeps it, points at distance 50
    assert np.allclose(dirs[:-1], np.tile([-1.0, 0.0, 0.0], (h - 1, 1)), atol=1e-12)
    assert np.allclose(pts[:-1], line[:-1] + [50.0, 0.0, 0.0], atol=1e-9)
    # the reference never sets the last row
    assert np.array_equal(pts[-1], [0.0, 0.0, 0.0])
    assert np.array_equal(dirs[-1], [1.0, 0.0, 0.0])


def test_curved_line_invariants():
    dims = (128, 128, 128)
    line = W.centerline(dims, 200, seed=2)
    info = G.CPRInfo()
    info.set_spacing((0.7, 0.7, 1.0))
    info.set_cpr_line_voxel(line)
    info.rotate_cpr(33.0, G.PLANE_STRETCHED_CPR)
    pts, dirs, w, h = info.get_cpr_info(G.PLANE_STRETCHED_CPR)
    assert w % 2 == 0 and 2 <= h <= len(line)
    # all points on one plane whose normal is the common direction
    d = dirs[0]
    assert np.allclose(dirs, d)
    off = pts @ d
    assert np.ptp(off) < 1e-9 * max(1.0, np.abs(off).max())
    # consecutive kept points at least min spacing apart (patient units)
    sp = np.array(info.spacing)
    gaps = np.linalg.norm(np.diff(pts, axis=0) * sp, axis=1)
    assert gaps.min() >= info.min_spacing - 1e-12
    info.rotate_cpr(-20.0, G.PLANE_STRAIGHTENED_CPR)
    pts, dirs, w, h = info.get_cpr_info(G.PLANE_STRAIGHTENED_CPR)
    assert h == len(line) and w == 102
    n = np.linalg.norm(dirs[:-1], axis=1)
    assert np.allclose(n, 1.0)
    # each offset stays perpendicular to the path to first order and at radius 50
    r = np.linalg.norm(pts[1:-1] - line[1:-1], axis=1)
    assert np.allclose(r, 50.0, rtol=1e-3)


def test_too_short_line():
    info = G.CPRInfo()
    assert not info.set_cpr_line_voxel(np.zeros((1, 3)))
    assert info.get_cpr_info(G.PLANE_STRETCHED_CPR) is None
    assert info.get_cpr_info("mpr") is None
