"""MPR plane state (monkeygl_amd/mpr.py) against the reference's helpers and
closed forms.

core/MPRInfo.cpp includes DataManager.h -> CuDataManager.h -> cuda_runtime.h
and does not compile here. Its helpers do: the golden vectors in
tests/golden/mpr_geometry.json come from the reference's Methods.cpp and
PlaneInfo compiled unmodified. The stateful parts (frames, sizes, Browse,
RotateCrossHair, plane index, crosshair/image mapping) are pinned by closed
forms and round trips."""
import json
import math
from pathlib import Path

import numpy as np
import pytest

from monkeygl_amd import cpr as G
from monkeygl_amd import mpr as M
from monkeygl_amd import params as P

GOLD = json.loads((Path(__file__).resolve().parent / "golden" / "mpr_geometry.json").read_text())
CARDIAC = ((512, 512, 361), (0.351, 0.351, 0.3))


def _h(xs):
    return tuple(float.fromhex(x) for x in xs)


def test_plane_helpers_match_reference_bits():
    for c in GOLD["dist_plane_hv"]:
        h = G.direction(_h(c["h0"]), _h(c["h1"]))
        v = G.direction(_h(c["v0"]), _h(c["v1"]))
        assert M.distance_point_to_plane_hv(_h(c["pt"]), h, v, _h(c["pp"])) == _h(c["out"])[0]
    for c in GOLD["project_hv"]:
        h = G.direction(_h(c["h0"]), _h(c["h1"]))
        v = G.direction(_h(c["v0"]), _h(c["v1"]))
        assert M.project_point_to_plane_hv(_h(c["pt"]), h, v, _h(c["pp"])) == _h(c["out"])
    for c in GOLD["transfer"]:
        m9 = _h(c["m"])
        m = [m9[0:3], m9[3:6], m9[6:9]]
        assert M.transfer_point(m, _h(c["pt"])) == _h(c["out"])
    for c in GOLD["plane_info"]:
        info = M.PlaneInfo(0)
        info.dir_h = G.direction(_h(c["h0"]), _h(c["h1"]))
        info.dir_v = G.direction(_h(c["v0"]), _h(c["v1"]))
        info.width, info.height, info.pixel_spacing = c["w"], c["h"], float.fromhex(c["ps"])
        assert info.norm() == _h(c["norm"])
        assert info.left_top(_h(c["center"])) == _h(c["left_top"])


def test_init_sizes_closed_form():
    dims, sp = CARDIAC
    # min spacing 0.3: x and y stretch by 0.351 / 0.3
    assert M.plane_init_size(dims, sp, M.PLANE_AXIAL) == (599, 599, 361)
    assert M.plane_init_size(dims, sp, M.PLANE_SAGITTAL) == (599, 361, 599)
    assert M.plane_init_size(dims, sp, M.PLANE_CORONAL_OBLIQUE) == (599, 361, 599)
    assert M.plane_init_size(dims, sp, M.PLANE_VR) == (512, 512, 1)
    assert M.plane_init_size((0, 1, 1), sp, M.PLANE_AXIAL) is None
    assert M.plane_init_size(dims, sp, 8) is None
    # the function the earlier default-frame helper used agrees
    for pt in M.MPR_PLANES:
        assert M.plane_init_size(dims, sp, pt) == P.plane_init_size(dims, sp, pt)


def test_default_frames_match_plane_params():
    dims, sp = CARDIAC
    m = M.MPRInfo(dims, sp)
    for pt in (M.PLANE_AXIAL, M.PLANE_SAGITTAL, M.PLANE_CORONAL):
        a = m.plane_params(pt)
        b = P.plane_params(dims, sp, pt)
        assert bytes(a) == bytes(b), pt
    # thick slab: halfNum = (int(thick / ps) - 1) / 2
    m.update_thickness(2.0)
    assert m.plane_params(M.PLANE_AXIAL).half_num == np.float32(2.5)


def test_browse_moves_along_normal():
    dims, sp = CARDIAC
    m = M.MPRInfo(dims, sp)
    c0 = m.crosshair
    m.browse(3.0, M.PLANE_AXIAL)          # normal (1,0,0) x (0,1,0) = +z
    assert m.crosshair == (c0[0], c0[1], c0[2] + 3.0)
    m.browse(-1.5, M.PLANE_CORONAL)       # normal (1,0,0) x (0,0,-1) = (0, 1, 0)
    assert m.crosshair == (c0[0], c0[1] - 1.5, c0[2] + 3.0)
    # the axial plane moved with the crosshair: its left-top corner's z follows
    p = m.plane_params(M.PLANE_AXIAL)
    assert abs(p.pt_left_top[2] - np.float32(c0[2] + 3.0)) == 0.0
    # its index counts slices of min spacing from the centre
    assert m.get_plane_index(M.PLANE_AXIAL) == int(3.0 / 0.3) + (361 - 1) // 2


def test_plane_index_round_trip():
    dims, sp = CARDIAC
    m = M.MPRInfo(dims, sp)
    for idx in (0, 17, 180, 359):
        m.set_plane_index(idx, M.PLANE_AXIAL)
        assert m.get_plane_index(M.PLANE_AXIAL) in (idx, idx - 1)   # truncation of a rounded distance
        z = m.crosshair[2]
        assert abs(z - (0.5 * 361 * 0.3 + (idx - 180) * 0.3)) < 1e-9


def test_rotate_crosshair_quarter_turn_and_back():
    dims, sp = CARDIAC
    m = M.MPRInfo(dims, sp)
    h0, v0 = m.get_direction_3d(M.PLANE_SAGITTAL)
    m.rotate_crosshair(90.0, M.PLANE_AXIAL)
    # a quarter turn about +z through the crosshair: (0,1,0) -> (-1,0,0)
    h, v = m.get_direction_3d(M.PLANE_SAGITTAL)
    assert np.allclose(h, (-1.0, 0.0, 0.0), atol=1e-7) and np.allclose(v, v0, atol=1e-12)
    hc, _ = m.get_direction_3d(M.PLANE_CORONAL)
    assert np.allclose(hc, (0.0, 1.0, 0.0), atol=1e-7)
    # sizes follow the projected corners: the turned sagittal plane spans x
    w, hh = m.get_plane_size(M.PLANE_SAGITTAL)
    assert w == int(512 * 0.351 / 0.3 + 1e-6) or w == int(512 * 0.351 / 0.3) - 1
    assert hh == 361
    # the angle is absolute: back to 0 restores the frames
    m.rotate_crosshair(0.0, M.PLANE_AXIAL)
    h, v = m.get_direction_3d(M.PLANE_SAGITTAL)
    assert np.allclose(h, h0, atol=1e-7) and np.allclose(v, v0, atol=1e-12)
    # a general angle keeps the frames orthonormal and the planes through the crosshair
    m.rotate_crosshair(37.0, M.PLANE_AXIAL)
    for pt in (M.PLANE_SAGITTAL, M.PLANE_CORONAL):
        h, v = m.get_direction_3d(pt)
        assert abs(np.dot(h, v)) < 1e-12 and abs(np.linalg.norm(h) - 1) < 1e-12
        n = np.cross(h, v)
        assert abs(np.dot(n, np.subtract(m.center_point_plane(tuple(n)), m.crosshair))) < 1e-9


def test_crosshair_image_round_trip():
    dims, sp = CARDIAC
    m = M.MPRInfo(dims, sp)
    m.browse(2.0, M.PLANE_SAGITTAL)
    m.rotate_crosshair(20.0, M.PLANE_CORONAL)
    for pt in (M.PLANE_AXIAL, M.PLANE_SAGITTAL, M.PLANE_CORONAL):
        x, y = m.get_crosshair_point(pt)
        q = m.transfer_image_to_voxel(x, y, pt)
        # TransferImage2Voxel maps image (x, y) from the left-top corner;
        # GetCrossHairPoint measures from the pixel centre (W-1)/2: half a pixel apart
        info = m.planes[pt]
        off = np.array(info.dir_h) * 0.5 * info.pixel_spacing + np.array(info.dir_v) * 0.5 * info.pixel_spacing
        assert np.allclose(np.array(q) + off, m.crosshair, atol=1e-9)
    # PanCrossHair moves the crosshair to an image point of the plane
    m.pan_crosshair(10.0, 20.0, M.PLANE_AXIAL)
    assert np.allclose(m.crosshair, m.transfer_image_to_voxel(10.0, 20.0, M.PLANE_AXIAL), atol=1e-12)


def test_oblique_2d_directions():
    dims, sp = CARDIAC
    m = M.MPRInfo(dims, sp)
    assert m.get_direction(M.PLANE_AXIAL) == ((1.0, 0.0), (0.0, 1.0))
    m.rotate_crosshair(30.0, M.PLANE_AXIAL_OBLIQUE)
    dh, dv = m.get_direction(M.PLANE_AXIAL_OBLIQUE)
    # the sagittal-oblique normal (0,1,0) x (0,0,-1) = (-1,0,0) and the
    # coronal-oblique normal (0,1,0), turned by 30 degrees about +z, seen in
    # the axial-oblique image
    a = math.radians(30.0)
    assert np.allclose(dh, (-math.cos(a), -math.sin(a)), atol=1e-6)
    assert np.allclose(dv, (-math.sin(a), math.cos(a)), atol=1e-6)


def test_rotate_matrix_rows():
    dims, sp = CARDIAC
    m = M.MPRInfo(dims, sp)
    r = m.get_plane_rotate_matrix(M.PLANE_CORONAL)
    assert np.array_equal(r, np.array([1, 0, 0, 0, 0, -1, 0, 1, 0], np.float32) + 0.0)
    assert m.get_plane_rotate_matrix(7) is None
