"""The orientation figure's checker (no GPU): the reference's own
FillHoleInImage_Ch1 under the restated lines of MergeOrientationBox /
GetMeshPoints / LoadBinarySTL (oracle/_ref), pinned by closed forms: the STL
normalisation, a facet facing the viewer lighting its interior with the
reference's colour, back faces culled, the first of two facets at equal depth
winning, and no corner for images of at most 64 pixels."""
import ctypes as C
import struct
from pathlib import Path

import numpy as np
import pytest

from monkeygl_amd.workloads import figure, stl_bytes

REF = Path(__file__).resolve().parent.parent / "oracle" / "_ref" / "libmkref.so"
pytestmark = pytest.mark.skipif(not REF.exists(), reason="oracle/_ref not built")


def _lib():
    L = C.CDLL(str(REF))
    L.ref_stl_normalize.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    L.ref_orientation_box.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int]
    return L


def _normalize(stl):
    n = struct.unpack_from("<I", stl, 80)[0]
    body = np.frombuffer(stl, np.uint8, 50 * n, 84).copy()
    out = np.zeros(n * 9, np.float32)
    _lib().ref_stl_normalize(body.ctypes.data, n, out.ctypes.data)
    return out.reshape(n, 3, 3)


def _box(fac9, rot, w=100, h=90, fill=0):
    img = np.full((h, w, 3), fill, np.uint8)
    f = np.ascontiguousarray(fac9, np.float32)
    r = np.ascontiguousarray(rot, np.float32)
    _lib().ref_orientation_box(f.ctypes.data, f.shape[0], r.ctypes.data, img.ctypes.data, w, h)
    return img


def test_stl_normalisation():
    f = figure(0)
    n = _normalize(stl_bytes(f))
    lo, hi = f.reshape(-1, 3).min(0), f.reshape(-1, 3).max(0)
    scale = max(hi[0] - lo[0], hi[2] - lo[2])
    exp = (f - (lo + hi) / 2) / scale
    exp[..., 2] *= -1
    assert np.allclose(n, exp, atol=1e-6)
    m = n.reshape(-1, 3)
    assert max(np.ptp(m[:, 0]), np.ptp(m[:, 2])) == pytest.approx(1.0, abs=1e-6)


EYE = np.eye(3, dtype=np.float32).reshape(9)


def test_front_facet_lights_its_interior():
    # one facet in the x-z plane, wound so that its normal is -y (towards the viewer)
    tri = np.array([[[-0.3, 0.0, -0.3], [0.3, 0.0, -0.3], [0.0, 0.0, 0.35]]], np.float32)
    nrm = np.cross(tri[0, 0] - tri[0, 2], tri[0, 1] - tri[0, 2])
    if nrm[1] > 0:
        tri = tri[:, ::-1].copy()
    img = _box(tri, EYE)
    corner = img[-64:, -64:]
    # the figure image is 512 px for [-0.5, 0.5]: sample (r, c) sits at 8c - 256, 8r - 256
    lit = corner.sum(-1) > 0
    assert lit[32, 32]                        # the facet's centroid region
    assert not lit[2, 2] and not lit[62, 62]  # far outside
    assert tuple(corner[32, 32]) == (242, 242, 89)   # int((0.902 + 0.05) * 255), int((0.302 + 0.05) * 255)
    assert (img[:-64].sum() == 0) and (img[:, :-64].sum() == 0)   # nothing outside the corner
    # the back face draws nothing
    assert _box(tri[:, ::-1].copy(), EYE).sum() == 0
    # images of at most 64 pixels are left alone
    assert _box(tri, EYE, 64, 200, fill=7).min() == 7


def test_first_facet_wins_at_equal_depth():
    a = np.array([[-0.3, 0.0, -0.3], [0.3, 0.0, -0.3], [0.0, 0.0, 0.35]], np.float32)
    if np.cross(a[0] - a[2], a[1] - a[2])[1] > 0:
        a = a[::-1].copy()
    # b: the same triangle tilted about x so its diffuse differs, with the
    # same mean depth (y sum) as a
    b = a.copy()
    b[:, 1] = b[:, 2] * 0.5
    b[:, 1] -= b[:, 1].mean()
    ab, ba = _box(np.stack([a, b]), EYE), _box(np.stack([b, a]), EYE)
    ra, rb = _box(a[None], EYE), _box(b[None], EYE)
    both = (ra.sum(-1) > 0) & (rb.sum(-1) > 0)
    assert both.any() and not np.array_equal(ra[both], rb[both])
    assert np.array_equal(ab[both], ra[both]) and np.array_equal(ba[both], rb[both])


def test_cpr_line_projection_matches_the_reference():
    """The drop-in's TransferVoxel2ImageInVR against the reference's own
    Point3d / Methods::MatrixMul arithmetic, bit for bit."""
    from monkeygl_amd import pyMonkeyGL as mk

    L = C.CDLL(str(REF))
    L.ref_voxel_to_vr.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_float,
                                  C.c_float, C.c_int, C.c_int, C.c_void_p]
    rng = np.random.default_rng(2)
    for dims, sp, dirz in (((64, 48, 40), (0.7, 0.7, 1.3), -1.0), ((33, 70, 21), (1.0, 0.5, 2.0), 1.0)):
        hm = mk.HelloMonkey()
        hm.SetVolumeArray(np.zeros(dims, np.int16))
        hm.SetSpacing(*sp)
        hm.SetDirection(mk.Direction3d(1, 0, 0), mk.Direction3d(0, 1, 0), mk.Direction3d(0, 0, dirz))
        hm.Rotate(33.0, -20.0)
        hm.Zoom(1.3)
        hm.Pan(4.5, -7.25)
        pts = rng.uniform(0, 1, (25, 3)) * np.array(dims)
        got = hm._voxel_to_vr([tuple(p) for p in pts], 300, 220)
        exp = np.zeros((25, 2), np.float32)
        d = np.array(dims, np.int32)
        s = np.array(sp, np.float64)
        mt = np.asarray(hm._ri.m_t, np.float32)
        L.ref_voxel_to_vr(np.ascontiguousarray(pts).ctypes.data, 25, d.ctypes.data, s.ctypes.data,
                          int(hm._need_invert_z()), mt.ctypes.data, float(hm._ri.x_translate),
                          float(hm._ri.y_translate), 300, 220, exp.ctypes.data)
        assert np.array_equal(got, exp)
