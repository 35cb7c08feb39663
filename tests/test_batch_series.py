"""Batch series geometry (no GPU): DataManager::GetBatchDirection3D
(core/DataManager.cpp:613-644) against closed forms of the axis-angle
rotation, and the series definition (core/BatchInfo.h:29-54; the renderer is
a stub in the reference, core/Render.cpp:117-120): image sizes, centres,
spacing along the normal, slab half count."""
import math

import numpy as np
import pytest

from monkeygl_amd import mpr as M
from monkeygl_amd import pyMonkeyGL as mk
from monkeygl_amd._capi import MK_MPR_MIP, MK_SAMPLE_NEAREST

FRAMES = [((1, 0, 0), (0, 1, 0)), ((0, 1, 0), (0, 0, -1)), ((1, 0, 0), (0, 0, -1)),
          (M._dir(1, 1, 0), M._dir(-1, 1, 0.5))]


def _cross(a, b):
    return np.cross(np.asarray(a, float), np.asarray(b, float))


@pytest.mark.parametrize("h,v", FRAMES)
def test_batch_direction_closed_forms(h, v):
    h, v = M._dir(*h), M._dir(*v)
    n = _cross(h, v)
    n /= np.linalg.norm(n)
    # 0 degrees: dirH unchanged; dirV is always the plane normal
    dh, dv = M.batch_direction_3d(h, v, 0.0)
    assert np.allclose(dh, h, atol=1e-12) and np.allclose(dv, n, atol=1e-12)
    # 90 degrees: n x h (counter-clockwise about the normal); 180: -h
    dh, _ = M.batch_direction_3d(h, v, 90.0)
    assert np.allclose(dh, _cross(n, h), atol=1e-8)
    dh, _ = M.batch_direction_3d(h, v, 180.0)
    assert np.allclose(dh, -np.asarray(h), atol=1e-8)
    # any angle: unit length, in the plane, at that angle from dirH
    for ang in (17.0, -63.5, 251.0):
        dh, dv = M.batch_direction_3d(h, v, ang)
        assert abs(np.linalg.norm(dh) - 1) < 1e-12 and abs(np.dot(dh, dv)) < 1e-9
        c = np.dot(dh, h)
        s = np.dot(np.cross(h, dh), n)
        assert math.isclose(math.degrees(math.atan2(s, c)) % 360, ang % 360, abs_tol=1e-6)


def test_batch_direction_through_the_api():
    hm = mk.HelloMonkey()
    hm.SetVolumeArray(np.zeros((16, 12, 10), np.int16))
    dh, dv = hm.GetBatchDirection3D(90.0, mk.PlaneAxial)
    assert np.allclose(dh.tuple(), (0, 1, 0), atol=1e-8) and np.allclose(dv.tuple(), (0, 0, 1))
    dh, dv = hm.GetBatchDirection3D(0.0, mk.PlaneCoronal)
    assert np.allclose(dh.tuple(), (1, 0, 0)) and np.allclose(dv.tuple(), (0, 1, 0))
    assert hm.GetBatchDirection3D(0.0, mk.PlaneVR) is None


def _centre(p):
    lt = np.array(list(p.pt_left_top), float)
    return lt + 0.5 * p.width * p.pixel_spacing * np.array(list(p.dir_h)) + \
        0.5 * p.height * p.pixel_spacing * np.array(list(p.dir_v))


@pytest.mark.parametrize("num", [1, 2, 7])
def test_series_geometry(num):
    h, v = M.batch_direction_3d((1.0, 0.0, 0.0), (0.0, 0.0, -1.0), 30.0)
    info = M.BatchInfo(h, v, (40.0, 35.0, 30.0), 60.5, 45.0, 0.5, 3.2, 1.25, num, MK_MPR_MIP)
    ps = M.batch_plane_params(info, (0.8, 0.9, 1.2), sampling=MK_SAMPLE_NEAREST)
    assert len(ps) == num
    n = _cross(h, v)
    for k, p in enumerate(ps):
        assert (p.width, p.height) == (121, 90)   # int(60.5 / 0.5), int(45 / 0.5)
        assert p.mpr_type == MK_MPR_MIP and p.sampling == MK_SAMPLE_NEAREST
        assert p.half_num == 2.5                  # int(3.2 / 0.5) = 6 samples
        exp = np.array([40.0, 35.0, 30.0]) + (k - (num - 1) / 2) * 1.25 * n
        assert np.allclose(_centre(p), exp, atol=1e-4), k
        assert np.allclose(list(p.dir_n), n, atol=1e-7)
        assert list(p.spacing) == pytest.approx([0.8, 0.9, 1.2])
    if num > 1:   # consecutive images one slice distance apart along the normal
        d = np.diff([_centre(p) for p in ps], axis=0)
        assert np.allclose(d @ n, 1.25, atol=1e-4) and np.allclose(np.cross(d, n), 0, atol=1e-4)


def test_degenerate_series():
    base = dict(center=(1.0, 1.0, 1.0), length_h=10.0, length_v=10.0, pixel_spacing=1.0, slice_thickness=1.0,
                slice_distance=1.0)
    assert M.batch_plane_params(M.BatchInfo((1, 0, 0), (0, 1, 0), num=0, **base), (1, 1, 1)) is None
    assert M.batch_plane_params(M.BatchInfo((1, 0, 0), (1, 0, 0), num=3, **base), (1, 1, 1)) is None
    assert M.batch_plane_params(M.BatchInfo((0, 0, 0), (0, 1, 0), num=3, **base), (1, 1, 1)) is None
    b = dict(base, length_h=0.5)
    assert M.batch_plane_params(M.BatchInfo((1, 0, 0), (0, 1, 0), num=3, **b), (1, 1, 1)) is None
    # the API: a zero BatchInfo (the reference's memset constructor) yields no images, without a device
    hm = mk.HelloMonkey()
    hm.SetVolumeArray(np.zeros((16, 12, 10), np.int16))
    bi = mk.BatchInfo()
    bi.fPixelSpacing = 1.0
    assert (bi.Width(), bi.Height()) == (0, 0)
    assert hm.GetBatchData(bi) == []
