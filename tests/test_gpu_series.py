"""Batch series (GetBatchData) on the GPU: every image of a series rendered by
the batcher equals mk_render_plane of its parameters alone (bit-identical),
nearest images equal the float oracle bit for bit, trilinear images are within
1 HU of the FP64 oracle on >= 99.9% of pixels; series longer than one batch
are cut into batches; the pyMonkeyGL entry point returns the same images."""
import numpy as np
import pytest

from monkeygl_amd import mpr as M
from monkeygl_amd import pyMonkeyGL as mk
from monkeygl_amd import workloads as W
from monkeygl_amd._capi import MK_MPR_AVERAGE, MK_MPR_MINIP, MK_MPR_MIP, MK_SAMPLE_LINEAR, MK_SAMPLE_NEAREST
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SPEC = W.VolumeSpec("series", (96, 80, 72), (0.8, 0.9, 1.2), 8, (3, 12), (300, 2000), n_vessels=3, seed=17)


@pytest.fixture(scope="module")
def vol():
    return W.make_volume(SPEC)


def _series(num, angle, mpr_type, thick, sampling, lh=60.0, lv=50.0, dist=1.5):
    h, v = M.batch_direction_3d((1.0, 0.0, 0.0), (0.0, 0.0, -1.0), angle)
    c = tuple(0.5 * d * s for d, s in zip(SPEC.dims, SPEC.spacing))
    info = M.BatchInfo(h, v, c, lh, lv, 0.8, thick, dist, num, mpr_type)
    return M.batch_plane_params(info, SPEC.spacing, sampling=sampling)


@pytest.mark.parametrize("mpr_type,thick", [(MK_MPR_AVERAGE, 0.8), (MK_MPR_MIP, 4.0), (MK_MPR_MINIP, 2.5)])
@pytest.mark.parametrize("sampling", [MK_SAMPLE_NEAREST, MK_SAMPLE_LINEAR])
def test_series_matches_single_planes_and_oracle(gpu_ctx, vol, mpr_type, thick, sampling):
    from monkeygl_amd.series import SeriesRenderer

    gpu_ctx.set_volume(vol)
    params = _series(9, 35.0, mpr_type, thick, sampling)
    sr = SeriesRenderer(gpu_ctx)
    try:
        got = sr.render(params)
    finally:
        sr.close()
    assert got.shape == (9, params[0].height, params[0].width)
    for k, p in enumerate(params):
        assert np.array_equal(got[k], gpu_ctx.render_plane(p)), k
        if sampling == MK_SAMPLE_NEAREST:
            assert np.array_equal(got[k], O.render_plane(p, vol, precision="f32")), k
        else:
            d = np.abs(got[k].astype(np.int32) - O.render_plane(p, vol).astype(np.int32))
            assert (d <= 1).mean() >= 0.999, (k, int(d.max()))


def test_long_series_is_cut_into_batches(gpu_ctx, vol):
    from monkeygl_amd.series import SeriesRenderer

    gpu_ctx.set_volume(vol)
    params = _series(37, -20.0, MK_MPR_MIP, 3.0, MK_SAMPLE_NEAREST, lh=70.0, lv=41.0, dist=0.9)
    sr = SeriesRenderer(gpu_ctx, max_items=8, max_tiles=64)   # 6x4 = 24 tiles per image: 2 images per batch
    try:
        got = sr.render(params)
    finally:
        sr.close()
    for k, p in enumerate(params):
        assert np.array_equal(got[k], O.render_plane(p, vol, precision="f32")), k


def test_get_batch_data_through_the_api(vol):
    hm = mk.HelloMonkey()
    hm.SetVolumeArray(vol.swapaxes(2, 0))
    hm.SetSpacing(*SPEC.spacing)
    dh, dv = hm.GetBatchDirection3D(25.0, mk.PlaneSagittal)
    bi = mk.BatchInfo()
    bi.dirH, bi.dirV = dh, dv
    bi.ptCenter = mk.Point3d(38.4, 36.0, 43.2)
    bi.fLengthH, bi.fLengthV, bi.fPixelSpacing = 64.0, 48.0, 0.8
    bi.fSliceThickness, bi.fSliceDistance, bi.nNum = 2.4, 2.0, 6
    bi.mprType = mk.MPRTypeMIP
    imgs = hm.GetBatchData(bi)
    assert len(imgs) == 6 and all(im.shape == (60, 80) for im in imgs)
    params = M.batch_plane_params(bi._series(), SPEC.spacing, invert_z=hm._need_invert_z())
    for k, p in enumerate(params):
        assert np.array_equal(imgs[k], hm._ctx.render_plane(p)), k
        d = np.abs(imgs[k].astype(np.int32) - O.render_plane(p, vol).astype(np.int32))
        assert (d <= 1).mean() >= 0.999, k
