"""Framing extremes: zoom, pan, exact axis views and cropped VOIs.

The kernels start each ray at the float-lattice point below the first
accumulated length that can pass the VOI test and stop after the last one
(a conservative interval with a margin), and a sample's brick decides what is
fetched. Those are the only parts of the mapping that are not the reference's
loop (core/kernel.cu:161-211), and framing moves every ray's interval: zoom
and pan come from RenderInfo (core/RenderInfo.cpp:82-99), the VOI from
cu_render's argument. Each frame here is compared whole against the FP64
oracle: every RGB channel within 1/255 on >= 99.9% of pixels; frames that
show no part of the volume must be exactly the background.
"""
import numpy as np
import pytest

from monkeygl_amd import params as P
from monkeygl_amd import workloads as W
from monkeygl_amd._capi import MK_RENDER_MIP, MK_RENDER_SURFACE, MK_RENDER_VR
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TYPES = [MK_RENDER_VR, MK_RENDER_MIP, MK_RENDER_SURFACE]
BG = (0.1, 0.2, 0.3, 1.0)


@pytest.fixture(scope="module")
def small():
    return W.make_volume(W.SMALL), P.transfer_function(W.tf_points(0))


@pytest.fixture(scope="module")
def cardiac():
    return W.make_volume(W.CARDIAC), P.transfer_function(W.tf_points(1))


def _info(size, view=None, preset=None, zoom=None, pan=None):
    info = P.RenderInfo()
    info.set_vr_size(size, size)
    if preset:
        getattr(info, preset)()
    if view:
        info.rotate(*view)
    if zoom:
        info.zoom(zoom)
    if pan:
        info.pan(*pan)
    return info


def _check(ctx, p, vol, luts, tag):
    got = ctx.render(p)
    ref = O.render(p, vol, luts, precision="f64")
    d = np.abs(got.astype(int) - ref.astype(int)).max(-1)
    frac = float((d <= 1).mean())
    print(f"{tag}: {frac:.6f} within 1/255, max {int(d.max())}", flush=True)
    assert frac >= 0.999, (tag, frac, int(d.max()))
    return got, ref


def _small_params(rt, info, labels=None, **kw):
    labels = labels or {0: P.LabelState(1.0, 2000.0, 0.0)}
    return P.render_params((128, 128, 128), (1.0, 1.0, 1.0), info, labels, render_type=rt, background=BG, **kw)


@pytest.mark.parametrize("rt", TYPES)
@pytest.mark.parametrize("zoom", [0.25, 4.0, 16.0])
def test_zoom_extremes(gpu_ctx, small, rt, zoom):
    vol, lut = small
    gpu_ctx.set_volume(vol)
    gpu_ctx.set_lut(0, lut)
    for view in [(20.0, 35.0), (0.0, 0.0)]:
        p = _small_params(rt, _info(256, view=view, zoom=zoom))
        _check(gpu_ctx, p, vol, {0: lut}, f"rt={rt} zoom={zoom} view={view}")


@pytest.mark.parametrize("rt", TYPES)
def test_pan_half_off_screen(gpu_ctx, small, rt):
    vol, lut = small
    gpu_ctx.set_volume(vol)
    gpu_ctx.set_lut(0, lut)
    for pan, zoom in [((128.0, 0.0), None), ((0.0, -128.0), None), ((-128.0, 96.0), None), ((-170.0, -140.0), None),
                      ((200.0, 150.0), 4.0), ((-60.0, 40.0), 0.25)]:
        p = _small_params(rt, _info(256, view=(15.0, 25.0), zoom=zoom, pan=pan))
        got, _ = _check(gpu_ctx, p, vol, {0: lut}, f"rt={rt} pan={pan} zoom={zoom}")
        # part of the frame shows the volume, part only background (Surface
        # may find no surface above 400 HU in the part that shows)
        bg = np.all(got == np.array([int(c * 255) for c in BG[:3]], np.uint8), axis=-1)
        if rt != MK_RENDER_SURFACE:
            assert 0.0 < float(bg.mean()) < 1.0, (pan, zoom, float(bg.mean()))


@pytest.mark.parametrize("rt", TYPES)
def test_pan_fully_off_screen_is_background(gpu_ctx, small, rt):
    vol, lut = small
    gpu_ctx.set_volume(vol)
    gpu_ctx.set_lut(0, lut)
    expect = np.array([int(c * 255) for c in BG[:3]], np.uint8)
    for pan, zoom in [((400.0, 0.0), None), ((0.0, -420.0), None), ((-300.0, 300.0), None), ((3000.0, 0.0), 16.0)]:
        p = _small_params(rt, _info(256, view=(33.0, 12.0), zoom=zoom, pan=pan))
        got, ref = _check(gpu_ctx, p, vol, {0: lut}, f"rt={rt} pan={pan} zoom={zoom}")
        assert np.all(ref == expect) and np.all(got == expect), (pan, zoom)


@pytest.mark.parametrize("rt", TYPES)
@pytest.mark.parametrize("invert_z", [False, True])
def test_exact_axis_views_small(gpu_ctx, small, rt, invert_z):
    vol, lut = small
    gpu_ctx.set_volume(vol)
    gpu_ctx.set_lut(0, lut)
    for preset in ["anterior", "posterior", "left", "right", "head", "foot"]:
        p = _small_params(rt, _info(256, preset=preset), invert_z=invert_z)
        _check(gpu_ctx, p, vol, {0: lut}, f"rt={rt} invz={invert_z} {preset}")
    for view in [(90.0, 0.0), (180.0, 0.0), (0.0, 90.0), (0.0, 180.0), (-90.0, 90.0)]:
        p = _small_params(rt, _info(256, view=view), invert_z=invert_z)
        _check(gpu_ctx, p, vol, {0: lut}, f"rt={rt} invz={invert_z} view={view}")


@pytest.mark.parametrize("rt", [MK_RENDER_VR, MK_RENDER_MIP])
def test_exact_axis_views_cardiac(gpu_ctx, cardiac, rt):
    O.set_threads(16)
    vol, lut = cardiac
    gpu_ctx.set_volume(vol)
    gpu_ctx.set_lut(0, lut)
    labels = {0: P.LabelState(1.0, 900.0, 250.0)}
    cases = [dict(preset="anterior"), dict(preset="left"), dict(preset="head"), dict(view=(90.0, 0.0)),
             dict(view=(180.0, 0.0)), dict(view=(0.0, 90.0)), dict(preset="foot", zoom=3.0, pan=(40.0, -30.0))]
    for c in cases:
        p = P.render_params(W.CARDIAC.dims, W.CARDIAC.spacing, _info(384, **c), labels, render_type=rt,
                            background=BG)
        _check(gpu_ctx, p, vol, {0: lut}, f"cardiac rt={rt} {c}")


@pytest.mark.parametrize("rt", TYPES)
def test_cropped_voi(gpu_ctx, small, rt):
    """cu_render's VOI (core/kernel.cu:208-211) cropping the volume, with a
    mask-free label palette; a VOI of one slab and one of a single voxel
    column included."""
    vol, lut = small
    gpu_ctx.set_volume(vol)
    gpu_ctx.set_lut(0, lut)
    vois = [(20, 100, 10, 90, 30, 110), (64, 64, 0, 127, 0, 127), (0, 127, 40, 41, 0, 127), (50, 50, 60, 60, 10, 120),
            (100, 127, 0, 30, 90, 127)]
    for voi in vois:
        for view, zoom in [((0.0, 0.0), None), ((30.0, 20.0), 2.0), ((90.0, 0.0), None), ((-40.0, 65.0), 0.5)]:
            p = _small_params(rt, _info(256, view=view, zoom=zoom))
            (p.voi.left, p.voi.right, p.voi.posterior, p.voi.anterior, p.voi.head, p.voi.foot) = voi
            _check(gpu_ctx, p, vol, {0: lut}, f"rt={rt} voi={voi} view={view} zoom={zoom}")
