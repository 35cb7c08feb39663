"""A request with bad parameters gets the same status in a batch as alone.

The batched launch runs the one-request entry points' argument checks on each
item: non-finite pixel spacing, half slab width or step scale, unknown render
type, MPR type or sampling, and bad sizes are refused with the same status by
mk_batch_submit as by mk_render_async / mk_render_plane_async /
mk_render_cpr_async, and a good batch still runs afterwards.
"""
import math

import numpy as np
import pytest

from monkeygl_amd import params as P
from monkeygl_amd import workloads as W
from monkeygl_amd._capi import (MK_ITEM_CPR, MK_ITEM_PLANE, MK_ITEM_RENDER, MK_RENDER_VR, MK_SAMPLE_NEAREST,
                                Batcher, BatchItem, Context, MkError)

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

DIMS = (48, 40, 36)


@pytest.fixture(scope="module")
def setup():
    spec = W.VolumeSpec("bval", DIMS, (1.0, 1.0, 1.2), 4, (3, 8), (300, 1500), seed=7)
    ctx = Context(0)
    ctx.set_volume(W.make_volume(spec))
    ctx.set_lut(0, P.transfer_function(W.tf_points(0)))
    b = Batcher([ctx], max_items=4, max_tiles=1024, nslots=1)
    arena = torch.zeros(1 << 20, dtype=torch.uint8, device="cuda")
    yield ctx, b, arena
    b.close()
    ctx.close()


def _render(**kw):
    info = P.RenderInfo()
    info.set_vr_size(32, 32)
    p = P.render_params(DIMS, (1.0, 1.0, 1.2), info, {0: P.LabelState(1.0, 2000.0, 0.0)}, render_type=MK_RENDER_VR)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _plane(**kw):
    q = P.plane_params(DIMS, (1.0, 1.0, 1.2), P.PLANE_AXIAL, sampling=MK_SAMPLE_NEAREST)
    for k, v in kw.items():
        setattr(q, k, v)
    return q


def _status(fn):
    try:
        fn()
    except MkError as e:
        return e.status
    return 0


RENDER_BAD = [dict(step_scale=math.nan), dict(step_scale=math.inf), dict(step_scale=-1.0), dict(render_type=7),
              dict(width=0), dict(height=20000)]
PLANE_BAD = [dict(pixel_spacing=math.inf), dict(pixel_spacing=math.nan), dict(pixel_spacing=0.0),
             dict(half_num=math.nan), dict(half_num=-1.0), dict(sampling=5), dict(mpr_type=9), dict(width=-3)]
CPR_BAD = [(0, 10, 0), (16, 10, 3), (16, 70000, 0)]


@pytest.mark.parametrize("bad", RENDER_BAD)
def test_render_item_status_matches_alone(setup, bad):
    ctx, b, arena = setup
    p = _render(**bad)
    alone = _status(lambda: ctx.render_async(p, arena, 0))
    it = BatchItem(kind=MK_ITEM_RENDER, context=0, out_offset=0, render=p)
    batched = _status(lambda: b.submit(0, [it], arena, arena.numel(), 0))
    assert alone != 0 and batched == alone, (bad, alone, batched)


@pytest.mark.parametrize("bad", PLANE_BAD)
def test_plane_item_status_matches_alone(setup, bad):
    ctx, b, arena = setup
    q = _plane(**bad)
    alone = _status(lambda: ctx.render_plane_async(q, arena, 0))
    it = BatchItem(kind=MK_ITEM_PLANE, context=0, out_offset=0, plane=q)
    batched = _status(lambda: b.submit(0, [it], arena, arena.numel(), 0))
    assert alone != 0 and batched == alone, (bad, alone, batched)


@pytest.mark.parametrize("w,h,sampling", CPR_BAD)
def test_cpr_item_status_matches_alone(setup, w, h, sampling):
    ctx, b, arena = setup
    geo = 4096
    alone = _status(lambda: ctx.render_cpr_async(w, h, arena[geo:].data_ptr(), arena[geo:].data_ptr(), False,
                                                 sampling, arena, 0))
    it = BatchItem(kind=MK_ITEM_CPR, context=0, out_offset=0, cpr_width=w, cpr_height=h, cpr_invert_z=0,
                   cpr_sampling=sampling, cpr_points_offset=geo, cpr_dirs_offset=geo)
    batched = _status(lambda: b.submit(0, [it], arena, arena.numel(), 0))
    assert alone != 0 and batched == alone, ((w, h, sampling), alone, batched)


def test_good_batch_runs_after_refusals(setup):
    ctx, b, arena = setup
    q = _plane()
    it = BatchItem(kind=MK_ITEM_PLANE, context=0, out_offset=0, plane=q)
    b.submit(0, [it], arena, arena.numel(), 0)
    b.wait(0)
    got = arena[:q.width * q.height * 2].view(torch.int16).cpu().numpy().reshape(q.height, q.width)
    assert np.array_equal(got, ctx.render_plane(q))
