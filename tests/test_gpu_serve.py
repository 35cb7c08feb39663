"""Served rendering on the GPU: each result is the same alone and in the mix.

Requests of every kind (VR with and without a label mask, MIP, slab MPR with
both samplings, stretched and straightened CPR) and ragged viewport sizes run
through the continuous-batching server in a mix of batches in flight on two
slots; each is then rendered alone (a batch of one) and through the plain
one-request entry points. Bar: bit-identical images in all three; the plain
entry points are held to the oracle by the render/plane tests, and a sample
of the served VR frames is checked against the FP64 oracle here as well
(1/255 on >= 99.9% of pixels)."""
import dataclasses

import numpy as np
import pytest

from monkeygl_amd import params as P
from monkeygl_amd import workloads as W
from monkeygl_amd._capi import MK_RENDER_VR, Context
from monkeygl_amd.serve import (KIND_PLANE, KIND_RENDER, RenderServer, Request, latency_stats, render_one,
                                run_sequential)
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SPECS = [
    (W.VolumeSpec("srv_a", (72, 64, 56), (1.0, 1.0, 1.0), 5, (3, 10), (300, 1800), seed=31), 0),
    (W.VolumeSpec("srv_b", (64, 72, 80), (0.7, 0.7, 1.0), 5, (3, 10), (300, 1800), n_vessels=3, seed=32), 3),
    (W.VolumeSpec("srv_c", (48, 48, 96), (0.5, 0.5, 0.8), 4, (3, 10), (300, 1800), seed=33), 2),
]
SIZES = (24, 40, 64, 100)


@pytest.fixture(scope="module")
def served():
    scene = W.serving_scene(SPECS)
    ctxs = []
    for sv in scene:
        c = Context(0)
        c.set_volume(sv.vol)
        if sv.mask is not None:
            c.set_mask(sv.mask)
        for l, lut in sv.luts.items():
            c.set_lut(l, lut)
        c.synchronize()
        ctxs.append(c)
    yield scene, ctxs
    for c in ctxs:
        c.close()


def _fresh(r):
    return dataclasses.replace(r, t_arrival=0.0, result=None, error=None, t_admit=0.0, t_done=0.0, solo=False)


def _requests(scene, n=48, seed=5, qps=0.0):
    return W.serving_requests(scene, n, qps, seed=seed, sizes=SIZES)


def test_mix_equals_alone_equals_plain(served):
    scene, ctxs = served
    reqs = _requests(scene)
    kinds = {r.tag[1] for r in reqs}
    assert kinds == {"vr", "mip", "mpr", "cpr"}, kinds
    srv = RenderServer(ctxs, hbm_budget=16 << 20, nslots=2, max_items=8, max_tiles=4096)
    try:
        print(f"[serve] mix of {len(reqs)} requests", flush=True)
        srv.run(reqs)
        assert all(r.error is None for r in reqs), [r.error for r in reqs if r.error]
        assert srv.batches >= 4 and max(srv.batch_sizes) > 1
        for r in reqs:
            alone = _fresh(r)
            srv.run([alone])
            assert alone.error is None, alone.error
            assert np.array_equal(alone.result, r.result), r.tag
            plain = render_one(ctxs[r.volume], r)
            assert np.array_equal(plain, r.result), r.tag
    finally:
        srv.close()


def test_served_vr_matches_oracle(served):
    scene, ctxs = served
    reqs = [r for r in _requests(scene, n=40, seed=9) if r.kind == KIND_RENDER and
            r.render.render_type == MK_RENDER_VR][:4]
    assert reqs
    srv = RenderServer(ctxs, hbm_budget=8 << 20, nslots=2, max_items=8, max_tiles=4096)
    try:
        srv.run(reqs)
    finally:
        srv.close()
    for r in reqs:
        sv = scene[r.volume]
        ref = O.render(r.render, sv.vol, sv.luts, mask_zyx=sv.mask)
        d = np.abs(r.result.astype(int) - ref.astype(int)).max(-1)
        assert (d <= 1).mean() >= 0.999, (r.tag, (d <= 1).mean(), d.max())


def test_bad_request_fails_alone(served):
    scene, ctxs = served
    reqs = _requests(scene, n=12, seed=11)
    bad_plane = next(r for r in reqs if r.kind == KIND_PLANE)
    bad_plane.plane.pixel_spacing = 0.0          # rejected by the batch validation
    bad_ctx = dataclasses.replace(reqs[0], volume=7)  # no such context
    mix = reqs + [bad_ctx]
    srv = RenderServer(ctxs, hbm_budget=16 << 20, nslots=2, max_items=16, max_tiles=4096)
    try:
        srv.run(mix)
    finally:
        srv.close()
    # the plane entry point's own check, run by the batch on the item
    assert bad_plane.error is not None and "pixel spacing" in bad_plane.error
    assert bad_ctx.error is not None and "context" in bad_ctx.error
    for r in reqs:
        if r is not bad_plane:
            assert r.error is None, (r.tag, r.error)
            assert np.array_equal(r.result, render_one(ctxs[r.volume], r)), r.tag


def test_admission_rejects_oversized(served):
    scene, ctxs = served
    reqs = _requests(scene, n=10, seed=13)
    info = P.RenderInfo()
    info.set_vr_size(512, 512)
    sv = scene[0]
    big = Request(KIND_RENDER, 0, render=P.render_params(sv.spec.dims, sv.spec.spacing, info, sv.labels, width=512,
                                                         height=512))
    srv = RenderServer(ctxs, hbm_budget=2 * 400_000, nslots=2, max_items=16, max_tiles=4096)
    try:
        srv.run(reqs + [big])
    finally:
        srv.close()
    assert big.error is not None and "arena bytes" in big.error
    assert all(r.error is None for r in reqs)


def test_open_loop_poisson(served):
    scene, ctxs = served
    reqs = _requests(scene, n=60, seed=17, qps=400.0)
    seq = [_fresh(r) for r in reqs]
    for a, b in zip(seq, reqs):
        a.t_arrival = b.t_arrival
    srv = RenderServer(ctxs, hbm_budget=16 << 20, nslots=2, max_items=16, max_tiles=4096)
    try:
        srv.run(reqs)
    finally:
        srv.close()
    st = latency_stats(reqs)
    run_sequential(ctxs, seq)
    st_seq = latency_stats(seq)
    print(f"[serve] served {st} sequential {st_seq}", flush=True)
    assert st["n"] == len(reqs)
    for a, b in zip(reqs, seq):
        assert b.t_done >= b.t_arrival and a.t_done >= a.t_arrival
        assert np.array_equal(a.result, b.result), a.tag


def test_encoded_mix_decodes_to_plain(served):
    """encode=True: every result is base64 PNG text (RGB frames; int16 images
    as RGBA of width W/2) that decodes to the plain entry point's image."""
    import base64

    from tests.png_util import read_png

    scene, ctxs = served
    reqs = _requests(scene, n=30, seed=23)
    srv = RenderServer(ctxs, hbm_budget=64 << 20, nslots=2, max_items=8, max_tiles=4096, encode=True)
    try:
        srv.run(reqs)
    finally:
        srv.close()
    for r in reqs:
        assert r.error is None, r.error
        assert isinstance(r.result, str)
        pix, info = read_png(base64.b64decode(r.result, validate=True))
        plain = render_one(ctxs[r.volume], r)
        if r.kind == KIND_RENDER:
            assert info["channels"] == 3 and np.array_equal(pix, plain), r.tag
        else:
            assert info["channels"] == 4 and info["width"] * 2 == r.size[0]
            assert np.array_equal(pix.reshape(-1).view(np.int16).reshape(plain.shape), plain), r.tag
