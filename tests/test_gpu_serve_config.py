"""Render serving at its stated configuration.

Three resident volumes (the 128^3 blob set, the 512x512x361 cardiac-sized set
and the 512x512x1559 body-sized set with a 4-label mask), each with its own
TFs; 120 seeded requests with Poisson arrivals at 500 requests/s, 60% VR /
20% MIP / 15% MPR / 5% CPR, square viewports from {256, 512, 768, 1024}^2 and
random cameras, served by the continuous-batching server (2 slots, 512 MiB
HBM budget), the shape of examples/server.py:303-318,497-530 served many at a
time. Bars:
* every served result equals the same request served alone (a batch of one)
  and rendered through the plain one-request entry points, bit for bit;
* VR and MIP: a band of 8 rows through the middle of each image within 1/255
  of the FP64 oracle on >= 99.9% of its pixels;
* MPR and CPR: nearest sampling bit-exact against the float oracle, trilinear
  within 1 HU of FP64 on >= 99.9% of pixels (whole images).
"""
import dataclasses

import numpy as np
import pytest

from monkeygl_amd import workloads as W
from monkeygl_amd._capi import MK_SAMPLE_NEAREST, Context
from monkeygl_amd.serve import KIND_CPR, KIND_PLANE, KIND_RENDER, RenderServer, render_one
from oracle import oracle as O

pytestmark = pytest.mark.gpu

N_REQUESTS = 120
QPS = 500.0


@pytest.fixture(scope="module")
def scene():
    sc = W.serving_scene()
    ctxs = []
    for sv in sc:
        c = Context(0)
        c.set_volume(sv.vol)
        if sv.mask is not None:
            c.set_mask(sv.mask)
        for l, lut in sv.luts.items():
            c.set_lut(l, lut)
        c.synchronize()
        ctxs.append(c)
    yield sc, ctxs
    for c in ctxs:
        c.close()


def _fresh(r):
    return dataclasses.replace(r, t_arrival=0.0, result=None, error=None, t_admit=0.0, t_done=0.0, solo=False)


def test_served_mix_at_config_size(scene):
    sc, ctxs = scene
    O.set_threads(16)
    reqs = W.serving_requests(sc, N_REQUESTS, QPS, seed=3)
    sizes = {r.size[0] for r in reqs if r.kind == KIND_RENDER}
    assert sizes == {256, 512, 768, 1024}, sizes
    assert {r.tag[1] for r in reqs} == {"vr", "mip", "mpr", "cpr"}
    assert {r.volume for r in reqs} == {0, 1, 2}
    srv = RenderServer(ctxs, hbm_budget=512 << 20, nslots=2, max_items=64, max_tiles=65536)
    try:
        print(f"[serve] {len(reqs)} requests at {QPS} qps over three volumes", flush=True)
        srv.run(reqs)
        assert all(r.error is None for r in reqs), [(r.tag, r.error) for r in reqs if r.error]
        assert max(srv.batch_sizes) > 1
        for r in reqs:
            alone = _fresh(r)
            srv.run([alone])
            assert alone.error is None, alone.error
            assert np.array_equal(alone.result, r.result), ("alone", r.tag)
            assert np.array_equal(render_one(ctxs[r.volume], r), r.result), ("plain", r.tag)
    finally:
        srv.close()
    print(f"[serve] {len(reqs)} results equal alone and plain, {srv.batches} batches", flush=True)

    worst = {}
    for r in reqs:
        sv = sc[r.volume]
        kind = r.tag[1]
        if r.kind == KIND_RENDER:
            p = r.render
            y0 = p.height // 2 - 4
            ref = O.render_rows(p, sv.vol, sv.luts, y0, y0 + 8, mask_zyx=sv.mask, precision="f64")
            d = np.abs(r.result[y0:y0 + 8].astype(int) - ref.astype(int)).max(-1)
            frac = float((d <= 1).mean())
            assert frac >= 0.999, (r.tag, r.volume, r.size, frac, int(d.max()))
        elif r.kind == KIND_PLANE:
            q = r.plane
            if q.sampling == MK_SAMPLE_NEAREST:
                assert np.array_equal(r.result, O.render_plane(q, sv.vol, precision="f32")), r.tag
                frac = 1.0
            else:
                d = np.abs(r.result.astype(np.int32) - O.render_plane(q, sv.vol).astype(np.int32))
                frac = float((d <= 1).mean())
                assert frac >= 0.999, (r.tag, frac, int(d.max()))
        else:
            assert r.kind == KIND_CPR
            prec = "f32" if r.cpr_sampling == MK_SAMPLE_NEAREST else "f64"
            ref = O.render_cpr(r.cpr_width, r.cpr_points, r.cpr_dirs, r.cpr_invert_z, sv.vol, r.cpr_sampling,
                               precision=prec)
            if r.cpr_sampling == MK_SAMPLE_NEAREST:
                assert np.array_equal(r.result, ref), r.tag
                frac = 1.0
            else:
                d = np.abs(r.result.astype(np.int32) - ref.astype(np.int32))
                frac = float((d <= 1).mean())
                assert frac >= 0.999, (r.tag, frac, int(d.max()))
        worst[kind] = min(worst.get(kind, 1.0), frac)
    print(f"[serve] oracle parity, worst fraction within the bar per kind: {worst}", flush=True)
