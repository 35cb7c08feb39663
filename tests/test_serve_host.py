"""Serving admission rule and request generator, on the CPU."""
import numpy as np

from monkeygl_amd import params as P
from monkeygl_amd import workloads as W
from monkeygl_amd._capi import MK_RENDER_MIP, MK_RENDER_VR
from monkeygl_amd.serve import KIND_CPR, KIND_PLANE, KIND_RENDER, Request, latency_stats, plan_batch

DIMS, SP = (64, 64, 64), (1.0, 1.0, 1.0)


def _vr(size, volume=0, rt=MK_RENDER_VR, step=1.0, alpha=1.0):
    info = P.RenderInfo()
    info.set_vr_size(size, size)
    p = P.render_params(DIMS, SP, info, {0: P.LabelState(alpha, 400.0, 40.0)}, render_type=rt, width=size,
                        height=size, step_scale=step)
    return Request(KIND_RENDER, volume, render=p)


def _mpr(size, volume=0):
    return Request(KIND_PLANE, volume, plane=P.plane_params(DIMS, SP, P.PLANE_AXIAL, size=(size, size)))


def _cpr(width, h, volume=0):
    return Request(KIND_CPR, volume, cpr_width=width, cpr_points=np.zeros((h, 3)), cpr_dirs=np.ones((h, 3)))


def test_request_sizes():
    r = _vr(100)
    assert r.size == (100, 100) and r.out_bytes() == 30000 and r.tiles() == 49
    assert r.arena_bytes() == 30208
    m = _mpr(40)
    assert m.out_bytes() == 3200 and m.tiles() == 9
    c = _cpr(102, 30)
    assert c.size == (102, 30) and c.out_bytes() == 102 * 30 * 2 and c.geo_bytes() == 2 * 30 * 24
    assert c.tiles() == 7 * 2


def test_plan_fifo_prefix_by_items_tiles_bytes():
    q = [_vr(64) for _ in range(10)]
    b, rej = plan_batch(q, max_items=4, max_tiles=10_000, arena_bytes=1 << 30)
    assert b == q[:4] and rej == []
    b, _ = plan_batch(q, max_items=64, max_tiles=16 * 3, arena_bytes=1 << 30)
    assert b == q[:3]
    per = q[0].arena_bytes()
    b, _ = plan_batch(q, max_items=64, max_tiles=10_000, arena_bytes=per * 5 + 10)
    assert b == q[:5]


def test_plan_stops_at_first_misfit_keeps_order():
    q = [_vr(64), _vr(1024), _vr(32)]
    b, _ = plan_batch(q, max_items=64, max_tiles=4096 + 8, arena_bytes=1 << 30)
    # the 1024^2 frame (4096 tiles) fits a slot alone but cannot follow the
    # first; the small one behind it waits too
    assert b == q[:1]


def test_plan_rejects_requests_that_never_fit():
    big = _vr(1024)
    q = [big, _vr(32), _mpr(32)]
    b, rej = plan_batch(q, max_items=64, max_tiles=10_000, arena_bytes=big.arena_bytes() - 1)
    assert rej == [big] and b == q[1:]
    b, rej = plan_batch(q, max_items=64, max_tiles=4095, arena_bytes=1 << 30)
    assert rej == [big]


def test_plan_splits_on_shared_table_conflicts():
    q = [_vr(32), _vr(32, rt=MK_RENDER_MIP), _vr(32, alpha=0.5)]
    b, _ = plan_batch(q, 64, 10_000, 1 << 30)
    # VR with another alpha table on the same volume starts a new batch;
    # MIP has its own table
    assert b == q[:2]
    q = [_vr(32), _vr(32, volume=1, alpha=0.5), _vr(32, step=0.5)]
    b, _ = plan_batch(q, 64, 10_000, 1 << 30)
    assert b == q[:2]
    # planes and CPRs carry no shared tables
    q = [_vr(32), _mpr(32), _cpr(102, 20), _mpr(48)]
    assert plan_batch(q, 64, 10_000, 1 << 30)[0] == q


def test_plan_solo_requests_go_alone():
    q = [_vr(32), _vr(32), _vr(32)]
    q[1].solo = True
    assert plan_batch(q, 64, 10_000, 1 << 30)[0] == q[:1]
    assert plan_batch(q[1:], 64, 10_000, 1 << 30)[0] == q[1:2]


def test_serving_requests_mix_and_rate():
    small = [(W.VolumeSpec("a", (24, 20, 18), (1.0, 1.0, 1.0), 2, (2, 4), (300, 900), seed=1), 0),
             (W.VolumeSpec("b", (20, 24, 22), (0.7, 0.7, 1.0), 2, (2, 4), (300, 900), seed=2), 2)]
    scene = W.serving_scene(small)
    reqs = W.serving_requests(scene, 3000, qps=500.0, seed=3, sizes=(16, 32))
    kinds = np.array([r.tag[1] for r in reqs])
    for k, p in W.SERVE_MIX:
        assert abs((kinds == k).mean() - p) < 0.03, k
    t = np.array([r.t_arrival for r in reqs])
    assert np.all(np.diff(t) >= 0)
    gaps = np.diff(t)
    assert abs(gaps.mean() - 1 / 500.0) < 0.1 / 500.0
    # exponential gaps: coefficient of variation ~ 1
    assert 0.9 < gaps.std() / gaps.mean() < 1.1
    assert {r.volume for r in reqs} == {0, 1}
    sizes = {r.size for r in reqs if r.kind != KIND_CPR}
    assert sizes == {(16, 16), (32, 32)}


def test_latency_stats():
    rs = []
    for i in range(100):
        r = _mpr(16)
        r.t_arrival = float(i)
        r.t_done = float(i) + (i + 1) * 1e-3
        rs.append(r)
    bad = _mpr(16)
    bad.error = "x"
    st = latency_stats(rs + [bad])
    assert st["n"] == 100
    assert abs(st["p50_ms"] - 50.5) < 1e-9
    assert abs(st["max_ms"] - 100.0) < 1e-9
