"""Marching cubes: the triangle table, the oracle's extraction, and the GPU
kernel (bit-exact triangle lists)."""
import re
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O

REF_MC = Path("/root/reference/core/MarchingCube.cpp")
CORNER = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
EDGE = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]


def test_table_matches_reference_table():
    if not REF_MC.exists():
        pytest.skip("reference not mounted")
    src = REF_MC.read_text()
    start = src.index("char table[256][16] = {")
    end = src.index("};", start)
    rows = re.findall(r"\{([^{}]*)\}", src[start + len("char table[256][16] = {"):end])
    ref = np.array([[int(x) for x in r.split(",")] for r in rows], np.int8)
    tri, cnt = O.mc_table()
    assert np.array_equal(tri, ref)
    assert np.array_equal(cnt, (ref >= 0).sum(1) // 3)


def _cube_volume(case):
    v = np.zeros((2, 2, 2), np.int16)
    for c in range(8):
        x, y, z = CORNER[c]
        v[z, y, x] = 0 if (case >> c) & 1 else 1000
    return v


@pytest.mark.parametrize("case", range(256))
def test_single_cube_cases(case):
    """Each case uses exactly the edges whose corners straddle the iso value,
    and with corner values 0/1000 at iso 500 every vertex is an edge midpoint."""
    tri, cnt = O.mc_table()
    v = _cube_volume(case)
    t = O.marching_cubes(v, (0, 0, 0, 2, 2, 2), 500.0)
    assert t.shape[0] == cnt[case]
    inside = [(case >> c) & 1 for c in range(8)]
    active = {e for e, (a, b) in enumerate(EDGE) if inside[a] != inside[b]}
    used = {int(e) for e in tri[case] if e >= 0}
    assert used == active
    mids = {tuple((CORNER[a] + CORNER[b]) / 2.0) for a, b in EDGE}
    for p in t.reshape(-1, 3):
        assert tuple(p.astype(np.float64)) in mids


def _sphere(n=40, r=13.3, c=(19.7, 20.2, 18.9)):
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    d = np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2)
    return np.round(100.0 * (r - d)).astype(np.int16)


def test_sphere_is_closed_with_euler_characteristic_2():
    vol = _sphere()
    t = O.marching_cubes(vol, (0, 0, 0, 40, 40, 40), 0.5)
    assert t.shape[0] > 1000
    verts, inv = np.unique(t.reshape(-1, 3), axis=0, return_inverse=True)
    inv = inv.reshape(-1, 3)
    edges = {}
    for a, b, c in inv:
        for e in ((a, b), (b, c), (c, a)):
            key = (min(e), max(e))
            edges[key] = edges.get(key, 0) + 1
    assert all(n == 2 for n in edges.values())          # watertight
    assert len(verts) - len(edges) + len(inv) == 2      # genus 0


def test_vertices_sit_on_edges_at_the_iso_value():
    vol = _sphere()
    iso = 0.5
    t = O.marching_cubes(vol, (0, 0, 0, 40, 40, 40), iso).reshape(-1, 3).astype(np.float64)
    fl = np.floor(t)
    frac = t - fl
    moving = (frac > 0).sum(1)
    assert np.all(moving <= 1)  # on an axis-aligned edge
    for p, f, m in zip(t[:500], fl[:500], moving[:500]):
        if m == 0:
            continue
        ax = int(np.argmax(p - f > 0))
        lo = f.astype(int)
        hi = lo.copy()
        hi[ax] += 1
        v0, v1 = float(vol[lo[2], lo[1], lo[0]]), float(vol[hi[2], hi[1], hi[0]])
        val = v0 + (p[ax] - f[ax]) * (v1 - v0)
        assert abs(val - iso) < 1e-3 * max(1.0, abs(v1 - v0))


@pytest.mark.gpu
@pytest.mark.parametrize("iso", [200.0, -300.0])
def test_gpu_marching_cubes_bit_exact(gpu_ctx, iso):
    from monkeygl_amd import workloads as W

    vol = W.make_volume(W.SMALL)
    gpu_ctx.set_volume(vol)
    # rows of 127, 84, 1 and 30 cubes: a wave holds parts of several rows, and
    # the last box ends at the volume's far x and z faces
    for box in [(0, 0, 0, 128, 128, 128), (5, 17, 3, 90, 101, 77), (60, 0, 0, 62, 128, 128), (97, 30, 64, 128, 70, 128)]:
        got = gpu_ctx.marching_cubes(box, iso)
        ref = O.marching_cubes(vol, box, iso)
        assert got.shape == ref.shape and ref.shape[0] > 0
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


@pytest.mark.gpu
def test_gpu_marching_cubes_rejects_bad_box(gpu_ctx):
    from monkeygl_amd._capi import MkError

    gpu_ctx.set_volume(np.zeros((8, 8, 8), np.int16))
    with pytest.raises(MkError):
        gpu_ctx.marching_cubes((0, 0, 0, 9, 8, 8), 1.0)
    with pytest.raises(MkError):
        gpu_ctx.marching_cubes((3, 0, 0, 4, 8, 8), 1.0)
