"""Work enqueued without waiting on streams other than the context's, while the
tables it reads change between calls.

A context rebuilds its LUTs, per-render-type skip tables and step lattice on
its own stream when a frame asks for other ones. Frames still running on a
caller's stream read the old tables, so the rebuild must wait for them on the
device. Each asynchronous frame here must equal the synchronous render of the
same parameters, bit for bit: one non-context stream, then two streams taking
turns. Marching-cubes calls on two streams share the context's scratch and
must each equal their synchronous result.
"""
import numpy as np
import pytest

from monkeygl_amd import params as P
from monkeygl_amd import workloads as W
from monkeygl_amd._capi import MK_RENDER_MIP, MK_RENDER_SURFACE, MK_RENDER_VR

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

DIMS = (128, 128, 128)
SIZE = 512


def _labels(k):
    """Palette k: label alphas and windows that change the skip tables."""
    return {0: P.LabelState([0.4, 0.8, 1.0][k % 3], 2000.0 - 300.0 * (k % 2), 50.0 * (k % 3)),
            1: P.LabelState([1.0, 0.3, 0.6][k % 3], 1500.0, -100.0),
            2: P.LabelState([0.8, 1.0, 0.0][k % 3], 1700.0 + 100.0 * k, -200.0),
            3: P.LabelState(0.6, 1300.0, 100.0 * (k % 2))}


def _luts(k):
    """LUT set k: label l reads the seeded control points k * 10 + l."""
    return {l: P.transfer_function(W.tf_points(10 * k + l)) for l in range(4)}


def _frames():
    """(render type, palette, LUT set, step scale, view) per frame."""
    out = []
    for i in range(12):
        rt = [MK_RENDER_VR, MK_RENDER_VR, MK_RENDER_MIP, MK_RENDER_VR, MK_RENDER_SURFACE, MK_RENDER_VR][i % 6]
        out.append((rt, i % 3, (i // 2) % 3, [1.0, 0.5, 1.0, 0.75][i % 4], (17.0 * i, 9.0 * (i % 5))))
    return out


def _params(rt, pal, step, view):
    info = P.RenderInfo()
    info.set_vr_size(SIZE, SIZE)
    info.rotate(*view)
    return P.render_params(DIMS, (1.0, 1.0, 1.0), info, _labels(pal), render_type=rt, step_scale=step)


@pytest.fixture(scope="module")
def scene():
    vol = W.make_volume(W.SMALL)
    mask = W.ellipsoid_mask(DIMS, 3, seed=5)
    return vol, mask


def _set_luts(ctx, k):
    for l, a in _luts(k).items():
        ctx.set_lut(l, a)


def _reference(ctx, frames):
    refs = []
    for rt, pal, lk, step, view in frames:
        _set_luts(ctx, lk)
        refs.append(ctx.render(_params(rt, pal, step, view)))
    return refs


@pytest.mark.parametrize("nstreams", [1, 2])
def test_async_frames_across_palette_lut_and_step_changes(gpu_ctx, scene, nstreams):
    vol, mask = scene
    gpu_ctx.set_volume(vol)
    gpu_ctx.set_mask(mask)
    frames = _frames()
    refs = _reference(gpu_ctx, frames)
    streams = [torch.cuda.Stream() for _ in range(nstreams)]
    outs = [torch.empty((SIZE, SIZE, 3), dtype=torch.uint8, device="cuda") for _ in frames]
    torch.cuda.synchronize()
    for i, (rt, pal, lk, step, view) in enumerate(frames):
        # new LUTs are uploaded by the next launch, after the frames before it
        _set_luts(gpu_ctx, lk)
        gpu_ctx.render_async(_params(rt, pal, step, view), outs[i], streams[i % nstreams].cuda_stream)
    torch.cuda.synchronize()
    bad = [i for i, (o, r) in enumerate(zip(outs, refs)) if not np.array_equal(o.cpu().numpy(), r)]
    print(f"nstreams={nstreams}: {len(frames) - len(bad)}/{len(frames)} async frames equal their synchronous render")
    assert not bad, bad


def test_async_marching_cubes_on_two_streams(gpu_ctx, scene):
    vol, _ = scene
    gpu_ctx.set_volume(vol)
    cases = [((0, 0, 0, 128, 128, 128), 300.0), ((10, 20, 5, 100, 90, 127), 800.0),
             ((0, 0, 0, 128, 128, 128), -200.0), ((40, 0, 30, 128, 64, 128), -200.0)]
    refs = [gpu_ctx.marching_cubes(b, iso) for b, iso in cases]
    assert all(len(r) > 0 for r in refs), [len(r) for r in refs]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    outs, counts = [], []
    torch.cuda.synchronize()
    for i, (b, iso) in enumerate(cases):
        cap = len(refs[i])
        out = torch.empty((max(cap, 1), 3, 3), dtype=torch.float32, device="cuda")
        cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
        gpu_ctx.marching_cubes_async(b, iso, out, cap, cnt, streams[i % 2].cuda_stream)
        outs.append(out)
        counts.append(cnt)
    torch.cuda.synchronize()
    for i, r in enumerate(refs):
        assert int(counts[i].item()) == len(r), (i, int(counts[i].item()), len(r))
        assert np.array_equal(outs[i][:len(r)].cpu().numpy(), r), i
