"""mk_marching_cubes_noalloc: the allocation-free marching-cubes entry point.

It runs the same three kernels as mk_marching_cubes_async, with the per-block
counts, offsets and total in a fixed device array per GPU instead of the
context's grown scratch. Two calls in a row (the second over a smaller box)
must each give the same triangle list and count as mk_marching_cubes_async,
bit for bit, with the full capacity and with a capacity below the total; a
count-only call gives the count; a box beyond the array's limit is refused.

The file is named to be collected last: it creates its own contexts and must
pass alone in a fresh process as well as after every other test.
"""
import numpy as np
import pytest

from monkeygl_amd import _capi
from monkeygl_amd import workloads as W
from monkeygl_amd._capi import MK_ERR_ARG, MkError

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

LIMIT_CUBES = (1 << 20) * 256   # the scratch holds 2^20 blocks of 256 cubes


def _run(fn, box, iso, cap, stream):
    out = torch.empty((max(cap, 1), 3, 3), dtype=torch.float32, device="cuda")
    cnt = torch.full((1,), -1, dtype=torch.int64, device="cuda")
    fn(box, iso, out, cap, cnt, stream)
    return out, cnt


def test_noalloc_matches_async_on_two_calls_second_smaller():
    ctx = _capi.Context(0)
    try:
        vol = W.make_volume(W.SMALL)
        ctx.set_volume(vol)
        stream = torch.cuda.Stream()
        # the second box is smaller than the first (568 and 7816 triangles)
        cases = [((0, 0, 0, 128, 128, 128), 300.0), ((10, 20, 5, 100, 90, 127), 800.0)]
        for box, iso in cases:
            n = len(ctx.marching_cubes(box, iso))
            assert n > 0
            for cap in (n, n // 3):
                ref_out, ref_cnt = _run(ctx.marching_cubes_async, box, iso, cap, stream.cuda_stream)
                got_out, got_cnt = _run(ctx.marching_cubes_noalloc, box, iso, cap, stream.cuda_stream)
                torch.cuda.synchronize()
                assert int(got_cnt.item()) == int(ref_cnt.item()) == n, (box, iso, cap)
                assert torch.equal(got_out[:cap], ref_out[:cap]), (box, iso, cap)
                print(f"box={box} iso={iso} cap={cap}: {n} triangles, noalloc == async")
        # count only (no triangle buffer)
        cnt = torch.full((1,), -1, dtype=torch.int64, device="cuda")
        ctx.marching_cubes_noalloc(cases[1][0], cases[1][1], None, 0, cnt, stream.cuda_stream)
        torch.cuda.synchronize()
        assert int(cnt.item()) == len(ctx.marching_cubes(*cases[1]))
    finally:
        ctx.close()


def test_noalloc_rejects_box_beyond_scratch():
    ctx = _capi.Context(0)
    try:
        dims_zyx = (1030, 512, 512)     # 1029 * 511 * 511 cubes > 2^28
        assert (dims_zyx[0] - 1) * (dims_zyx[1] - 1) * (dims_zyx[2] - 1) > LIMIT_CUBES
        dev = torch.zeros(dims_zyx, dtype=torch.int16, device="cuda")
        ctx.set_volume_device(dev, dims_zyx)
        del dev
        with pytest.raises(MkError) as e:
            ctx.marching_cubes_noalloc((0, 0, 0, 512, 512, 1030), 100.0, None, 0, None, 0)
        assert e.value.status == MK_ERR_ARG and "scratch limit" in str(e.value)
        # one z-slab less fits: an all-zero volume has no crossing at iso 100
        cnt = torch.full((1,), -1, dtype=torch.int64, device="cuda")
        ctx.marching_cubes_noalloc((0, 0, 0, 512, 512, 1028), 100.0, None, 0, cnt, 0)
        torch.cuda.synchronize()
        assert int(cnt.item()) == 0
    finally:
        ctx.close()
