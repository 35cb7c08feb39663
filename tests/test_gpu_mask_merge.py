"""Incremental mask edits on the device (mk_merge_object_mask): the merged
mask read back equals the reference's merge rule applied to the same edits
(VolumeInfo::AddNewObjectMask / UpdateObjectMask, core/VolumeInfo.cpp:94-133:
add sets the label where the object is nonzero; update also clears the label
where it is zero), on volumes with partial bricks; renders after the merges are
bit-identical to renders after a full upload of that mask (same brick label
statistics); the origin slice (GetOriginData) is the volume slice with -2048
outside the mask; argument and state errors."""
import numpy as np
import pytest

from monkeygl_amd import params as P
from monkeygl_amd import workloads as W
from monkeygl_amd._capi import (MK_ERR_RANGE, MK_ERR_STATE, MK_MASK_ADD, MK_MASK_UPDATE, MK_RENDER_MIP,
                                MK_RENDER_VR, MkError)

pytestmark = pytest.mark.gpu


def _objects(shape, seed):
    rng = np.random.default_rng(seed)
    objs = []
    for k in range(3):
        o = np.zeros(shape, np.uint8)
        lo = [int(rng.integers(0, s // 2)) for s in shape]
        hi = [int(l + rng.integers(s // 4, s // 2 + 1)) for l, s in zip(lo, shape)]
        o[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = rng.integers(1, 255, dtype=np.uint8)
        o[rng.random(shape) < 0.02] = 1   # scattered voxels in every brick
        objs.append(o)
    return objs


def _add(mask, obj, label):
    mask[obj > 0] = label


def _update(mask, obj, label):
    mask[(obj == 0) & (mask == label)] = 0
    mask[obj > 0] = label


@pytest.mark.parametrize("dims", [(37, 29, 19), (64, 48, 40)])
def test_merges_match_the_rule_and_a_full_upload(gpu_ctx, dims):
    from monkeygl_amd._capi import Context

    spec = W.VolumeSpec("merge", dims, (0.9, 0.8, 1.1), 5, (2, 6), (300, 1500), n_vessels=2, seed=3)
    vol = W.make_volume(spec)
    o1, o2, o3 = _objects(vol.shape, seed=sum(dims))
    gpu_ctx.set_volume(vol)
    exp = np.zeros(vol.shape, np.uint8)
    for obj, lab, mode in ((o1, 1, MK_MASK_ADD), (o2, 2, MK_MASK_ADD), (o3, 1, MK_MASK_UPDATE),
                           (o1, 64, MK_MASK_ADD), (o2, 64, MK_MASK_UPDATE)):
        gpu_ctx.merge_mask(obj, lab, mode)
        (_add if mode == MK_MASK_ADD else _update)(exp, obj, lab)
        assert np.array_equal(gpu_ctx.get_mask(), exp), (lab, mode)
    # renders through the merged mask equal renders through an uploaded one
    ref = Context(0)
    try:
        ref.set_volume(vol)
        ref.set_mask(exp)
        for c in (gpu_ctx, ref):
            for lab in (0, 1, 2, 64):
                c.set_lut(lab, P.transfer_function(W.tf_points(lab % 4)))
        info = P.RenderInfo()
        info.set_vr_size(96, 80)
        info.rotate(25.0, 15.0)
        labels = {l: P.LabelState(1.0, 500.0, 80.0) for l in (0, 1, 2, 64)}
        for rt in (MK_RENDER_VR, MK_RENDER_MIP):
            p = P.render_params(dims, spec.spacing, info, labels, rt)
            assert np.array_equal(gpu_ctx.render(p), ref.render(p)), rt
    finally:
        ref.close()


def test_origin_slice(gpu_ctx):
    import torch

    spec = W.VolumeSpec("origin", (40, 30, 12), (1.0, 1.0, 1.0), 3, (2, 5), (300, 1500), n_vessels=1, seed=8)
    vol = W.make_volume(spec)
    gpu_ctx.set_volume(vol)
    out = torch.empty(40 * 30, dtype=torch.int16, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    for z, zc in ((5, 5), (-3, 0), (40, 11)):
        gpu_ctx.origin_slice_async(z, out, stream=s)
        assert np.array_equal(out.cpu().numpy().reshape(30, 40), vol[zc]), z
    obj = np.zeros(vol.shape, np.uint8)
    obj[:, 5:20, 10:33] = 1
    gpu_ctx.merge_mask(obj, 3, MK_MASK_ADD)
    gpu_ctx.origin_slice_async(7, out, stream=s)
    assert np.array_equal(out.cpu().numpy().reshape(30, 40), np.where(obj[7] > 0, vol[7], -2048))


def test_errors(gpu_ctx):
    vol = W.make_volume(W.VolumeSpec("err", (16, 16, 8), (1.0, 1.0, 1.0), 2, (2, 4), (300, 900), n_vessels=1,
                                     seed=2))
    gpu_ctx.set_volume(vol)   # drops any mask
    obj = np.ones(vol.shape, np.uint8)
    with pytest.raises(MkError) as e:
        gpu_ctx.get_mask()
    assert e.value.status == MK_ERR_STATE
    with pytest.raises(MkError) as e:
        gpu_ctx.merge_mask(obj, 1, MK_MASK_UPDATE)
    assert e.value.status == MK_ERR_STATE
    with pytest.raises(MkError) as e:
        gpu_ctx.merge_mask(obj, 65, MK_MASK_ADD)
    assert e.value.status == MK_ERR_RANGE
    with pytest.raises(MkError):
        gpu_ctx.merge_mask(obj[:4], 1, MK_MASK_ADD)
