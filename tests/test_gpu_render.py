"""GPU parity of VR / MIP / Surface against the FP64 oracle, through the C ABI.

Bar: every RGB channel within 1/255 of the FP64 oracle on at least 99.9% of
pixels (the tolerance absorbs float-vs-double rounding and the truncating
x255 conversion of core/kernel.cu:40-47).
"""
import numpy as np
import pytest

from monkeygl_amd import params as P
from monkeygl_amd import workloads as W
from monkeygl_amd._capi import MK_RENDER_MIP, MK_RENDER_SURFACE, MK_RENDER_VR, MkError
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1          # 1/255 per channel
MIN_FRAC = 0.999

VIEWS = [(0, 0), (30, 0), (0, 45), (20, 35), (-75, 110)]


def frac_within(a, b, tol=TOL):
    d = np.abs(a.astype(np.int32) - b.astype(np.int32)).max(axis=-1)
    return float((d <= tol).mean()), int(d.max())


@pytest.fixture(scope="module")
def small():
    vol = W.make_volume(W.SMALL)
    lut = P.transfer_function(W.tf_points(0))
    return vol, lut


def _params(rt, ang, size=256, zoom=None, pan=None, labels=None, invert_z=False, **kw):
    info = P.RenderInfo()
    info.set_vr_size(size, size)
    info.rotate(*ang)
    if zoom:
        info.zoom(zoom)
    if pan:
        info.pan(*pan)
    labels = labels or {0: P.LabelState(1.0, 2000.0, 0.0)}
    return P.render_params((128, 128, 128), (1.0, 1.0, 1.0), info, labels, render_type=rt, invert_z=invert_z, **kw)


@pytest.mark.parametrize("rt", [MK_RENDER_VR, MK_RENDER_MIP, MK_RENDER_SURFACE])
@pytest.mark.parametrize("ang", VIEWS)
def test_render_small_matches_oracle(gpu_ctx, small, rt, ang):
    vol, lut = small
    gpu_ctx.set_volume(vol)
    gpu_ctx.set_lut(0, lut)
    p = _params(rt, ang)
    got = gpu_ctx.render(p)
    ref = O.render(p, vol, {0: lut}, precision="f64")
    frac, dmax = frac_within(got, ref)
    print(f"rt={rt} ang={ang} frac={frac:.6f} max={dmax}")
    assert frac >= MIN_FRAC, (frac, dmax)


@pytest.mark.parametrize("rt", [MK_RENDER_VR, MK_RENDER_MIP])
def test_zoom_pan_invert_z(gpu_ctx, small, rt):
    vol, lut = small
    gpu_ctx.set_volume(vol)
    gpu_ctx.set_lut(0, lut)
    p = _params(rt, (15, 25), size=192, zoom=1.7, pan=(13.0, -21.0), invert_z=True)
    got = gpu_ctx.render(p)
    ref = O.render(p, vol, {0: lut})
    frac, dmax = frac_within(got, ref)
    assert frac >= MIN_FRAC, (frac, dmax)


def test_nonsquare_and_odd_sizes(gpu_ctx, small):
    vol, lut = small
    gpu_ctx.set_volume(vol)
    gpu_ctx.set_lut(0, lut)
    for w, h in [(1, 1), (17, 3), (130, 77), (300, 128)]:
        p = _params(MK_RENDER_VR, (10, 20), width=w, height=h)
        got = gpu_ctx.render(p)
        ref = O.render(p, vol, {0: lut})
        frac, dmax = frac_within(got, ref)
        assert got.shape == (h, w, 3)
        assert frac >= MIN_FRAC or dmax <= 1, (w, h, frac, dmax)


def test_build_mode_step_exit_4096_lut(gpu_ctx, small):
    """Config-0 build parameters: half-voxel step, exit at 0.99, 4096-entry LUT."""
    vol, _ = small
    pts = W.tf_points(0)
    lut = P.transfer_function({int(round(k * 4095 / 99)): v for k, v in pts.items()}, n=4096)
    gpu_ctx.set_volume(vol)
    gpu_ctx.set_lut(0, lut)
    for ang in [(0, 0), (33, 12)]:
        p = _params(MK_RENDER_VR, ang, step_scale=0.5, exit_alpha=0.99)
        got = gpu_ctx.render(p)
        ref = O.render(p, vol, {0: lut})
        frac, dmax = frac_within(got, ref)
        assert frac >= MIN_FRAC, (ang, frac, dmax)


@pytest.mark.parametrize("rt", [MK_RENDER_VR, MK_RENDER_MIP])
@pytest.mark.parametrize("step", [0.05, 0.3, 3.0])
def test_step_scale_sweep(gpu_ctx, small, rt, step):
    """Fine and coarse sampling steps: the empty-cell walk jumps over the float
    accumulation acc += L1 by per-binade closed forms whose size and caps
    depend on L1 (core/kernel.cu:133-264 and 266-364 add L1 once per lattice
    point), so frames at 1/20, 3/10 and 3 times the voxel step against FP64."""
    vol, lut = small
    gpu_ctx.set_volume(vol)
    gpu_ctx.set_lut(0, lut)
    p = _params(rt, (20, 35), size=128, step_scale=step)
    got = gpu_ctx.render(p)
    ref = O.render(p, vol, {0: lut}, precision="f64")
    frac, dmax = frac_within(got, ref)
    print(f"rt={rt} step={step} frac={frac:.6f} max={dmax}")
    assert frac >= MIN_FRAC, (step, frac, dmax)


def test_long_volume_composes_address_terms(gpu_ctx):
    """A 24x20x8200 volume (40 x 40 x 0.1 mm voxels, about cubic in space):
    its per-axis address-term tables would take 33 KiB, above the 32 KiB the
    kernels stage in LDS, so VR (with and without a mask) and MIP compose the
    terms from bit fields instead. Both paths against the FP64 oracle."""
    dims, spacing = (24, 20, 8200), (40.0, 40.0, 0.1)
    nx, ny, nz = dims
    zz, yy, xx = np.meshgrid(np.arange(nz) * spacing[2], np.arange(ny) * spacing[1], np.arange(nx) * spacing[0],
                             indexing="ij")
    rng = np.random.default_rng(11)
    vol = np.full((nz, ny, nx), -1000.0, np.float32)
    ext = np.array([nx * spacing[0], ny * spacing[1], nz * spacing[2]])
    for _ in range(5):
        c = rng.uniform(0.2, 0.8, 3) * ext
        r = rng.uniform(0.12, 0.25) * ext.min()
        d2 = ((xx - c[0]) ** 2 + (yy - c[1]) ** 2 + (zz - c[2]) ** 2) / r ** 2
        vol += rng.uniform(800, 2000) * np.exp(-d2).astype(np.float32)
    vol = np.clip(vol, -1024, 3000).astype(np.int16)
    mask = np.zeros(vol.shape, np.uint8)
    mask[: nz // 2, :, : nx // 2] = 1
    mask[nz // 3:, ny // 2:, :] = 2
    lut0 = P.transfer_function(W.tf_points(0))
    luts = {0: lut0, 1: P.transfer_function(W.tf_points(31)), 2: P.transfer_function(W.tf_points(32))}
    labels = {0: P.LabelState(1.0, 2000.0, 0.0), 1: P.LabelState(0.5, 1600.0, 100.0), 2: P.LabelState(0.8, 1800.0, -50.0)}
    gpu_ctx.set_volume(vol)
    for l, a in luts.items():
        gpu_ctx.set_lut(l, a)

    def check(rt, ang, lab, use_mask):
        info = P.RenderInfo()
        info.set_vr_size(160, 160)
        info.rotate(*ang)
        p = P.render_params(dims, spacing, info, lab, render_type=rt)
        got = gpu_ctx.render(p)
        ref = O.render(p, vol, {l: luts[l] for l in lab}, mask_zyx=mask if use_mask else None)
        frac, dmax = frac_within(got, ref)
        assert frac >= MIN_FRAC, (rt, ang, use_mask, frac, dmax)
        assert (ref.sum(-1) > 0).mean() > 0.2, "the view should show the volume"

    for ang in [(0, 0), (30, 40)]:
        check(MK_RENDER_VR, ang, {0: labels[0]}, False)
        check(MK_RENDER_MIP, ang, {0: labels[0]}, False)
    gpu_ctx.set_mask(mask)
    try:
        for ang in [(0, 0), (30, 40)]:
            check(MK_RENDER_VR, ang, labels, True)
            check(MK_RENDER_MIP, ang, labels, True)
    finally:
        gpu_ctx.clear_mask()


def test_all_65_labels_vr_and_mip(gpu_ctx, small):
    """Labels 0..64 all present and visible. VR keeps one accumulator per
    label and thread in LDS, so its tiles shrink to 8x8; MIP keeps none and
    must render every pixel with its own 16x16 tiles."""
    vol, lut0 = small
    z, y, x = np.meshgrid(np.arange(128), np.arange(128), np.arange(128), indexing="ij")
    mask = (1 + (x // 9 + 3 * (y // 11) + 7 * (z // 13)) % 64).astype(np.uint8)
    mask[:, :, :20] = 0
    gpu_ctx.set_volume(vol)
    gpu_ctx.set_mask(mask)
    luts, labels = {}, {}
    for l in range(65):
        luts[l] = lut0 if l == 0 else P.transfer_function(W.tf_points(200 + l))
        labels[l] = P.LabelState(0.3 + 0.7 * ((l * 37) % 11) / 10.0, 1400.0 + 10.0 * l, -150.0 + 5.0 * l)
    for l, a in luts.items():
        gpu_ctx.set_lut(l, a)
    try:
        for rt in (MK_RENDER_VR, MK_RENDER_MIP):
            for ang in [(0, 0), (35, 20)]:
                p = _params(rt, ang, size=160, labels=labels)
                got = gpu_ctx.render(p)
                ref = O.render(p, vol, luts, mask_zyx=mask)
                frac, dmax = frac_within(got, ref)
                assert frac >= MIN_FRAC, (rt, ang, frac, dmax)
    finally:
        gpu_ctx.clear_mask()


def test_multilabel_mask(gpu_ctx, small):
    """Four ellipsoid labels, each with its own LUT, window and alpha cap."""
    vol, lut0 = small
    mask = W.ellipsoid_mask((128, 128, 128), 4, seed=3)
    gpu_ctx.set_volume(vol)
    gpu_ctx.set_mask(mask)
    luts = {0: lut0}
    labels = {0: P.LabelState(0.4, 2000.0, 0.0)}
    for l in range(1, 5):
        luts[l] = P.transfer_function(W.tf_points(10 + l))
        labels[l] = P.LabelState([1.0, 0.8, 0.3, 0.6][l - 1], 1500.0 + 100 * l, -100.0 * l)
    for l, a in luts.items():
        gpu_ctx.set_lut(l, a)
    for rt in (MK_RENDER_VR, MK_RENDER_MIP):
        for ang in [(0, 0), (40, 10)]:
            p = _params(rt, ang, labels=labels)
            got = gpu_ctx.render(p)
            ref = O.render(p, vol, luts, mask_zyx=mask)
            frac, dmax = frac_within(got, ref)
            assert frac >= MIN_FRAC, (rt, ang, frac, dmax)
    gpu_ctx.clear_mask()


def test_empty_tf_gives_background(gpu_ctx, small):
    vol, _ = small
    gpu_ctx.set_volume(vol)
    gpu_ctx.set_lut(0, np.zeros((100, 4), np.float32))
    p = _params(MK_RENDER_VR, (0, 0), background=(0.2, 0.4, 0.6, 1.0))
    got = gpu_ctx.render(p)
    assert (got[..., 0] == int(0.2 * 255)).all() and (got[..., 2] == int(np.float32(0.6) * 255)).all()


def test_errors_are_reported(gpu_ctx, small):
    vol, lut = small
    gpu_ctx.set_volume(vol)
    p = _params(MK_RENDER_VR, (0, 0))
    p.render_type = 7
    with pytest.raises(MkError):
        gpu_ctx.render(p)
    bad = np.full((128, 128, 128), 65, np.uint8)
    with pytest.raises(MkError):
        gpu_ctx.set_mask(bad)
    with pytest.raises(MkError):
        gpu_ctx.set_lut(65, lut)


@pytest.mark.parametrize("rt", [MK_RENDER_VR, MK_RENDER_MIP, MK_RENDER_SURFACE])
def test_footprint_matches_oracle(gpu_ctx, small, rt):
    """The measurement variant marks the voxels the reference rule reads; the
    oracle marks them independently. The sets agree up to samples whose
    footprint flips at a texel boundary between float and double positions."""
    import torch

    vol, lut = small
    gpu_ctx.set_volume(vol)
    gpu_ctx.set_lut(0, lut)
    p = _params(rt, (20, 35))
    out = torch.empty(p.width * p.height * 3, dtype=torch.uint8, device="cuda")
    tv = torch.zeros(vol.size, dtype=torch.uint8, device="cuda")
    gpu_ctx.render_footprint(p, out, tv, None, stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    g = tv.cpu().numpy().reshape(vol.shape).astype(bool)
    img, t = O.render(p, vol, {0: lut}, precision="f32", touched=True)
    r = t.astype(bool)
    assert r.sum() > 0
    assert np.logical_xor(g, r).sum() <= 0.002 * r.sum(), (g.sum(), r.sum(), np.logical_xor(g, r).sum())
    # the measurement variant renders the same frame
    frac, dmax = frac_within(out.cpu().numpy().reshape(p.height, p.width, 3), img)
    assert frac >= MIN_FRAC, (frac, dmax)


def _capped_scene(gpu_ctx, vol, n_labels=3, seed=5):
    """Label 0 with alpha 0.05 everywhere and a cap of 0.3, plus ellipsoids of
    labels with ordinary LUTs and their own caps."""
    mask = W.ellipsoid_mask((128, 128, 128), n_labels, seed=seed)
    gpu_ctx.set_volume(vol)
    gpu_ctx.set_mask(mask)
    lut0 = np.zeros((100, 4), np.float32)
    lut0[:, 0] = np.linspace(0.2, 1.0, 100)
    lut0[:, 1] = 0.5
    lut0[:, 2] = np.linspace(1.0, 0.1, 100)
    lut0[:, 3] = 0.05
    luts = {0: lut0}
    labels = {0: P.LabelState(0.3, 2000.0, 0.0)}
    for l in range(1, n_labels + 1):
        luts[l] = P.transfer_function(W.tf_points(20 + l))
        labels[l] = P.LabelState([1.0, 0.5, 0.8, 0.9][(l - 1) % 4], 1600.0, -50.0 * l)
    for l, a in luts.items():
        gpu_ctx.set_lut(l, a)
    return mask, luts, labels


@pytest.mark.parametrize("masked", [True, False])
def test_two_pass_vr_is_exact_for_every_budget(gpu_ctx, small, masked):
    """Rays still marching after the first pass's trip budget finish in the
    second pass, eight consecutive samples at a time; the frame must be the
    same, bit for bit, for every budget (0 = one pass, -1 = chosen per
    frame). Budgets 1 and 3 put
    nearly all of the work in the second pass, with backtracking, step
    changes, VOI exits, label caps and early exit inside its batches."""
    vol, lut = small
    if masked:
        _, _, labels = _capped_scene(gpu_ctx, vol)
    else:
        gpu_ctx.set_volume(vol)
        gpu_ctx.set_lut(0, lut)
        labels = None
    try:
        for ang in [(0, 0), (40, 10), (-75, 110)]:
            kw = dict(labels=labels) if labels else {}
            p = _params(MK_RENDER_VR, ang, **kw)
            if ang == (40, 10):
                p.voi.left, p.voi.right = 20, 100   # rays leave and re-enter the VOI
            gpu_ctx.set_vr_trip_budget(0)
            one = gpu_ctx.render(p)
            assert (one != 0).any()
            for budget in (1, 3, 16, 128, -1):
                gpu_ctx.set_vr_trip_budget(budget)
                two = gpu_ctx.render(p)
                bad = int((one != two).any(axis=-1).sum())
                assert bad == 0, (masked, ang, budget, bad)
    finally:
        gpu_ctx.set_vr_trip_budget(-1)
        if masked:
            gpu_ctx.clear_mask()


def test_two_pass_vr_on_two_streams(gpu_ctx, small):
    """VR renders of one context share its second-pass queue; frames enqueued
    alternately on two streams (the queue growing between them) must each
    equal the frame rendered alone in one pass."""
    import torch

    vol, lut = small
    gpu_ctx.set_volume(vol)
    gpu_ctx.set_lut(0, lut)
    ps = [_params(MK_RENDER_VR, ang, size=sz) for ang, sz in
          [((0, 0), 256), ((40, 10), 384), ((-75, 110), 256), ((20, 35), 512), ((10, 20), 320)]]
    try:
        gpu_ctx.set_vr_trip_budget(0)
        want = [gpu_ctx.render(p) for p in ps]
        gpu_ctx.set_vr_trip_budget(3)
        streams = [torch.cuda.Stream(), torch.cuda.Stream()]
        outs = [torch.empty(p.width * p.height * 3, dtype=torch.uint8, device="cuda") for p in ps]
        for k, (p, o) in enumerate(zip(ps, outs)):
            gpu_ctx.render_async(p, o, stream=streams[k % 2].cuda_stream)
        torch.cuda.synchronize()
        for k, (p, o, w) in enumerate(zip(ps, outs, want)):
            got = o.cpu().numpy().reshape(p.height, p.width, 3)
            assert np.array_equal(got, w), (k, int((got != w).any(axis=-1).sum()))
    finally:
        gpu_ctx.set_vr_trip_budget(-1)


def test_two_pass_vr_queue_grows_and_is_reused(gpu_ctx, small):
    """The queue has a slot per pixel and grows with the frame: a small frame,
    a larger one, then the small one again, all equal to one pass."""
    vol, lut = small
    gpu_ctx.set_volume(vol)
    gpu_ctx.set_lut(0, lut)
    try:
        for sz in (64, 300, 1024, 64, 300):
            p = _params(MK_RENDER_VR, (25, 15), size=sz)
            gpu_ctx.set_vr_trip_budget(0)
            one = gpu_ctx.render(p)
            for budget in (1, -1):
                gpu_ctx.set_vr_trip_budget(budget)
                two = gpu_ctx.render(p)
                assert np.array_equal(one, two), (sz, budget)
    finally:
        gpu_ctx.set_vr_trip_budget(-1)


def test_vr_trip_budget_rejects_values_below_auto(gpu_ctx):
    with pytest.raises(MkError):
        gpu_ctx.set_vr_trip_budget(-2)
    gpu_ctx.set_vr_trip_budget(-1)


def test_skipping_is_exact_with_capped_label(gpu_ctx, small):
    """Empty-cell skipping and the skip tables change nothing: the frame
    equals, bit for bit, the measurement variant's, which samples every
    lattice point. Label 0's LUT has alpha 0.05 everywhere and a cap of 0.3,
    so rays reach the cap early and then march at L1/8 through label-0
    tissue, between ellipsoids of labels with their own caps."""
    import torch

    vol, _ = small
    mask, luts, labels = _capped_scene(gpu_ctx, vol)
    s = torch.cuda.current_stream().cuda_stream
    for ang in [(0, 0), (40, 10), (-75, 110)]:
        p = _params(MK_RENDER_VR, ang, labels=labels)
        got = gpu_ctx.render(p)
        out = torch.empty(p.width * p.height * 3, dtype=torch.uint8, device="cuda")
        tv = torch.zeros(vol.size, dtype=torch.uint8, device="cuda")
        gpu_ctx.render_footprint(p, out, tv, None, stream=s)
        torch.cuda.synchronize()
        full = out.cpu().numpy().reshape(p.height, p.width, 3)
        assert (got != 0).any()
        assert np.array_equal(got, full), (ang, int((got != full).any(axis=-1).sum()))
        if ang == (40, 10):
            ref = O.render(p, vol, luts, mask_zyx=mask)
            frac, dmax = frac_within(got, ref)
            assert frac >= MIN_FRAC, (frac, dmax)
    gpu_ctx.clear_mask()


@pytest.mark.parametrize("rt", [MK_RENDER_VR, MK_RENDER_MIP, MK_RENDER_SURFACE])
def test_output_at_any_byte_offset(gpu_ctx, small, rt):
    """The RGB output may start at any byte address: the same frame rendered
    to an aligned buffer and to one a byte off alignment must be equal, bit
    for bit, and neither render may write outside its width * height * 3
    bytes (the sizes leave partial 8x8 pixel blocks at the right and bottom
    edges)."""
    import torch

    vol, lut = small
    gpu_ctx.set_volume(vol)
    gpu_ctx.set_lut(0, lut)
    s = torch.cuda.current_stream().cuda_stream
    for w, h in [(260, 132), (256, 256), (516, 300)]:
        p = _params(rt, (20, 35), width=w, height=h)
        n = w * h * 3
        frames = []
        for off in (0, 1):
            buf = torch.full((n + 16,), 0xA5, dtype=torch.uint8, device="cuda")
            gpu_ctx.render_async(p, buf[off:off + n], stream=s)
            torch.cuda.synchronize()
            host = buf.cpu().numpy()
            assert (host[:off] == 0xA5).all() and (host[off + n:] == 0xA5).all(), (w, h, off)
            frames.append(host[off:off + n].reshape(h, w, 3))
        assert np.array_equal(frames[0], frames[1]), (w, h, int((frames[0] != frames[1]).any(-1).sum()))
        ref = O.render(p, vol, {0: lut})
        frac, dmax = frac_within(frames[0], ref)
        assert frac >= MIN_FRAC, (w, h, frac, dmax)


def test_volume_dims_above_the_limit_are_refused(gpu_ctx, small):
    """A dimension above MK_MAX_DIM is an argument error before anything is
    allocated, and the context keeps rendering its volume."""
    from monkeygl_amd._capi import MK_MAX_DIM

    vol, lut = small
    gpu_ctx.set_volume(vol)
    gpu_ctx.set_lut(0, lut)
    p = _params(MK_RENDER_VR, (10, 20), size=64)
    before = gpu_ctx.render(p)
    with pytest.raises(MkError):
        gpu_ctx.set_volume(np.zeros((MK_MAX_DIM + 1, 1, 1), np.int16))
    assert np.array_equal(gpu_ctx.render(p), before)
