"""Differentiable transfer function: forward parity against FP64, the adjoint
against FP64 central finite differences, and a short Adam fit."""
import numpy as np
import pytest

from monkeygl_amd import params as P
from monkeygl_amd import workloads as W
from oracle import oracle as O

SPEC = W.VolumeSpec("fit64", (64, 64, 64), (1.0, 1.0, 1.0), 6, (3, 9), (300, 1800), seed=21)
N_LUT = 64


def _views(size=96, angles=((0.0, 0.0), (35.0, 20.0))):
    out = []
    for a in angles:
        info = P.RenderInfo()
        info.set_vr_size(size, size)
        info.rotate(*a)
        out.append(P.render_params(SPEC.dims, SPEC.spacing, info, {0: P.LabelState(1.0, 2200.0, 0.0)},
                                   width=size, height=size))
    return out


def _lut(seed, lo=0.02, hi=0.3):
    rng = np.random.default_rng(seed)
    lut = np.empty((N_LUT, 4), np.float32)
    lut[:, :3] = rng.uniform(0.1, 1.0, (N_LUT, 3))
    lut[:, 3] = rng.uniform(lo, hi, N_LUT)
    return lut


def test_fit_oracle_constant_volume_closed_form():
    """Constant volume, flat LUT: zero gradient so k = 0.35, alpha a at every
    in-VOI sample of the fixed lattice: C = 0.35 c (1 - (1-a)^n), n counted on
    the float lattice."""
    dims = (16, 16, 16)
    vol = np.full(dims[::-1], 100, np.int16)
    info = P.RenderInfo()
    info.set_vr_size(8, 8)
    p = P.render_params(dims, (1.0, 1.0, 1.0), info, {0: P.LabelState(1.0, 400.0, 40.0)})
    lut = np.tile(np.array([0.5, 0.25, 1.0, 0.2]), (8, 1))
    img = O.fit_render(p, vol, lut, 1.5)
    # count in-VOI lattice points of the ray through pixel (4, 4): u = v = 0
    L1 = np.float32(1.0) / np.float32(16)
    acc, n = np.float32(0.0), 0
    while float(acc) < 1.732:
        fy = (float(acc) - 0.866) * 1.0
        py = fy * 1.0 + 0.5
        if 0 <= int(py * 16) <= 15 and int(0.5 * 16) <= 15:
            n += 1
        acc = np.float32(acc + L1)
    a = 0.2 * 1.5
    exp = 0.35 * np.array([0.5, 0.25, 1.0]) * (1 - (1 - a) ** n)
    assert np.allclose(img[4, 4], exp, rtol=1e-12, atol=1e-12), (img[4, 4], exp, n)


@pytest.mark.gpu
def test_tf_forward_matches_fp64(gpu_ctx):
    import torch

    vol = W.make_volume(SPEC)
    gpu_ctx.set_volume(vol)
    lut = _lut(1)
    dl = torch.tensor(lut, device="cuda")
    for p in _views():
        out = torch.empty((p.height, p.width, 3), dtype=torch.float32, device="cuda")
        gpu_ctx.tf_forward(p, dl, N_LUT, 1.5, out, stream=torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        ref = O.fit_render(p, vol, lut.astype(np.float64), 1.5)
        err = np.abs(out.cpu().numpy().astype(np.float64) - ref)
        assert err.max() < 2e-4 * max(1.0, ref.max()), err.max()


def _assert_fd(r, n_check=64):
    print(f"checked {r['big']} of {r['checked']} entries ({r['live']} live, |g| from {r['g_range'][0]:.2e} to "
          f"{r['g_range'][1]:.2e}), max rel err {r['max_rel_err']:.2e}, scale rel err {r['scale_rel_err']:.2e}")
    assert r["checked"] == n_check, r
    assert r["big"] >= n_check - 8, r
    assert r["max_rel_err"] <= 1e-3, r
    assert r["small_abs_err_ok"], r
    assert r["scale_rel_err"] <= 1e-3, r


@pytest.mark.gpu
def test_tf_gradient_matches_fp64_finite_differences(gpu_ctx):
    """dLoss/dLUT and dLoss/dscale from the adjoint vs FP64 central differences
    (h = 1e-4) on 64 random live LUT entries: relative error <= 1e-3."""
    from oracle.fit_check import fd_check

    vol = W.make_volume(SPEC)
    gpu_ctx.set_volume(vol)
    _assert_fd(fd_check(gpu_ctx, vol, _views(), _lut(2), _lut(3), 1.2, 1.5))


@pytest.mark.gpu
def test_tf_gradient_2048_entry_lut(gpu_ctx):
    """A 2048-entry LUT: its gradient table does not fit beside the LUT in the
    workgroup's LDS, so the adjoint adds to device memory directly; the same
    finite-difference bar holds, and the forward matches FP64."""
    import torch

    from oracle.fit_check import fd_check

    vol = W.make_volume(SPEC)
    gpu_ctx.set_volume(vol)
    n = 2048

    def lut(seed):
        r = np.random.default_rng(seed)
        a = np.empty((n, 4), np.float32)
        a[:, :3] = r.uniform(0.1, 1.0, (n, 3))
        a[:, 3] = r.uniform(0.02, 0.3, n)
        return a

    l0 = lut(12)
    views = _views()
    dl = torch.tensor(l0, device="cuda")
    out = torch.empty((views[1].height, views[1].width, 3), dtype=torch.float32, device="cuda")
    gpu_ctx.tf_forward(views[1], dl, n, 1.5, out, stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    ref = O.fit_render(views[1], vol, l0.astype(np.float64), 1.5)
    assert np.abs(out.cpu().numpy() - ref).max() < 2e-4 * max(1.0, ref.max())
    _assert_fd(fd_check(gpu_ctx, vol, views, lut(101), l0, 1.2, 1.5))


@pytest.mark.gpu
def test_tf_backward_from_kept_colours_equals_remarch(gpu_ctx):
    """The adjoint given the forward pass's float64 colours equals the adjoint
    that re-marches each ray for them (up to the order of the double atomic
    sums), and those colours round to the float image."""
    import torch

    vol = W.make_volume(SPEC)
    gpu_ctx.set_volume(vol)
    lut = torch.tensor(_lut(4), device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    for p in _views(angles=((0.0, 0.0), (35.0, 20.0), (-60.0, 75.0))):
        rgb = torch.empty((p.height, p.width, 3), dtype=torch.float32, device="cuda")
        rgb64 = torch.full((p.height, p.width, 3), float("nan"), dtype=torch.float64, device="cuda")
        gpu_ctx.tf_forward(p, lut, N_LUT, 1.4, rgb, stream=stream, dev_rgb64=rgb64)
        gen = torch.Generator(device="cuda").manual_seed(5)
        g = torch.randn((p.height, p.width, 3), generator=gen, device="cuda")
        out = []
        for kept in (rgb64, None):
            g_lut = torch.zeros(N_LUT * 4, dtype=torch.float64, device="cuda")
            g_s = torch.zeros(1, dtype=torch.float64, device="cuda")
            gpu_ctx.tf_backward(p, lut, N_LUT, 1.4, g, g_lut, g_s, stream=stream, dev_rgb64=kept)
            out.append((g_lut, g_s))
        torch.cuda.synchronize()
        assert torch.equal(rgb64.to(torch.float32), rgb)
        (a, sa), (b, sb) = out
        tol = 1e-10 * float(b.abs().max())
        assert float(b.abs().max()) > 0.0
        assert float((a - b).abs().max()) <= tol, float((a - b).abs().max())
        assert abs(float(sa) - float(sb)) <= 1e-10 * abs(float(sb)), (float(sa), float(sb))


@pytest.mark.gpu
def test_tf_ragged_image_forward_and_adjoint(gpu_ctx):
    """A 75 x 53 view (tiles cut at both edges, width and height unequal) with
    a pan and zoom: the forward against FP64, and the adjoint from kept
    colours equal to the re-marching adjoint, with every gradient finite."""
    import torch

    vol = W.make_volume(SPEC)
    gpu_ctx.set_volume(vol)
    info = P.RenderInfo()
    info.set_vr_size(75, 53)
    info.rotate(-20.0, 50.0)
    info.zoom(1.3)
    info.pan(4.0, -3.0)
    p = P.render_params(SPEC.dims, SPEC.spacing, info, {0: P.LabelState(1.0, 2200.0, 0.0)}, width=75, height=53)
    lut = _lut(6)
    dl = torch.tensor(lut, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    rgb = torch.empty((53, 75, 3), dtype=torch.float32, device="cuda")
    rgb64 = torch.empty((53, 75, 3), dtype=torch.float64, device="cuda")
    gpu_ctx.tf_forward(p, dl, N_LUT, 1.3, rgb, stream=stream, dev_rgb64=rgb64)
    g = torch.randn((53, 75, 3), generator=torch.Generator(device="cuda").manual_seed(9), device="cuda")
    grads = []
    for kept in (rgb64, None):
        g_lut = torch.zeros(N_LUT * 4, dtype=torch.float64, device="cuda")
        g_s = torch.zeros(1, dtype=torch.float64, device="cuda")
        gpu_ctx.tf_backward(p, dl, N_LUT, 1.3, g, g_lut, g_s, stream=stream, dev_rgb64=kept)
        grads.append((g_lut, g_s))
    torch.cuda.synchronize()
    ref = O.fit_render(p, vol, lut.astype(np.float64), 1.3)
    err = np.abs(rgb.cpu().numpy().astype(np.float64) - ref)
    assert ref.max() > 0.0
    assert err.max() < 2e-4 * max(1.0, ref.max()), err.max()
    (a, sa), (b, sb) = grads
    assert bool(torch.isfinite(a).all()) and bool(torch.isfinite(sa).all())
    assert float(b.abs().max()) > 0.0
    assert float((a - b).abs().max()) <= 1e-10 * float(b.abs().max())
    assert abs(float(sa) - float(sb)) <= 1e-10 * abs(float(sb))


@pytest.mark.gpu
def test_short_fit_reduces_loss(gpu_ctx):
    import torch

    from monkeygl_amd.tf_fit import fit, render_tf

    vol = W.make_volume(SPEC)
    gpu_ctx.set_volume(vol)
    views = _views()
    lut_gt = torch.tensor(_lut(2), device="cuda")
    with torch.no_grad():
        target = render_tf(gpu_ctx, views, lut_gt, torch.tensor(1.2))
    lut, s, losses = fit(gpu_ctx, views, target, torch.tensor(_lut(3), device="cuda"), 1.5, steps=60, lr=0.02)
    assert losses[-1] < 0.1 * losses[0], (losses[0], losses[-1])
