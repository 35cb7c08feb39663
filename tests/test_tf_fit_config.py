"""Transfer-function fitting at its stated configuration: the 256^3 16-blob
volume, a 256-entry RGBA LUT plus an opacity scale, 8 views at 512^2.

* forward: all 8 whole views of the fitting rule against the FP64 oracle,
  every pixel within 2e-4 of the image maximum;
* adjoint: against FP64 central differences (h = 1e-4) on 64 random live LUT
  entries and the scale, relative error <= 1e-3. The loss is the mean squared
  error over a fixed band of 8 rows of every view (rows 252..259, through the
  middle of the volume), so that the 130 FP64 loss evaluations stay within
  seconds; the GPU still renders and differentiates whole views;
* 24 Adam steps from the start LUT toward the ground truth's images lower
  the loss.
"""
import numpy as np
import pytest

from monkeygl_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

BAND = (252, 260)


@pytest.fixture(scope="module")
def fit_scene(gpu_ctx):
    vol = W.make_volume(W.TF_FIT)
    gpu_ctx.set_volume(vol)
    return vol, W.fit_views(512), W.fit_lut(1), W.fit_lut(2)


def test_fit_config_forward_all_views_match_fp64(gpu_ctx, fit_scene):
    from monkeygl_amd.tf_fit import render_tf

    vol, views, _, lut0 = fit_scene
    gpu_ctx.set_volume(vol)
    with torch.no_grad():
        img = render_tf(gpu_ctx, views, torch.tensor(lut0, device="cuda"),
                        torch.tensor(W.FIT_SCALE_START)).cpu().numpy()
    assert img.shape == (8, 512, 512, 3)
    for k, p in enumerate(views):
        ref = O.fit_render(p, vol, lut0.astype(np.float64), W.FIT_SCALE_START)
        assert ref.max() > 0.05, k
        err = np.abs(img[k].astype(np.float64) - ref).max() / max(1.0, ref.max())
        print(f"view {k}: max relative error {err:.2e}", flush=True)
        assert err < 2e-4, (k, err)


def test_fit_config_adjoint_matches_fp64_finite_differences(gpu_ctx, fit_scene):
    from oracle.fit_check import fd_check

    vol, views, lut_gt, lut0 = fit_scene
    gpu_ctx.set_volume(vol)
    r = fd_check(gpu_ctx, vol, views, lut_gt, lut0, W.FIT_SCALE_TRUE, W.FIT_SCALE_START, rows=BAND)
    print(f"checked {r['big']} of {r['checked']} entries ({r['live']} live), max rel err {r['max_rel_err']:.2e}, "
          f"scale rel err {r['scale_rel_err']:.2e}")
    assert r["checked"] == 64 and r["big"] >= 56, r
    assert r["max_rel_err"] <= 1e-3, r
    assert r["small_abs_err_ok"], r
    assert r["scale_rel_err"] <= 1e-3, r


def test_fit_config_adam_lowers_the_loss(gpu_ctx, fit_scene):
    from monkeygl_amd.tf_fit import fit, render_tf

    vol, views, lut_gt, lut0 = fit_scene
    gpu_ctx.set_volume(vol)
    with torch.no_grad():
        target = render_tf(gpu_ctx, views, torch.tensor(lut_gt, device="cuda"), torch.tensor(W.FIT_SCALE_TRUE))
    _, s, losses = fit(gpu_ctx, views, target, torch.tensor(lut0, device="cuda"), W.FIT_SCALE_START, steps=24,
                       lr=0.01)
    print(f"24 Adam steps: loss {losses[0]:.4g} -> {losses[-1]:.4g}, scale {s:.4f}")
    assert all(np.isfinite(losses))
    assert losses[-1] < 0.5 * losses[0], (losses[0], losses[-1])
