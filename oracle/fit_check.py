"""TEST INFRASTRUCTURE ONLY: the transfer-function adjoint checked against
FP64 central finite differences of the oracle's fitting rule (h = 1e-4).

Used by tests/ and by bench.py's fitting block, as the checker."""
from __future__ import annotations

import numpy as np

from . import oracle as O


def fd_check(ctx, vol, views, lut_gt, lut0, s_gt: float, s0: float, rows=None, n_check: int = 64, seed: int = 7,
             h: float = 1e-4) -> dict:
    """Gradient of the mean squared error to the ground truth's FP64 images,
    from the GPU adjoint (mk_tf_backward through render_tf) and from FP64
    central differences on n_check random LUT entries the loss depends on,
    plus the opacity scale. rows = (y0, y1) restricts the loss to that band
    of rows of every view (the GPU still renders whole views; rays outside the
    band get a zero image gradient). Returns the comparison; the caller
    asserts on it."""
    import torch

    from monkeygl_amd.tf_fit import render_tf

    y0, y1 = rows if rows is not None else (0, views[0].height)

    def fp64_images(l, s):
        if rows is None:
            return np.stack([O.fit_render(p, vol, l, s) for p in views])
        return np.stack([O.fit_render_rows(p, vol, l, s, y0, y1) for p in views])

    target = fp64_images(lut_gt.astype(np.float64), s_gt)
    lut = torch.tensor(lut0, device="cuda", requires_grad=True)
    scale = torch.tensor(s0, dtype=torch.float32, requires_grad=True)
    img = render_tf(ctx, views, lut, scale)
    loss = torch.mean((img[:, y0:y1] - torch.tensor(target, dtype=torch.float32, device="cuda")) ** 2)
    loss.backward()
    g_an = lut.grad.detach().cpu().numpy().astype(np.float64).reshape(-1)
    gs_an = float(scale.grad)

    def fp64_loss(l, s):
        return np.mean((fp64_images(l, s) - target) ** 2)

    base = lut0.astype(np.float64)
    # entries the loss depends on (most of the LUT is read by no sample and
    # has an exactly zero gradient)
    live = np.flatnonzero(np.abs(g_an) > 1e-5 * np.abs(g_an).max())
    rng = np.random.default_rng(seed)
    idx = rng.choice(live, size=min(n_check, live.size), replace=False)
    g_fd = np.zeros(idx.size)
    for j, flat in enumerate(idx):
        lp, lm = base.copy(), base.copy()
        lp.flat[flat] += h
        lm.flat[flat] -= h
        g_fd[j] = (fp64_loss(lp, s0) - fp64_loss(lm, s0)) / (2 * h)
    gs_fd = (fp64_loss(base, s0 + h) - fp64_loss(base, s0 - h)) / (2 * h)
    an = g_an[idx]
    scale_ref = np.abs(g_fd).max()
    # relative error wherever the difference quotient stands clear of its own
    # rounding noise (loss ~1e-2, h = 1e-4: noise ~1e-13)
    big = np.abs(g_fd) > 1e-9 * max(scale_ref, 1e-12) + 1e-11
    rel = np.abs(an - g_fd)[big] / np.abs(g_fd)[big]
    return {
        "live": int(live.size), "checked": int(idx.size), "big": int(big.sum()),
        "max_rel_err": float(rel.max()) if rel.size else float("nan"),
        "small_abs_err_ok": bool(np.all(np.abs(an - g_fd)[~big] <= 1e-9 * scale_ref + 1e-11)),
        "scale_rel_err": float(abs(gs_an - gs_fd) / abs(gs_fd)),
        "g_range": (float(np.abs(g_fd)[big].min()) if big.any() else 0.0, float(scale_ref)),
    }
