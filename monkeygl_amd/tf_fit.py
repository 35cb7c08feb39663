"""Transfer-function fitting through the hand-written adjoint.

``render_tf(ctx, views, lut, scale)`` renders the fitting rule of
``mk_tf_forward`` (see include/monkeygl_hip.h) for a list of views and is
differentiable with respect to the LUT (n x 4, float32, on the GPU) and the
opacity scale (a 0-d tensor). When a gradient is wanted the forward pass also
keeps the colours in float64 (24 bytes per pixel), and the backward pass
calls ``mk_tf_backward`` with them, which re-marches every ray once;
gradients accumulate in float64 on the device.
"""
from __future__ import annotations

import torch

from ._capi import Context, RenderParams


class _TFRender(torch.autograd.Function):
    @staticmethod
    def forward(ctx_, lut: torch.Tensor, scale: torch.Tensor, mk: Context, views: list) -> torch.Tensor:
        if lut.device.type != "cuda" or lut.dtype != torch.float32 or lut.dim() != 2 or lut.shape[1] != 4:
            raise ValueError("lut must be a (n, 4) float32 CUDA tensor")
        lut_c = lut.contiguous()
        s = float(scale.detach().cpu())
        h, w = views[0].height, views[0].width
        if any(p.height != h or p.width != w for p in views):
            raise ValueError("all views must have the same image size")
        out = torch.empty((len(views), h, w, 3), dtype=torch.float32, device=lut.device)
        keep = ctx_.needs_input_grad[0] or ctx_.needs_input_grad[1]
        rgb64 = torch.empty((len(views), h, w, 3), dtype=torch.float64, device=lut.device) if keep else None
        stream = torch.cuda.current_stream(lut.device).cuda_stream
        for i, p in enumerate(views):
            mk.tf_forward(p, lut_c, lut_c.shape[0], s, out[i], stream=stream,
                          dev_rgb64=None if rgb64 is None else rgb64[i])
        ctx_.save_for_backward(lut_c)
        ctx_.mk, ctx_.views, ctx_.s, ctx_.scale_like, ctx_.rgb64 = mk, views, s, scale, rgb64
        return out

    @staticmethod
    def backward(ctx_, grad: torch.Tensor):
        (lut,) = ctx_.saved_tensors
        n = lut.shape[0]
        g = grad.contiguous().to(torch.float32)
        g_lut = torch.zeros(n * 4, dtype=torch.float64, device=lut.device)
        g_s = torch.zeros(1, dtype=torch.float64, device=lut.device)
        stream = torch.cuda.current_stream(lut.device).cuda_stream
        for i, p in enumerate(ctx_.views):
            ctx_.mk.tf_backward(p, lut, n, ctx_.s, g[i], g_lut, g_s, stream=stream, dev_rgb64=ctx_.rgb64[i])
        sl = ctx_.scale_like
        return (g_lut.view(n, 4).to(torch.float32), g_s.view(()).to(sl.dtype).to(sl.device), None, None)


def render_tf(mk: Context, views: list, lut: torch.Tensor, scale: torch.Tensor) -> torch.Tensor:
    """(len(views), H, W, 3) float32 images of the fitting rule."""
    return _TFRender.apply(lut, scale, mk, list(views))


def fit(mk: Context, views: list, target: torch.Tensor, lut0: torch.Tensor, scale0: float, steps: int = 200,
        lr: float = 0.01, log_every: int = 0):
    """Adam on (LUT, opacity scale) minimising the mean squared error to the
    target images. Returns (lut, scale, losses)."""
    lut = lut0.detach().clone().requires_grad_(True)
    scale = torch.tensor(float(scale0), dtype=torch.float32, requires_grad=True)
    opt = torch.optim.Adam([lut, scale], lr=lr)
    losses = []
    for k in range(steps):
        opt.zero_grad()
        img = render_tf(mk, views, lut, scale)
        loss = torch.mean((img - target) ** 2)
        loss.backward()
        opt.step()
        with torch.no_grad():
            lut.clamp_(0.0, 1.0)
            scale.clamp_(0.0, 8.0)
        losses.append(float(loss.detach()))
        if log_every and k % log_every == 0:
            print(f"[fit] step {k} loss {losses[-1]:.6g}", flush=True)
    return lut.detach(), float(scale.detach()), losses
