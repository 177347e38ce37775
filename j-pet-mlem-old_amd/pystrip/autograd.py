"""torch.autograd Functions over the strip MLEM gradients of the C ABI.

Optional: needs torch with a ROCm device.  The HIP library computes values and
gradients in one call on its own stream (torch's stream is synchronised
first); torch only supplies device memory.

  model = StripModel(geometry, events)            # events: (n, 3) float32
  L = loglik(rho, sigma, model)                    # rho: (n_y, n_z) cuda
  loss = unrolled_mlem_loss(rho0, sigma, target, model, iterations=8)
  loss.backward()   # rho.grad / rho0.grad, sigma.grad (sigma_z, sigma_dl)

The event supports (bbox and the 3-sigma ellipse) are those of the sigma the
context was built with, so sigma gradients are the almost-everywhere
derivatives of the pixel sums; a new sigma value builds a new context.

Import this module (or torch) before anything loads the HIP library: torch
ships its own HIP runtime under the same soname, and the library then binds
to it, so both share one runtime and device pointers pass between them.
"""
from __future__ import annotations

import numpy as np
import torch

from . import MODE_CPU, Reconstruction
from . import _library_loaded

if _library_loaded() and not torch.cuda.is_available():
    raise ImportError(
        "pystrip.autograd must be imported before the HIP library is loaded "
        "(a second HIP runtime would not share torch's device memory)")
if torch.cuda.is_available():
    torch.cuda.init()


class StripModel:
    """Scanner geometry and event list, with the context of the last sigma."""

    def __init__(self, geometry, events, device: int = 0):
        self.geometry = geometry
        self.events = np.ascontiguousarray(events, dtype=np.float32)
        self.device = device
        self._key = None
        self._rec = None

    def context(self, sigma) -> Reconstruction:
        key = (float(sigma[0]), float(sigma[1]))
        if key != self._key:
            g = _with_sigma(self.geometry, key)
            if self._rec is not None:
                self._rec.close()
            rec = Reconstruction(g, MODE_CPU, device=self.device)
            rec.set_events(self.events)
            self._rec, self._key = rec, key
        return self._rec


def _with_sigma(geometry, key):
    """Copy of a Geometry structure with other sigmas."""
    g = type(geometry)()
    for name, _ in geometry._fields_:
        setattr(g, name, getattr(geometry, name))
    g.sigma_z, g.sigma_dl = key
    return g


def _prepare(t: torch.Tensor, shape):
    if not t.is_cuda:
        raise ValueError("images must be device tensors")
    t = t.detach().to(torch.float32).contiguous()
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"image shape {tuple(t.shape)} != {tuple(shape)}")
    return t


class _LogLikelihood(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rho, sigma, model):
        rec = model.context(sigma.detach().cpu().tolist())
        r = _prepare(rho, rec.shape)
        grad = torch.empty_like(r)
        torch.cuda.current_stream(r.device).synchronize()
        value, gs = rec.loglik_gradient_device(r.data_ptr(), grad.data_ptr())
        ctx.save_for_backward(grad, torch.as_tensor(gs, dtype=sigma.dtype,
                                                     device=sigma.device))
        return torch.tensor(value, dtype=torch.float64, device=rho.device)

    @staticmethod
    def backward(ctx, g):
        grad, gs = ctx.saved_tensors
        return (g.to(grad.dtype) * grad).to(torch.float32), \
            g.to(gs.dtype).to(gs.device) * gs, None


class _UnrolledMLEMLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rho0, sigma, target, model, iterations):
        rec = model.context(sigma.detach().cpu().tolist())
        r0 = _prepare(rho0, rec.shape)
        t = _prepare(target, rec.shape)
        grad = torch.empty_like(r0)
        torch.cuda.current_stream(r0.device).synchronize()
        loss, gs = rec.unrolled_gradient_device(r0.data_ptr(), int(iterations),
                                                t.data_ptr(), grad.data_ptr())
        rho_n = torch.empty_like(r0)
        rec.get_rho_device(rho_n.data_ptr())
        # d loss / d target = -2 (rho_N - target) / n_pixels
        grad_t = (t - rho_n) * (2.0 / t.numel())
        ctx.save_for_backward(grad, torch.as_tensor(gs, dtype=sigma.dtype,
                                                    device=sigma.device),
                              grad_t)
        return torch.tensor(loss, dtype=torch.float64, device=rho0.device)

    @staticmethod
    def backward(ctx, g):
        grad, gs, grad_t = ctx.saved_tensors
        return ((g.to(grad.dtype) * grad).to(torch.float32),
                g.to(gs.dtype).to(gs.device) * gs,
                (g.to(grad_t.dtype) * grad_t).to(torch.float32), None, None)


def loglik(rho: torch.Tensor, sigma: torch.Tensor, model: StripModel):
    """sum_{e: D_e > 0} log D_e - sum_p rho(p), differentiable in rho and
    sigma = (sigma_z, sigma_dl)."""
    return _LogLikelihood.apply(rho, sigma, model)


def unrolled_mlem_loss(rho0: torch.Tensor, sigma: torch.Tensor,
                       target: torch.Tensor, model: StripModel,
                       iterations: int = 8):
    """mean((MLEM^iterations(rho0) - target)^2), differentiable in rho0,
    sigma and target; every iterate is kept as a checkpoint on the device."""
    return _UnrolledMLEMLoss.apply(rho0, sigma, target, model, iterations)
