"""Checks of pystrip.autograd, run as its own process by
tests/test_gpu_adjoint.py::test_torch_autograd: torch is imported before the
HIP library is loaded, so both use one HIP runtime.  Exits non-zero on a
failed check."""
import os
import sys

import torch  # first: the library then binds to torch's HIP runtime

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "j-pet-mlem-old_amd"))

import numpy as np  # noqa: E402

from pystrip import autograd as sag  # noqa: E402
from pystrip import phantom  # noqa: E402


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def main():
    if not torch.cuda.is_available():
        print("no device visible to torch", flush=True)
        return 2
    w = phantom.WORKLOADS["strip-256x160-2mm-1e6"]
    ev = w.events(3000, seed=8)
    model = sag.StripModel(w.geometry(), ev)
    rng = np.random.default_rng(9)
    rho_np = rng.uniform(0.5, 1.5, size=(w.n_y, w.n_z)).astype(np.float32)
    rho = torch.tensor(rho_np, device="cuda", requires_grad=True)
    sigma = torch.tensor([w.sigma_z * 1.2, w.sigma_dl * 1.2],
                         dtype=torch.float64, requires_grad=True)
    L = sag.loglik(rho, sigma, model)
    (2.0 * L).backward()
    rec = model.context(sigma.detach().tolist())
    v, g, gs = rec.loglik_gradient(rho_np)
    assert abs(L.item() - v) <= 1e-7 * abs(v), (L.item(), v)
    assert rel_l2(rho.grad.cpu().numpy(), 2 * g) <= 1e-6
    assert np.allclose(sigma.grad.numpy(), 2 * gs, rtol=1e-6), (sigma.grad, gs)
    rho0 = torch.tensor(rho_np, device="cuda", requires_grad=True)
    target = torch.tensor(rng.uniform(0, 2, size=rho_np.shape),
                          dtype=torch.float32, device="cuda",
                          requires_grad=True)
    sigma.grad = None
    loss = sag.unrolled_mlem_loss(rho0, sigma, target, model, iterations=3)
    loss.backward()
    l2, g2, s2 = rec.unrolled_gradient(rho_np, 3,
                                       target.detach().cpu().numpy())
    # a second run of the same computation: the back-projection adds with
    # atomics, so the two agree to rounding, not bit for bit
    assert abs(loss.item() - l2) <= 1e-5 * abs(l2), (loss.item(), l2)
    assert rel_l2(rho0.grad.cpu().numpy(), g2) <= 1e-4
    assert np.allclose(sigma.grad.numpy(), s2, rtol=1e-4), (sigma.grad, s2)
    rho_n = rec.rho
    gt = -2.0 * (rho_n - target.detach().cpu().numpy()) / rho_n.size
    assert rel_l2(target.grad.cpu().numpy(), gt) <= 1e-4
    # gradient descent on sigma moves towards the generating values
    print("torch autograd checks passed", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
