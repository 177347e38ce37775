"""GPU gradients (C ABI) against the FP64 differentiation checker.

The checker (oracle_loglik_grad_f64, oracle_mlem_loss_f64) is itself pinned
by FP64 central finite differences in tests/test_oracle_diff.py.

Tolerances (FP32 GPU against FP64):
  - log-likelihood value: 2e-6 relative;
  - image gradients: relative L2 <= 1e-4 (log-likelihood) and <= 1e-3
    (through unrolled MLEM iterations, a product of per-iteration errors);
  - sigma gradients: relative 1e-3 (sums over all events of terms of both
    signs, accumulated in FP32);
  - unrolled loss: 1e-4 relative; directional derivatives along random image
    directions: relative 1e-3.
"""
import numpy as np
import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu

CFG = "strip-256x160-2mm-1e6"


def _case(pystrip_mod, phantom_mod, oracle_mod, n, sigma_scale=1.0, seed=21):
    w = phantom_mod.WORKLOADS[CFG]
    ev = w.events(n, seed=seed)
    g = w.geometry()
    g.sigma_z *= sigma_scale
    g.sigma_dl *= sigma_scale
    t = list(w.oracle_geometry_tuple())
    t[6] *= sigma_scale
    t[7] *= sigma_scale
    og = oracle_mod.geometry(*t)
    rec = pystrip_mod.Reconstruction(g, pystrip_mod.MODE_CPU)
    rec.set_events(ev)
    return w, ev, g, og, rec


@pytest.mark.parametrize("sigma_scale", [1.0, 1.4])
def test_loglik_gradient(pystrip_mod, phantom_mod, oracle_mod, sigma_scale):
    print_case("log-likelihood gradient, sigma scale", sigma_scale)
    w, ev, g, og, rec = _case(pystrip_mod, phantom_mod, oracle_mod, 20000,
                              sigma_scale)
    rng = np.random.default_rng(2)
    rho = rng.uniform(0.5, 1.5, size=rec.shape).astype(np.float32)
    value, grad, gs = rec.loglik_gradient(rho)
    rv, rg, rs = oracle_mod.loglik_grad(og, ev.astype(np.float64),
                                        rho.astype(np.float64))
    print("value", value, rv, "sigma grad", gs, rs)
    assert abs(value - rv) <= 2e-6 * abs(rv)
    assert oracle_mod.rel_l2(grad, rg) <= 1e-4
    assert np.all(np.abs(gs - rs) <= 1e-3 * np.abs(rs))


def test_loglik_gradient_current_image(pystrip_mod, phantom_mod, oracle_mod):
    print_case("log-likelihood gradient of the context image")
    w, ev, g, og, rec = _case(pystrip_mod, phantom_mod, oracle_mod, 5000)
    rec.iterate(2)
    rho = rec.rho
    v0, g0, s0 = rec.loglik_gradient()
    v1, g1, s1 = rec.loglik_gradient(rho)
    # same computation twice (the back-projection adds with atomics)
    assert abs(v0 - v1) <= 1e-7 * abs(v0)
    assert oracle_mod.rel_l2(g0, g1) <= 1e-6
    assert np.allclose(s0, s1, rtol=1e-6)


@pytest.mark.parametrize("iterations", [1, 4])
def test_unrolled_gradient(pystrip_mod, phantom_mod, oracle_mod, iterations):
    print_case("unrolled MLEM gradient, iterations", iterations)
    w, ev, g, og, rec = _case(pystrip_mod, phantom_mod, oracle_mod, 3000,
                              sigma_scale=1.4)
    rng = np.random.default_rng(5)
    rho0 = rng.uniform(0.5, 1.5, size=rec.shape).astype(np.float32)
    target = rng.uniform(0.0, 2.0, size=rec.shape).astype(np.float32)
    loss, grad, gs = rec.unrolled_gradient(rho0, iterations, target)
    ev64 = ev.astype(np.float64)
    r64, t64 = rho0.astype(np.float64), target.astype(np.float64)
    rl, _, rho_n = oracle_mod.mlem_loss(og, ev64, r64, iterations, t64)
    print("loss", loss, rl)
    assert abs(loss - rl) <= 1e-4 * abs(rl)
    # the context image is rho_N
    assert oracle_mod.rel_l2(rec.rho, rho_n) <= 1e-4
    # sigma gradients: forward-mode derivatives along the two axes
    for i in range(2):
        d = np.zeros(2)
        d[i] = 1.0
        _, j, _ = oracle_mod.mlem_loss(og, ev64, r64, iterations, t64,
                                       d_sigma=d)
        print("sigma", i, gs[i], j)
        assert abs(gs[i] - j) <= 1e-3 * abs(j)
    # image gradient: directional derivatives along random directions
    for k in range(2):
        v = rng.standard_normal(rec.shape)
        _, j, _ = oracle_mod.mlem_loss(og, ev64, r64, iterations, t64,
                                       d_rho0=v)
        an = float((grad.astype(np.float64) * v).sum())
        print("direction", k, an, j)
        assert abs(an - j) <= 1e-3 * abs(j)


def test_unsupported_modes(pystrip_mod, phantom_mod):
    print_case("gradients refuse the legacy semantics")
    w = phantom_mod.WORKLOADS[CFG]
    rec = pystrip_mod.Reconstruction(w.geometry(), pystrip_mod.MODE_LEGACY_GPU)
    rec.set_events(w.events(100, seed=1))
    with pytest.raises(pystrip_mod.StripMLEMError):
        rec.loglik_gradient()


def test_torch_autograd():
    # its own process: torch must load its HIP runtime before the library
    # (this test process loaded the library first)
    import os
    import subprocess
    import sys
    print_case("torch.autograd Functions (tests/torch_autograd_check.py)")
    script = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                          "torch_autograd_check.py")
    r = subprocess.run([sys.executable, script], capture_output=True,
                       text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
