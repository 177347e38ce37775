"""Pins of the FP64 differentiation checker (oracle_loglik_grad_f64,
oracle_mlem_loss_f64) without the GPU.

 - at the geometry's own sigmas its objective values equal the plain oracle's
   (log-likelihood of oracle_loglik_f64, MLEM images of oracle_mlem_f64),
   which pins the re-derived kernel;
 - its gradients equal FP64 central finite differences (h = 1e-3 relative) of
   the same objective, the support held at the geometry's sigmas;
 - the unrolled-MLEM directional derivative (forward mode) equals central
   finite differences along random directions in (rho0, sigma).
"""
import numpy as np
import pytest

W = "strip-256x160-2mm-1e6"


@pytest.fixture(scope="module")
def case(oracle_mod, phantom_mod):
    w = phantom_mod.WORKLOADS[W]
    g = oracle_mod.geometry(*w.oracle_geometry_tuple())
    ev = w.events(300, seed=11).astype(np.float64)
    rng = np.random.default_rng(3)
    rho = rng.uniform(0.5, 1.5, size=(w.n_y, w.n_z))
    return w, g, ev, rho


def test_loglik_value_matches_plain_oracle(oracle_mod, case):
    w, g, ev, rho = case
    L, _, _ = oracle_mod.loglik_grad(g, ev, rho)
    ref = oracle_mod.lib().oracle_loglik_f64(
        g, oracle_mod.OracleMode(*oracle_mod.MODE_CPU), oracle_mod._dp(ev),
        ev.shape[0],
        oracle_mod._dp(np.ascontiguousarray(rho)), None, 1)
    assert abs(L - ref) <= 1e-12 * abs(ref)


@pytest.mark.parametrize("scale", [1.0, 1.3])
def test_loglik_sigma_gradient_fd(oracle_mod, case, scale):
    w, g, ev, rho = case
    s0 = np.array([w.sigma_z, w.sigma_dl]) * scale
    _, _, gs = oracle_mod.loglik_grad(g, ev, rho, s0)
    for i in range(2):
        h = 1e-3 * s0[i]
        sp, sm = s0.copy(), s0.copy()
        sp[i] += h
        sm[i] -= h
        fd = (oracle_mod.loglik_grad(g, ev, rho, sp)[0] -
              oracle_mod.loglik_grad(g, ev, rho, sm)[0]) / (2 * h)
        assert abs(fd - gs[i]) <= 1e-5 * max(abs(gs[i]), 1e-3 * abs(fd) + 1)


def test_loglik_rho_gradient_fd(oracle_mod, case):
    w, g, ev, rho = case
    _, gr, _ = oracle_mod.loglik_grad(g, ev, rho)
    rng = np.random.default_rng(4)
    direction = rng.standard_normal(rho.shape)
    h = 1e-4
    fd = (oracle_mod.loglik_grad(g, ev, rho + h * direction)[0] -
          oracle_mod.loglik_grad(g, ev, rho - h * direction)[0]) / (2 * h)
    an = float((gr * direction).sum())
    assert abs(fd - an) <= 1e-6 * np.abs(gr * direction).sum()
    # pixels no event reaches: gradient -1 exactly
    _, gr1, _ = oracle_mod.loglik_grad(g, ev[:1], rho)
    assert np.sum(gr1 == -1.0) > 0.9 * gr1.size


def test_mlem_loss_image_matches_plain_oracle(oracle_mod, case):
    w, g, ev, rho = case
    _, _, rho_n = oracle_mod.mlem_loss(g, ev, np.ones_like(rho), 3,
                                       np.zeros_like(rho))
    ref = oracle_mod.mlem(g, ev, iterations=3, dtype=np.float64)
    assert np.max(np.abs(rho_n - ref)) <= 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("iterations", [1, 3])
def test_mlem_loss_directional_derivative_fd(oracle_mod, case, iterations):
    w, g, ev, rho = case
    rng = np.random.default_rng(7 + iterations)
    target = rng.uniform(0, 2, size=rho.shape)
    s0 = np.array([w.sigma_z, w.sigma_dl])
    d_rho = rng.standard_normal(rho.shape) * 0.1
    d_sig = rng.standard_normal(2) * s0
    loss, jvp, _ = oracle_mod.mlem_loss(g, ev, rho, iterations, target, s0,
                                        d_rho, d_sig)
    h = 1e-3
    lp = oracle_mod.mlem_loss(g, ev, rho + h * d_rho, iterations, target,
                              s0 + h * d_sig)[0]
    lm = oracle_mod.mlem_loss(g, ev, rho - h * d_rho, iterations, target,
                              s0 - h * d_sig)[0]
    fd = (lp - lm) / (2 * h)
    assert abs(fd - jvp) <= 2e-5 * abs(jvp) + 1e-12
