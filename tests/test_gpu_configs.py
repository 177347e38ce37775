"""Parity of every benchmark configuration at its own workload (GPU, C ABI).

Each configuration is run on its own grid, phantom, sigmas and iteration
count, against the CPU oracle (the C restatement of the reference's
Reconstruction<F, GaussianKernel<F>>, reconstruction.h:110-154) or, for the
gradients, the FP64 differentiation checker (itself pinned by central finite
differences in tests/test_oracle_diff.py).  The seeded inputs are checked
against their pinned SHA-256 first (tests/golden/event_digests.json).

  strip-256x160-2mm-1e6    plain MLEM, 1e6 events, 10 iterations (all events)
  64-member batch          batched form, 2 % of the events, autotuned,
                           graph-replayed 10 iterations
  strip-512x256-1mm-5e6    differentiation at sigma (14, 60) mm: 8 unrolled
                           iterations at 2e4 events against the FP64 checker;
                           the log-likelihood sigma gradient at all 5e6
                           events against the FP64 checker; the unrolled image
                           gradient at 5e6 events against central differences
                           of the GPU's own loss
  strip-2000x800-0.5mm-2e8 16 ordered subsets at 2e5 events against the
                           oracle's OSEM (the full 2e8 streamed set is checked
                           through invariants in tests/test_gpu_osem.py)

Tolerances: images rel-L2 <= 1e-4 against the FP32 oracle (the error against
FP64 is printed); log-likelihood value 2e-6 relative; sigma gradients 1e-4
relative (per-event terms in FP32, sums in FP64); unrolled loss 1e-4, its
sigma and image directional derivatives 1e-3 (products of per-iteration FP32
errors); central differences of the FP32 loss 5e-3.
"""
import numpy as np
import pytest

from conftest import check_events, print_case

pytestmark = pytest.mark.gpu

CFG1 = "strip-256x160-2mm-1e6"
CFG4 = "strip-512x256-1mm-5e6"
CFG5 = "strip-2000x800-0.5mm-2e8"
SIGMA0 = (14.0, 60.0)  # the differentiation workload's initial sigma guess


def test_cfg1_full_workload(pystrip_mod, phantom_mod, oracle_mod):
    w = phantom_mod.WORKLOADS[CFG1]
    print_case(CFG1, w.n_events, "events,", w.iterations, "iterations")
    ev = w.events()
    check_events(phantom_mod, ev, CFG1)
    rec = pystrip_mod.Reconstruction(w.geometry())
    rec.set_events(ev)
    rec.iterate(w.iterations)
    gpu = rec.rho
    og = oracle_mod.geometry(*w.oracle_geometry_tuple())
    ref32 = oracle_mod.mlem(og, ev, iterations=w.iterations)
    e32 = oracle_mod.rel_l2(gpu, ref32)
    ref64 = oracle_mod.mlem(og, ev.astype(np.float64),
                            iterations=w.iterations, dtype=np.float64)
    print("rel-L2 vs FP32 oracle", e32, "GPU vs FP64",
          oracle_mod.rel_l2(gpu, ref64), "FP32 oracle vs FP64",
          oracle_mod.rel_l2(ref32, ref64), flush=True)
    assert e32 <= 1e-4
    assert np.all(np.isfinite(gpu))


def test_cfg3_batch_of_64(pystrip_mod, phantom_mod, oracle_mod):
    members = phantom_mod.batched_workload(64, seed=0, event_scale=0.02)
    n = sum(w.n_events for w in members)
    print_case("batch of", len(members), "members,", n, "events, autotuned,",
               "10 graph-replayed iterations")
    b = pystrip_mod.Batch()
    evs = []
    for w in members:
        ev = w.events()
        evs.append(ev)
        b.add(w.geometry(), ev)
    b.iterate(1)
    ms = b.autotune()
    print("autotuned tiling", b.tiling(), "groups", b.groups(), "ms", ms,
          flush=True)
    b.iterate(9)
    worst = 0.0
    for i, w in enumerate(members):
        og = oracle_mod.geometry(*w.oracle_geometry_tuple())
        ref = oracle_mod.mlem(og, evs[i], iterations=10)
        e = oracle_mod.rel_l2(b.rho(i), ref)
        worst = max(worst, e)
        assert e <= 1e-4, (w.name, e)
    print("worst member rel-L2 vs FP32 oracle", worst)
    b.close()


def _cfg4(pystrip_mod, phantom_mod, oracle_mod, n, seed=None):
    w = phantom_mod.WORKLOADS[CFG4]
    ev = w.events(n, seed=seed)
    check_events(phantom_mod, ev, CFG4, seed)
    g = w.geometry()
    g.sigma_z, g.sigma_dl = SIGMA0
    t = list(w.oracle_geometry_tuple())
    t[6], t[7] = SIGMA0
    og = oracle_mod.geometry(*t)
    rec = pystrip_mod.Reconstruction(g)
    rec.set_events(ev)
    return w, ev, og, rec


def test_cfg4_unrolled_gradient(pystrip_mod, phantom_mod, oracle_mod):
    # MSE to the ground-truth phantom through 8 unrolled MLEM iterations from
    # rho = 1 at the sigma guess, against the FP64 checker's forward-mode
    # derivatives
    print_case(CFG4, "2e4 events, sigma", SIGMA0, "8 unrolled iterations")
    w, ev, og, rec = _cfg4(pystrip_mod, phantom_mod, oracle_mod, 20_000, 41)
    target = phantom_mod.phantom_image(w.phantom, w.n_z, w.n_y,
                                       w.pixel).astype(np.float32)
    rho0 = np.ones(rec.shape, np.float32)
    loss, grad, gs = rec.unrolled_gradient(rho0, w.iterations, target)
    ev64 = ev.astype(np.float64)
    r64, t64 = rho0.astype(np.float64), target.astype(np.float64)
    for i in range(2):
        d = np.zeros(2)
        d[i] = 1.0
        rl, j, rho_n = oracle_mod.mlem_loss(og, ev64, r64, w.iterations, t64,
                                            d_sigma=d)
        if i == 0:
            print("loss", loss, rl, flush=True)
            assert abs(loss - rl) <= 1e-4 * abs(rl)
            assert oracle_mod.rel_l2(rec.rho, rho_n) <= 1e-4
        print("d loss / d sigma", i, gs[i], j, flush=True)
        assert abs(gs[i] - j) <= 1e-3 * abs(j)
    rng = np.random.default_rng(4)
    for k in range(2):
        v = rng.standard_normal(rec.shape)
        _, j, _ = oracle_mod.mlem_loss(og, ev64, r64, w.iterations, t64,
                                       d_rho0=v)
        an = float((grad.astype(np.float64) * v).sum())
        print("image direction", k, an, j, flush=True)
        assert abs(an - j) <= 1e-3 * abs(j)


def test_cfg4_loglik_gradient(pystrip_mod, phantom_mod, oracle_mod):
    print_case(CFG4, "2e4 events, log-likelihood gradient at sigma", SIGMA0)
    w, ev, og, rec = _cfg4(pystrip_mod, phantom_mod, oracle_mod, 20_000, 41)
    target = phantom_mod.phantom_image(w.phantom, w.n_z, w.n_y, w.pixel)
    rho = (target + 0.1).astype(np.float32)
    value, grad, gs = rec.loglik_gradient(rho)
    rv, rg, rs = oracle_mod.loglik_grad(og, ev.astype(np.float64),
                                        rho.astype(np.float64))
    print("value", value, rv, "sigma gradient", gs, rs, flush=True)
    assert abs(value - rv) <= 2e-6 * abs(rv)
    assert oracle_mod.rel_l2(grad, rg) <= 1e-4
    assert np.all(np.abs(gs - rs) <= 1e-4 * np.abs(rs))


def test_cfg4_full_loglik_sigma_gradient(pystrip_mod, phantom_mod,
                                         oracle_mod):
    # all 5e6 events: the per-event terms summed in FP64 on the device
    # against the FP64 checker over the same events and image
    w = phantom_mod.WORKLOADS[CFG4]
    print_case(CFG4, w.n_events, "events, log-likelihood sigma gradient")
    w, ev, og, rec = _cfg4(pystrip_mod, phantom_mod, oracle_mod, w.n_events)
    rec.iterate(2)
    rho = rec.rho
    value, _, gs = rec.loglik_gradient(rho, want_grad_rho=False)
    rv, _, rs = oracle_mod.loglik_grad(og, ev.astype(np.float64),
                                       rho.astype(np.float64),
                                       want_grad_rho=False)
    print("value", value, rv, "sigma gradient", gs, rs, flush=True)
    assert abs(value - rv) <= 2e-6 * abs(rv)
    assert np.all(np.abs(gs - rs) <= 1e-4 * np.abs(rs))


def test_cfg4_full_unrolled_image_gradient(pystrip_mod, phantom_mod,
                                           oracle_mod):
    # all 5e6 events, 8 unrolled iterations: the image gradient along a
    # random direction against central differences of the GPU's own loss
    # (the support does not depend on the image, so the difference quotient
    # converges to the same derivative)
    w = phantom_mod.WORKLOADS[CFG4]
    print_case(CFG4, w.n_events, "events, unrolled image gradient vs central "
               "differences")
    w, ev, og, rec = _cfg4(pystrip_mod, phantom_mod, oracle_mod, w.n_events)
    target = phantom_mod.phantom_image(w.phantom, w.n_z, w.n_y,
                                       w.pixel).astype(np.float32)
    rho0 = np.ones(rec.shape, np.float32)
    loss, grad, _ = rec.unrolled_gradient(rho0, w.iterations, target)
    rng = np.random.default_rng(6)
    v = rng.uniform(-1.0, 1.0, size=rec.shape)
    h = 1e-2
    lp, _, _ = rec.unrolled_gradient((rho0 + h * v).astype(np.float32),
                                     w.iterations, target,
                                     want_grad_rho=False)
    lm, _, _ = rec.unrolled_gradient((rho0 - h * v).astype(np.float32),
                                     w.iterations, target,
                                     want_grad_rho=False)
    fd = (lp - lm) / (2 * h)
    an = float((grad.astype(np.float64) * v).sum())
    print("loss", loss, "directional derivative", an, "central difference",
          fd, flush=True)
    assert np.isfinite(loss) and np.all(np.isfinite(grad))
    assert abs(an - fd) <= 5e-3 * abs(fd)


def test_cfg5_osem_against_oracle(pystrip_mod, phantom_mod, oracle_mod):
    w = phantom_mod.WORKLOADS[CFG5]
    print_case(CFG5, "2e5 events, 16 subsets, 2 full iterations")
    ev = w.events(200_000, seed=43)
    check_events(phantom_mod, ev, CFG5, 43)
    og = oracle_mod.geometry(*w.oracle_geometry_tuple())
    rec = pystrip_mod.Reconstruction(w.geometry())
    rec.set_events(ev)
    rec.set_subsets(16)
    rec.iterate(2)
    gpu = rec.rho
    ref32 = oracle_mod.osem(og, ev, 16, iterations=2)
    e32 = oracle_mod.rel_l2(gpu, ref32)
    ref64 = oracle_mod.osem(og, ev.astype(np.float64), 16, iterations=2,
                            dtype=np.float64)
    print("rel-L2 vs FP32 oracle", e32, "GPU vs FP64",
          oracle_mod.rel_l2(gpu, ref64), "FP32 oracle vs FP64",
          oracle_mod.rel_l2(ref32, ref64), "span table", rec.span_table(),
          flush=True)
    assert e32 <= 1e-4
