"""Ordered subsets, streamed uploads and the FP64 image total on the GPU.

 - OSEM images match the oracle's ordered-subset iterations (contiguous
   subsets by event index, update times N / N_s) within 1e-4 relative L2;
 - appending events in chunks (pinned double-buffered uploads) gives the same
   resident events as one upload: equal support counts and forward
   projections;
 - the log-likelihood over subset plans equals the one over a single plan;
 - the FP64 image total after an MLEM iteration equals #{e : D_e > 0}.
"""
import numpy as np
import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu

CFG1 = "strip-256x160-2mm-1e6"
CFG2 = "strip-1000x400-1mm-2e7"


@pytest.mark.parametrize("subsets", [2, 5])
def test_osem_parity(pystrip_mod, phantom_mod, oracle_mod, subsets):
    print_case("OSEM parity, subsets", subsets)
    w = phantom_mod.WORKLOADS[CFG1]
    ev = w.events(60_000, seed=31)
    og = oracle_mod.geometry(*w.oracle_geometry_tuple())
    rec = pystrip_mod.Reconstruction(w.geometry())
    rec.set_events(ev)
    rec.set_subsets(subsets)
    rec.iterate(2)
    ref = oracle_mod.osem(og, ev, subsets, iterations=2)
    e = oracle_mod.rel_l2(rec.rho, ref)
    print("rel-L2", e)
    assert e <= 1e-4
    # back to plain MLEM on the same context
    rec.set_subsets(1)
    rec.rho = np.ones(rec.shape, np.float32)
    rec.iterate(2)
    assert oracle_mod.rel_l2(rec.rho, oracle_mod.mlem(og, ev, iterations=2)) \
        <= 1e-4


def test_osem_loglik_matches_single_plan(pystrip_mod, phantom_mod):
    print_case("log-likelihood over subset plans")
    w = phantom_mod.WORKLOADS[CFG1]
    ev = w.events(40_000, seed=33)
    rng = np.random.default_rng(3)
    rho = rng.uniform(0.5, 1.5, size=(w.n_y, w.n_z)).astype(np.float32)
    a = pystrip_mod.Reconstruction(w.geometry())
    a.set_events(ev)
    va, ga, sa = a.loglik_gradient(rho)
    b = pystrip_mod.Reconstruction(w.geometry())
    b.set_events(ev)
    b.set_subsets(4)
    vb, gb, sb = b.loglik_gradient(rho)
    assert abs(va - vb) <= 1e-6 * abs(va)
    assert np.linalg.norm(ga - gb) <= 1e-6 * np.linalg.norm(ga)
    assert np.allclose(sa, sb, rtol=1e-4)


def test_streamed_append_equals_single_upload(pystrip_mod, phantom_mod):
    # 2.05e7 events in three appends: every append spans several 2^23-event
    # chunks, so both staging buffers cycle
    print_case("streamed append", CFG2)
    w = phantom_mod.WORKLOADS[CFG2]
    ev = w.events(20_500_000, seed=35)
    a = pystrip_mod.Reconstruction(w.geometry())
    a.set_events(ev)
    b = pystrip_mod.Reconstruction(w.geometry())
    b.reserve_events(9_000_000)
    for lo, hi in ((0, 9_000_000), (9_000_000, 17_000_000),
                   (17_000_000, 20_500_000)):
        b.append_events(ev[lo:hi])
    assert b.n_events == a.n_events == len(ev)
    assert a.kernel_evaluations() == b.kernel_evaluations()
    rng = np.random.default_rng(5)
    rho = rng.uniform(0.5, 2.0, size=a.shape).astype(np.float32)
    a.rho = rho
    b.rho = rho
    da, db = a.forward(), b.forward()
    assert np.array_equal(da, db)


def test_image_total_fp64(pystrip_mod, phantom_mod):
    print_case("FP64 image total = #{D > 0}")
    w = phantom_mod.WORKLOADS[CFG1]
    ev = w.events(50_000, seed=37)
    rec = pystrip_mod.Reconstruction(w.geometry())
    rec.set_events(ev)
    n_pos = int((rec.forward() > 0).sum())
    rec.iterate(1)
    total = rec.image_total()
    ref = float(rec.rho.astype(np.float64).sum())
    assert abs(total - ref) <= 1e-9 * ref
    assert abs(total - n_pos) <= 1e-5 * n_pos


@pytest.mark.parametrize("plan_events", [20_000, 33_333])
def test_split_plans_equal_single_plan(pystrip_mod, phantom_mod, oracle_mod,
                                       plan_events):
    # plain MLEM over several plans (contiguous event ranges accumulating one
    # update): same support, forward projections and images as one plan
    print_case("plain MLEM split over plans of", plan_events, "events")
    w = phantom_mod.WORKLOADS[CFG1]
    ev = w.events(70_000, seed=39)
    og = oracle_mod.geometry(*w.oracle_geometry_tuple())
    a = pystrip_mod.Reconstruction(w.geometry())
    a.set_events(ev)
    b = pystrip_mod.Reconstruction(w.geometry())
    b.set_plan_events(plan_events)
    b.set_events(ev)
    assert a.kernel_evaluations() == b.kernel_evaluations()
    # the plan queries read the plans the iterations run (several here)
    assert a.band_groups() == b.band_groups()
    assert a.span_table() == b.span_table()
    ta, tb = a.band_tiling(), b.band_tiling()
    print("tiling one plan", ta, "split", tb)
    assert ta[:2] == tb[:2] and ta[3:] == tb[3:]
    rng = np.random.default_rng(7)
    rho = rng.uniform(0.5, 2.0, size=a.shape).astype(np.float32)
    a.rho = rho
    b.rho = rho
    da, db = a.forward(), b.forward()
    nz = da > 0
    assert np.array_equal(nz, db > 0)
    assert np.max(np.abs(da[nz] - db[nz]) / da[nz]) < 1e-6
    a.rho = np.ones(a.shape, np.float32)
    b.rho = np.ones(b.shape, np.float32)
    a.iterate(3)
    b.iterate(3)
    assert oracle_mod.rel_l2(b.rho, a.rho) <= 1e-5
    assert oracle_mod.rel_l2(b.rho, oracle_mod.mlem(og, ev, iterations=3)) \
        <= 1e-4
    # mass after one more update: sum rho = #{e : D_e > 0} (FP64 total)
    n_pos = int((b.forward() > 0).sum())
    b.iterate(1)
    assert abs(b.image_total() - n_pos) <= 1e-5 * n_pos


def test_osem_largest_grid_invariants(pystrip_mod, phantom_mod):
    # the whole largest workload: 2e8 events on the 2000 x 800 grid at
    # 0.5 mm, generated on the device in 1e7-event chunks and streamed back
    # in through the double-buffered appends, 16 subsets.  After a full
    # iteration (its last subset update) the FP64 image total equals
    # N / N_s x #{e in s : D_e > 0} for the last subset s, and the
    # log-likelihood gradient over the final image is finite.
    w = phantom_mod.WORKLOADS["strip-2000x800-0.5mm-2e8"]
    n, chunk, S = w.n_events, 10_000_000, 16
    print_case("OSEM 2000x800 @0.5 mm,", n, "streamed events, 16 subsets")
    rec = pystrip_mod.Reconstruction(w.geometry())
    rec.reserve_events(n)
    for k in range(0, n, chunk):
        out = pystrip_mod.generate_events(w.phantom, w.radius, w.length,
                                          w.sigma_z, w.sigma_dl, chunk,
                                          seed=1000 + k // chunk)
        if k == 0:
            # the device generator's stream is pinned by its first chunk
            print("first chunk SHA-256",
                  phantom_mod.events_digest(out["responses"][:chunk]),
                  flush=True)
        rec.append_events(out["responses"][:chunk])
    assert rec.n_events == n
    rec.set_subsets(S)
    rec.iterate(1)
    total = rec.image_total()
    # the last subset's events with D > 0 (the update keeps rho > 0 on their
    # supports, so the set is the one before the update)
    last = slice(n * (S - 1) // S, n)
    n_pos = int((rec.forward()[last] > 0).sum())
    expect = n / (n - n * (S - 1) // S) * n_pos
    print("image total", total, "expected", expect, "span table",
          rec.span_table(), flush=True)
    assert abs(total - expect) <= 1e-4 * expect
    v, g, gs = rec.loglik_gradient(want_grad_rho=False)
    print("loglik", v, "d/d sigma", gs, flush=True)
    assert np.isfinite(v) and np.all(np.isfinite(gs))
