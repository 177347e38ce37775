"""Batched reconstructions (strip_mlem_batch_*) on the GPU.

Members with different grids, pixel sizes, phantoms and event counts are
iterated together (one launch per pass over all members, graph-replayed
iterations); each member must match the FP32 oracle (rel-L2 <= 1e-4) and the
plain single-reconstruction form.
"""
import numpy as np
import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu


def _members(phantom_mod, n=5, seed=3, scale=0.01):
    return phantom_mod.batched_workload(n, seed=seed, event_scale=scale)


def _build(pystrip_mod, members, graph=True):
    b = pystrip_mod.Batch()
    b.set_graph(graph)
    evs = []
    for w in members:
        ev = w.events()
        evs.append(ev)
        b.add(w.geometry(), ev)
    return b, evs


def test_batch_matches_oracle(pystrip_mod, phantom_mod, oracle_mod):
    members = _members(phantom_mod)
    print_case("batch parity", [w.name for w in members])
    b, evs = _build(pystrip_mod, members)
    b.iterate(3)
    for i, w in enumerate(members):
        og = oracle_mod.geometry(*w.oracle_geometry_tuple())
        ref = oracle_mod.mlem(og, evs[i], iterations=3)
        e = oracle_mod.rel_l2(b.rho(i), ref)
        print(w.name, e)
        assert e <= 1e-4


def test_batch_graph_and_eager_agree(pystrip_mod, phantom_mod):
    members = _members(phantom_mod, n=4, seed=5)
    print_case("batch graph vs eager, 5 iterations")
    b1, _ = _build(pystrip_mod, members, graph=True)
    b2, _ = _build(pystrip_mod, members, graph=False)
    b1.iterate(5)
    b2.iterate(2)
    b2.iterate(3)
    for i in range(len(members)):
        a, c = b1.rho(i), b2.rho(i)
        assert np.linalg.norm(a - c) <= 1e-5 * np.linalg.norm(c)


def test_batch_and_plain_agree(pystrip_mod, phantom_mod, oracle_mod):
    members = _members(phantom_mod, n=4, seed=9)
    print_case("batch vs plain contexts")
    b, evs = _build(pystrip_mod, members)
    b.iterate(4)
    for i, w in enumerate(members):
        rec = pystrip_mod.Reconstruction(w.geometry())
        rec.set_events(evs[i])
        rec.iterate(4)
        assert oracle_mod.rel_l2(b.rho(i), rec.rho) <= 1e-5
    # a member driven alone, then with the batch
    m = b.member(1)
    m.iterate(1)
    b.iterate(1)
    og = oracle_mod.geometry(*members[1].oracle_geometry_tuple())
    ref = oracle_mod.mlem(og, evs[1], iterations=6)
    assert oracle_mod.rel_l2(b.rho(1), ref) <= 1e-4


@pytest.mark.parametrize("tiling,groups", [((64, 32, 32, 1024, 256, 2), (1, 1)),
                                           ((16, 16, 16, 2048, 512, 4), (4, 2))])
def test_batch_tiling_takes_effect(pystrip_mod, phantom_mod, oracle_mod,
                                   tiling, groups):
    # a new tiling after the first iterations rebuilds every member's plan
    members = _members(phantom_mod, n=3, seed=13)
    print_case("batch re-tiling", tiling, groups)
    b, evs = _build(pystrip_mod, members)
    b.iterate(1)
    b.set_tiling(*tiling)
    b.set_groups(*groups)
    b.iterate(2)
    for i, w in enumerate(members):
        m = b.member(i)
        assert m.band_tiling()[:2] == tiling[:2]
        assert m.band_groups() == groups
        og = oracle_mod.geometry(*w.oracle_geometry_tuple())
        ref = oracle_mod.mlem(og, evs[i], iterations=3)
        assert oracle_mod.rel_l2(b.rho(i), ref) <= 1e-4


def test_batch_autotune_keeps_images(pystrip_mod, phantom_mod, oracle_mod):
    members = _members(phantom_mod, n=3, seed=11)
    print_case("batch autotune")
    b, evs = _build(pystrip_mod, members)
    b.iterate(1)
    before = [b.rho(i) for i in range(len(members))]
    best = b.autotune()
    t = b.tiling()
    gr = b.groups()
    print("tiling", t, "groups", gr, "ms", best)
    assert best > 0
    assert t[0] in (16, 32, 64) and t[1] in (16, 32, 64)
    assert t[2] in (8, 16, 32, 64) and t[5] >= 1
    assert gr[0] in (1, 2, 4) and gr[1] in (1, 2, 4)
    for i in range(len(members)):
        assert np.array_equal(b.rho(i), before[i])
    b.iterate(2)
    for i, w in enumerate(members):
        og = oracle_mod.geometry(*w.oracle_geometry_tuple())
        ref = oracle_mod.mlem(og, evs[i], iterations=3)
        assert oracle_mod.rel_l2(b.rho(i), ref) <= 1e-4


def test_batch_empty_member(pystrip_mod, phantom_mod):
    print_case("batch with an empty member")
    members = _members(phantom_mod, n=2, seed=13)
    b, _ = _build(pystrip_mod, members)
    w = members[0]
    b.add(w.geometry(), np.zeros((0, 3), np.float32))
    b.iterate(2)
    assert np.all(b.rho(2) == 0)
    assert np.all(np.isfinite(b.rho(0)))
