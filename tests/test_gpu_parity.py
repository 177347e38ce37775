"""GPU parity of the HIP path against the CPU oracle (calls through the C ABI).

Tolerances (floating point, FP32 path):
  - per-event forward projections D_e: max relative error 2e-5
  - MLEM images after N iterations: relative L2 error <= 1e-4 against the FP32
    oracle (the graded bar), and the error of both against the FP64 oracle is
    printed;
  - kernel-evaluation counts (the 3-sigma support) are integers: exact.
"""
import numpy as np
import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu

CFG1 = "strip-256x160-2mm-1e6"
CFG2 = "strip-1000x400-1mm-2e7"


# hot-path variants: bbox wavefront kernel (1), row-band kernels with the
# per-upload span table (2), row-band kernels solving the row runs in every
# iteration (-2)
ALGORITHMS = [1, 2, -2]


def _setup(pystrip_mod, phantom_mod, oracle_mod, name, n, seed=0,
           mode=None, algorithm=0):
    w = phantom_mod.WORKLOADS[name]
    ev = w.events(n, seed=seed)
    g = w.geometry()
    og = oracle_mod.geometry(*w.oracle_geometry_tuple())
    rec = pystrip_mod.Reconstruction(g, mode or pystrip_mod.MODE_CPU)
    if algorithm:
        rec.set_algorithm(abs(algorithm))
        if algorithm < 0:
            rec.set_span_table(False)
    rec.set_events(ev)
    if abs(algorithm) == 2:
        assert rec.span_table() == (algorithm > 0)
    return w, ev, rec, og


def test_device_present(pystrip_mod, phantom_mod, oracle_mod):
    print_case("device name")
    w, ev, rec, og = _setup(pystrip_mod, phantom_mod, oracle_mod, CFG1, 100)
    name = rec.device_name()
    print(name)
    assert "gfx950" in name


@pytest.mark.parametrize("algorithm", ALGORITHMS)
@pytest.mark.parametrize("name", [CFG1, CFG2])
def test_support_counts_exact(pystrip_mod, phantom_mod, oracle_mod, name,
                              algorithm):
    # the bbox kernel tests every bbox pixel; the row-band kernels walk
    # analytic row spans re-tested exactly at their ends: both must visit
    # exactly the reference's support (reconstruction.h:255-264)
    print_case("support counts", name, "algorithm", algorithm)
    w, ev, rec, og = _setup(pystrip_mod, phantom_mod, oracle_mod, name, 20000,
                            algorithm=algorithm)
    nb, ne = oracle_mod.support_counts(og, ev)
    assert rec.algorithm == abs(algorithm)
    assert rec.kernel_evaluations() == int(ne.sum())


@pytest.mark.parametrize("algorithm", ALGORITHMS)
@pytest.mark.parametrize("name", [CFG1, CFG2])
def test_forward_projection(pystrip_mod, phantom_mod, oracle_mod, name,
                            algorithm):
    print_case("forward", name, "algorithm", algorithm)
    w, ev, rec, og = _setup(pystrip_mod, phantom_mod, oracle_mod, name, 20000,
                            algorithm=algorithm)
    rng = np.random.default_rng(1)
    rho = rng.uniform(0.5, 2.0, size=rec.shape).astype(np.float32)
    rec.rho = rho
    d_gpu = rec.forward().astype(np.float64)
    d_ref = oracle_mod.forward(og, ev, rho).astype(np.float64)
    nz = d_ref > 0
    assert np.array_equal(d_gpu > 0, nz)
    rel = np.abs(d_gpu[nz] - d_ref[nz]) / d_ref[nz]
    print("max rel", rel.max())
    assert rel.max() < 2e-5


@pytest.mark.parametrize("algorithm", ALGORITHMS)
def test_mlem_parity_plain_small(pystrip_mod, phantom_mod, oracle_mod,
                                 algorithm):
    print_case("mlem parity", CFG1, "1e5 events, 10 iterations, algorithm",
               algorithm)
    w, ev, rec, og = _setup(pystrip_mod, phantom_mod, oracle_mod, CFG1,
                            100_000, algorithm=algorithm)
    rec.iterate(10)
    gpu = rec.rho
    ref32 = oracle_mod.mlem(og, ev, iterations=10)
    ref64 = oracle_mod.mlem(og, ev.astype(np.float64), iterations=10,
                            dtype=np.float64)
    e32 = oracle_mod.rel_l2(gpu, ref32)
    print("rel-L2 vs FP32 oracle", e32, "GPU vs FP64", oracle_mod.rel_l2(gpu, ref64),
          "FP32 oracle vs FP64", oracle_mod.rel_l2(ref32, ref64))
    assert e32 <= 1e-4


@pytest.mark.parametrize("algorithm", ALGORITHMS)
def test_mlem_parity_scale_grid(pystrip_mod, phantom_mod, oracle_mod,
                                algorithm):
    print_case("mlem parity", CFG2, "1e5 events, 5 iterations, algorithm",
               algorithm)
    w, ev, rec, og = _setup(pystrip_mod, phantom_mod, oracle_mod, CFG2,
                            100_000, algorithm=algorithm)
    rec.iterate(5)
    gpu = rec.rho
    ref32 = oracle_mod.mlem(og, ev, iterations=5)
    e32 = oracle_mod.rel_l2(gpu, ref32)
    print("rel-L2 vs FP32 oracle", e32)
    assert e32 <= 1e-4


@pytest.mark.parametrize("algorithm", ALGORITHMS)
def test_mass_conservation(pystrip_mod, phantom_mod, oracle_mod, algorithm):
    # sum_p rho_1(p) = #{e : D_e > 0} (reconstruction.h:290-296)
    print_case("mass conservation, algorithm", algorithm)
    w, ev, rec, og = _setup(pystrip_mod, phantom_mod, oracle_mod, CFG1, 50_000,
                            algorithm=algorithm)
    d = rec.forward()
    n_pos = int((d > 0).sum())
    rec.iterate(1)
    total = float(rec.rho.astype(np.float64).sum())
    assert abs(total - n_pos) / n_pos < 1e-5


def test_legacy_mode(pystrip_mod, phantom_mod, oracle_mod):
    print_case("legacy --gpu semantics")
    w, ev, rec, og = _setup(pystrip_mod, phantom_mod, oracle_mod, CFG1,
                            50_000, mode=pystrip_mod.MODE_LEGACY_GPU)
    rec.iterate(3)
    gpu = rec.rho
    ref = oracle_mod.mlem(og, ev, iterations=3, mode=oracle_mod.MODE_LEGACY_GPU)
    e = oracle_mod.rel_l2(gpu, ref)
    print("legacy rel-L2", e)
    assert e <= 1e-4


@pytest.mark.parametrize("kind", [0, 1])
def test_sensitivity_maps(pystrip_mod, phantom_mod, oracle_mod, kind):
    print_case("sensitivity map", kind)
    w, ev, rec, og = _setup(pystrip_mod, phantom_mod, oracle_mod, CFG1, 10)
    s = rec.sensitivity(kind).astype(np.float64)
    r = oracle_mod.sensitivity_map(og, kind, dtype=np.float64)
    assert np.max(np.abs(s - r)) < 2e-6


@pytest.mark.parametrize("algorithm", ALGORITHMS)
def test_empty_and_outside_events(pystrip_mod, phantom_mod, oracle_mod,
                                  algorithm):
    print_case("empty / outside events, algorithm", algorithm)
    w = phantom_mod.WORKLOADS[CFG1]
    rec = pystrip_mod.Reconstruction(w.geometry())
    rec.set_algorithm(abs(algorithm))
    rec.set_span_table(algorithm > 0)
    rec.set_events(np.zeros((0, 3), np.float32))
    rec.iterate(2)
    assert np.all(rec.rho == 0)
    # events far outside the grid: empty clamped boxes, D = 0, guarded
    far = np.array([[2000.0, 2000.0, 0.0], [-3000.0, -3000.0, 10.0]],
                   np.float32)
    rec2 = pystrip_mod.Reconstruction(w.geometry())
    rec2.set_algorithm(abs(algorithm))
    rec2.set_span_table(algorithm > 0)
    rec2.set_events(far)
    d = rec2.forward()
    og = oracle_mod.geometry(*w.oracle_geometry_tuple())
    assert np.allclose(d, oracle_mod.forward(og, far, np.ones(rec2.shape, np.float32)))
    rec2.iterate(1)
    assert np.all(np.isfinite(rec2.rho))


def test_run_protocol(pystrip_mod, phantom_mod, oracle_mod):
    # output/progress protocol of GPU::Reconstruction::run
    # (cuda/reconstruction.cu:79-113)
    print_case("run protocol")
    w = phantom_mod.WORKLOADS[CFG1]
    ev = w.events(20000)
    outs, prog = [], []
    final = pystrip_mod.run(w.geometry(), ev, np.ones((160, 256), np.float32),
                            2, 3, output=lambda it, img: outs.append((it, img)),
                            progress=lambda c, f: prog.append((c, f)))
    assert [o[0] for o in outs] == [-1, 1, 2, 3, 6]
    assert prog[0] == (0, False) and prog[1] == (0, True)
    assert prog[-1] == (6, False)
    og = oracle_mod.geometry(*w.oracle_geometry_tuple())
    ref = oracle_mod.mlem(og, ev, iterations=6)
    assert oracle_mod.rel_l2(final, ref) <= 1e-4
    assert oracle_mod.rel_l2(outs[-1][1], ref) <= 1e-4
    s = oracle_mod.sensitivity_map(og, 1)
    assert np.max(np.abs(outs[0][1] - s)) < 2e-6


def test_algorithms_agree_numerator_sensitivity(pystrip_mod, phantom_mod,
                                                oracle_mod):
    # sensitivity dividing the back-projection (thread_granularity.cuh:92-100
    # with the clamped bbox): both algorithms and the oracle agree
    print_case("numerator sensitivity mode, both algorithms")
    mode = (0, 2, 0, 1)
    imgs = []
    for algorithm in ALGORITHMS:
        w, ev, rec, og = _setup(pystrip_mod, phantom_mod, oracle_mod, CFG1,
                                30_000, mode=mode, algorithm=algorithm)
        rec.iterate(3)
        imgs.append(rec.rho)
    ref = oracle_mod.mlem(og, ev, iterations=3, mode=mode)
    for img in imgs:
        assert oracle_mod.rel_l2(img, ref) <= 1e-4


def test_algorithm_selection(pystrip_mod, phantom_mod):
    print_case("algorithm selection")
    w = phantom_mod.WORKLOADS[CFG1]
    rec = pystrip_mod.Reconstruction(w.geometry())
    assert rec.algorithm == 2
    legacy = pystrip_mod.Reconstruction(w.geometry(),
                                        pystrip_mod.MODE_LEGACY_GPU)
    assert legacy.algorithm == 1
    with pytest.raises(pystrip_mod.StripMLEMError):
        legacy.set_algorithm(2)


@pytest.mark.parametrize("name", [CFG1, CFG2])
def test_recurrence_mode_parity(pystrip_mod, phantom_mod, oracle_mod, name):
    # the restarted exp recurrence along rows (exp mode 1) against the oracle
    print_case("exp recurrence mode", name)
    w, ev, rec, og = _setup(pystrip_mod, phantom_mod, oracle_mod, name,
                            50_000, algorithm=2)
    rec.set_exp_mode(1)
    rng = np.random.default_rng(5)
    rho = rng.uniform(0.5, 2.0, size=rec.shape).astype(np.float32)
    rec.rho = rho
    d_gpu = rec.forward().astype(np.float64)
    d_ref = oracle_mod.forward(og, ev, rho).astype(np.float64)
    nz = d_ref > 0
    assert np.max(np.abs(d_gpu[nz] - d_ref[nz]) / d_ref[nz]) < 2e-5
    rec.rho = np.ones(rec.shape, np.float32)
    rec.iterate(5)
    e = oracle_mod.rel_l2(rec.rho, oracle_mod.mlem(og, ev, iterations=5))
    print("rel-L2", e)
    assert e <= 1e-4


@pytest.mark.parametrize("tiling,waves", [((64, 16, 32, 4096, 1024), 0),
                                          ((64, 16, 16, 4096, 512), 2),
                                          ((32, 32, 16, 2048, 1024), 1),
                                          ((64, 64, 64, 8192, 2048), 0)])
def test_band_tilings(pystrip_mod, phantom_mod, oracle_mod, tiling, waves):
    # every band height / bucket / task size / workgroup shape is the same
    # reconstruction
    print_case("band tiling", tiling, "waves", waves)
    w, ev, rec, og = _setup(pystrip_mod, phantom_mod, oracle_mod, CFG2,
                            60_000, algorithm=2)
    rec.set_band_tiling(*tiling)
    rec.set_backprojection_waves(waves)
    nb, ne = oracle_mod.support_counts(og, ev)
    assert rec.kernel_evaluations() == int(ne.sum())
    rec.iterate(3)
    e = oracle_mod.rel_l2(rec.rho, oracle_mod.mlem(og, ev, iterations=3))
    print("rel-L2", e, "tiling", rec.band_tiling())
    assert e <= 1e-4


@pytest.mark.parametrize("rows,groups,spans,exp_mode", [
    ((16, 16), (4, 2), True, 1), ((16, 16), (2, 4), True, 0),
    ((32, 32), (2, 2), True, 1), ((16, 16), (4, 4), False, 1),
    ((32, 16), (2, 2), False, 0)])
def test_band_groups(pystrip_mod, phantom_mod, oracle_mod, rows, groups,
                     spans, exp_mode):
    # several pairs per wavefront step (lane groups with their own pairs and
    # back-projection windows): forward D per event, support count and the
    # reconstruction against the oracle
    print_case("band groups", rows, groups, "spans", spans, "exp", exp_mode)
    w, ev, rec, og = _setup(pystrip_mod, phantom_mod, oracle_mod, CFG2,
                            60_000, algorithm=2)
    rec.set_span_table(spans)
    rec.set_exp_mode(exp_mode)
    rec.set_band_tiling(rows[0], rows[1], 16, 2048, 512)
    rec.set_band_groups(*groups)
    nb, ne = oracle_mod.support_counts(og, ev)
    assert rec.kernel_evaluations() == int(ne.sum())
    assert rec.band_groups() == groups
    rng = np.random.default_rng(9)
    rho = rng.uniform(0.5, 2.0, size=rec.shape).astype(np.float32)
    rec.rho = rho
    d_gpu = rec.forward().astype(np.float64)
    d_ref = oracle_mod.forward(og, ev, rho).astype(np.float64)
    nz = d_ref > 0
    assert np.array_equal(d_gpu > 0, nz)
    assert np.max(np.abs(d_gpu[nz] - d_ref[nz]) / d_ref[nz]) < 2e-5
    rec.rho = np.ones(rec.shape, np.float32)
    rec.iterate(3)
    e = oracle_mod.rel_l2(rec.rho, oracle_mod.mlem(og, ev, iterations=3))
    print("rel-L2", e, "tiling", rec.band_tiling())
    assert e <= 1e-4


@pytest.mark.parametrize("cu_waves,bucket", [(4, 8), (8, 16), (12, 8),
                                             (32, 8), (32, 64)])
def test_backprojection_occupancy(pystrip_mod, phantom_mod, oracle_mod,
                                  cu_waves, bucket):
    # automatic back-projection groups: each task takes 4, 2 or 1 pairs per
    # step as its window fits the per-wavefront LDS budget of the occupancy
    # (4 wavefronts per CU: one 160 KiB workgroup, 4 everywhere; 8: mostly
    # 4; 32: a mix down to 1), so one
    # launch runs tasks of every lane layout
    print_case("back-projection occupancy", cu_waves, "bucket", bucket)
    w, ev, rec, og = _setup(pystrip_mod, phantom_mod, oracle_mod, CFG2,
                            60_000, algorithm=2)
    rec.set_band_tiling(32, 16, bucket, 2048, 512)
    rec.set_band_groups(0, 0)
    rec.set_backprojection_occupancy(cu_waves)
    assert rec.band_groups() == (2, 4)
    rec.iterate(3)
    e = oracle_mod.rel_l2(rec.rho, oracle_mod.mlem(og, ev, iterations=3))
    print("rel-L2", e)
    assert e <= 1e-4


def test_full_size_against_reference_image(pystrip_mod, phantom_mod,
                                           oracle_mod):
    # the headline workload at full size (2e7 seeded events, 20 iterations)
    # against the reference's own CPU reconstruction of the same events
    # (tests/golden/strip_1000x400_2e7_20it_ref.npz, make_full_image.py)
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                        "strip_1000x400_2e7_20it_ref.npz")
    gold = np.load(path)
    print_case("full size", CFG2, int(gold["n_events"]), "events,",
               int(gold["iterations"]), "iterations")
    w = phantom_mod.WORKLOADS[CFG2]
    assert w.seed == int(gold["seed"])
    ev = w.events(int(gold["n_events"]))
    # the events the reference image was made from
    assert phantom_mod.events_digest(ev) == str(gold["events_sha256"])
    rec = pystrip_mod.Reconstruction(w.geometry())
    rec.set_events(ev)
    rec.iterate(int(gold["iterations"]))
    e32 = oracle_mod.rel_l2(rec.rho, gold["rho_f32"])
    print("rel-L2 vs the reference FP32 image", e32)
    if "rho_f64" in gold:
        print("rel-L2 vs the reference FP64 image",
              oracle_mod.rel_l2(rec.rho, gold["rho_f64"]),
              "reference FP32 vs FP64",
              oracle_mod.rel_l2(gold["rho_f32"], gold["rho_f64"]))
    assert e32 <= 1e-4


def test_legacy_constructed_cases(pystrip_mod, oracle_mod):
    # the legacy warp kernel's float-reciprocal row split on a bbox 41
    # columns wide (columns 0 of rows 26..32 skipped, column 41 of rows
    # 28..31 visited) and its unguarded 1/D (rho0 = 0 on the support: NaN
    # there), against the oracle's legacy mode, whose both behaviours are
    # pinned in tests/test_oracle.py
    print_case("legacy float split and unguarded 1/D")
    G = pystrip_mod.Geometry.make(350, 500, 128, 64, 2, 20, 42)
    og = oracle_mod.geometry(350, 500, 128, 64, 2, 2, 20, 42)
    ev = np.array([[-87.0, -87.0, 0.0]], np.float32)
    rec = pystrip_mod.Reconstruction(G, pystrip_mod.MODE_LEGACY_GPU)
    rec.set_events(ev)
    rec.iterate(1)
    ref = oracle_mod.mlem(og, ev, iterations=1, mode=oracle_mod.MODE_LEGACY_GPU)
    gpu = rec.rho
    assert np.array_equal(gpu > 0, ref > 0)
    assert np.all(gpu[26:33, 0] == 0) and np.all(gpu[28:32, 41] > 0)
    assert oracle_mod.rel_l2(gpu, ref) <= 1e-5
    zero = np.zeros(rec.shape, np.float32)
    rec.rho = zero
    rec.iterate(1)
    ref0 = oracle_mod.mlem(og, ev, rho0=zero, iterations=1,
                           mode=oracle_mod.MODE_LEGACY_GPU)
    nan = np.isnan(rec.rho)
    assert nan.sum() > 100
    assert np.array_equal(nan, np.isnan(ref0))
    assert np.all(rec.rho[~nan] == 0)
