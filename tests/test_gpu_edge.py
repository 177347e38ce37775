"""Edge geometries on the GPU: the reference command line's own defaults and
sigma_z at or below one pixel.

  - cli-defaults: lengths in metres as `2d_strip_reconstruction` defaults
    them (options.cpp:28-34): R 0.5, L 1, 0.01 pixels, sigma_z 0.015,
    sigma_dl 0.06 (sigma_z = 1.5 pixels), grid ceil(L/p) x ceil(2R/p) =
    100 x 100 (reconstruction_cmd.cpp option completion);
  - sigma_z 0.75 and 0.4 pixels (2 mm pixels, sigma_z 1.5 / 0.8 mm), where a
    row's in-ellipse run is a few pixels and the exp recurrence's column
    ratios are far from 1.

Every hot-path variant (bbox kernel; row-band kernels with and without the
span table, exp per pixel and by recurrence) must visit exactly the oracle's
support, match its image after 5 iterations (rel-L2 <= 1e-4) and keep the
image finite.  Per-event forward projections are held to 1e-5 relative to
the FP32 oracle: the row-band kernels evaluate the exponent in the vertex
form C (dz - dy tan)^2 + 4 id sec^2 dy^2 (no cancelling terms) and round the
pixel offset dz as the reference does ((column centre - centre) - dy tan,
gaussian_kernel.h:51-79), so where sigma_z is 0.4 pixels the GPU and the
FP32 oracle stay within 1e-6 of each other although both are 2e-5 from the
FP64 oracle; the errors of the GPU and of the FP32 oracle against the FP64
oracle are printed.
"""
import numpy as np
import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu


def _workloads(phantom_mod):
    Ellipse, Workload = phantom_mod.Ellipse, phantom_mod.Workload
    metres = [Ellipse(0.12, -0.045, 0.06, 0.075, -20, 4),
              Ellipse(-0.105, 0.03, 0.09, 0.06, 30, 2),
              Ellipse(0.0, 0.0, 0.3, 0.25, 0, 1)]
    return {
        "cli-defaults": Workload("cli-defaults", 0.5, 1.0, 100, 100, 0.01,
                                 0.015, 0.06, metres, 20_000, 5),
        "sigma-z-0.75px": Workload("sigma-z-0.75px", 350.0, 500.0, 256, 160,
                                   2.0, 1.5, 6.0, phantom_mod.PHANTOM_3,
                                   20_000, 5),
        "sigma-z-0.4px": Workload("sigma-z-0.4px", 350.0, 500.0, 256, 160,
                                  2.0, 0.8, 3.2, phantom_mod.PHANTOM_3,
                                  20_000, 5),
    }


# (algorithm, span table, exp mode); exp mode 1 = the row recurrence
VARIANTS = [(1, False, 0), (2, True, 1), (2, True, 0), (2, False, 1),
            (2, False, 0)]


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("name", ["cli-defaults", "sigma-z-0.75px",
                                  "sigma-z-0.4px"])
def test_edge_geometry_parity(pystrip_mod, phantom_mod, oracle_mod, name,
                              variant):
    algorithm, spans, exp_mode = variant
    print_case("edge geometry", name, "algorithm", algorithm, "spans", spans,
               "exp mode", exp_mode)
    w = _workloads(phantom_mod)[name]
    ev = w.events(seed=51)
    og = oracle_mod.geometry(*w.oracle_geometry_tuple())
    rec = pystrip_mod.Reconstruction(w.geometry())
    rec.set_algorithm(algorithm)
    if algorithm == 2:
        rec.set_span_table(spans)
        rec.set_exp_mode(exp_mode)
    rec.set_events(ev)
    if algorithm == 2:
        assert rec.span_table() == spans
    nb, ne = oracle_mod.support_counts(og, ev)
    assert rec.kernel_evaluations() == int(ne.sum())
    rng = np.random.default_rng(3)
    rho = rng.uniform(0.5, 2.0, size=rec.shape).astype(np.float32)
    rec.rho = rho
    d_gpu = rec.forward().astype(np.float64)
    d_ref = oracle_mod.forward(og, ev, rho).astype(np.float64)
    nz = d_ref > 0
    assert np.array_equal(d_gpu > 0, nz)
    rel = np.abs(d_gpu[nz] - d_ref[nz]) / d_ref[nz]
    d64 = oracle_mod.forward(og, ev.astype(np.float64), rho.astype(np.float64),
                             dtype=np.float64)
    print("max rel D: GPU vs FP64", np.max(np.abs(d_gpu[nz] - d64[nz]) /
                                           d64[nz]),
          "FP32 oracle vs FP64", np.max(np.abs(d_ref[nz] - d64[nz]) /
                                        d64[nz]), flush=True)
    rec.rho = np.ones(rec.shape, np.float32)
    rec.iterate(w.iterations)
    gpu = rec.rho
    ref = oracle_mod.mlem(og, ev, iterations=w.iterations)
    e = oracle_mod.rel_l2(gpu, ref)
    print("mean support", ne.mean(), "max rel D", rel.max(), "rel-L2", e,
          flush=True)
    assert rel.max() < 1e-5
    assert np.all(np.isfinite(gpu))
    assert e <= 1e-4


def test_wide_bbox_falls_back_to_the_bbox_kernel(pystrip_mod, phantom_mod,
                                                 oracle_mod):
    # sigma_z of 500 pixels: a 16-row window of the widest bbox (up to 2784
    # columns) would exceed the 160 KiB of LDS, so the automatic choice is
    # the bbox kernel and the row-band algorithm is refused
    print_case("wide-bbox geometry: automatic fallback")
    w = phantom_mod.Workload("wide", 350.0, 500.0, 3000, 40, 0.1, 50.0,
                             60.0, phantom_mod.PHANTOM_3, 300, 1)
    ev = w.events(seed=53)
    rec = pystrip_mod.Reconstruction(w.geometry())
    assert rec.algorithm == 1
    with pytest.raises(pystrip_mod.StripMLEMError):
        rec.set_algorithm(2)
    rec.set_events(ev)
    rec.iterate(1)
    og = oracle_mod.geometry(*w.oracle_geometry_tuple())
    e = oracle_mod.rel_l2(rec.rho, oracle_mod.mlem(og, ev, iterations=1))
    print("rel-L2", e)
    assert e <= 1e-4
