"""The oracle's 3D hybrid reconstruction (oracle/hybrid_oracle_impl.inc)
against the reference's own PET3D::Hybrid::Reconstruction built from its
sources (oracle/_ref, src/3d/hybrid/reconstruction.h:32-549), and against
what the mathematics fixes.

Geometry and events come from pystrip.hybrid: a 48-detector ring, an 80 x 80
pixel grid of 5 mm, 80 planes, LOR tubes 2 cm wide.
"""
import numpy as np
import pytest

from conftest import print_case


@pytest.fixture(scope="module")
def hybrid_mod(pystrip_mod):
    from pystrip import hybrid
    return hybrid


@pytest.fixture(scope="module")
def ring(hybrid_mod):
    return hybrid_mod.ring_geometry()


def test_geometry_layout(hybrid_mod, ring):
    geo = ring
    n = geo.n_detectors
    assert geo.n_lors == n * (n + 1) // 2
    assert np.all(geo.info_begin <= geo.info_end)
    assert np.all(geo.info_begin[1:] == geo.info_end[:-1])
    # the padding info follows the last LOR's infos
    assert geo.info_end[-1] == geo.n_infos - 1
    assert geo.weights[0, -1] == 0
    for li in range(0, geo.n_lors, 97):
        p = geo.positions[geo.info_begin[li]:geo.info_end[li]]
        assert np.all(np.diff(p) >= 0)
        assert np.all((p > 0) & (p < 1))
    ev = hybrid_mod.hybrid_events(geo, 5000, seed=1)
    assert np.all(ev[:, 0] > ev[:, 1])
    assert np.all(np.abs(ev[:, 2:4]) < geo.length / 2 + 5 * geo.sigma_z)


@pytest.mark.parametrize("weights_mode,sens_mode", [(1, 0), (0, 0), (1, 1)])
def test_against_reference_build(oracle_mod, hybrid_mod, ring, weights_mode,
                                  sens_mode):
    if not oracle_mod.ref_available():
        pytest.skip("oracle/_ref not built (reference tree absent)")
    print_case("hybrid oracle vs reference build", weights_mode, sens_mode)
    ev = hybrid_mod.hybrid_events(ring, 20000, seed=3)
    a, na, sa, wa = oracle_mod.hybrid_mlem(ring, ev, iterations=3,
                                           weights_mode=weights_mode,
                                           sens_mode=sens_mode)
    b, nb, sb, wb = oracle_mod.ref_hybrid_mlem(ring, ev, iterations=3,
                                               weights_mode=weights_mode,
                                               sens_mode=sens_mode)
    assert na == nb
    # the reference build contracts a*b + c into FMA (g++ default), the
    # oracle does not; the pixel's distance from the LOR's line,
    # c . n - distance_from_origin, cancels, so single weights differ by up
    # to ~1e-5 relative and their sums per pixel by ~1e-6
    assert np.max(np.abs(sa - sb[0]) / np.abs(sb[0]).max()) < 5e-6
    nz = wb[0] > 0
    assert np.array_equal(wa > 0, nz)
    assert np.max(np.abs(wa - wb[0])[nz] / wb[0][nz]) < 5e-5
    e = oracle_mod.rel_l2(a, b)
    print("rel-L2", e)
    assert e < 2e-6
    # the FP64 flavour agrees with both to FP32 rounding
    c, nc, _, _ = oracle_mod.hybrid_mlem(ring, ev, iterations=3,
                                         weights_mode=weights_mode,
                                         sens_mode=sens_mode,
                                         dtype=np.float64)
    assert nc == na
    assert oracle_mod.rel_l2(c, b) < 1e-5


def test_external_sensitivity(oracle_mod, hybrid_mod, ring):
    if not oracle_mod.ref_available():
        pytest.skip("oracle/_ref not built (reference tree absent)")
    ev = hybrid_mod.hybrid_events(ring, 5000, seed=4)
    rng = np.random.default_rng(2)
    sens = rng.uniform(0.5, 2.0, size=(ring.n_rows, ring.n_columns)
                       ).astype(np.float32)
    a, na, sa, _ = oracle_mod.hybrid_mlem(ring, ev, iterations=2,
                                          sens_mode=2, sens_in=sens)
    b, nb, sb, _ = oracle_mod.ref_hybrid_mlem(ring, ev, iterations=2,
                                              sens_mode=2, sens_in=sens)
    assert na == nb
    assert np.array_equal(sa, sens) and np.array_equal(sb[0], sens)
    assert oracle_mod.rel_l2(a, b) < 2e-6


def test_calculate_weight_and_sensitivity(oracle_mod, ring):
    # calculate_weight: gauss_norm_w exp(-d^2 / (2 sigma_w^2)) with
    # sigma_w = 0.3 width, d the pixel centre's distance from the LOR's
    # line; sensitivity: the sum of each pixel's weights; the normalised
    # weights divide by it (reconstruction.h:218-271)
    geo = ring
    _, _, sens, w = oracle_mod.hybrid_mlem(geo, np.zeros((0, 5), np.float32),
                                           iterations=0, dtype=np.float64)
    ps = geo.pixel_size
    cx = -geo.n_columns * ps / 2 + ps / 2 + geo.pixels[:, 0] * ps
    cy = -geo.n_rows * ps / 2 + ps / 2 + geo.pixels[:, 1] * ps
    lor = np.repeat(np.arange(geo.n_lors), geo.info_end - geo.info_begin)
    n = len(lor)
    s = geo.seg[lor].astype(np.float64)
    d = np.stack([s[:, 2] - s[:, 0], s[:, 3] - s[:, 1]], 1)
    d /= np.linalg.norm(d, axis=1)[:, None]
    dist = np.abs((cx[:n] - s[:, 0]) * d[:, 1] - (cy[:n] - s[:, 1]) * d[:, 0])
    sw = 0.3 * geo.widths[lor].astype(np.float64)
    raw = np.exp(-dist ** 2 / (2 * sw ** 2)) / (sw * np.sqrt(2 * np.pi))
    expect_sens = np.zeros(geo.n_rows * geo.n_columns)
    pix = geo.pixels[:n, 1].astype(np.int64) * geo.n_columns + geo.pixels[:n, 0]
    np.add.at(expect_sens, pix, raw)
    assert np.allclose(sens.ravel(), expect_sens, rtol=1e-5, atol=0)
    assert np.allclose(w[:n], raw / expect_sens[pix], rtol=1e-5, atol=0)


def test_update_preserves_event_count(oracle_mod, hybrid_mod, ring):
    # from rho = 1, one update gives sum_v rho'[v] s[pixel(v)] = number of
    # events with a positive denominator: each event spreads 1 over its
    # voxels in proportion to weight x sensitivity
    ev = hybrid_mod.hybrid_events(ring, 20000, seed=5)
    r, n, s, _ = oracle_mod.hybrid_mlem(ring, ev, iterations=1,
                                        dtype=np.float64)
    mass = (r * s[None]).sum()
    assert abs(mass - n) <= 1e-9 * n
    r2, _, _, _ = oracle_mod.hybrid_mlem(ring, ev, iterations=2,
                                         dtype=np.float64)
    assert abs((r2 * s[None]).sum() - n) <= 1e-9 * n


def test_plane_and_info_ranges_follow_the_box(oracle_mod, hybrid_mod, ring):
    # bb_intersects_grid_with_positive_weight (reconstruction.h:151-162):
    # an event whose box lies beyond the axial grid is dropped, one inside
    # is kept
    geo = ring
    inside = np.array([[30, 6, 0.01, -0.01, 0.0]], np.float32)
    beyond = np.array([[30, 6, 0.45, 0.45, 0.0]], np.float32)
    _, n_in, _, _ = oracle_mod.hybrid_mlem(geo, inside, iterations=1)
    _, n_out, _, _ = oracle_mod.hybrid_mlem(geo, beyond, iterations=1)
    assert (n_in, n_out) == (1, 0)


def test_repeated_visit_follows_the_kernel_cache(oracle_mod, hybrid_mod,
                                                  ring):
    # an event visits the pixel of its LOR's last info twice (the extra info
    # past the box is the next LOR's first and repeats it): the reference's
    # CPU path adds the voxel's last cached weight on both visits
    from hybrid_cases import repeated_pixel_events, repeated_pixel_geometry
    geo, li, first, second = repeated_pixel_geometry(ring)
    ev = repeated_pixel_events(geo, li, first, second, 200, seed=7)
    ints, _, keep = oracle_mod.hybrid_frames(
        geo, ev, oracle_mod.hybrid_mlem(geo, ev[:0], iterations=0)[3])
    assert keep.all()
    assert np.all(ints[:, 2] == geo.info_end[li] + 1)
    assert np.all(ints[:, 1] < geo.info_end[li])
    a, na, _, w = oracle_mod.hybrid_mlem(geo, ev, iterations=1)
    b, nb, _, _ = oracle_mod.hybrid_mlem(geo, ev, iterations=1, legacy=True)
    assert na == nb == len(ev)
    if oracle_mod.ref_available():
        r, nr, _, _ = oracle_mod.ref_hybrid_mlem(geo, ev, iterations=1)
        assert nr == na
        assert oracle_mod.rel_l2(a, r) < 2e-6
        assert oracle_mod.rel_l2(b, r) > 1e-3
    # the two updates differ only on the repeated pixel's voxels, where the
    # cached update is the legacy one with the first visit's weight
    # replaced by the extra info's: a / b = 2 w_x / (w_j + w_x)
    x, y = geo.pixels[geo.info_end[li]]
    j = geo.info_end[li] - 1
    diff = np.argwhere(np.abs(a - b) > 1e-6 * np.abs(b).max())
    assert len(diff) and np.all(diff[:, 1] == y) and np.all(diff[:, 2] == x)
    wx, wj = float(w[j + 1]), float(w[j])
    ratio = a[:, y, x][b[:, y, x] > 0] / b[:, y, x][b[:, y, x] > 0]
    assert np.allclose(ratio, 2 * wx / (wj + wx), rtol=1e-5)


def test_legacy_update_equals_the_cached_one_without_repeats(
        oracle_mod, hybrid_mod, ring):
    ev = hybrid_mod.hybrid_events(ring, 5000, seed=8)
    a = oracle_mod.hybrid_mlem(ring, ev, iterations=2)[0]
    b = oracle_mod.hybrid_mlem(ring, ev, iterations=2, legacy=True)[0]
    assert oracle_mod.rel_l2(a, b) < 1e-6
