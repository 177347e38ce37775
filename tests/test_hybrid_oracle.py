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


def _geometry_file(hybrid_mod, path, geo, order=None, extra=()):
    """geo's LOR records in the given slot order (default: all, by index),
    then the extra records (first, second, seg, width, infos)."""
    n = geo.n_detectors
    head = np.zeros(1, hybrid_mod._GRID_HEADER)
    head["n_columns"], head["n_rows"] = geo.n_columns, geo.n_rows
    head["pixel_size"] = np.float32(geo.pixel_size)
    head["llx"] = -np.float32(geo.pixel_size) * geo.n_columns / 2
    head["lly"] = -np.float32(geo.pixel_size) * geo.n_rows / 2
    first = np.repeat(np.arange(n), np.arange(1, n + 1))
    second = np.concatenate([np.arange(f + 1) for f in range(n)])
    recs = []
    for li in (range(geo.n_lors) if order is None else order):
        b, e = geo.info_begin[li], geo.info_end[li]
        infos = np.zeros(e - b, hybrid_mod._PIXEL_INFO)
        infos["x"], infos["y"] = geo.pixels[b:e, 0], geo.pixels[b:e, 1]
        infos["t"], infos["weight"] = geo.positions[b:e], geo.weights[0, b:e]
        recs.append((first[li], second[li], geo.seg[li], geo.widths[li],
                     infos))
    with open(path, "wb") as out:
        out.write(b"PETg" + np.int16(n).tobytes() + head.tobytes())
        for f, s, sg, w, infos in list(recs) + list(extra):
            h = np.zeros(1, hybrid_mod._LOR_HEADER)
            h["first"], h["second"], h["seg"], h["width"] = f, s, sg, w
            h["n_pixels"] = len(infos)
            out.write(h.tobytes() + infos.tobytes())


def test_geometry_file_round_trip(hybrid_mod, ring, tmp_path):
    """write_geometry then read_geometry gives back the arrays (the padding
    info, in no LOR, is not written, and the reader appends it), and
    writing what was read gives the same bytes."""
    geo, p = ring, tmp_path / "ring.geom"
    hybrid_mod.write_geometry(geo, p)
    back = hybrid_mod.read_geometry(p, geo.n_planes, geo.z_left)
    n = geo.n_infos - 1
    assert back.n_detectors == geo.n_detectors and back.n_infos == n + 1
    assert (back.n_columns, back.n_rows, back.n_planes) == \
        (geo.n_columns, geo.n_rows, geo.n_planes)
    assert back.pixel_size == np.float32(geo.pixel_size)
    assert abs(back.ring_radius - geo.ring_radius) < 1e-6
    for name in ("seg", "widths", "info_begin", "info_end"):
        assert np.array_equal(getattr(back, name), getattr(geo, name)), name
    assert np.array_equal(back.pixels, geo.pixels)
    assert np.array_equal(back.positions, geo.positions)
    assert np.array_equal(back.weights, geo.weights)
    # 4 + 2 + 16 bytes of header, 32 per LOR record, 20 per info
    size = 22 + 32 * geo.n_lors + 20 * n
    assert p.stat().st_size == size
    q = tmp_path / "again.geom"
    hybrid_mod.write_geometry(back, q)
    assert p.read_bytes() == q.read_bytes()


def test_geometry_file_refusals(hybrid_mod, ring, tmp_path):
    p = tmp_path / "bad.geom"
    hybrid_mod.write_geometry(ring, p)
    raw = bytearray(p.read_bytes())
    raw[3] = ord("G")  # the double-precision magic
    p.write_bytes(bytes(raw))
    with pytest.raises(ValueError, match="not a binary float"):
        hybrid_mod.read_geometry(p, 10)
    raw[3] = ord("g")
    raw[14:18] = np.float32(0.0).tobytes()  # lower-left x at 0: off-centre
    p.write_bytes(bytes(raw))
    with pytest.raises(ValueError, match="not centred"):
        hybrid_mod.read_geometry(p, 10)


def test_geometry_file_against_reference_reader(oracle_mod, hybrid_mod, ring,
                                                tmp_path):
    """read_geometry against the reference's own Geometry(ibstream) +
    GeometrySOA(geometry, 1) (oracle/_ref) on a file whose records come in
    shuffled order, with a later record replacing an earlier one for the
    same LOR, a missing slot (which the flattening turns into an empty
    LOR 0) and a last record cut short (which ends the stream)."""
    if not oracle_mod.ref_available():
        pytest.skip("oracle/_ref not built")
    geo = ring
    rng = np.random.default_rng(3)
    order = rng.permutation(geo.n_lors)
    order = order[order != 700]
    li = int(order[5])
    f = int((np.sqrt(8 * li + 1) - 1) // 2)
    s = li - f * (f + 1) // 2
    infos = np.zeros(3, hybrid_mod._PIXEL_INFO)
    infos["x"], infos["y"] = [1, 2, 3], [4, 5, 6]
    infos["t"], infos["weight"] = [0.1, 0.2, 0.3], [0.5, 0.25, 0.125]
    seg = np.array([0.1, 0.2, 0.3, 0.4], np.float32)
    short = (f, s, seg, np.float32(0.02), infos)
    p = tmp_path / "shuffled.geom"
    _geometry_file(hybrid_mod, p, geo, order,
                   extra=[(f, s, seg, np.float32(0.03), infos), short])
    raw = p.read_bytes()
    p.write_bytes(raw[:-10])  # the last record loses part of its infos
    mine = hybrid_mod.read_geometry(p, geo.n_planes)
    ref = oracle_mod.ref_read_geometry(p)
    assert ref["n_detectors"] == mine.n_detectors
    assert (ref["n_columns"], ref["n_rows"]) == (mine.n_columns, mine.n_rows)
    assert ref["grid"][0] == np.float32(mine.pixel_size)
    n = len(ref["positions"])
    assert mine.n_infos == n + 1 and mine.weights[0, n] == 0
    for name in ("pixels", "positions"):
        assert np.array_equal(ref[name], getattr(mine, name)[:n]), name
    for name in ("info_begin", "info_end"):
        assert np.array_equal(ref[name], getattr(mine, name)), name
    # the reference leaves a missing slot's segment and width unset
    kept = np.arange(geo.n_lors) != 700
    for name in ("seg", "widths"):
        assert np.array_equal(ref[name][kept], getattr(mine, name)[kept])
    assert mine.info_begin[700] == mine.info_end[700] == 0
    assert np.array_equal(ref["weights"], mine.weights[0, :n])
    # the replaced LOR holds the later record, the missing slot emptied LOR 0
    assert mine.widths[li] == np.float32(0.03)
    assert mine.info_end[li] - mine.info_begin[li] == 3
    assert mine.info_end[0] == mine.info_begin[0] and mine.widths[0] == 0
    assert mine.n_infos == geo.n_infos - (geo.info_end[li] -
                                             geo.info_begin[li]) - \
        (geo.info_end[700] - geo.info_begin[700]) + 3
