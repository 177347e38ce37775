"""3D hybrid reconstruction on the GPU (include/hybrid_mlem.h) against the
oracle's restatement (oracle/hybrid_oracle_impl.inc) and, where it was
built, the reference's own PET3D::Hybrid::Reconstruction (oracle/_ref).

Geometry: a 48-detector ring, 80 x 80 pixels of 5 mm, 80 planes, LOR tubes
2 cm wide (pystrip.hybrid.ring_geometry).  Bounds: weights and sensitivity
within 1e-5 / 1e-6 of the oracle (single expf of the weight and the same
per-pixel summation order); frame events bit-exact (integer box and the
FP32 frame parameters, evaluated without contraction on both sides); images
rel-L2 <= 1e-4 (the GPU sums each event's denominator over a wavefront and
adds its updates with float atomics, in another order than the reference's
sequential loop).
"""
import ctypes
import os

import numpy as np
import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hybrid_mod(pystrip_mod):
    from pystrip import hybrid
    return hybrid


@pytest.fixture(scope="module")
def ring(hybrid_mod):
    return hybrid_mod.ring_geometry()


def _rec(hybrid_mod, geo, semantics=0, calculate_weight=True, sens_mode=0,
         sens=None, algorithm=0):
    rec = hybrid_mod.HybridReconstruction(geo, semantics)
    rec.prepare_weights(calculate_weight, sens_mode, sens)
    rec.set_algorithm(algorithm)
    return rec


@pytest.mark.parametrize("weights_mode,sens_mode",
                         [(1, 0), (0, 0), (1, 1), (0, 2)])
def test_weights_and_sensitivity(oracle_mod, hybrid_mod, ring, weights_mode,
                                 sens_mode):
    print_case("hybrid weights/sensitivity", weights_mode, sens_mode)
    sens_in = None
    if sens_mode == 2:
        sens_in = np.random.default_rng(4).uniform(
            0.5, 2.0, (ring.n_rows, ring.n_columns)).astype(np.float32)
    rec = _rec(hybrid_mod, ring, 0, weights_mode, sens_mode, sens_in)
    _, _, s_ref, w_ref = oracle_mod.hybrid_mlem(
        ring, np.zeros((0, 5), np.float32), iterations=0,
        weights_mode=weights_mode, sens_mode=sens_mode, sens_in=sens_in)
    s, w = rec.sensitivity, rec.weights
    print("sensitivity max rel", np.max(np.abs(s - s_ref)) / s_ref.max(),
          flush=True)
    assert np.max(np.abs(s - s_ref)) <= 1e-6 * s_ref.max()
    nz = w_ref != 0
    assert np.array_equal(w != 0, nz)
    assert np.max(np.abs(w - w_ref)[nz] / np.abs(w_ref[nz])) <= 1e-5
    rec.close()


def test_frame_events_bit_exact(oracle_mod, hybrid_mod, ring):
    print_case("hybrid frame events")
    ev = hybrid_mod.hybrid_events(ring, 20000, seed=11)
    # boxes beyond the axial grid are filtered out; a box far before the
    # LOR's first info still holds that info (upper_bound + 1 of a box
    # below every position is the first info) and is kept
    ev[:500, 2:4] += 0.6
    ev[500:1000, 4] = 3.0
    rec = _rec(hybrid_mod, ring)
    rec.add_responses(ev)
    ints, floats = rec.frame_events()
    o_ints, o_floats, keep = oracle_mod.hybrid_frames(ring, ev, rec.weights)
    print("kept", keep.sum(), "of", len(ev), flush=True)
    assert rec.n_events == int(keep.sum())
    assert not keep[:500].any()
    far = o_ints[500:1000]
    assert np.all(far[:, 2] - far[:, 1] == 1)
    assert np.array_equal(far[:, 1], ring.info_begin[far[:, 0]])
    k = o_ints[keep]
    assert np.array_equal(ints[:, 0], k[:, 0])
    assert np.array_equal(ints[:, 1], k[:, 1])
    assert np.array_equal(ints[:, 2], np.minimum(k[:, 2], ring.n_infos))
    assert np.array_equal(ints[:, 3] & 0xffff, k[:, 3])
    assert np.array_equal(ints[:, 3] >> 16, k[:, 4])
    assert np.array_equal(floats, o_floats[keep])
    assert rec.voxel_visits() == int(((k[:, 2] - k[:, 1]) *
                                      (k[:, 4] - k[:, 3])).sum())
    # appending keeps the order of the kept events
    rec.add_responses(ev[:3000])
    assert rec.n_events == int(keep.sum() + keep[:3000].sum())
    ints2, _ = rec.frame_events()
    assert np.array_equal(ints2[len(ints):, :3],
                          o_ints[:3000][keep[:3000]][:, :3])
    rec.close()


@pytest.mark.parametrize("algorithm", [1, 2])
@pytest.mark.parametrize("weights_mode,sens_mode", [(1, 0), (0, 1)])
def test_mlem_parity(oracle_mod, hybrid_mod, ring, weights_mode, sens_mode,
                     algorithm):
    print_case("hybrid MLEM", weights_mode, sens_mode, "algorithm",
               algorithm)
    ev = hybrid_mod.hybrid_events(ring, 50000, seed=12)
    rec = _rec(hybrid_mod, ring, 0, weights_mode, sens_mode,
               algorithm=algorithm)
    rec.add_responses(ev)
    rec.iterate(3)
    assert rec.algorithm == algorithm
    gpu = rec.rho
    ref, n_used, _, _ = oracle_mod.hybrid_mlem(
        ring, ev, iterations=3, weights_mode=weights_mode,
        sens_mode=sens_mode)
    assert rec.n_events == n_used
    e = oracle_mod.rel_l2(gpu, ref)
    msg = f"rel-L2 vs oracle {e:.3e}"
    if oracle_mod.ref_available():
        r, nr, _, _ = oracle_mod.ref_hybrid_mlem(
            ring, ev, iterations=3, weights_mode=weights_mode,
            sens_mode=sens_mode)
        assert nr == n_used
        er = oracle_mod.rel_l2(gpu, r)
        msg += f", vs reference build {er:.3e}"
        assert er <= 1e-4
    print(msg, "event kernel ms", rec.last_iteration_ms(), flush=True)
    assert e <= 1e-4
    assert np.all(np.isfinite(gpu))
    rec.close()


@pytest.mark.parametrize("algorithm", [1, 2])
def test_no_events_and_empty_lors(oracle_mod, hybrid_mod, ring, algorithm):
    """Edge inputs: no events at all (the update gives a zero image, as the
    reference's rho * 0 / sensitivity does), then responses on every LOR
    with no pixel infos (first > second), which keep or drop the event by
    the info after the range, the next LOR's first (the upper_bound + 1 of
    the reference's event reader): event count and image against the
    oracle and the reference build."""
    print_case("hybrid edge events", "algorithm", algorithm)
    rec = _rec(hybrid_mod, ring, algorithm=algorithm)
    rec.add_responses(np.zeros((0, 5), np.float32))
    rec.iterate(2)
    assert rec.n_events == 0
    img = rec.rho
    assert np.all(img == 0.0)
    ref0, n0, _, _ = oracle_mod.hybrid_mlem(
        ring, np.zeros((0, 5), np.float32), iterations=2)
    assert n0 == 0 and np.all(ref0 == 0.0)
    rec.close()
    li = np.nonzero(ring.info_end == ring.info_begin)[0]
    f = ((np.sqrt(8 * li + 1) - 1) // 2).astype(np.int64)
    s = li - f * (f + 1) // 2
    sel = f > s
    f, s = f[sel], s[sel]
    rng = np.random.default_rng(2)
    ev = np.stack([f, s, rng.uniform(-0.1, 0.1, len(f)),
                   rng.uniform(-0.1, 0.1, len(f)),
                   rng.uniform(-0.1, 0.1, len(f))], 1).astype(np.float32)
    rec = _rec(hybrid_mod, ring, algorithm=algorithm)
    rec.add_responses(ev)
    rec.iterate(2)
    gpu = rec.rho
    ref, n_used, _, _ = oracle_mod.hybrid_mlem(ring, ev, iterations=2)
    assert 0 < n_used < len(ev)
    assert rec.n_events == n_used
    e = oracle_mod.rel_l2(gpu, ref)
    msg = f"{len(ev)} events on empty LORs, {n_used} kept, rel-L2 {e:.3e}"
    if oracle_mod.ref_available():
        r, nr, _, _ = oracle_mod.ref_hybrid_mlem(ring, ev, iterations=2)
        assert nr == n_used
        assert oracle_mod.rel_l2(gpu, r) <= 1e-4
    print(msg, flush=True)
    assert e <= 1e-4 and np.all(np.isfinite(gpu))
    rec.close()


@pytest.mark.parametrize("algorithm", [1, 2])
def test_geometry_from_file(oracle_mod, hybrid_mod, ring, algorithm,
                            tmp_path):
    """The reconstruction of a geometry read from a file in the reference's
    binary barrel geometry format (write_geometry, read_geometry) matches
    the oracle on the arrays read, and the file holds the same geometry."""
    print_case("hybrid geometry file", "algorithm", algorithm)
    path = tmp_path / "ring.geom"
    hybrid_mod.write_geometry(ring, path)
    geo = hybrid_mod.read_geometry(path, ring.n_planes, ring.z_left,
                                   ring.sigma_z, ring.sigma_dl, ring.length)
    assert np.array_equal(geo.positions, ring.positions)
    assert np.array_equal(geo.info_end, ring.info_end)
    ev = hybrid_mod.hybrid_events(geo, 30000, seed=21)
    rec = _rec(hybrid_mod, geo, algorithm=algorithm)
    rec.add_responses(ev)
    rec.iterate(2)
    gpu = rec.rho
    ref, n_used, _, _ = oracle_mod.hybrid_mlem(geo, ev, iterations=2)
    assert rec.n_events == n_used
    e = oracle_mod.rel_l2(gpu, ref)
    print(f"rel-L2 vs oracle {e:.3e}", flush=True)
    assert e <= 1e-4 and np.all(np.isfinite(gpu))
    rec.close()


@pytest.mark.parametrize("algorithm", [1, 2])
@pytest.mark.parametrize("semantics", [0, 1])
def test_repeated_visit(oracle_mod, hybrid_mod, ring, semantics, algorithm):
    from hybrid_cases import repeated_pixel_events, repeated_pixel_geometry
    print_case("hybrid repeated voxel visit, semantics", semantics,
               "algorithm", algorithm)
    geo, li, first, second = repeated_pixel_geometry(ring)
    ev = np.concatenate([
        repeated_pixel_events(geo, li, first, second, 300, seed=5),
        hybrid_mod.hybrid_events(geo, 3000, seed=6)])
    rec = _rec(hybrid_mod, geo, semantics, algorithm=algorithm)
    rec.add_responses(ev)
    rec.iterate(2)
    assert rec.algorithm == algorithm
    gpu = rec.rho
    ref = oracle_mod.hybrid_mlem(geo, ev, iterations=2,
                                 legacy=bool(semantics))[0]
    other = oracle_mod.hybrid_mlem(geo, ev, iterations=2,
                                   legacy=not semantics)[0]
    e, e_other = oracle_mod.rel_l2(gpu, ref), oracle_mod.rel_l2(gpu, other)
    print("rel-L2", e, "against the other semantics", e_other, flush=True)
    assert e <= 1e-5
    assert e_other > 1e-3
    rec.close()


@pytest.mark.parametrize("algorithm", [1, 2])
def test_start_image_and_layout(oracle_mod, pystrip_mod, hybrid_mod, ring,
                                algorithm):
    print_case("hybrid start image, algorithm", algorithm)
    rng = np.random.default_rng(9)
    rho0 = rng.uniform(0.5, 2.0, ring.shape).astype(np.float32)
    ev = hybrid_mod.hybrid_events(ring, 20000, seed=13)
    rec = _rec(hybrid_mod, ring, algorithm=algorithm)
    rec.rho = rho0
    assert np.array_equal(rec.rho, rho0)
    rec.add_responses(ev)
    rec.iterate(2)
    ref = oracle_mod.hybrid_mlem(ring, ev, iterations=2, rho0=rho0)[0]
    assert oracle_mod.rel_l2(rec.rho, ref) <= 1e-4
    # a LOR index outside the geometry is refused, the events held stay
    n = rec.n_events
    bad = ev[:10].copy()
    bad[3, 0] = 400
    with pytest.raises(pystrip_mod.StripMLEMError):
        rec.add_responses(bad)
    assert rec.n_events == n
    rec.close()


@pytest.mark.parametrize("algorithm", [1, 2])
def test_update_preserves_event_count_at_scale(hybrid_mod, ring, algorithm):
    # from rho = 1, one update spreads each event's unit over its voxels in
    # proportion to weight x sensitivity: sum_v rho'(v) s(pixel(v)) = number
    # of events (none of these repeats a voxel), at 2e6 events
    print_case("hybrid mass identity, 2e6 events, algorithm", algorithm)
    ev = hybrid_mod.hybrid_events(ring, 2_000_000, seed=14)
    rec = _rec(hybrid_mod, ring, algorithm=algorithm)
    rec.add_responses(ev)
    rec.iterate(1)
    assert rec.algorithm == algorithm
    mass = (rec.rho.astype(np.float64) *
            rec.sensitivity.astype(np.float64)[None]).sum()
    print("events", rec.n_events, "mass", mass, "visits", rec.voxel_visits(),
          "event kernel ms", rec.last_iteration_ms(), flush=True)
    assert abs(mass - rec.n_events) <= 1e-4 * rec.n_events
    rec.close()


def test_full_ring_workload(oracle_mod, hybrid_mod):
    # the bench workload's geometry (384 detectors, 9.2e6 pixel infos,
    # 160 x 160 x 100 voxels) with 2e5 of its events: both algorithms
    # against the oracle
    w = hybrid_mod.WORKLOADS["hybrid-ring384-160x160x100-5mm-1e7"]
    print_case("hybrid full ring", w.name, "2e5 events")
    geo = w.geometry()
    ev = w.events(geo, 200_000, seed=15)
    ref, n_used, _, _ = oracle_mod.hybrid_mlem(geo, ev, iterations=2)
    images = {}
    for algorithm in (1, 2):
        rec = _rec(hybrid_mod, geo, algorithm=algorithm)
        rec.add_responses(ev)
        rec.iterate(2)
        assert rec.algorithm == algorithm and rec.n_events == n_used
        images[algorithm] = rec.rho
        e = oracle_mod.rel_l2(images[algorithm], ref)
        print("algorithm", algorithm, "rel-L2", e, "event kernels ms",
              rec.last_iteration_ms(), flush=True)
        assert e <= 1e-4
        rec.close()
    assert oracle_mod.rel_l2(images[1], images[2]) <= 1e-4


@pytest.mark.parametrize("algorithm", [1, 2])
def test_frame_events_from_the_reader(oracle_mod, pystrip_mod, hybrid_mod,
                                      ring, algorithm):
    # the reference's GPU run receives the normalised weights, the
    # sensitivity and the frame events of its CPU reader
    # (reconstruction_cmd.cpp:596-600): sensitivity mode 3 and
    # hybrid_mlem_set_frame_events, as include/hybrid_gpu_reconstruction.hpp
    # drives them
    from dataclasses import replace
    print_case("hybrid frame events from the reader, algorithm", algorithm)
    ev = hybrid_mod.hybrid_events(ring, 30000, seed=16)
    _, _, sens, w = oracle_mod.hybrid_mlem(ring, ev[:0], iterations=0)
    ints, floats, keep = oracle_mod.hybrid_frames(ring, ev, w)
    frames = np.zeros(int(keep.sum()), hybrid_mod.FRAME_EVENT_DTYPE)
    frames["first"] = ev[keep, 0]
    frames["second"] = ev[keep, 1]
    for c, name in enumerate(["up", "right", "tan", "sec"]):
        frames[name] = floats[keep, c]
    frames["info_begin"] = ints[keep, 1]
    frames["info_end"] = ints[keep, 2]
    frames["plane_begin"] = ints[keep, 3]
    frames["plane_end"] = ints[keep, 4]
    geo = replace(ring, weights=w[None].astype(np.float32))
    rec = hybrid_mod.HybridReconstruction(geo)
    rec.prepare_weights(False, 3, sens)
    rec.set_algorithm(algorithm)
    assert np.array_equal(rec.weights, geo.weights[0])
    assert np.array_equal(rec.sensitivity, sens.astype(np.float32))
    rec.set_frame_events(frames)
    assert rec.n_events == len(frames)
    assert "gfx950" in rec.device_name()
    rec.iterate(3)
    assert rec.algorithm == algorithm
    ref = oracle_mod.hybrid_mlem(ring, ev, iterations=3)[0]
    e = oracle_mod.rel_l2(rec.rho, ref)
    print("rel-L2", e, flush=True)
    assert e <= 1e-4
    bad = frames[:4].copy()
    bad["plane_end"][2] = ring.n_planes + 1
    with pytest.raises(pystrip_mod.StripMLEMError):
        rec.set_frame_events(bad)
    rec.close()


@pytest.mark.parametrize("semantics", [0, 1])
def test_lor_space_in_plane_chunks(oracle_mod, hybrid_mod, semantics):
    # 200 planes: the ring's largest LOR (327 infos) needs 327 x 201 floats,
    # more than a workgroup's LDS, so the LOR-space iteration holds the
    # slabs in chunks of planes, summing each event's D_e over the chunks;
    # against the oracle and the wavefront-per-event algorithm
    geo = hybrid_mod.ring_geometry(n_planes=200)
    print_case("hybrid plane chunks", "200 planes", "semantics", semantics)
    ev = hybrid_mod.hybrid_events(geo, 20000, seed=31)
    legacy = semantics == 1
    ref = oracle_mod.hybrid_mlem(geo, ev, iterations=3, legacy=legacy)[0]
    images = {}
    for algorithm in (0, 1):
        rec = _rec(hybrid_mod, geo, semantics, algorithm=algorithm)
        rec.add_responses(ev)
        rec.iterate(3)
        assert rec.algorithm == (2 if algorithm == 0 else 1)
        images[rec.algorithm] = rec.rho
        e = oracle_mod.rel_l2(rec.rho, ref)
        print("algorithm", rec.algorithm, "rel-L2", e, flush=True)
        assert e <= 1e-4
        rec.close()
    assert oracle_mod.rel_l2(images[1], images[2]) <= 1e-4


def test_tables_beyond_2_32_cells(oracle_mod, hybrid_mod):
    # the full ring with 480 planes of 5 mm: 9.2e6 infos x 480 planes = 4.4e9
    # cells per LOR-space table, more than one launch has work-items (a
    # dispatch counts them in 32 bits); LOR space against the
    # wavefront-per-event algorithm on 2e5 events
    w = hybrid_mod.WORKLOADS["hybrid-ring384-160x160x100-5mm-1e7"]
    print_case("hybrid tables beyond 2^32 cells", "480 planes")
    geo = hybrid_mod.ring_geometry(w.n_detectors, w.ring_radius, w.lor_width,
                                   w.n_pixels, w.pixel_size, 480,
                                   w.fov_radius, w.sigma_z, w.sigma_dl,
                                   w.length)
    assert geo.n_infos * 480 > 2 ** 32
    ev = w.events(geo, 200_000, seed=17)
    images = {}
    for algorithm in (1, 2):
        rec = _rec(hybrid_mod, geo, algorithm=algorithm)
        rec.add_responses(ev)
        rec.iterate(2)
        assert rec.algorithm == algorithm
        images[algorithm] = rec.rho
        rec.close()
    e = oracle_mod.rel_l2(images[2], images[1])
    print("rel-L2 LOR space vs wavefront per event", e, flush=True)
    assert e <= 1e-4


DROPIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(
    __file__))), "tests", "dropin", "_build", "libhybrid_dropin.so")


@pytest.mark.skipif(not os.path.exists(DROPIN),
                    reason="drop-in run harness not built (it is compiled "
                           "from the reference's headers at build time)")
@pytest.mark.parametrize("semantics", [0, 1])
def test_dropin_run(oracle_mod, hybrid_mod, ring, semantics):
    # the reference's own preparation and event reader, then the C++ drop-in
    # run() (include/hybrid_gpu_reconstruction.hpp) as the command line's
    # --gpu branch calls it: 2 blocks of 3 iterations, so the run outputs
    # iterations 1, 2 (early outputs of block 0), 3 and 6 (block ends)
    print_case("hybrid drop-in run", "semantics", semantics)
    ev = hybrid_mod.hybrid_events(ring, 20000, seed=21)
    lib = ctypes.CDLL(DROPIN)
    P = ctypes.c_void_p
    f = ctypes.c_float
    i64 = ctypes.c_int64
    fn = lib.dropin_hybrid_run
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_int, P, P, i64, P, P, P, P, P, ctypes.c_int,
                   ctypes.c_int, f, ctypes.c_int, f, f, f, f, P, i64,
                   ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, P,
                   ctypes.c_int, P, P, P, P, P, ctypes.c_int, P,
                   ctypes.c_int]
    seg, widths, pixels, positions, weights, begin, end = \
        oracle_mod._geometry_args(ring)
    evf = np.ascontiguousarray(ev, np.float32)
    rho = np.ones(ring.shape, np.float32)
    nv = rho.size
    max_out = 8
    iters = np.zeros(max_out, np.int32)
    outs = np.zeros((max_out, nv), np.float32)
    n_out = ctypes.c_int()
    n_frames = ctypes.c_int64()
    name = ctypes.create_string_buffer(256)
    err = ctypes.create_string_buffer(512)
    rc = fn(ring.n_detectors, seg.ctypes.data, widths.ctypes.data,
            ring.n_infos, pixels.ctypes.data, positions.ctypes.data,
            weights.ctypes.data, begin.ctypes.data, end.ctypes.data,
            ring.n_columns, ring.n_rows, ring.pixel_size, ring.n_planes,
            ring.z_left, ring.sigma_z, ring.sigma_dl, ring.length,
            evf.ctypes.data, evf.shape[0], 2, 3, 0, semantics,
            rho.ctypes.data, max_out, iters.ctypes.data, outs.ctypes.data,
            ctypes.byref(n_out), ctypes.byref(n_frames), name, 256, err, 512)
    assert rc == 0, err.value.decode()
    assert b"gfx950" in name.value
    assert list(iters[:n_out.value]) == [1, 2, 3, 6]
    legacy = semantics == 1
    ref6, n_used, _, _ = oracle_mod.hybrid_mlem(ring, ev, iterations=6,
                                                legacy=legacy)
    ref3 = oracle_mod.hybrid_mlem(ring, ev, iterations=3, legacy=legacy)[0]
    assert n_frames.value == n_used
    e6 = oracle_mod.rel_l2(rho, ref6)
    e3 = oracle_mod.rel_l2(outs[2].reshape(rho.shape), ref3)
    assert np.array_equal(outs[3].reshape(rho.shape), rho)
    msg = f"rel-L2 vs oracle: iteration 3 {e3:.3e}, iteration 6 {e6:.3e}"
    if oracle_mod.ref_available() and not legacy:
        r6 = oracle_mod.ref_hybrid_mlem(ring, ev, iterations=6)[0]
        er = oracle_mod.rel_l2(rho, r6)
        msg += f", vs reference build {er:.3e}"
        assert er <= 1e-4
    print(msg, flush=True)
    assert e3 <= 1e-4 and e6 <= 1e-4
