#!/usr/bin/env python3
"""Throughput benchmark of the strip list-mode TOF MLEM reconstruction.

A step is one MLEM iteration over every resident event of the workload
(forward projection + back-projection + image update).  The metric is
event-updates/s = events x iterations / time, with the events already
resident in HBM when the timed region starts.

Prints one JSON line.  Default workload: 2e7 seeded events on a 1000 x 400
grid of 1 mm pixels (R = 400 mm, L = 1000 mm, sigma_z = 10 mm,
sigma_dl = 42 mm, 10-ellipse head phantom), 20 iterations.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "j-pet-mlem-old_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

# published MI355X peak for FP32 (TFLOP/s); the kernels use non-packed FP32,
# whose issue rate is half of it
PEAK_FP32_TFLOPS = 157.3
# algorithmic FP32 operations per kernel evaluation (forward + back-projection
# of one in-ellipse pixel): Q = rowA + dz*(rowB + C*dz) (2 FMA = 4 flop),
# exp2 (1), K = pre*e (1), D += K*rho (2); again Q (4), exp2 (1), K (1),
# K*inv (1) and the image add (1)
FLOPS_PER_KERNEL_EVAL = 16
# HBM bytes per launch (read + written) of the band kernels on the default
# workload, from two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) each scaled
# by the counter's reading for a plain 256 MiB copy kernel (x2.0 and x1.0);
# recorded in profiles/band_kernels_strip-1000x400-1mm-2e7_counters.md
MEASURED_TRAFFIC_BYTES = {
    "back-projection": (41.51 + 0.90) * 1e9,
    "forward projection": (22.49 + 2.24) * 1e9,
}
TRAFFIC_WORKLOAD = "strip-1000x400-1mm-2e7"


def usable_cores():
    """CPUs this process may run on: its affinity set, reduced to the
    cgroup's CPU quota where one is set (cgroup v2 cpu.max)."""
    n = len(os.sched_getaffinity(0))
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            quota, period = f.read().split()[:2]
        if quota != "max":
            n = max(1, min(n, int(float(quota) / float(period))))
    except (OSError, ValueError):
        pass
    return n


def cpu_baseline(workload, n_sample):
    """Times the reference CPU reconstruction (oracle/_ref, the reference's own
    Reconstruction<float, GaussianKernel<float>>, in the widest ISA build the
    host CPU runs) -- or the oracle port where the reference build is absent
    -- on a bounded sample of the workload, one iteration, on every usable
    core and on 16 threads."""
    import oracle
    events = workload.events(n_sample, seed=workload.seed + 1)
    g = oracle.geometry(*workload.oracle_geometry_tuple())
    kind = "reference" if oracle.ref_available() else "port"
    isa = oracle.ref_native_isa() if kind == "reference" else "x86-64-v3"
    model, _ = oracle.host_cpu()

    def timed(threads):
        if kind == "reference":
            def fn(ev):
                oracle.ref_mlem(g, ev, iterations=1, n_threads=threads, isa=isa)
        else:
            def fn(ev):
                oracle.mlem(g, ev, iterations=1, n_threads=threads)
        fn(events[: min(2000, n_sample)])
        t0 = time.perf_counter()
        fn(events)
        return time.perf_counter() - t0

    cores = usable_cores()
    dt = timed(cores)
    out = {
        "value": n_sample / dt,
        "unit": "event-updates/s",
        "cores": cores,
        "kind": kind,
        "sample": f"{n_sample} events of the same workload (seed "
                  f"{workload.seed + 1}), 1 iteration, {dt:.2f} s wall",
        "cpu_model": model,
        "host_logical_cpus": os.cpu_count(),
        "affinity_cpus": len(os.sched_getaffinity(0)),
        "isa": isa,
    }
    if cores != 16:
        dt16 = timed(16)
        out["value_16_threads"] = n_sample / dt16
    else:
        out["value_16_threads"] = out["value"]
    return out


def reference_parity(rec, workload, events):
    """Error of the GPU image after the fixture's iteration count from
    rho = 1 on the fixture's events against the reference's own FP32 and
    FP64 CPU images of the same events (tests/golden/
    strip_1000x400_2e7_20it_ref.npz, made by make_full_image.py); None where
    the bench workload is not the fixture's."""
    path = os.path.join(ROOT, "tests", "golden",
                        "strip_1000x400_2e7_20it_ref.npz")
    if not os.path.exists(path):
        return None
    from pystrip import phantom
    gold = np.load(path)
    if (str(gold["workload"]) != workload.name or
            int(gold["n_events"]) != len(events) or
            int(gold["seed"]) != workload.seed):
        return None

    def rel_l2(a, b):
        a = np.asarray(a, np.float64)
        b = np.asarray(b, np.float64)
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    iterations = int(gold["iterations"])
    rec.rho = np.ones(rec.shape, np.float32)
    rec.iterate(iterations)
    img = rec.rho
    return {
        "iterations": iterations,
        "fixture": "tests/golden/strip_1000x400_2e7_20it_ref.npz",
        "events_sha256_match": (phantom.events_digest(events) ==
                                str(gold["events_sha256"])),
        "vs_reference_fp32_rel_l2": rel_l2(img, gold["rho_f32"]),
        "vs_reference_fp64_rel_l2": rel_l2(img, gold["rho_f64"]),
        "reference_fp32_vs_fp64_rel_l2": rel_l2(gold["rho_f32"],
                                                gold["rho_f64"]),
        "bar": 1e-4,
        "image_total_fp64": rec.image_total(),
    }


def adjoint_timing(pystrip, phantom, name, sigma, iterations):
    """Cost of the differentiation form against plain MLEM iterations on the
    same context: unrolled-MLEM loss gradient through `iterations` updates
    (forward with checkpoints + reverse steps) and the log-likelihood
    gradient, at the initial sigma guess."""
    w = phantom.WORKLOADS[name]
    g = w.geometry()
    g.sigma_z, g.sigma_dl = sigma
    rec = pystrip.Reconstruction(g, pystrip.MODE_CPU, device=0)
    rec.set_events(w.events())
    target = phantom.phantom_image(w.phantom, w.n_z, w.n_y,
                                   w.pixel).astype(np.float32)
    rho0 = np.ones(rec.shape, np.float32)
    rec.iterate(1)
    rec.unrolled_gradient(rho0, 1, target)
    rec.synchronize()
    t = time.perf_counter()
    rec.iterate(iterations)
    rec.synchronize()
    t_iter = (time.perf_counter() - t) / iterations
    t = time.perf_counter()
    loss, _, gs = rec.unrolled_gradient(rho0, iterations, target)
    t_unrolled = time.perf_counter() - t
    t = time.perf_counter()
    _, _, gl = rec.loglik_gradient()
    t_loglik = time.perf_counter() - t
    reverse = (t_unrolled - iterations * t_iter) / iterations
    spans = rec.span_table()
    rec.close()
    return {
        "workload": name,
        "n_events": w.n_events,
        "sigma_mm": list(sigma),
        "iterations": iterations,
        "mlem_iteration_ms": t_iter * 1e3,
        "unrolled_loss_gradient_ms": t_unrolled * 1e3,
        "reverse_step_ms": reverse * 1e3,
        "reverse_step_over_iteration": reverse / t_iter,
        "loglik_gradient_ms": t_loglik * 1e3,
        "loglik_gradient_over_iteration": t_loglik / t_iter,
        "loss": loss,
        "d_loss_d_sigma": [float(x) for x in gs],
        "d_loglik_d_sigma": [float(x) for x in gl],
        "span_table": spans,
    }


def osem_timing(pystrip, phantom, name, n_subsets, iterations, chunk):
    """Ordered subsets at the largest size: the workload's events generated
    on the device in `chunk`-event pieces and streamed back in through the
    pinned double-buffered appends, n_subsets subsets by event index, one
    untimed and `iterations` - 1 timed full iterations, then the
    log-likelihood gradient over the final image."""
    w = phantom.WORKLOADS[name]
    rec = pystrip.Reconstruction(w.geometry(), device=0)
    rec.reserve_events(w.n_events)
    t_gen = t_app = 0.0
    for k in range(0, w.n_events, chunk):
        m = min(chunk, w.n_events - k)
        t = time.perf_counter()
        out = pystrip.generate_events(w.phantom, w.radius, w.length,
                                      w.sigma_z, w.sigma_dl, m,
                                      seed=1000 + k // chunk)
        t_gen += time.perf_counter() - t
        t = time.perf_counter()
        rec.append_events(out["responses"][:m])
        t_app += time.perf_counter() - t
    rec.set_subsets(n_subsets)
    t = time.perf_counter()
    n_evals = rec.kernel_evaluations()
    t_plan = time.perf_counter() - t
    rec.iterate(1)
    rec.synchronize()
    t = time.perf_counter()
    rec.iterate(iterations - 1)
    rec.synchronize()
    t_it = (time.perf_counter() - t) / (iterations - 1)
    t = time.perf_counter()
    _, _, gs = rec.loglik_gradient(want_grad_rho=False)
    t_grad = time.perf_counter() - t
    out = {
        "workload": name,
        "events": w.n_events,
        "subsets": n_subsets,
        "chunk_events": chunk,
        "generation_s": t_gen,
        "streamed_append_s": t_app,
        "plans_s": t_plan,
        "kernel_evaluations_per_iteration": n_evals,
        "full_iteration_s": t_it,
        "event_updates_per_s": w.n_events / t_it,
        "kernel_evaluations_per_s": n_evals / t_it,
        "image_total_fp64": rec.image_total(),
        "loglik_sigma_gradient_s": t_grad,
        "d_loglik_d_sigma": [float(x) for x in gs],
        "span_table": rec.span_table(),
    }
    rec.close()
    return out


def batched_timing(pystrip, phantom, n_members, iterations):
    """Batched form against the plain form on the same members: n_members
    reconstructions with uneven event counts and grids, `iterations` MLEM
    iterations each, after the autotuner has chosen the tiling."""
    members = phantom.batched_workload(n_members, seed=0)
    b = pystrip.Batch(device=0)
    for w in members:
        b.add(w.geometry(), w.events())
    b.iterate(1)
    tuned_ms = b.autotune()
    b.iterate(4)  # untimed: records the iteration graph of the chosen tiling
    t = time.perf_counter()
    b.iterate(iterations)
    batch_s = time.perf_counter() - t
    # the plain form at the library's default tiling (its own best; the
    # members' plans otherwise carry the batch's tuned tiling)
    views = [b.member(i) for i in range(len(members))]
    for m in views:
        m.set_band_tiling()
        m.set_band_groups()
        m.iterate(1)
        m.synchronize()
    t = time.perf_counter()
    for m in views:
        m.iterate(iterations)
    plain_s = time.perf_counter() - t
    n_events = sum(w.n_events for w in members)
    out = {
        "members": len(members),
        "events": n_events,
        "grids": sorted({f"{w.n_z}x{w.n_y}@{w.pixel:g}mm" for w in members}),
        "iterations": iterations,
        "tiling": list(b.tiling()),
        "groups": list(b.groups()),
        "backprojection_cu_waves": b.backprojection_occupancy(),
        "autotuned_iteration_ms": tuned_ms,
        "batched_ms_per_iteration": batch_s * 1e3 / iterations,
        "plain_ms_per_iteration": plain_s * 1e3 / iterations,
        "batched_event_updates_per_s": n_events * iterations / batch_s,
        "plain_event_updates_per_s": n_events * iterations / plain_s,
        "speedup_over_plain": plain_s / batch_s,
        "members_with_span_table": sum(bool(m.span_table()) for m in views),
    }
    b.close()
    return out


# Algorithmic work of one hybrid voxel visit in the LOR-space iteration:
# forward b = (z - right) - dy tan, b^2 kexp, exp2, A K, x P, += (8), and
# back-projection b, b^2 kexp, exp2, F K, += (7)
HYBRID_FLOPS_PER_VISIT = 15


def hybrid_roofline(geo, visits, event_ms, gather_ms, algorithm):
    """The event kernels against the FP32 peak (algorithmic FLOPs per
    visit x visits per iteration / their time), and the gather against HBM
    (G and bp rows of every info, rho, the extra sums and the new image)."""
    out = {
        "event_kernels": {
            "bound": "valu",
            "achieved": visits * HYBRID_FLOPS_PER_VISIT / (event_ms * 1e-3)
            / 1e12 if event_ms > 0 else None,
            "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "avg_ms": event_ms,
            "flops_per_visit": HYBRID_FLOPS_PER_VISIT,
        },
    }
    e = out["event_kernels"]
    e["frac"] = e["achieved"] / PEAK_FP32_TFLOPS if e["achieved"] else None
    if algorithm == 2 and gather_ms > 0:
        nv = geo.n_rows * geo.n_columns * geo.n_planes
        nbytes = 4 * (2 * geo.n_infos * geo.n_planes + 3 * nv)
        gbps = nbytes / (gather_ms * 1e-3) / 1e9
        out["gather"] = {"bound": "hbm", "achieved": gbps, "peak": 8000.0,
                         "unit": "GB/s", "frac": gbps / 8000.0,
                         "avg_ms": gather_ms, "bytes_per_launch": nbytes}
    return out


def hybrid_timing(name, steps, warmup, cpu_sample, with_cpu, algorithm=0):
    """The 3D hybrid reconstruction (include/hybrid_mlem.h) on a named
    workload: weights and sensitivity prepared on the device, events
    translated and filtered on the device, then `steps` timed MLEM
    iterations after `warmup` untimed ones.  The CPU baseline is the
    reference's own Reconstruction (oracle/_ref) on `cpu_sample` events of
    the same workload: one iteration's time, the difference of runs with
    one and with zero iterations."""
    from pystrip import hybrid
    w = hybrid.WORKLOADS[name]
    t = time.perf_counter()
    geo = w.geometry()
    events = w.events(geo)
    t_gen = time.perf_counter() - t
    rec = hybrid.HybridReconstruction(geo)
    rec.set_algorithm(algorithm)
    t = time.perf_counter()
    rec.prepare_weights(True, 0)
    rec.add_responses(events)
    t_setup = time.perf_counter() - t
    visits = rec.voxel_visits()
    t = time.perf_counter()
    rec.iterate(warmup)
    t_first = time.perf_counter() - t
    t = time.perf_counter()
    rec.iterate(steps)
    elapsed = time.perf_counter() - t
    kernel_ms = rec.last_iteration_ms()
    event_ms, gather_ms, timed = rec.timing()
    n = rec.n_events
    used_algorithm = rec.algorithm
    mass = float((rec.rho.astype(np.float64)).sum())
    rec.close()
    out = {
        "workload": name,
        "n_events": w.n_events,
        "n_events_kept": n,
        "n_pixel_infos": geo.n_infos,
        "voxels": list(geo.shape),
        "algorithm": used_algorithm,
        "steps": steps,
        "warmup": warmup,
        "event_updates_per_s": n * steps / elapsed,
        "ms_per_iteration": elapsed * 1e3 / steps,
        "event_kernel_ms_last_iteration": kernel_ms,
        "voxel_visits_per_iteration": visits,
        "voxel_visits_per_s": visits / (kernel_ms * 1e-3),
        "roofline": hybrid_roofline(geo, visits, event_ms / max(timed, 1),
                                    gather_ms / max(timed, 1),
                                    used_algorithm),
        "timing_s": {"geometry_and_events": t_gen,
                     "weights_sensitivity_and_frames": t_setup,
                     "warmup_iterations_and_plan": t_first},
        "check": {"image_mass": mass},
        "cpu_baseline": None,
    }
    if with_cpu:
        import oracle
        if oracle.ref_available():
            sample = events[:cpu_sample]
            cores = usable_cores()
            t = time.perf_counter()
            oracle.ref_hybrid_mlem(geo, sample, iterations=0, n_threads=cores)
            t0 = time.perf_counter() - t
            t = time.perf_counter()
            _, used, _, _ = oracle.ref_hybrid_mlem(geo, sample, iterations=1,
                                                   n_threads=cores)
            t1 = time.perf_counter() - t
            out["cpu_baseline"] = {
                "value": used / max(t1 - t0, 1e-9),
                "unit": "event-updates/s", "cores": cores,
                "kind": "reference",
                "sample": f"{len(sample)} events of {name}, one iteration "
                          f"(run with 1 iteration {t1:.2f} s minus run "
                          f"with 0 {t0:.2f} s)"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="strip-1000x400-1mm-2e7")
    ap.add_argument("--events", type=int, default=0,
                    help="override the event count (0 = the workload's)")
    ap.add_argument("--algorithm", type=int, default=0)
    ap.add_argument("--cpu-sample", type=int, default=4_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--adjoint-workload", default="strip-512x256-1mm-5e6")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--batch-members", type=int, default=64)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--osem-workload", default="strip-2000x800-0.5mm-2e8")
    ap.add_argument("--no-osem", action="store_true")
    ap.add_argument("--no-parity", action="store_true",
                    help="skip the comparison with the reference's image")
    ap.add_argument("--hybrid-workload",
                    default="hybrid-ring384-160x160x100-5mm-1e7")
    ap.add_argument("--hybrid-steps", type=int, default=5)
    ap.add_argument("--hybrid-cpu-sample", type=int, default=200_000)
    ap.add_argument("--hybrid-algorithm", type=int, default=0)
    ap.add_argument("--no-hybrid", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: runs on one GPU only (asked for {args.gpus}); "
              f"event sharding over several GPUs is not implemented",
              file=sys.stderr)
        sys.exit(2)

    import pystrip
    from pystrip import phantom

    w = phantom.WORKLOADS[args.workload]
    n_events = args.events or w.n_events
    t_gen = time.perf_counter()
    events = w.events(n_events)
    t_gen = time.perf_counter() - t_gen

    rec = pystrip.Reconstruction(w.geometry(), pystrip.MODE_CPU, device=0)
    if args.algorithm:
        rec.set_algorithm(args.algorithm)
    t_up = time.perf_counter()
    rec.set_events(events)
    t_up = time.perf_counter() - t_up
    # the first query builds the row-band plan (sorts, records, span table,
    # pair lists), then counts the support
    t_plan = time.perf_counter()
    n_evals = rec.kernel_evaluations()
    t_plan = time.perf_counter() - t_plan

    # the context's stream carries all device work: synchronising it brackets
    # the timed iterations
    rec.iterate(args.warmup)
    rec.synchronize()
    rec.set_timing(True)
    t0 = time.perf_counter()
    rec.iterate_async(args.steps)
    rec.synchronize()
    elapsed = time.perf_counter() - t0
    fwd_ms, bwd_ms, iter_ms, n_timed = rec.get_timing_detail()
    rec.set_timing(False)
    kernel_ms = fwd_ms + bwd_ms

    value = n_events * args.steps / elapsed
    ms_per_step = elapsed * 1e3 / args.steps
    n_timed = max(n_timed, 1)
    fwd_avg, bwd_avg = fwd_ms / n_timed, bwd_ms / n_timed
    two_kernels = bwd_avg > 0
    if two_kernels:
        # the dominant kernel of the two carries half the algorithmic work
        dom_name = ("back-projection" if bwd_avg >= fwd_avg
                    else "forward projection")
        dom_ms = max(fwd_avg, bwd_avg)
        dom_flops = n_evals * FLOPS_PER_KERNEL_EVAL / 2
    else:
        dom_name = "event pass (forward + back-projection)"
        dom_ms = fwd_avg
        dom_flops = n_evals * FLOPS_PER_KERNEL_EVAL
    kernel_avg_ms = kernel_ms / n_timed
    achieved_tflops = dom_flops / (dom_ms * 1e-3) / 1e12

    rho = rec.rho
    mass = float(rho.astype(np.float64).sum())
    spans = rec.span_table()
    parity = None if args.no_parity else reference_parity(rec, w, events)

    line = {
        "metric": "MLEM event-updates/s on 1 GPU vs host CPU",
        "value": value,
        "unit": "event-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": w.name,
            "n_events": n_events,
            "grid": f"{w.n_z}x{w.n_y}",
            "pixel_mm": w.pixel,
            "R_mm": w.radius,
            "L_mm": w.length,
            "sigma_z_mm": w.sigma_z,
            "sigma_dl_mm": w.sigma_dl,
            "algorithm": rec.algorithm,
            "span_table": spans,
            "kernel_evaluations_per_iteration": n_evals,
            "mean_kernel_evaluations_per_event": n_evals / n_events,
            "device": rec.device_name(),
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved_tflops,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / PEAK_FP32_TFLOPS,
            "traffic": (MEASURED_TRAFFIC_BYTES.get(dom_name)
                        if w.name == TRAFFIC_WORKLOAD and
                        n_events == w.n_events and args.algorithm in (0, 2)
                        else None),
            "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + "
                            "WRITE_SIZE, copy-calibrated)",
            "kernel": dom_name,
            "kernel_avg_ms": dom_ms,
            "forward_avg_ms": fwd_avg,
            "backprojection_avg_ms": bwd_avg,
            "event_kernels_avg_ms": kernel_avg_ms,
            "iteration_avg_ms": iter_ms / n_timed,
            "kernel_evaluations_per_s": n_evals / (kernel_avg_ms * 1e-3),
            "flops_per_kernel_evaluation": FLOPS_PER_KERNEL_EVAL,
        },
        "cpu_baseline": None,
        "timing_s": {"event_generation": t_gen, "upload_and_setup": t_up,
                     "plan_and_support_count": t_plan},
        "check": {"image_mass": mass},
        "parity": parity,
    }
    if not args.no_adjoint:
        rec.close()
        line["adjoint"] = adjoint_timing(pystrip, phantom,
                                         args.adjoint_workload, (14.0, 60.0),
                                         8)
    if not args.no_batch:
        rec.close()
        line["batched"] = batched_timing(pystrip, phantom, args.batch_members,
                                         10)
    if not args.no_osem:
        rec.close()
        line["osem_largest"] = osem_timing(pystrip, phantom, args.osem_workload,
                                           16, 5, 10_000_000)
    if not args.no_hybrid:
        rec.close()
        line["hybrid"] = hybrid_timing(args.hybrid_workload,
                                       args.hybrid_steps, 1,
                                       args.hybrid_cpu_sample,
                                       not args.no_cpu, args.hybrid_algorithm)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(w, args.cpu_sample)
    rec.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
