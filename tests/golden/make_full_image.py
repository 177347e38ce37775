#!/usr/bin/env python3
"""Writes the full-size golden image of the 1000 x 400 @1 mm, 2e7-event head
workload: the REFERENCE's own Reconstruction<float, GaussianKernel<float>>
(oracle/_ref/libstrip_ref.so, built by `make -C oracle ref` from the headers
under /root/reference/src), 20 MLEM iterations from rho = 1 on the
workload's seeded events (pystrip/phantom.py), and with --fp64 also the
Reconstruction<double, ...> image.  The events are regenerated from the seed
by the test, so only the images and the events' SHA-256 (which the test
checks before it compares images) are stored.  About 45 min (FP32) and 70 min
(FP64) on 8 host threads.

  python tests/golden/make_full_image.py [--fp64] [--threads N]
"""
import argparse
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "j-pet-mlem-old_amd"))

import oracle  # noqa: E402
from pystrip import phantom  # noqa: E402

NAME = "strip-1000x400-1mm-2e7"
ITERATIONS = 20
OUT = os.path.join(HERE, "strip_1000x400_2e7_20it_ref.npz")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--fp64", action="store_true")
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    args = ap.parse_args()
    w = phantom.WORKLOADS[NAME]
    ev = w.events()
    g = oracle.geometry(*w.oracle_geometry_tuple())
    out = {"workload": np.array(NAME), "n_events": np.int64(len(ev)),
           "seed": np.int64(w.seed), "iterations": np.int32(ITERATIONS),
           "events_sha256": np.array(phantom.events_digest(ev))}
    out["rho_f32"] = oracle.ref_mlem(g, ev, iterations=ITERATIONS,
                                     n_threads=args.threads)
    if args.fp64:
        out["rho_f64"] = oracle.ref_mlem(g, ev.astype(np.float64),
                                         iterations=ITERATIONS,
                                         n_threads=args.threads,
                                         dtype=np.float64).astype(np.float32)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
