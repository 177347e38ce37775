#!/usr/bin/env python3
"""Writes tests/golden/event_digests.json: the SHA-256 of the seeded event
sets the parity tests and the bench generate (pystrip/phantom.py, numpy's
Philox stream), so a test can tell a changed input stream from a parity
failure before it compares images.

  python tests/golden/make_event_digests.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "j-pet-mlem-old_amd"))

from pystrip import phantom  # noqa: E402

# (workload, events, seed); seed None = the workload's own
CASES = [
    ("strip-256x160-2mm-1e6", 1_000_000, None),
    ("strip-1000x400-1mm-2e7", 20_000_000, None),
    ("strip-512x256-1mm-5e6", 5_000_000, None),
    ("strip-512x256-1mm-5e6", 20_000, 41),
    ("strip-2000x800-0.5mm-2e8", 200_000, 43),
]


def key(name, n, seed):
    return f"{name}:{n}:{'default' if seed is None else seed}"


def main():
    out = {}
    for name, n, seed in CASES:
        w = phantom.WORKLOADS[name]
        out[key(name, n, seed)] = phantom.events_digest(w.events(n, seed=seed))
        print(key(name, n, seed), out[key(name, n, seed)], flush=True)
    with open(os.path.join(HERE, "event_digests.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main()
