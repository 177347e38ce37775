#!/usr/bin/env python3
"""Writes the golden fixtures of the strip MLEM path from the REFERENCE's own
code (oracle/_ref/libstrip_ref.so, built from /root/reference/src headers by
`make -C oracle ref`).  Run once in a container where the reference tree is
present; the fixtures (inputs and expected outputs, .npz) are committed so the
tests can pin the oracle without the reference.

  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "j-pet-mlem-old_amd"))

import oracle  # noqa: E402
from pystrip import phantom  # noqa: E402


def mlem_case(name, geometry_tuple, events, iterations):
    g = oracle.geometry(*geometry_tuple)
    f32 = oracle.ref_mlem(g, events, iterations=iterations, n_threads=1)
    f64 = oracle.ref_mlem(g, events.astype(np.float64), iterations=iterations,
                          n_threads=1, dtype=np.float64)
    sens = np.empty((g.n_y_pixels, g.n_z_pixels), np.float32)
    oracle.ref().ref_sensitivity_map_f32(
        __import__("ctypes").byref(g),
        sens.ctypes.data_as(__import__("ctypes").POINTER(__import__("ctypes").c_float)))
    psens = np.empty_like(sens)
    oracle.ref().ref_pixel_sensitivity_map_f32(
        __import__("ctypes").byref(g),
        psens.ctypes.data_as(__import__("ctypes").POINTER(__import__("ctypes").c_float)))
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, geometry=np.array(geometry_tuple, np.float64),
                        events=events.astype(np.float32),
                        iterations=np.int32(iterations), rho_f32=f32,
                        rho_f64=f64, sensitivity=sens,
                        pixel_sensitivity=psens)
    print("wrote", path, os.path.getsize(path), "bytes")


def kernel_case():
    rng = np.random.default_rng(2024)
    n = 200
    sz = rng.uniform(5, 15, n)
    sdl = rng.uniform(20, 70, n)
    y = rng.uniform(-150, 150, n)
    ang = rng.uniform(-0.9, 0.9, n)
    tan = np.tan(ang)
    sec = np.sqrt(1 + tan * tan)
    R = rng.uniform(300, 450, n)
    dz = rng.uniform(-40, 40, n)
    dy = rng.uniform(-60, 60, n)
    args = np.stack([sz, sdl, y, tan, sec, R, dz, dy], axis=1)
    r = oracle.ref()
    g32 = np.array([r.ref_gaussian_kernel_f32(*map(float, a)) for a in args],
                   np.float32)
    g64 = np.array([r.ref_gaussian_kernel_f64(*map(float, a)) for a in args])
    a32 = np.array([r.ref_analytic_kernel_f32(*map(float, a)) for a in args],
                   np.float32)
    a64 = np.array([r.ref_analytic_kernel_f64(*map(float, a)) for a in args])
    path = os.path.join(HERE, "kernels.npz")
    np.savez_compressed(path, args=args, gaussian_f32=g32, gaussian_f64=g64,
                        analytic_f32=a32, analytic_f64=a64)
    print("wrote", path, os.path.getsize(path), "bytes")


def event_params_case():
    w = phantom.WORKLOADS["strip-1000x400-1mm-2e7"]
    ev = w.events(2000, seed=11)
    g = oracle.geometry(*w.oracle_geometry_tuple())
    out = np.empty((len(ev), 9), np.float32)
    import ctypes
    buf = (ctypes.c_float * 9)()
    for i, (a, b, c) in enumerate(ev):
        oracle.ref().ref_event_params_f32(ctypes.byref(g), float(a), float(b),
                                          float(c), buf)
        out[i] = list(buf)
    path = os.path.join(HERE, "event_params.npz")
    np.savez_compressed(path, geometry=np.array(w.oracle_geometry_tuple()),
                        events=ev, params=out)
    print("wrote", path, os.path.getsize(path), "bytes")


def main():
    if not oracle.ref_available():
        oracle.build(reference=True)
    w1 = phantom.WORKLOADS["strip-256x160-2mm-1e6"]
    mlem_case("strip_256x160_5000ev_3it", w1.oracle_geometry_tuple(),
              w1.events(5000, seed=3), 3)
    # tiny grid: 24 x 16 pixels of 8 mm, events partly outside the grid
    tiny = (120.0, 300.0, 24, 16, 8.0, 8.0, 10.0, 42.0, 0.0, 0.0)
    ev = phantom.generate_events(phantom.PHANTOM_3, 120.0, 300.0, 10.0, 42.0,
                                 400, seed=5)
    mlem_case("strip_24x16_400ev_5it", tiny, ev, 5)
    kernel_case()
    event_params_case()


if __name__ == "__main__":
    main()
