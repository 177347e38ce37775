#!/usr/bin/env python3
"""Writes tests/golden/io/: an image and start-image inputs, and what the
REFERENCE's own writers and readers make of them (oracle/_ref, built from
/root/reference/src by `make -C oracle ref`):

  image_in.bin                 24 x 16 float image (zeros, values below and
                               at the 1e-12 tuple cutoff, large values)
  matrix.txt tuples.txt image.bin raw.nrrd ascii.nrrd
                               ref_write_images (PixelMap text stream,
                               output_tuples, obstream, nrrd_writer)
  start.txt start_float.bin start_int.bin
                               --rho inputs
  start_expected.npz           ref_read_start_image of each (rho after the
                               read; the text read leaves rho = 1)

  python tests/golden/make_io_golden.py
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

OUT = os.path.join(HERE, "io")
N_Z, N_Y = 24, 16


def image():
    rng = np.random.default_rng(2024)
    img = rng.uniform(0.0, 3.0, size=(N_Y, N_Z)).astype(np.float32)
    img[rng.random((N_Y, N_Z)) < 0.3] = 0.0
    img[0, :6] = [1e-13, 9.99e-13, 1e-12, 1.0000001e-12, 5e-7, 12345.678]
    img[1, :4] = [1e5, 3.14159274, 0.1, 2.5e-5]
    return img


def main():
    os.makedirs(OUT, exist_ok=True)
    L = oracle.ref()
    L.ref_write_images.argtypes = [ctypes.POINTER(oracle.OracleGeometry),
                                   ctypes.POINTER(ctypes.c_float),
                                   ctypes.c_char_p]
    L.ref_read_start_image.argtypes = [
        ctypes.POINTER(oracle.OracleGeometry), ctypes.c_char_p, ctypes.c_int,
        ctypes.c_int, ctypes.POINTER(ctypes.c_float)]
    g = oracle.geometry(350, 500, N_Z, N_Y, 2, 2, 10, 42)
    img = image()
    img.tofile(os.path.join(OUT, "image_in.bin"))
    L.ref_write_images(ctypes.byref(g),
                       img.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                       OUT.encode())
    rng = np.random.default_rng(7)
    with open(os.path.join(OUT, "start.txt"), "w") as f:
        for row in rng.uniform(0.5, 2.0, size=(N_Y, N_Z)):
            f.write(" ".join(f"{v:g}" for v in row) + "\n")
    rng.uniform(0.5, 2.0, size=(N_Y, N_Z)).astype(np.float32).tofile(
        os.path.join(OUT, "start_float.bin"))
    rng.integers(0, 5000, size=(N_Y, N_Z)).astype(np.int32).tofile(
        os.path.join(OUT, "start_int.bin"))
    expected = {}
    for key, fn, txt, is_int in (("txt", "start.txt", 1, 0),
                                 ("float", "start_float.bin", 0, 0),
                                 ("int", "start_int.bin", 0, 1)):
        out = np.empty((N_Y, N_Z), np.float32)
        L.ref_read_start_image(
            ctypes.byref(g), os.path.join(OUT, fn).encode(), txt, is_int,
            out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        expected[key] = out
    np.savez(os.path.join(OUT, "start_expected.npz"), **expected)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
