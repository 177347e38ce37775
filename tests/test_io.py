"""Image files and start images of 2d_strip_reconstruction (CPU).

The writers and the --rho reader of j-pet-mlem-old_amd/host/io.hpp, driven
by tests/io/io_writers.cpp, against the files the reference's own code
writes and reads (tests/golden/io/, made by tests/golden/make_io_golden.py
through oracle/_ref):
  - full-matrix text of the --gpu protocol (pixel_map.h:134-147);
  - `x y v` tuples with 12 fixed decimals of the values >= 1e-12, the CPU
    protocol's text (reconstruction.h:203-213, reconstruction_cmd.cpp:281-285);
  - raw floats and the NRRD headers, raw and ascii (nrrd_writer.h:19-28);
  - --rho start images: raw floats, int32 counts + 1000 (--rho-is-int) and
    text, which the reference streams into copies of the pixels, leaving
    rho = 1 (pixel_map.h:127-132).
Every output must equal the reference's byte for byte.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "io")
N_Z, N_Y = 24, 16

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")


@pytest.fixture(scope="module")
def io_tool(tmp_path_factory):
    exe = tmp_path_factory.mktemp("io") / "io_writers"
    cmd = ["g++", "-std=c++17", "-O1", "-Wall",
           "-I", os.path.join(ROOT, "j-pet-mlem-old_amd", "host"),
           "-o", str(exe), os.path.join(ROOT, "tests", "io", "io_writers.cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    return str(exe)


@pytest.mark.parametrize("name", ["matrix.txt", "tuples.txt", "image.bin",
                                  "raw.nrrd", "ascii.nrrd"])
def test_image_files_equal_the_reference_writers(io_tool, tmp_path, name):
    r = subprocess.run([io_tool, "write", os.path.join(GOLD, "image_in.bin"),
                        str(N_Z), str(N_Y), str(tmp_path)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    with open(tmp_path / name, "rb") as a, \
            open(os.path.join(GOLD, name), "rb") as b:
        assert a.read() == b.read()


def test_tuples_cutoff_and_format():
    # the golden tuples themselves: 1e-12 (as float32) kept, 9.99e-13 not
    lines = open(os.path.join(GOLD, "tuples.txt")).read().splitlines()
    img = np.fromfile(os.path.join(GOLD, "image_in.bin"),
                      np.float32).reshape(N_Y, N_Z)
    assert len(lines) == int((img >= np.float32(1e-12)).sum())
    assert lines[0] == "2 0 0.000000000001"


@pytest.mark.parametrize("kind,fn,is_int", [("txt", "start.txt", 0),
                                            ("float", "start_float.bin", 0),
                                            ("int", "start_int.bin", 1)])
def test_start_images_read_like_the_reference(io_tool, tmp_path, kind, fn,
                                              is_int):
    out = tmp_path / "rho.bin"
    r = subprocess.run([io_tool, "read", os.path.join(GOLD, fn), str(N_Z),
                        str(N_Y), str(is_int), str(out)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(out, np.float32).reshape(N_Y, N_Z)
    expect = np.load(os.path.join(GOLD, "start_expected.npz"))[kind]
    assert np.array_equal(got, expect)
    if kind == "txt":
        assert np.all(got == 1.0)
