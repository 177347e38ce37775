"""The drop-in header compiles against the reference's own headers.

tests/dropin/dropin_call.cpp makes the `--gpu` call the way the reference CLI
does (src/2d/strip/reconstruction_cmd.cpp:223-262): reference Scanner,
Response<float> vector, PixelMap image and lambdas, with
include/strip_gpu_reconstruction.hpp standing where cuda/reconstruction.h
stood.  Compile only (the CPU container has no GPU); skipped where the
reference tree is absent.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "src", "2d", "strip")),
                    reason="reference tree not present")
@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_dropin_header_compiles_with_reference_headers(tmp_path):
    cmd = ["g++", "-std=c++11", "-fsyntax-only", "-Wall",
           "-I", os.path.join(REF, "src"),
           "-I", os.path.join(REF, "lib", "cmdline"),
           "-I", os.path.join(REF, "lib", "json"),
           "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "dropin", "dropin_call.cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0, r.stderr[-2000:]


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "src", "3d", "hybrid")),
                    reason="reference tree not present")
@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_hybrid_dropin_header_compiles_with_reference_headers(tmp_path):
    # tests/dropin/hybrid_dropin_call.cpp makes the hybrid `--gpu` call of
    # src/3d/hybrid/reconstruction_cmd.cpp:596-640 with
    # include/hybrid_gpu_reconstruction.hpp in place of cuda/reconstruction.h;
    # MAX_DETECTORS as the reference's build defines it (CMakeLists.txt:19)
    cmd = ["g++", "-std=c++11", "-fsyntax-only", "-Wall",
           "-DMAX_DETECTORS=512",
           "-I", os.path.join(REF, "src"),
           "-I", os.path.join(REF, "lib", "cmdline"),
           "-I", os.path.join(REF, "lib", "json"),
           "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "dropin", "hybrid_dropin_call.cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0, r.stderr[-2000:]
