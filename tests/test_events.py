"""Event ingestion (include/strip_events.h) against the reference's own
readers: the fast text parser (src/util/text_parser.h:140-175, through
oracle/_ref) must give bit-identical floats, including its quirks."""
import ctypes
import os

import numpy as np
import pytest

from conftest import ROOT

LIB = os.path.join(ROOT, "j-pet-mlem-old_amd", "lib", "libstrip_mlem.so")


@pytest.fixture(scope="module")
def ev_lib():
    if not os.path.exists(LIB):
        pytest.skip("library not built")
    L = ctypes.CDLL(LIB)
    L.strip_events_read.restype = ctypes.c_int
    L.strip_events_read.argtypes = [ctypes.c_char_p, ctypes.c_int,
                                    ctypes.POINTER(ctypes.POINTER(ctypes.c_float)),
                                    ctypes.POINTER(ctypes.c_int64)]
    L.strip_events_free.argtypes = [ctypes.POINTER(ctypes.c_float)]
    L.strip_events_last_error.restype = ctypes.c_char_p
    L.strip_events_scale.argtypes = [ctypes.POINTER(ctypes.c_float),
                                     ctypes.c_int64, ctypes.c_int,
                                     ctypes.c_double, ctypes.c_int,
                                     ctypes.c_double, ctypes.c_double]
    return L


def read(L, path, is_3d=0):
    buf = ctypes.POINTER(ctypes.c_float)()
    n = ctypes.c_int64()
    rc = L.strip_events_read(path.encode(), is_3d, ctypes.byref(buf),
                             ctypes.byref(n))
    if rc:
        raise ValueError(L.strip_events_last_error().decode())
    out = np.ctypeslib.as_array(buf, shape=(n.value * 3,)).copy() \
        if n.value else np.zeros(0, np.float32)
    L.strip_events_free(buf)
    return out.reshape(-1, 3)


def ref_read(oracle_mod, path, is_3d=0):
    out = np.zeros((10000, 3), np.float32)
    n = oracle_mod.ref().ref_load_txt_events(
        path.encode(), is_3d, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
        10000)
    if n < 0:
        raise ValueError("reference parser threw")
    return out[:n]


TEXT = """0.0740503 0.121193 0.221124
-0.0885979 0.230476 0.186496
1.5e-3 -2.25E-2 +7.125
12 -3. .5
123.456789 -98765.4321 0.000001
1e2 2.5e1 -3e+1
  4.5   -6.25 7
"""


def test_text_parser_matches_reference(tmp_path, ev_lib, oracle_mod):
    if not oracle_mod.ref_available():
        pytest.skip("oracle/_ref not built")
    p = tmp_path / "events.txt"
    p.write_text(TEXT)
    ours = read(ev_lib, str(p))
    ref = ref_read(oracle_mod, str(p))
    assert ours.shape == ref.shape == (7, 3)
    assert np.array_equal(ours.view(np.uint32), ref.view(np.uint32))


def test_text_parser_quirks(tmp_path, ev_lib):
    # the reference's positive powers of ten are 10 * 10^-(e-1): "1e2" -> 1
    # (text_parser.h:83-90); a last line without '\n' is not read
    p = tmp_path / "q.txt"
    p.write_text("1e2 2.5e1 -3e+1\n1 2 3")
    ours = read(ev_lib, str(p))
    assert ours.shape == (1, 3)
    assert ours[0, 0] == np.float32(1.0)


def test_text_3d_and_errors(tmp_path, ev_lib, oracle_mod):
    p = tmp_path / "e3d.txt"
    p.write_text("3 4 1.0 2.0 3.0\n-1 7 0.5 -0.5 0.25\n")
    ours = read(ev_lib, str(p), 1)
    assert np.allclose(ours, [[1, 2, 3], [0.5, -0.5, 0.25]])
    if oracle_mod.ref_available():
        assert np.array_equal(ours, ref_read(oracle_mod, str(p), 1))
    bad = tmp_path / "bad.txt"
    bad.write_text("1.0 . 2.0\n")
    with pytest.raises(ValueError, match="period"):
        read(ev_lib, str(bad))
    with pytest.raises(ValueError):
        read(ev_lib, str(tmp_path / "missing.txt"))
    with pytest.raises(ValueError, match="3D"):
        read(ev_lib, str(tmp_path / "x.bin"), 1)


def test_binary_roundtrip_and_scaling(tmp_path, ev_lib):
    ev = np.random.default_rng(0).normal(size=(1001, 3)).astype(np.float32)
    p = tmp_path / "events.bin"
    ev.tofile(p)
    back = read(ev_lib, str(p))
    assert np.array_equal(back, ev)
    # a trailing partial record is ignored like the reference's stream read
    with open(p, "ab") as f:
        f.write(b"\0" * 8)
    assert read(ev_lib, str(p)).shape == (1001, 3)
    x = ev.copy()
    ev_lib.strip_events_scale(x.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                              len(x), 1, 10.0, 1, 2.0, 3.0)
    assert np.allclose(x[:, :2], ev[:, :2] * 10)
    assert np.allclose(x[:, 2], ev[:, 2] * 6.0)
