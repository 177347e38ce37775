"""The C-ABI library loads without a GPU and exports every entry point that
include/*.h declares; host-side argument checking (no GPU compute)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADERS = [os.path.join(ROOT, "include", h)
           for h in sorted(os.listdir(os.path.join(ROOT, "include")))
           if h.endswith(".h")]


def declared_functions(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = set()
    for m in re.finditer(r"\b([a-z_][a-z0-9_]*)\s*\(", text):
        name = m.group(1)
        if name.startswith(("strip_mlem_", "strip_phantom_", "strip_events_",
                             "hybrid_mlem_")):
            # skip typedef'd function-pointer names
            if not re.search(r"\(\s*\*\s*" + name + r"\b", text):
                names.add(name)
    return names


def library_for(header):
    base = os.path.basename(header)
    lib = {"strip_mlem.h": "libstrip_mlem.so",
           "strip_events.h": "libstrip_mlem.so",
           "strip_phantom.h": "libstrip_mlem.so",
           "hybrid_mlem.h": "libstrip_mlem.so"}.get(base)
    return os.path.join(ROOT, "j-pet-mlem-old_amd", "lib", lib) if lib else None


@pytest.mark.parametrize("header", HEADERS, ids=os.path.basename)
def test_library_exports_declared_symbols(header):
    lib_path = library_for(header)
    assert lib_path, f"no library mapped for {header}"
    if not os.path.exists(lib_path):
        pytest.skip(f"{lib_path} not built")
    lib = ctypes.CDLL(lib_path)
    names = declared_functions(header)
    assert names, header
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_abi(pystrip_mod):
    from pystrip import hybrid
    names = set()
    for h in ("strip_mlem.h", "strip_phantom.h"):
        names |= declared_functions(os.path.join(ROOT, "include", h))
    assert set(pystrip_mod.exported_symbols()) == names
    assert set(hybrid.exported_symbols()) == declared_functions(
        os.path.join(ROOT, "include", "hybrid_mlem.h"))


def test_argument_checks_without_gpu(pystrip_mod):
    lib = pystrip_mod.load_library()
    ctx = ctypes.c_void_p()
    # NULL geometry, bad pixel counts and bad sigmas are rejected before any
    # device is touched
    assert lib.strip_mlem_create(None, None, 0, ctypes.byref(ctx)) != 0
    g = pystrip_mod.Geometry.make(350, 500, 0, 160, 2, 10, 42)
    assert lib.strip_mlem_create(ctypes.byref(g), None, 0,
                                 ctypes.byref(ctx)) != 0
    assert b"pixel" in lib.strip_mlem_last_error()
    g = pystrip_mod.Geometry.make(350, 500, 256, 160, 2, -1, 42)
    assert lib.strip_mlem_create(ctypes.byref(g), None, 0,
                                 ctypes.byref(ctx)) != 0
    assert b"sigma" in lib.strip_mlem_last_error()
    m = pystrip_mod.Mode(5, 0, 0, 1)
    g = pystrip_mod.Geometry.make(350, 500, 256, 160, 2, 10, 42)
    assert lib.strip_mlem_create(ctypes.byref(g), ctypes.byref(m), 0,
                                 ctypes.byref(ctx)) != 0
    assert lib.strip_mlem_set_launch(None, 0, 0) != 0
    assert lib.strip_mlem_n_events(None) == -1
    # tiling knobs on a NULL context or batch are refused, not dereferenced
    assert lib.strip_mlem_set_band_groups(None, 0, 0) != 0
    assert lib.strip_mlem_set_backprojection_occupancy(None, 16) != 0
    assert lib.strip_mlem_batch_set_backprojection_occupancy(None, 16) != 0
    assert lib.strip_mlem_batch_get_backprojection_occupancy(None, None) != 0


def test_hybrid_argument_checks_without_gpu(pystrip_mod):
    # the geometry checks of hybrid_mlem_create run on the host before any
    # device is touched
    import ctypes as C

    import numpy as np
    from pystrip import hybrid
    lib = hybrid.load_library()
    geo = hybrid.ring_geometry(n_detectors=16, n_pixels=20, pixel_size=0.02)
    ctx = C.c_void_p()

    def create(g, **over):
        arrays = [np.ascontiguousarray(g.seg, np.float32),
                  np.ascontiguousarray(g.widths, np.float32),
                  np.ascontiguousarray(g.pixels, np.int16),
                  np.ascontiguousarray(g.positions, np.float32),
                  np.ascontiguousarray(g.weights[0], np.float32),
                  np.ascontiguousarray(over.get("begin", g.info_begin),
                                       np.int64),
                  np.ascontiguousarray(over.get("end", g.info_end),
                                       np.int64)]
        s = hybrid.GeometryC(g.n_detectors, g.n_infos,
                             over.get("n_planes_half", 1),
                             *[a.ctypes.data for a in arrays], g.n_columns,
                             g.n_rows, g.n_planes, g.pixel_size, g.z_left,
                             g.sigma_z, g.sigma_dl, g.length)
        return lib.hybrid_mlem_create(C.byref(s), 0, 0, C.byref(ctx))

    assert lib.hybrid_mlem_create(None, 0, 0, C.byref(ctx)) != 0
    assert create(geo, n_planes_half=2) == -5
    assert b"n_planes_half" in lib.strip_mlem_last_error()
    end = geo.info_end.copy()
    end[-1] = geo.n_infos + 1
    assert create(geo, end=end) == -1
    assert b"range" in lib.strip_mlem_last_error()
    # a LOR whose infos repeat a pixel
    li = int(np.argmax(geo.info_end - geo.info_begin))
    b = int(geo.info_begin[li])
    geo.pixels[b + 1] = geo.pixels[b]
    assert create(geo) == -1
    assert b"twice" in lib.strip_mlem_last_error()
    assert lib.hybrid_mlem_n_events(None) == -1
    assert lib.hybrid_mlem_iterate(None, 1) != 0
