"""The allocation-free forward projection (strip_mlem_forward_noalloc)
against strip_mlem_forward, through the C ABI.

Device buffers are made here with the HIP runtime's own hipMalloc (the test
may allocate; the entry point may not).  Two calls in a row on the same
GPU, the second with fewer events and another image, each against a
context's strip_mlem_forward on the same events and image: bit for bit
against the bbox kernel (the same arithmetic), within 2e-5 relative against
the default row-band kernels, and the output beyond the second call's events
untouched.  Sizes above the scratch maximum and sensitivity modes are
refused.
"""
import ctypes

import numpy as np
import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu

CFG1 = "strip-256x160-2mm-1e6"


class Hip:
    """hipMalloc / hipMemcpy / hipFree / hipDeviceSynchronize of the HIP
    runtime the library is linked against."""

    def __init__(self):
        self.rt = ctypes.CDLL("libamdhip64.so")
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p),
                                      ctypes.c_size_t]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_size_t, ctypes.c_int]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.bufs = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), nbytes) == 0
        self.bufs.append(p)
        return p.value

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.alloc(arr.nbytes)
        assert self.rt.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1) == 0
        return p

    def download(self, p, shape, dtype=np.float32):
        out = np.empty(shape, dtype)
        assert self.rt.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0
        return out

    def sync(self):
        assert self.rt.hipDeviceSynchronize() == 0

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def _context_forward(pystrip_mod, geometry, ev, rho, algorithm):
    rec = pystrip_mod.Reconstruction(geometry)
    rec.set_algorithm(algorithm)
    rec.set_events(ev)
    rec.rho = rho
    d = rec.forward()
    rec.close()
    return d


def test_forward_noalloc_matches_forward(pystrip_mod, phantom_mod):
    w = phantom_mod.WORKLOADS[CFG1]
    g = w.geometry()
    hip = Hip()
    rng = np.random.default_rng(17)
    n1, n2 = 200_000, 37_000
    out = hip.alloc(4 * n1)
    try:
        sentinel = np.full(n1, -7.0, np.float32)
        assert hip.rt.hipMemcpy(out, sentinel.ctypes.data, 4 * n1, 1) == 0
        for n, seed in ((n1, 71), (n2, 72)):
            print_case("forward_noalloc", n, "events, seed", seed)
            ev = w.events(n, seed=seed)
            rho = rng.uniform(0.5, 2.0, size=(w.n_y, w.n_z)).astype(np.float32)
            d_ev, d_rho = hip.upload(ev), hip.upload(rho)
            pystrip_mod.forward_noalloc(g, d_ev, n, d_rho, out)
            hip.sync()
            got = hip.download(out, n1)
            exact = _context_forward(pystrip_mod, g, ev, rho, 1)
            band = _context_forward(pystrip_mod, g, ev, rho, 2)
            assert np.array_equal(got[:n], exact)
            nz = band > 0
            assert np.array_equal(got[:n] > 0, nz)
            rel = np.abs(got[:n][nz].astype(np.float64) - band[nz]) / band[nz]
            print("max rel vs row-band", rel.max(), flush=True)
            assert rel.max() < 2e-5
            if n < n1:
                # the earlier call's outputs beyond this call's events stay
                assert np.all(got[n:] != -7.0)
        # refused: above the scratch maximum, a sensitivity mode
        big = pystrip_mod.NOALLOC_MAX_EVENTS + 1
        with pytest.raises(pystrip_mod.StripMLEMError):
            pystrip_mod.forward_noalloc(g, out, big, out, out)
        with pytest.raises(pystrip_mod.StripMLEMError):
            pystrip_mod.forward_noalloc(g, out, 10, out, out,
                                        mode=pystrip_mod.MODE_LEGACY_GPU)
    finally:
        hip.free()
