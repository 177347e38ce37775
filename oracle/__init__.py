"""ctypes front-end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this module, and only as the checker.  The product path never imports it.

``liboracle.so`` is the C restatement of the reference's CPU strip
reconstruction (strip_oracle.c).  ``_ref/libstrip_ref.so`` is the reference's
own code compiled from its headers (ref_harness.cpp); it exists only where
the reference tree was available when the oracle was built.

Parity status of each oracle function (what pins it) is listed in the header
of strip_oracle.c; every function is pinned by a test in tests/.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libstrip_ref.so")
# the same harness built with -march=x86-64-v4 (AVX-512)
REF_LIB_V4 = os.path.join(HERE, "_ref", "libstrip_ref_v4.so")
AVX512_FLAGS = ("avx512f", "avx512cd", "avx512bw", "avx512dq", "avx512vl")


class OracleGeometry(ctypes.Structure):
    _fields_ = [
        ("radius", ctypes.c_double),
        ("scintillator_length", ctypes.c_double),
        ("n_z_pixels", ctypes.c_int),
        ("n_y_pixels", ctypes.c_int),
        ("pixel_width", ctypes.c_double),
        ("pixel_height", ctypes.c_double),
        ("sigma_z", ctypes.c_double),
        ("sigma_dl", ctypes.c_double),
        ("center_y", ctypes.c_double),
        ("center_z", ctypes.c_double),
    ]


class OracleMode(ctypes.Structure):
    _fields_ = [
        ("kernel", ctypes.c_int),
        ("sensitivity", ctypes.c_int),
        ("sensitivity_map", ctypes.c_int),
        ("guard_zero", ctypes.c_int),
    ]


MODE_CPU = (0, 0, 0, 1)
MODE_LEGACY_GPU = (1, 1, 1, 0)

_lib = None
_ref = {}


def build(reference: bool = False):
    """Compiles liboracle.so (and _ref/ when asked and possible)."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    if reference and os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def geometry(radius, length, n_z, n_y, pixel_w, pixel_h, sigma_z, sigma_dl,
             center_y=0.0, center_z=0.0):
    return OracleGeometry(float(radius), float(length), int(n_z), int(n_y),
                          float(pixel_w), float(pixel_h), float(sigma_z),
                          float(sigma_dl), float(center_y), float(center_z))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        G = ctypes.POINTER(OracleGeometry)
        M = ctypes.POINTER(OracleMode)
        f = ctypes.c_float
        d = ctypes.c_double
        fp = ctypes.POINTER(f)
        dp = ctypes.POINTER(d)
        ip = ctypes.POINTER(ctypes.c_int)
        i64 = ctypes.c_int64
        i64p = ctypes.POINTER(i64)
        sig = {
            "oracle_event_params_f32": (None, [G, f, f, f, fp, ip]),
            "oracle_event_params_f64": (None, [G, d, d, d, dp, ip]),
            "oracle_gaussian_kernel_f32": (f, [f] * 8),
            "oracle_gaussian_kernel_f64": (d, [d] * 8),
            "oracle_analytic_kernel_f32": (f, [f] * 8),
            "oracle_analytic_kernel_f64": (d, [d] * 8),
            "oracle_bb_y_f64": (d, [d, d, d]),
            "oracle_bb_z_f64": (d, [d, d, d]),
            "oracle_pixel_at_f32": (None, [G, f, f, ip, ip]),
            "oracle_pixel_at_f64": (None, [G, d, d, ip, ip]),
            "oracle_pixel_center_f32": (None, [G, ctypes.c_int, ctypes.c_int, fp, fp]),
            "oracle_pixel_center_f64": (None, [G, ctypes.c_int, ctypes.c_int, dp, dp]),
            "oracle_sensitivity_f32": (f, [G, f, f]),
            "oracle_sensitivity_f64": (d, [G, d, d]),
            "oracle_sensitivity_map_f32": (None, [G, ctypes.c_int, fp]),
            "oracle_sensitivity_map_f64": (None, [G, ctypes.c_int, dp]),
            "oracle_to_projection_f64": (None, [d, d, d, d, dp, dp, dp]),
            "oracle_mlem_f32": (ctypes.c_int, [G, M, fp, i64, fp, ctypes.c_int, ctypes.c_int]),
            "oracle_mlem_f64": (ctypes.c_int, [G, M, dp, i64, dp, ctypes.c_int, ctypes.c_int]),
            "oracle_osem_f32": (ctypes.c_int, [G, M, fp, i64, fp, ctypes.c_int,
                                               ctypes.c_int, ctypes.c_int]),
            "oracle_osem_f64": (ctypes.c_int, [G, M, dp, i64, dp, ctypes.c_int,
                                               ctypes.c_int, ctypes.c_int]),
            "oracle_forward_f32": (ctypes.c_int, [G, M, fp, i64, fp, fp, ctypes.c_int]),
            "oracle_forward_f64": (ctypes.c_int, [G, M, dp, i64, dp, dp, ctypes.c_int]),
            "oracle_support_counts_f32": (ctypes.c_int, [G, fp, i64, i64p, i64p, ctypes.c_int]),
            "oracle_loglik_f64": (d, [G, M, dp, i64, dp, dp, ctypes.c_int]),
            "oracle_loglik_grad_f64": (ctypes.c_int, [G, dp, i64, dp, dp, dp, dp, dp]),
            "oracle_mlem_loss_f64": (ctypes.c_int, [G, dp, i64, dp, ctypes.c_int, dp, dp,
                                                     dp, dp, dp, dp, dp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_LIB)


def host_cpu():
    """(model name, set of ISA flags) of this host, from /proc/cpuinfo."""
    model, flags = "unknown", set()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name") and model == "unknown":
                    model = line.split(":", 1)[1].strip()
                elif line.startswith("flags") and not flags:
                    flags = set(line.split(":", 1)[1].split())
    except OSError:
        pass
    return model, flags


def ref_native_isa() -> str:
    """The widest build of the reference harness this host can run:
    "x86-64-v4" where the AVX-512 build exists and the CPU has its
    extensions, else "x86-64-v3"."""
    _, flags = host_cpu()
    if os.path.exists(REF_LIB_V4) and all(f in flags for f in AVX512_FLAGS):
        return "x86-64-v4"
    return "x86-64-v3"


def ref(isa="x86-64-v3"):
    """The reference's own code (raises if oracle/_ref was not built), in
    the x86-64-v3 build or, with isa="x86-64-v4", the AVX-512 build."""
    if isa not in _ref:
        path = REF_LIB_V4 if isa == "x86-64-v4" else REF_LIB
        if not os.path.exists(path):
            raise RuntimeError(f"{path} not built")
        L = ctypes.CDLL(path)
        G = ctypes.POINTER(OracleGeometry)
        f = ctypes.c_float
        d = ctypes.c_double
        fp = ctypes.POINTER(f)
        dp = ctypes.POINTER(d)
        ip = ctypes.POINTER(ctypes.c_int)
        i64 = ctypes.c_int64
        sig = {
            "ref_mlem_f32": (ctypes.c_int, [G, fp, i64, fp, ctypes.c_int, ctypes.c_int]),
            "ref_mlem_f64": (ctypes.c_int, [G, dp, i64, dp, ctypes.c_int, ctypes.c_int]),
            "ref_sensitivity_map_f32": (ctypes.c_int, [G, fp]),
            "ref_pixel_sensitivity_map_f32": (ctypes.c_int, [G, fp]),
            "ref_gaussian_kernel_f32": (f, [f] * 8),
            "ref_gaussian_kernel_f64": (d, [d] * 8),
            "ref_analytic_kernel_f32": (f, [f] * 8),
            "ref_analytic_kernel_f64": (d, [d] * 8),
            "ref_event_params_f32": (None, [G, f, f, f, fp]),
            "ref_pixel_at_f32": (None, [G, f, f, ip, ip]),
            "ref_sensitivity_f32": (f, [G, f, f]),
            "ref_load_txt_events": (i64, [ctypes.c_char_p, ctypes.c_int, fp, i64]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _ref[isa] = L
    return _ref[isa]


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def mlem(g: OracleGeometry, events, rho0=None, iterations=1, mode=MODE_CPU,
         n_threads=0, dtype=np.float32):
    """Runs the oracle MLEM; returns the image (n_y, n_z)."""
    ev = np.ascontiguousarray(events, dtype=dtype)
    shape = (g.n_y_pixels, g.n_z_pixels)
    rho = (np.ones(shape, dtype) if rho0 is None
           else np.ascontiguousarray(rho0, dtype=dtype).reshape(shape).copy())
    m = OracleMode(*mode)
    if dtype == np.float32:
        rc = lib().oracle_mlem_f32(ctypes.byref(g), ctypes.byref(m),
                                   _p(ev, ctypes.c_float), ev.shape[0],
                                   _p(rho, ctypes.c_float), iterations,
                                   n_threads)
    else:
        rc = lib().oracle_mlem_f64(ctypes.byref(g), ctypes.byref(m),
                                   _p(ev, ctypes.c_double), ev.shape[0],
                                   _p(rho, ctypes.c_double), iterations,
                                   n_threads)
    if rc != 0:
        raise RuntimeError("oracle_mlem failed")
    return rho


def forward(g, events, rho, mode=MODE_CPU, n_threads=0, dtype=np.float32):
    ev = np.ascontiguousarray(events, dtype=dtype)
    r = np.ascontiguousarray(rho, dtype=dtype)
    out = np.empty(ev.shape[0], dtype=dtype)
    m = OracleMode(*mode)
    t = ctypes.c_float if dtype == np.float32 else ctypes.c_double
    fn = lib().oracle_forward_f32 if dtype == np.float32 else lib().oracle_forward_f64
    fn(ctypes.byref(g), ctypes.byref(m), _p(ev, t), ev.shape[0], _p(r, t),
       _p(out, t), n_threads)
    return out


def support_counts(g, events, n_threads=0):
    ev = np.ascontiguousarray(events, dtype=np.float32)
    nb = np.empty(ev.shape[0], dtype=np.int64)
    ne = np.empty(ev.shape[0], dtype=np.int64)
    lib().oracle_support_counts_f32(ctypes.byref(g), _p(ev, ctypes.c_float),
                                    ev.shape[0], _p(nb, ctypes.c_int64),
                                    _p(ne, ctypes.c_int64), n_threads)
    return nb, ne


def sensitivity_map(g, kind=0, dtype=np.float32):
    out = np.empty((g.n_y_pixels, g.n_z_pixels), dtype=dtype)
    if dtype == np.float32:
        lib().oracle_sensitivity_map_f32(ctypes.byref(g), kind,
                                         _p(out, ctypes.c_float))
    else:
        lib().oracle_sensitivity_map_f64(ctypes.byref(g), kind,
                                         _p(out, ctypes.c_double))
    return out


def event_params(g, z_u, z_d, dl, dtype=np.float32):
    t = ctypes.c_float if dtype == np.float32 else ctypes.c_double
    out = (t * 9)()
    box = (ctypes.c_int * 4)()
    fn = lib().oracle_event_params_f32 if dtype == np.float32 else \
        lib().oracle_event_params_f64
    fn(ctypes.byref(g), z_u, z_d, dl, out, box)
    return np.array(list(out), dtype=dtype), tuple(box)


def ref_mlem(g, events, rho0=None, iterations=1, n_threads=0,
             dtype=np.float32, isa="x86-64-v3"):
    ev = np.ascontiguousarray(events, dtype=dtype)
    shape = (g.n_y_pixels, g.n_z_pixels)
    rho = (np.ones(shape, dtype) if rho0 is None
           else np.ascontiguousarray(rho0, dtype=dtype).reshape(shape).copy())
    if dtype == np.float32:
        ref(isa).ref_mlem_f32(ctypes.byref(g), _p(ev, ctypes.c_float),
                           ev.shape[0], _p(rho, ctypes.c_float), iterations,
                           n_threads)
    else:
        ref(isa).ref_mlem_f64(ctypes.byref(g), _p(ev, ctypes.c_double),
                           ev.shape[0], _p(rho, ctypes.c_double), iterations,
                           n_threads)
    return rho


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def _dp(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def loglik_grad(g, events, rho, sigma=None, want_grad_rho=True):
    """FP64 Poisson log-likelihood L = sum_{D>0} log D_e - sum rho and its
    gradients (d L / d rho, d L / d (sigma_z, sigma_dl)); the kernel at
    `sigma` (default: the geometry's), the support at the geometry's sigmas.
    With want_grad_rho False the image gradient is not formed (None)."""
    ev = np.ascontiguousarray(events, np.float64)
    r = np.ascontiguousarray(rho, np.float64)
    sg = np.array(sigma if sigma is not None else (g.sigma_z, g.sigma_dl),
                  np.float64)
    value = np.zeros(1)
    grad_rho = np.empty_like(r) if want_grad_rho else None
    grad_sigma = np.zeros(2)
    rc = lib().oracle_loglik_grad_f64(ctypes.byref(g), _dp(ev), ev.shape[0],
                                      _dp(r), _dp(sg), _dp(value),
                                      _dp(grad_rho) if want_grad_rho else None,
                                      _dp(grad_sigma))
    if rc:
        raise RuntimeError("oracle_loglik_grad_f64 failed")
    return float(value[0]), grad_rho, grad_sigma


def mlem_loss(g, events, rho0, iterations, target, sigma=None, d_rho0=None,
              d_sigma=None):
    """FP64 unrolled-MLEM loss mean((rho_N - target)^2), its directional
    derivative along (d_rho0, d_sigma) (forward mode) and rho_N."""
    ev = np.ascontiguousarray(events, np.float64)
    r0 = np.ascontiguousarray(rho0, np.float64)
    tg = np.ascontiguousarray(target, np.float64)
    sg = np.array(sigma if sigma is not None else (g.sigma_z, g.sigma_dl),
                  np.float64)
    dr = None if d_rho0 is None else np.ascontiguousarray(d_rho0, np.float64)
    ds = None if d_sigma is None else np.array(d_sigma, np.float64)
    loss = np.zeros(1)
    jvp = np.zeros(1)
    rho_n = np.empty_like(r0)
    rc = lib().oracle_mlem_loss_f64(ctypes.byref(g), _dp(ev), ev.shape[0],
                                    _dp(r0), int(iterations), _dp(tg), _dp(sg),
                                    _dp(dr), _dp(ds), _dp(loss), _dp(jvp),
                                    _dp(rho_n))
    if rc:
        raise RuntimeError("oracle_mlem_loss_f64 failed")
    return float(loss[0]), float(jvp[0]), rho_n


def osem(g, events, n_subsets, rho0=None, iterations=1, mode=MODE_CPU,
         n_threads=0, dtype=np.float32):
    """Ordered-subset iterations (contiguous subsets by event index)."""
    ev = np.ascontiguousarray(events, dtype=dtype)
    shape = (g.n_y_pixels, g.n_z_pixels)
    rho = (np.ones(shape, dtype) if rho0 is None
           else np.ascontiguousarray(rho0, dtype=dtype).reshape(shape).copy())
    m = OracleMode(*mode)
    if dtype == np.float32:
        rc = lib().oracle_osem_f32(ctypes.byref(g), ctypes.byref(m),
                                   _p(ev, ctypes.c_float), ev.shape[0],
                                   _p(rho, ctypes.c_float), iterations,
                                   n_subsets, n_threads)
    else:
        rc = lib().oracle_osem_f64(ctypes.byref(g), ctypes.byref(m),
                                   _p(ev, ctypes.c_double), ev.shape[0],
                                   _p(rho, ctypes.c_double), iterations,
                                   n_subsets, n_threads)
    if rc:
        raise RuntimeError("oracle_osem failed")
    return rho


# ---------------------------------------------------------------------------
# 3D hybrid reconstruction
# ---------------------------------------------------------------------------

def _geometry_args(geo):
    return (np.ascontiguousarray(geo.seg, np.float32),
            np.ascontiguousarray(geo.widths, np.float32),
            np.ascontiguousarray(geo.pixels, np.int16),
            np.ascontiguousarray(geo.positions, np.float32),
            np.ascontiguousarray(geo.weights, np.float32),
            np.ascontiguousarray(geo.info_begin, np.int64),
            np.ascontiguousarray(geo.info_end, np.int64))


def ref_hybrid_mlem(geo, events, iterations=1, rho0=None, weights_mode=1,
                    sens_mode=0, use_3d_sensitivity=False, sens_in=None,
                    n_threads=0):
    """The reference's own PET3D::Hybrid::Reconstruction (oracle/_ref) on
    the geometry arrays of pystrip.hybrid.HybridGeometry.  Returns
    (rho, n_used_events, sensitivity, normalised weights)."""
    L = ref()
    fn = L.ref_hybrid_mlem
    fp, i16p, i64p = (ctypes.POINTER(ctypes.c_float),
                      ctypes.POINTER(ctypes.c_int16),
                      ctypes.POINTER(ctypes.c_int64))
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_int, fp, fp, ctypes.c_int64, i16p, fp, fp,
                   ctypes.c_int, i64p, i64p, ctypes.c_int, ctypes.c_int,
                   ctypes.c_float, ctypes.c_int, ctypes.c_float,
                   ctypes.c_float, ctypes.c_float, ctypes.c_float,
                   ctypes.c_int, ctypes.c_int, ctypes.c_int, fp, ctypes.c_int,
                   fp, ctypes.c_int64, fp, ctypes.c_int, ctypes.c_int, i64p,
                   fp, fp]
    seg, widths, pixels, positions, weights, begin, end = _geometry_args(geo)
    ev = np.ascontiguousarray(events, np.float32)
    rho = (np.ones(geo.shape, np.float32) if rho0 is None
           else np.ascontiguousarray(rho0, np.float32).copy())
    depth = (geo.n_planes_half if geo.n_planes_half > 1
             else (geo.n_planes // 2 if use_3d_sensitivity else 1))
    sens = np.empty((depth, geo.n_rows, geo.n_columns), np.float32)
    sin = (np.ascontiguousarray(sens_in, np.float32) if sens_in is not None
           else None)
    wout = np.empty_like(weights)
    used = ctypes.c_int64()
    rc = fn(geo.n_detectors, _p(seg, ctypes.c_float),
            _p(widths, ctypes.c_float), geo.n_infos,
            _p(pixels, ctypes.c_int16), _p(positions, ctypes.c_float),
            _p(weights, ctypes.c_float), geo.n_planes_half,
            _p(begin, ctypes.c_int64), _p(end, ctypes.c_int64),
            geo.n_columns, geo.n_rows, geo.pixel_size, geo.n_planes,
            geo.z_left, geo.sigma_z, geo.sigma_dl, geo.length, weights_mode,
            sens_mode, int(bool(use_3d_sensitivity)),
            _p(sin, ctypes.c_float) if sin is not None else None, depth,
            _p(ev, ctypes.c_float), ev.shape[0], _p(rho, ctypes.c_float),
            iterations, n_threads, ctypes.byref(used),
            _p(sens, ctypes.c_float), _p(wout, ctypes.c_float))
    if rc:
        raise RuntimeError("ref_hybrid_mlem failed")
    return rho, used.value, sens, wout


def ref_read_geometry(path):
    """The reference's own reader of a binary barrel geometry file
    (PET2D::Barrel::Geometry<float, short>(util::ibstream&) flattened by
    GeometrySOA(geometry, 1), oracle/_ref).  Returns a dict of n_detectors,
    n_columns, n_rows, grid (pixel_size, lower-left x, y) and the arrays
    seg, widths, pixels, positions, weights, info_begin, info_end."""
    L = ref()
    fn = L.ref_read_geometry
    fp, i16p, i64p, ip = (ctypes.POINTER(ctypes.c_float),
                          ctypes.POINTER(ctypes.c_int16),
                          ctypes.POINTER(ctypes.c_int64),
                          ctypes.POINTER(ctypes.c_int))
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_char_p, ip, i64p, ip, ip, fp, fp, fp, i16p, fp,
                   fp, i64p, i64p]
    n_det, n_infos = ctypes.c_int(), ctypes.c_int64()
    n_columns, n_rows = ctypes.c_int(), ctypes.c_int()
    grid = np.zeros(4, np.float32)
    name = os.fsencode(path)
    if fn(name, ctypes.byref(n_det), ctypes.byref(n_infos),
          ctypes.byref(n_columns), ctypes.byref(n_rows),
          _p(grid, ctypes.c_float), None, None, None, None, None, None, None):
        raise RuntimeError(f"ref_read_geometry could not read {path}")
    n_lors = (n_det.value + 1) * n_det.value // 2
    out = dict(seg=np.empty((n_lors, 4), np.float32),
               widths=np.empty(n_lors, np.float32),
               pixels=np.empty((n_infos.value, 2), np.int16),
               positions=np.empty(n_infos.value, np.float32),
               weights=np.empty(n_infos.value, np.float32),
               info_begin=np.empty(n_lors, np.int64),
               info_end=np.empty(n_lors, np.int64))
    if fn(name, ctypes.byref(n_det), ctypes.byref(n_infos),
          ctypes.byref(n_columns), ctypes.byref(n_rows),
          _p(grid, ctypes.c_float), _p(out["seg"], ctypes.c_float),
          _p(out["widths"], ctypes.c_float), _p(out["pixels"], ctypes.c_int16),
          _p(out["positions"], ctypes.c_float),
          _p(out["weights"], ctypes.c_float),
          _p(out["info_begin"], ctypes.c_int64),
          _p(out["info_end"], ctypes.c_int64)):
        raise RuntimeError(f"ref_read_geometry could not read {path}")
    out.update(n_detectors=n_det.value, n_columns=n_columns.value,
               n_rows=n_rows.value, grid=grid[:3].copy())
    return out


class OracleHybrid(ctypes.Structure):
    _fields_ = [("n_detectors", ctypes.c_int), ("n_infos", ctypes.c_int64),
                ("n_planes_half", ctypes.c_int),
                ("seg", ctypes.c_void_p), ("widths", ctypes.c_void_p),
                ("pixels", ctypes.c_void_p), ("positions", ctypes.c_void_p),
                ("info_begin", ctypes.c_void_p), ("info_end", ctypes.c_void_p),
                ("n_columns", ctypes.c_int), ("n_rows", ctypes.c_int),
                ("pixel_size", ctypes.c_double), ("n_planes", ctypes.c_int),
                ("z_left", ctypes.c_double), ("sigma_z", ctypes.c_double),
                ("sigma_dl", ctypes.c_double), ("length", ctypes.c_double)]


def _oracle_hybrid(geo):
    keep = _geometry_args(geo)
    seg, widths, pixels, positions, weights, begin, end = keep
    h = OracleHybrid(geo.n_detectors, geo.n_infos, geo.n_planes_half,
                     seg.ctypes.data, widths.ctypes.data, pixels.ctypes.data,
                     positions.ctypes.data, begin.ctypes.data,
                     end.ctypes.data, geo.n_columns, geo.n_rows,
                     float(np.float32(geo.pixel_size)), geo.n_planes,
                     float(np.float32(geo.z_left)),
                     float(np.float32(geo.sigma_z)),
                     float(np.float32(geo.sigma_dl)),
                     float(np.float32(geo.length)))
    return h, keep


def hybrid_frames(geo, events, weights):
    """translate_to_frame + the box filter (FP32) of every event against
    normalised weights: (ints (n, 5): lor, info begin, info end, plane
    begin, plane end; floats (n, 4): up, right, tan, sec; keep (n,))."""
    h, keep_alive = _oracle_hybrid(geo)
    ev = np.ascontiguousarray(events, np.float32)
    w = np.ascontiguousarray(weights, np.float32)
    n = ev.shape[0]
    ints = np.empty((n, 5), np.int64)
    floats = np.empty((n, 4), np.float32)
    keep = np.empty(n, np.int8)
    fn = lib().oracle_hybrid_frames
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.POINTER(OracleHybrid), ctypes.c_void_p,
                   ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                   ctypes.c_void_p, ctypes.c_void_p]
    if fn(ctypes.byref(h), w.ctypes.data, ev.ctypes.data, n, ints.ctypes.data,
          floats.ctypes.data, keep.ctypes.data):
        raise RuntimeError("oracle_hybrid_frames failed")
    return ints, floats, keep.astype(bool)


def hybrid_mlem(geo, events, iterations=1, rho0=None, weights_mode=1,
                sens_mode=0, sens_in=None, n_threads=0, dtype=np.float32,
                legacy=False):
    """The C restatement of the hybrid reconstruction (strip_oracle.c,
    hybrid_oracle_impl.inc).  Returns (rho, n_used, sensitivity, weights).
    legacy: each visit of a voxel adds its own weight (the CUDA kernel's
    update) instead of the voxel's last cached weight (the CPU path's)."""
    h, keep = _oracle_hybrid(geo)
    seg, widths, pixels, positions, weights, begin, end = keep
    real = ctypes.c_float if dtype == np.float32 else ctypes.c_double
    fn = (lib().oracle_hybrid_mlem_f32 if dtype == np.float32
          else lib().oracle_hybrid_mlem_f64)
    fp = ctypes.POINTER(ctypes.c_float)
    rp = ctypes.POINTER(real)
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.POINTER(OracleHybrid), fp, ctypes.c_int,
                   ctypes.c_int, ctypes.c_int, fp, fp, ctypes.c_int64, rp,
                   ctypes.c_int, ctypes.c_int,
                   ctypes.POINTER(ctypes.c_int64), rp, rp, ctypes.c_int]
    ev = np.ascontiguousarray(events, np.float32)
    rho = (np.ones(geo.shape, dtype) if rho0 is None
           else np.ascontiguousarray(rho0, dtype).copy())
    sens = np.empty((geo.n_rows, geo.n_columns), dtype)
    wout = np.empty(geo.n_infos, dtype)
    sin = (np.ascontiguousarray(sens_in, np.float32) if sens_in is not None
           else None)
    used = ctypes.c_int64()
    rc = fn(ctypes.byref(h), _p(weights, ctypes.c_float), weights_mode,
            sens_mode, 0, _p(sin, ctypes.c_float) if sin is not None else None,
            _p(ev, ctypes.c_float), ev.shape[0], _p(rho, real), iterations,
            n_threads, ctypes.byref(used), _p(sens, real), _p(wout, real),
            int(bool(legacy)))
    if rc:
        raise RuntimeError(f"oracle_hybrid_mlem failed ({rc})")
    return rho, used.value, sens, wout
