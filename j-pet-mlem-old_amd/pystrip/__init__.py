"""Python binding over the C ABI of the MI355X strip MLEM library.

This is a thin ctypes layer over ``include/strip_mlem.h``; every call runs the
HIP kernels in ``lib/libstrip_mlem.so``.  There is no CPU fallback: if the
library is missing, importing :class:`Reconstruction` raises.

The class mirrors the reference's ``PET2D::Strip::Reconstruction`` usage
(src/2d/strip/reconstruction.h:57-154, reconstruction_cmd.cpp:136-292):
construct it from a scanner geometry, load responses (z_u, z_d, dl), run
iteration blocks, read ``rho``.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PACKAGE_ROOT = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PACKAGE_ROOT, "lib", "libstrip_mlem.so")

STRIP_MLEM_OK = 0


class StripMLEMError(RuntimeError):
    """A C-ABI call returned a non-zero status."""


class Geometry(ctypes.Structure):
    """strip_mlem_geometry: Scanner(R, L, n_y, n_z, h, w, sigma_z, sigma_dl)
    (src/2d/strip/scanner.h:39-67)."""

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

    @classmethod
    def make(cls, radius, length, n_z, n_y, pixel_size, sigma_z, sigma_dl,
             pixel_height=None, center_y=0.0, center_z=0.0):
        return cls(float(radius), float(length), int(n_z), int(n_y),
                   float(pixel_size),
                   float(pixel_size if pixel_height is None else pixel_height),
                   float(sigma_z), float(sigma_dl), float(center_y),
                   float(center_z))

    def as_tuple(self):
        return tuple(getattr(self, f) for f, _ in self._fields_)


class Mode(ctypes.Structure):
    """strip_mlem_mode: reconstruction semantics."""

    _fields_ = [
        ("kernel", ctypes.c_int),
        ("sensitivity", ctypes.c_int),
        ("sensitivity_map", ctypes.c_int),
        ("guard_zero", ctypes.c_int),
    ]


class PhantomRegion(ctypes.Structure):
    """strip_phantom_region: kind 0 ellipse (x y a b angle_deg intensity),
    1 rectangle (x y a b intensity; a, b half sizes), 2 point."""

    _fields_ = [("kind", ctypes.c_int), ("x", ctypes.c_double),
                ("y", ctypes.c_double), ("a", ctypes.c_double),
                ("b", ctypes.c_double), ("angle_deg", ctypes.c_double),
                ("intensity", ctypes.c_double)]


class PhantomParams(ctypes.Structure):
    """strip_phantom_params (include/strip_phantom.h)."""

    _fields_ = [("radius", ctypes.c_double),
                ("scintillator_length", ctypes.c_double),
                ("sigma_z", ctypes.c_double), ("sigma_dl", ctypes.c_double),
                ("seed", ctypes.c_uint64), ("additive", ctypes.c_int),
                ("only_detected", ctypes.c_int), ("n_z", ctypes.c_int),
                ("n_y", ctypes.c_int), ("pixel", ctypes.c_double),
                ("image_origin_z", ctypes.c_double),
                ("image_origin_y", ctypes.c_double)]


class PhantomOutputs(ctypes.Structure):
    """strip_phantom_outputs (include/strip_phantom.h)."""

    _fields_ = [("responses", ctypes.c_void_p), ("exact", ctypes.c_void_p),
                ("events", ctypes.c_void_p),
                ("image_emitted", ctypes.c_void_p),
                ("image_detected", ctypes.c_void_p),
                ("image_detected_w_error", ctypes.c_void_p),
                ("tan_bins", ctypes.c_int),
                ("image_tan_bins", ctypes.c_void_p),
                ("n_detected", ctypes.c_int64),
                ("n_emitted", ctypes.c_int64)]


# Reference CPU path (graded parity mode), reconstruction.h:220-298
MODE_CPU = (0, 0, 0, 1)
# Legacy CUDA warp kernel, cuda/reconstruction/warp_granularity.cuh:88-152
MODE_LEGACY_GPU = (1, 1, 1, 0)

_OUTPUT_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int,
                              ctypes.POINTER(ctypes.c_float), ctypes.c_int,
                              ctypes.c_int)
_PROGRESS_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int,
                                ctypes.c_int)
_NAME_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_char_p)

_lib = None


def library_path() -> str:
    return LIB_PATH


def load_library():
    """Loads lib/libstrip_mlem.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise StripMLEMError(
            f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    i64 = ctypes.c_int64
    fptr = ctypes.POINTER(ctypes.c_float)
    dptr = ctypes.POINTER(ctypes.c_double)
    sigs = {
        "strip_mlem_create": (ctypes.c_int, [ctypes.POINTER(Geometry),
                                             ctypes.POINTER(Mode), ctypes.c_int,
                                             ctypes.POINTER(P)]),
        "strip_mlem_destroy": (None, [P]),
        "strip_mlem_set_events": (ctypes.c_int, [P, fptr, i64]),
        "strip_mlem_set_events_device": (ctypes.c_int, [P, P, i64]),
        "strip_mlem_n_events": (i64, [P]),
        "strip_mlem_set_rho": (ctypes.c_int, [P, fptr]),
        "strip_mlem_get_rho": (ctypes.c_int, [P, fptr]),
        "strip_mlem_set_rho_device": (ctypes.c_int, [P, P]),
        "strip_mlem_get_rho_device": (ctypes.c_int, [P, P]),
        "strip_mlem_get_sensitivity": (ctypes.c_int, [P, ctypes.c_int, fptr]),
        "strip_mlem_iterate": (ctypes.c_int, [P, ctypes.c_int]),
        "strip_mlem_iterate_async": (ctypes.c_int, [P, ctypes.c_int]),
        "strip_mlem_synchronize": (ctypes.c_int, [P]),
        "strip_mlem_set_launch": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int]),
        "strip_mlem_set_algorithm": (ctypes.c_int, [P, ctypes.c_int]),
        "strip_mlem_get_algorithm": (ctypes.c_int, [P]),
        "strip_mlem_set_band_tiling": (ctypes.c_int, [P] + [ctypes.c_int] * 5),
        "strip_mlem_set_exp_mode": (ctypes.c_int, [P, ctypes.c_int]),
        "strip_mlem_set_span_table": (ctypes.c_int, [P, ctypes.c_int]),
        "strip_phantom_generate": (ctypes.c_int, [
            ctypes.c_int, ctypes.POINTER(PhantomParams),
            ctypes.POINTER(PhantomRegion), ctypes.c_int, i64, fptr, fptr,
            ctypes.POINTER(i64), ctypes.POINTER(i64), ctypes.c_void_p,
            ctypes.c_void_p, ctypes.c_void_p]),
        "strip_phantom_generate_full": (ctypes.c_int, [
            ctypes.c_int, ctypes.POINTER(PhantomParams),
            ctypes.POINTER(PhantomRegion), ctypes.c_int, i64,
            ctypes.POINTER(PhantomOutputs)]),
        "strip_phantom_tan_kernel_map": (ctypes.c_int, [
            ctypes.c_int, ctypes.POINTER(PhantomParams), ctypes.c_int,
            ctypes.c_double, ctypes.c_double, fptr]),
        "strip_mlem_generate_events": (ctypes.c_int, [
            P, ctypes.POINTER(PhantomParams), ctypes.POINTER(PhantomRegion),
            ctypes.c_int, i64]),
        "strip_phantom_random_blocks": (ctypes.c_int, [
            ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
            ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p]),
        "strip_phantom_last_error": (ctypes.c_char_p, []),
        "strip_mlem_set_backprojection_waves": (ctypes.c_int, [P, ctypes.c_int]),
        "strip_mlem_set_backprojection_occupancy": (ctypes.c_int,
                                                    [P, ctypes.c_int]),
        "strip_mlem_batch_set_backprojection_occupancy": (ctypes.c_int,
                                                          [P, ctypes.c_int]),
        "strip_mlem_batch_get_backprojection_occupancy": (
            ctypes.c_int, [P, ctypes.POINTER(ctypes.c_int)]),
        "strip_mlem_set_band_groups": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int]),
        "strip_mlem_get_band_groups": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_int),
                                                      ctypes.POINTER(ctypes.c_int)]),
        "strip_mlem_batch_set_groups": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int]),
        "strip_mlem_batch_get_groups": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_int),
                                                       ctypes.POINTER(ctypes.c_int)]),
        "strip_mlem_reserve_events": (ctypes.c_int, [P, i64]),
        "strip_mlem_append_events": (ctypes.c_int, [P, fptr, i64]),
        "strip_mlem_set_subsets": (ctypes.c_int, [P, ctypes.c_int]),
        "strip_mlem_set_plan_events": (ctypes.c_int, [P, i64]),
        "strip_mlem_image_total": (ctypes.c_int, [P, dptr]),
        "strip_mlem_batch_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(P)]),
        "strip_mlem_batch_destroy": (None, [P]),
        "strip_mlem_batch_add": (ctypes.c_int, [P, ctypes.POINTER(Geometry), fptr, i64,
                                                ctypes.POINTER(ctypes.c_int)]),
        "strip_mlem_batch_size": (ctypes.c_int, [P]),
        "strip_mlem_batch_context": (ctypes.c_int, [P, ctypes.c_int, ctypes.POINTER(P)]),
        "strip_mlem_batch_iterate": (ctypes.c_int, [P, ctypes.c_int]),
        "strip_mlem_batch_set_rho": (ctypes.c_int, [P, ctypes.c_int, fptr]),
        "strip_mlem_batch_get_rho": (ctypes.c_int, [P, ctypes.c_int, fptr]),
        "strip_mlem_batch_set_tiling": (ctypes.c_int, [P] + [ctypes.c_int] * 6),
        "strip_mlem_batch_get_tiling": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_int)]),
        "strip_mlem_batch_set_graph": (ctypes.c_int, [P, ctypes.c_int]),
        "strip_mlem_batch_last_ms": (ctypes.c_int, [P, dptr]),
        "strip_mlem_batch_autotune": (ctypes.c_int, [P, dptr]),
        "strip_mlem_loglik_gradient": (ctypes.c_int, [P, fptr, dptr, fptr, dptr]),
        "strip_mlem_loglik_gradient_device": (ctypes.c_int, [P, P, dptr, P, dptr]),
        "strip_mlem_unrolled_gradient": (ctypes.c_int, [P, fptr, ctypes.c_int, fptr,
                                                        dptr, fptr, dptr]),
        "strip_mlem_unrolled_gradient_device": (ctypes.c_int, [P, P, ctypes.c_int, P,
                                                               dptr, P, dptr]),
        "strip_mlem_get_span_table": (ctypes.c_int,
                                      [P, ctypes.POINTER(ctypes.c_int)]),
        "strip_mlem_get_band_tiling": (ctypes.c_int, [
            P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
            ctypes.POINTER(ctypes.c_int), ctypes.POINTER(i64),
            ctypes.POINTER(i64)]),
        "strip_mlem_kernel_evaluations": (ctypes.c_int,
                                          [P, ctypes.POINTER(i64)]),
        "strip_mlem_forward": (ctypes.c_int, [P, fptr]),
        "strip_mlem_forward_noalloc": (ctypes.c_int, [
            ctypes.POINTER(Geometry), ctypes.POINTER(Mode), ctypes.c_void_p,
            i64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
        "strip_mlem_set_timing": (ctypes.c_int, [P, ctypes.c_int]),
        "strip_mlem_get_timing": (ctypes.c_int, [P,
                                                 ctypes.POINTER(ctypes.c_double),
                                                 ctypes.POINTER(ctypes.c_double),
                                                 ctypes.POINTER(i64)]),
        "strip_mlem_get_timing_detail": (ctypes.c_int, [
            P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
            ctypes.POINTER(ctypes.c_double), ctypes.POINTER(i64)]),
        "strip_mlem_device_name": (ctypes.c_int, [P, ctypes.c_char_p,
                                                  ctypes.c_int]),
        "strip_mlem_last_error": (ctypes.c_char_p, []),
        "strip_mlem_run": (ctypes.c_int, [ctypes.POINTER(Geometry),
                                          ctypes.POINTER(Mode), fptr,
                                          ctypes.c_int, fptr, ctypes.c_int,
                                          ctypes.c_int, _OUTPUT_FN,
                                          _PROGRESS_FN, P, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int,
                                          _NAME_FN]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    """Names of every C-ABI entry point the binding expects."""
    return [
        "strip_mlem_create", "strip_mlem_destroy", "strip_mlem_set_events",
        "strip_mlem_set_events_device", "strip_mlem_n_events",
        "strip_mlem_set_rho", "strip_mlem_get_rho", "strip_mlem_set_rho_device",
        "strip_mlem_get_rho_device", "strip_mlem_get_sensitivity",
        "strip_mlem_iterate", "strip_mlem_iterate_async",
        "strip_mlem_synchronize", "strip_mlem_set_launch",
        "strip_mlem_set_algorithm", "strip_mlem_get_algorithm",
        "strip_mlem_set_band_tiling", "strip_mlem_get_band_tiling",
        "strip_mlem_set_exp_mode", "strip_mlem_set_span_table",
        "strip_mlem_set_backprojection_waves",
        "strip_mlem_set_backprojection_occupancy",
        "strip_mlem_batch_set_backprojection_occupancy",
        "strip_mlem_batch_get_backprojection_occupancy",
        "strip_mlem_set_band_groups", "strip_mlem_get_band_groups",
        "strip_mlem_batch_set_groups", "strip_mlem_batch_get_groups",
        "strip_mlem_get_span_table", "strip_mlem_loglik_gradient",
        "strip_mlem_loglik_gradient_device", "strip_mlem_unrolled_gradient",
        "strip_mlem_unrolled_gradient_device", "strip_mlem_batch_create",
        "strip_mlem_batch_destroy", "strip_mlem_batch_add",
        "strip_mlem_batch_size", "strip_mlem_batch_context",
        "strip_mlem_batch_iterate", "strip_mlem_batch_set_rho",
        "strip_mlem_batch_get_rho", "strip_mlem_batch_set_tiling",
        "strip_mlem_batch_get_tiling", "strip_mlem_batch_set_graph",
        "strip_mlem_batch_last_ms", "strip_mlem_batch_autotune",
        "strip_mlem_reserve_events", "strip_mlem_append_events",
        "strip_phantom_generate", "strip_mlem_generate_events",
        "strip_phantom_generate_full", "strip_phantom_tan_kernel_map",
        "strip_phantom_random_blocks", "strip_phantom_last_error",
        "strip_mlem_set_subsets", "strip_mlem_image_total",
        "strip_mlem_set_plan_events",
        "strip_mlem_kernel_evaluations", "strip_mlem_forward",
        "strip_mlem_forward_noalloc",
        "strip_mlem_set_timing", "strip_mlem_get_timing",
        "strip_mlem_get_timing_detail",
        "strip_mlem_device_name", "strip_mlem_last_error", "strip_mlem_run",
    ]


def _library_loaded() -> bool:
    return _lib is not None


def _check(rc):
    if rc != STRIP_MLEM_OK:
        msg = _lib.strip_mlem_last_error().decode(errors="replace")
        raise StripMLEMError(f"status {rc}: {msg}")


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _as_events(responses) -> np.ndarray:
    ev = np.ascontiguousarray(responses, dtype=np.float32)
    if ev.ndim != 2 or ev.shape[1] != 3:
        raise ValueError("responses must be an (N, 3) array of z_u, z_d, dl")
    return ev


class Reconstruction:
    """GPU reconstruction context (one HIP stream, resident events/images)."""

    def __init__(self, geometry: Geometry, mode=MODE_CPU, device: int = 0):
        self._lib = load_library()
        self.geometry = geometry
        self.mode = Mode(*mode)
        self._ctx = ctypes.c_void_p()
        _check(self._lib.strip_mlem_create(ctypes.byref(geometry),
                                           ctypes.byref(self.mode), device,
                                           ctypes.byref(self._ctx)))
        self.shape = (geometry.n_y_pixels, geometry.n_z_pixels)

    def close(self):
        if self._ctx:
            self._lib.strip_mlem_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- events / images --------------------------------------------------
    def set_events(self, responses):
        ev = _as_events(responses)
        _check(self._lib.strip_mlem_set_events(self._ctx, _fptr(ev),
                                               ev.shape[0]))

    def set_events_device(self, device_ptr: int, n: int):
        _check(self._lib.strip_mlem_set_events_device(
            self._ctx, ctypes.c_void_p(device_ptr), int(n)))

    @property
    def n_events(self) -> int:
        return int(self._lib.strip_mlem_n_events(self._ctx))

    @property
    def rho(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float32)
        _check(self._lib.strip_mlem_get_rho(self._ctx, _fptr(out)))
        return out

    @rho.setter
    def rho(self, value):
        v = np.ascontiguousarray(value, dtype=np.float32).reshape(self.shape)
        _check(self._lib.strip_mlem_set_rho(self._ctx, _fptr(v)))

    def set_rho_device(self, device_ptr: int):
        _check(self._lib.strip_mlem_set_rho_device(self._ctx,
                                                   ctypes.c_void_p(device_ptr)))

    def get_rho_device(self, device_ptr: int):
        _check(self._lib.strip_mlem_get_rho_device(self._ctx,
                                                   ctypes.c_void_p(device_ptr)))

    def sensitivity(self, kind: int = 0) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float32)
        _check(self._lib.strip_mlem_get_sensitivity(self._ctx, kind,
                                                    _fptr(out)))
        return out

    # -- iterations --------------------------------------------------------
    def iterate(self, n: int = 1):
        _check(self._lib.strip_mlem_iterate(self._ctx, int(n)))

    def iterate_async(self, n: int = 1):
        _check(self._lib.strip_mlem_iterate_async(self._ctx, int(n)))

    def synchronize(self):
        _check(self._lib.strip_mlem_synchronize(self._ctx))

    def set_launch(self, n_blocks: int = 0, threads_per_block: int = 0):
        _check(self._lib.strip_mlem_set_launch(self._ctx, n_blocks,
                                               threads_per_block))

    def set_algorithm(self, algorithm: int):
        _check(self._lib.strip_mlem_set_algorithm(self._ctx, algorithm))

    def set_band_tiling(self, forward_rows=32, backward_rows=16,
                        bucket_columns=8, forward_pairs=2048,
                        backward_pairs=512):
        _check(self._lib.strip_mlem_set_band_tiling(
            self._ctx, forward_rows, backward_rows, bucket_columns,
            forward_pairs, backward_pairs))

    def set_exp_mode(self, mode: int):
        """0 = exp2 per pixel, 1 = restarted exp recurrence along rows."""
        _check(self._lib.strip_mlem_set_exp_mode(self._ctx, int(mode)))

    def set_band_groups(self, forward_group: int = 0, backward_group: int = 0):
        """(Event, band) pairs one wavefront processes together in the
        forward / back-projection kernels (1, 2 or 4; 0 = automatic)."""
        _check(self._lib.strip_mlem_set_band_groups(
            self._ctx, int(forward_group), int(backward_group)))

    def band_groups(self):
        """(forward, back-projection) pairs per wavefront step in use."""
        f, b = ctypes.c_int(), ctypes.c_int()
        _check(self._lib.strip_mlem_get_band_groups(
            self._ctx, ctypes.byref(f), ctypes.byref(b)))
        return f.value, b.value

    def set_backprojection_occupancy(self, cu_waves: int):
        """Back-projection wavefronts per CU (4..32) whose LDS windows the
        automatic groups keep resident."""
        _check(self._lib.strip_mlem_set_backprojection_occupancy(
            self._ctx, int(cu_waves)))

    def set_backprojection_waves(self, waves: int):
        """Wavefronts per back-projection workgroup (0 = automatic)."""
        _check(self._lib.strip_mlem_set_backprojection_waves(self._ctx,
                                                             int(waves)))

    def set_span_table(self, enable: bool):
        """Store every event's per-row in-ellipse runs once per upload (True)
        or solve them in every iteration (False)."""
        _check(self._lib.strip_mlem_set_span_table(self._ctx, int(bool(enable))))

    def span_table(self) -> bool:
        """Whether the plan in use has the span table."""
        v = ctypes.c_int()
        _check(self._lib.strip_mlem_get_span_table(self._ctx, ctypes.byref(v)))
        return bool(v.value)

    def band_tiling(self):
        """(forward rows, back-projection rows, window columns,
        forward pairs, back-projection pairs)."""
        hf, hb, w = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        nf, nb = ctypes.c_int64(), ctypes.c_int64()
        _check(self._lib.strip_mlem_get_band_tiling(
            self._ctx, ctypes.byref(hf), ctypes.byref(hb), ctypes.byref(w),
            ctypes.byref(nf), ctypes.byref(nb)))
        return hf.value, hb.value, w.value, nf.value, nb.value

    @property
    def algorithm(self) -> int:
        return int(self._lib.strip_mlem_get_algorithm(self._ctx))

    def kernel_evaluations(self) -> int:
        v = ctypes.c_int64()
        _check(self._lib.strip_mlem_kernel_evaluations(self._ctx,
                                                       ctypes.byref(v)))
        return int(v.value)

    def forward(self) -> np.ndarray:
        out = np.empty(self.n_events, dtype=np.float32)
        _check(self._lib.strip_mlem_forward(self._ctx, _fptr(out)))
        return out

    # ---- streaming and ordered subsets ------------------------------------

    def reserve_events(self, n: int):
        _check(self._lib.strip_mlem_reserve_events(self._ctx, int(n)))

    def append_events(self, events):
        ev = np.ascontiguousarray(events, dtype=np.float32).reshape(-1, 3)
        _check(self._lib.strip_mlem_append_events(self._ctx, _fptr(ev),
                                                  ev.shape[0]))

    def set_subsets(self, n_subsets: int):
        """Ordered subsets by event index (1 = plain MLEM)."""
        _check(self._lib.strip_mlem_set_subsets(self._ctx, int(n_subsets)))

    def set_plan_events(self, n: int):
        """Events per row-band plan of plain MLEM; more are split over
        several plans that accumulate one update."""
        _check(self._lib.strip_mlem_set_plan_events(self._ctx, int(n)))

    def image_total(self) -> float:
        """Sum of the image's pixels, accumulated in FP64 on the device."""
        v = ctypes.c_double()
        _check(self._lib.strip_mlem_image_total(self._ctx, ctypes.byref(v)))
        return v.value

    def generate_events(self, regions, n_detected, seed=0, additive=False):
        """Appends n_detected events generated on the device from the
        phantom `regions` (phantom.Ellipse list or PhantomRegion list)."""
        arr = _regions_array(regions)
        p = PhantomParams()
        p.seed = int(seed)
        p.additive = int(bool(additive))
        p.only_detected = 1
        _check(self._lib.strip_mlem_generate_events(
            self._ctx, ctypes.byref(p), arr, len(arr), int(n_detected)))

    # ---- differentiation -------------------------------------------------

    def loglik_gradient(self, rho=None, want_grad_rho=True):
        """Poisson log-likelihood sum_{D>0} log D_e - sum rho of `rho`
        (default: the current image) with d/d rho and d/d (sigma_z, sigma_dl),
        the event supports held fixed.  Returns (value, grad_rho, grad_sigma).
        """
        value = ctypes.c_double()
        gs = (ctypes.c_double * 2)()
        r = None
        if rho is not None:
            r = np.ascontiguousarray(rho, dtype=np.float32).reshape(self.shape)
        g = np.empty(self.shape, np.float32) if want_grad_rho else None
        _check(self._lib.strip_mlem_loglik_gradient(
            self._ctx, _fptr(r) if r is not None else None,
            ctypes.byref(value), _fptr(g) if g is not None else None, gs))
        return value.value, g, np.array(list(gs))

    def unrolled_gradient(self, rho0, iterations, target, want_grad_rho=True):
        """Loss mean((rho_N - target)^2) after `iterations` MLEM updates from
        rho0 (None: the current image) and its gradients with respect to rho0
        and (sigma_z, sigma_dl).  Returns (loss, grad_rho0, grad_sigma); the
        image becomes rho_N."""
        loss = ctypes.c_double()
        gs = (ctypes.c_double * 2)()
        r = None
        if rho0 is not None:
            r = np.ascontiguousarray(rho0, dtype=np.float32).reshape(self.shape)
        t = np.ascontiguousarray(target, dtype=np.float32).reshape(self.shape)
        g = np.empty(self.shape, np.float32) if want_grad_rho else None
        _check(self._lib.strip_mlem_unrolled_gradient(
            self._ctx, _fptr(r) if r is not None else None, int(iterations),
            _fptr(t), ctypes.byref(loss), _fptr(g) if g is not None else None,
            gs))
        return loss.value, g, np.array(list(gs))

    def loglik_gradient_device(self, rho_ptr, grad_ptr):
        """Device-pointer form (row-major float32 images on this context's
        device; rho_ptr 0 = the current image, grad_ptr 0 = no image
        gradient).  Returns (value, grad_sigma)."""
        value = ctypes.c_double()
        gs = (ctypes.c_double * 2)()
        _check(self._lib.strip_mlem_loglik_gradient_device(
            self._ctx, ctypes.c_void_p(rho_ptr or None), ctypes.byref(value),
            ctypes.c_void_p(grad_ptr or None), gs))
        return value.value, np.array(list(gs))

    def unrolled_gradient_device(self, rho0_ptr, iterations, target_ptr,
                                 grad_ptr):
        """Device-pointer form of unrolled_gradient.  Returns
        (loss, grad_sigma)."""
        loss = ctypes.c_double()
        gs = (ctypes.c_double * 2)()
        _check(self._lib.strip_mlem_unrolled_gradient_device(
            self._ctx, ctypes.c_void_p(rho0_ptr or None), int(iterations),
            ctypes.c_void_p(target_ptr), ctypes.byref(loss),
            ctypes.c_void_p(grad_ptr or None), gs))
        return loss.value, np.array(list(gs))

    def set_timing(self, enable: bool):
        _check(self._lib.strip_mlem_set_timing(self._ctx, int(bool(enable))))

    def get_timing(self):
        k = ctypes.c_double()
        t = ctypes.c_double()
        n = ctypes.c_int64()
        _check(self._lib.strip_mlem_get_timing(self._ctx, ctypes.byref(k),
                                               ctypes.byref(t),
                                               ctypes.byref(n)))
        return k.value, t.value, n.value

    def get_timing_detail(self):
        """(forward_ms, backward_ms, iteration_ms, n_iterations) summed over
        the timed iterations since the last call."""
        f = ctypes.c_double()
        b = ctypes.c_double()
        t = ctypes.c_double()
        n = ctypes.c_int64()
        _check(self._lib.strip_mlem_get_timing_detail(
            self._ctx, ctypes.byref(f), ctypes.byref(b), ctypes.byref(t),
            ctypes.byref(n)))
        return f.value, b.value, t.value, n.value

    def device_name(self) -> str:
        buf = ctypes.create_string_buffer(256)
        _check(self._lib.strip_mlem_device_name(self._ctx, buf, 256))
        return buf.value.decode(errors="replace")


NOALLOC_MAX_EVENTS = 1 << 18  # STRIP_MLEM_NOALLOC_MAX_EVENTS


def forward_noalloc(geometry: Geometry, responses_ptr: int, n: int,
                    rho_ptr: int, denominators_ptr: int, mode=None,
                    stream: int = 0):
    """strip_mlem_forward_noalloc: forward projections of n device-resident
    responses through a device-resident row-major image into device memory,
    asynchronously on `stream` of the current device, allocating nothing."""
    lib = load_library()
    m = None if mode is None else ctypes.byref(Mode(*mode))
    _check(lib.strip_mlem_forward_noalloc(
        ctypes.byref(geometry), m, ctypes.c_void_p(responses_ptr), int(n),
        ctypes.c_void_p(rho_ptr), ctypes.c_void_p(denominators_ptr),
        ctypes.c_void_p(stream) if stream else None))


def run(geometry: Geometry, responses, rho, n_blocks: int,
        n_iterations_in_block: int, mode=MODE_CPU, device: int = 0,
        output=None, progress=None, device_name=None, cuda_blocks: int = 0,
        cuda_threads: int = 0):
    """The reference's GPU::Reconstruction::run protocol
    (src/2d/strip/cuda/reconstruction.cu:31-113) over strip_mlem_run.
    ``output(iteration, image)`` receives a copy of each output image."""
    lib = load_library()
    ev = _as_events(responses)
    r = np.ascontiguousarray(rho, dtype=np.float32).reshape(
        geometry.n_y_pixels, geometry.n_z_pixels).copy()

    def _out(_user, it, img, nz, ny):
        if output is not None:
            arr = np.ctypeslib.as_array(img, shape=(ny * nz,)).reshape(ny, nz)
            output(it, arr.copy())

    def _prog(_user, completed, finished):
        if progress is not None:
            progress(completed, bool(finished))

    def _name(_user, name):
        if device_name is not None:
            device_name(name.decode(errors="replace"))

    cb_out, cb_prog, cb_name = _OUTPUT_FN(_out), _PROGRESS_FN(_prog), \
        _NAME_FN(_name)
    m = Mode(*mode)
    _check(lib.strip_mlem_run(ctypes.byref(geometry), ctypes.byref(m),
                              _fptr(ev), ev.shape[0], _fptr(r), n_blocks,
                              n_iterations_in_block, cb_out, cb_prog, None,
                              device, cuda_blocks, cuda_threads, cb_name))
    return r


class _MemberView(Reconstruction):
    """A batch member's own context, owned by the batch (valid while the
    batch lives)."""

    def __init__(self, lib, ctx, geometry):  # noqa: D401 - no new context
        self._lib = lib
        self._ctx = ctx
        self.geometry = geometry
        self.shape = (geometry.n_y_pixels, geometry.n_z_pixels)

    def close(self):
        self._ctx = ctypes.c_void_p()


class Batch:
    """Many independent reconstructions iterated together
    (strip_mlem_batch_*): one launch per pass over all members, iterations
    replayed from a HIP graph."""

    def __init__(self, device: int = 0):
        self._lib = load_library()
        self._b = ctypes.c_void_p()
        _check(self._lib.strip_mlem_batch_create(device, ctypes.byref(self._b)))
        self.geometries = []

    def add(self, geometry: "Geometry", events) -> int:
        ev = np.ascontiguousarray(events, dtype=np.float32).reshape(-1, 3)
        idx = ctypes.c_int()
        _check(self._lib.strip_mlem_batch_add(self._b, ctypes.byref(geometry),
                                              _fptr(ev), ev.shape[0],
                                              ctypes.byref(idx)))
        self.geometries.append(geometry)
        return idx.value

    def __len__(self):
        return int(self._lib.strip_mlem_batch_size(self._b))

    def iterate(self, n: int = 1):
        _check(self._lib.strip_mlem_batch_iterate(self._b, int(n)))

    def last_ms(self) -> float:
        v = ctypes.c_double()
        _check(self._lib.strip_mlem_batch_last_ms(self._b, ctypes.byref(v)))
        return v.value

    def rho(self, index: int) -> np.ndarray:
        g = self.geometries[index]
        out = np.empty((g.n_y_pixels, g.n_z_pixels), np.float32)
        _check(self._lib.strip_mlem_batch_get_rho(self._b, index, _fptr(out)))
        return out

    def set_rho(self, index: int, value):
        g = self.geometries[index]
        v = np.ascontiguousarray(value, np.float32).reshape(
            g.n_y_pixels, g.n_z_pixels)
        _check(self._lib.strip_mlem_batch_set_rho(self._b, index, _fptr(v)))

    def member(self, index: int) -> Reconstruction:
        """The member's own context: its iterate() is the plain form."""
        c = ctypes.c_void_p()
        _check(self._lib.strip_mlem_batch_context(self._b, index,
                                                  ctypes.byref(c)))
        return _MemberView(self._lib, c, self.geometries[index])

    def set_tiling(self, forward_rows=32, backward_rows=16, bucket_columns=8,
                   forward_pairs=2048, backward_pairs=512, grid_factor=4):
        _check(self._lib.strip_mlem_batch_set_tiling(
            self._b, forward_rows, backward_rows, bucket_columns,
            forward_pairs, backward_pairs, grid_factor))

    def tiling(self):
        t = (ctypes.c_int * 6)()
        _check(self._lib.strip_mlem_batch_get_tiling(self._b, t))
        return tuple(t)

    def set_groups(self, forward_group=0, backward_group=0):
        _check(self._lib.strip_mlem_batch_set_groups(
            self._b, int(forward_group), int(backward_group)))

    def groups(self):
        f, b = ctypes.c_int(), ctypes.c_int()
        _check(self._lib.strip_mlem_batch_get_groups(
            self._b, ctypes.byref(f), ctypes.byref(b)))
        return f.value, b.value

    def set_backprojection_occupancy(self, cu_waves: int):
        _check(self._lib.strip_mlem_batch_set_backprojection_occupancy(
            self._b, int(cu_waves)))

    def backprojection_occupancy(self) -> int:
        v = ctypes.c_int()
        _check(self._lib.strip_mlem_batch_get_backprojection_occupancy(
            self._b, ctypes.byref(v)))
        return v.value

    def set_graph(self, enable: bool):
        _check(self._lib.strip_mlem_batch_set_graph(self._b, int(bool(enable))))

    def autotune(self) -> float:
        v = ctypes.c_double()
        _check(self._lib.strip_mlem_batch_autotune(self._b, ctypes.byref(v)))
        return v.value

    def close(self):
        if self._b:
            self._lib.strip_mlem_batch_destroy(self._b)
            self._b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _regions_array(regions):
    regions = list(regions)
    arr = (PhantomRegion * len(regions))()
    for i, r in enumerate(regions):
        if isinstance(r, PhantomRegion):
            arr[i] = r
        else:  # phantom.Ellipse
            arr[i] = PhantomRegion(0, r.x, r.y, r.a, r.b, r.angle, r.intensity)
    return arr


def generate_events(regions, radius, length, sigma_z, sigma_dl, n, seed=0,
                    additive=False, only_detected=True, exact=False,
                    image_grid=None, device=0):
    """Events generated on the GPU (strip_phantom_generate).

    Returns dict(responses=(m, 3) float32, exact=(m, 3) or None, n_emitted,
    images=(emitted, detected, detected_w_error) or None), m = detected
    events (= n with only_detected).  image_grid = (n_z, n_y, pixel[, origin_z,
    origin_y]) requests the three histograms."""
    lib = load_library()
    arr = _regions_array(regions)
    p = PhantomParams(float(radius), float(length), float(sigma_z),
                      float(sigma_dl), int(seed), int(bool(additive)),
                      int(bool(only_detected)), 0, 0, 0.0, 0.0, 0.0)
    imgs = None
    ptrs = [None, None, None]
    if image_grid is not None:
        g = list(image_grid) + [0.0] * (5 - len(image_grid))
        p.n_z, p.n_y, p.pixel = int(g[0]), int(g[1]), float(g[2])
        p.image_origin_z, p.image_origin_y = float(g[3]), float(g[4])
        imgs = [np.zeros((p.n_y, p.n_z), np.int32) for _ in range(3)]
        ptrs = [im.ctypes.data_as(ctypes.c_void_p) for im in imgs]
    resp = np.empty((max(int(n), 1), 3), np.float32)
    ex = np.empty_like(resp) if exact else None
    nd, ne = ctypes.c_int64(), ctypes.c_int64()
    rc = lib.strip_phantom_generate(
        device, ctypes.byref(p), arr, len(arr), int(n), _fptr(resp),
        _fptr(ex) if ex is not None else None, ctypes.byref(nd),
        ctypes.byref(ne), *ptrs)
    if rc != 0:
        raise StripMLEMError(
            f"status {rc}: {lib.strip_phantom_last_error().decode()}")
    m = nd.value
    return {"responses": resp[:m], "exact": None if ex is None else ex[:m],
            "n_emitted": ne.value, "images": imgs}


def _phantom_params(regions, radius, length, sigma_z, sigma_dl, seed,
                    additive, only_detected, image_grid):
    p = PhantomParams(float(radius), float(length), float(sigma_z),
                      float(sigma_dl), int(seed), int(bool(additive)),
                      int(bool(only_detected)), 0, 0, 0.0, 0.0, 0.0)
    if image_grid is not None:
        g = list(image_grid) + [0.0] * (5 - len(image_grid))
        p.n_z, p.n_y, p.pixel = int(g[0]), int(g[1]), float(g[2])
        p.image_origin_z, p.image_origin_y = float(g[3]), float(g[4])
    return p


def generate_full(regions, radius, length, sigma_z, sigma_dl, n, image_grid,
                  seed=0, additive=False, only_detected=True, tan_bins=0,
                  device=0):
    """Every output of the phantom tool (strip_phantom_generate_full):
    dict(responses, exact, events (m, 4): emission point and direction,
    images (emitted, detected, detected_w_error), tan_bins (tan_bins, n_y,
    n_z) int32 or None, n_emitted)."""
    lib = load_library()
    arr = _regions_array(regions)
    p = _phantom_params(regions, radius, length, sigma_z, sigma_dl, seed,
                        additive, only_detected, image_grid)
    m1 = max(int(n), 1)
    resp = np.empty((m1, 3), np.float32)
    ex = np.empty((m1, 3), np.float32)
    evs = np.empty((m1, 4), np.float32)
    imgs = [np.zeros((p.n_y, p.n_z), np.int32) for _ in range(3)]
    tan = (np.zeros((tan_bins, p.n_y, p.n_z), np.int32) if tan_bins > 0
           else None)
    o = PhantomOutputs()
    vp = ctypes.c_void_p
    o.responses = resp.ctypes.data_as(vp)
    o.exact = ex.ctypes.data_as(vp)
    o.events = evs.ctypes.data_as(vp)
    o.image_emitted = imgs[0].ctypes.data_as(vp)
    o.image_detected = imgs[1].ctypes.data_as(vp)
    o.image_detected_w_error = imgs[2].ctypes.data_as(vp)
    o.tan_bins = int(tan_bins)
    o.image_tan_bins = tan.ctypes.data_as(vp) if tan is not None else None
    rc = lib.strip_phantom_generate_full(device, ctypes.byref(p), arr,
                                         len(arr), int(n), ctypes.byref(o))
    if rc != 0:
        raise StripMLEMError(
            f"status {rc}: {lib.strip_phantom_last_error().decode()}")
    m = o.n_detected
    return {"responses": resp[:m], "exact": ex[:m], "events": evs[:m],
            "images": imgs, "tan_bins": tan, "n_emitted": o.n_emitted}


def tan_kernel_map(radius, length, sigma_z, sigma_dl, image_grid, tan_bins,
                   kernel_point=(0.0, 0.0), only_detected=True, device=0):
    """The phantom tool's --tan-bins kernel map (tan_bins, n_y, n_z) float32
    (strip_phantom_tan_kernel_map)."""
    lib = load_library()
    p = _phantom_params([], radius, length, sigma_z, sigma_dl, 0, False,
                        only_detected, image_grid)
    out = np.empty((tan_bins, p.n_y, p.n_z), np.float32)
    rc = lib.strip_phantom_tan_kernel_map(device, ctypes.byref(p),
                                          int(tan_bins),
                                          float(kernel_point[0]),
                                          float(kernel_point[1]), _fptr(out))
    if rc != 0:
        raise StripMLEMError(
            f"status {rc}: {lib.strip_phantom_last_error().decode()}")
    return out


def random_blocks(key0, key1, first, c1, c2, n_blocks, device=0):
    """Raw Philox4x64-10 blocks of the generator's stream."""
    lib = load_library()
    out = np.empty((n_blocks, 4), np.uint64)
    rc = lib.strip_phantom_random_blocks(device, key0, key1, first, c1, c2,
                                         n_blocks,
                                         out.ctypes.data_as(ctypes.c_void_p))
    if rc != 0:
        raise StripMLEMError(lib.strip_phantom_last_error().decode())
    return out
