"""Synthetic inputs of the 3D hybrid reconstruction: a ring scanner's LOR
geometry in the reference's GeometrySOA layout and seeded list-mode events.

Geometry (src/2d/barrel/geometry_soa.h:262-287): n_detectors detector
centres on a circle; every LOR (first >= second, index first (first + 1) / 2
+ second, src/2d/barrel/lor.h:21-38) is the segment from the second
detector's centre to the first's (geometry_soa.h:139-141); its pixel infos
are the pixels of the grid whose centres lie within half the LOR width of
the segment's line and inside the field of view, sorted by their scaled
projection t along the segment (geometry_soa.h:146-170), stored LOR after
LOR in index order, followed by one padding info (the grid's centre pixel,
weight 0, position 2) that belongs to no LOR: the reference's event setup
takes the info after the first position above the box's far end
(reconstruction.h:131-138), which for the last LOR with infos is one past
the arrays.  The reference builds this from a Monte-Carlo system
matrix or a Boost.Geometry LOR-pixel intersection (2d_barrel_matrix,
2d_barrel_geometry), neither of which is part of this build; the tube of
pixel centres is this module's own stand-in with the same layout.

Events (the Response of src/3d/hybrid/scanner.h: lor, z_up, z_dn, dl): an
emission point drawn in a cylinder phantom, a random transaxial direction
whose line meets the ring at the two detectors nearest to its crossings,
an axial slope uniform within the acceptance; projected like the strip
(scanner.h:107-113 of 2d/strip with y = the position along the LOR from its
middle): z_up = z + (R - up) tan at the first detector, z_dn = z - (R + up)
tan at the second, dl = -2 up sqrt(1 + tan^2), R half the LOR length; kept
when |z_up|, |z_dn| < L / 2, then Gaussian errors sigma_z, sigma_z,
sigma_dl.  Lengths in metres.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass
class HybridGeometry:
    n_detectors: int
    ring_radius: float
    n_columns: int
    n_rows: int
    pixel_size: float
    n_planes: int
    z_left: float
    sigma_z: float
    sigma_dl: float
    length: float
    seg: np.ndarray        # (n_lors, 4) float32: start x, y, end x, y
    widths: np.ndarray     # (n_lors,) float32
    pixels: np.ndarray     # (n_infos, 2) int16: x, y
    positions: np.ndarray  # (n_infos,) float32
    weights: np.ndarray    # (n_planes_half, n_infos) float32
    info_begin: np.ndarray  # (n_lors,) int64
    info_end: np.ndarray    # (n_lors,) int64

    @property
    def n_lors(self):
        return self.seg.shape[0]

    @property
    def n_infos(self):
        return self.pixels.shape[0]

    @property
    def n_planes_half(self):
        return self.weights.shape[0]

    @property
    def shape(self):
        """Voxel image shape (planes, rows, columns): index
        (z n_rows + y) n_columns + x (voxel_grid.h:54-56)."""
        return (self.n_planes, self.n_rows, self.n_columns)


def lor_index(first, second):
    return first * (first + 1) // 2 + second


def _tube_infos(seg, lor_width, n_pixels, pixel_size, fov, chunk=2048):
    """Pixel infos of the LORs seg (m, 4) float32 (start, end): for each LOR
    the pixels whose centre is inside the field of view and within half the
    width of the LOR's line, with their scaled projection t along the
    segment.  Candidates come from walking the LOR's main axis (at most
    ceil(width / (|cos| pixel)) + 3 pixels across per step); the test is
    applied in FP64.  Returns (lor, raster index, x, y, t float32) arrays."""
    n, ps = n_pixels, pixel_size
    ll = -n * ps / 2
    c0 = ll + ps / 2
    axis = np.arange(n)
    centre = ll + ps / 2 + axis * ps
    out = []
    for lo in range(0, len(seg), chunk):
        sg = seg[lo:lo + chunk]
        sx, sy = sg[:, 0].astype(np.float64), sg[:, 1].astype(np.float64)
        d = (sg[:, 2:] - sg[:, :2]).astype(np.float64)
        length = np.hypot(d[:, 0], d[:, 1])
        ok = length > 0
        length = np.where(ok, length, 1.0)
        ux, uy = d[:, 0] / length, d[:, 1] / length
        horiz = np.abs(ux) >= np.abs(uy)
        # main-axis coordinate a (column for horizontal LORs, row else);
        # the line's cross coordinate there, and the half band across
        main = np.where(horiz, ux, uy)
        safe = np.where(np.abs(main) > 0, main, 1.0)
        cross_at = np.where(horiz[:, None],
                            sy[:, None] + (centre[None] - sx[:, None]) *
                            (uy / safe)[:, None],
                            sx[:, None] + (centre[None] - sy[:, None]) *
                            (ux / safe)[:, None])
        half = lor_width / (2 * np.abs(safe))
        k = int(np.ceil(2 * half.max() / ps)) + 3
        b0 = np.floor((cross_at - half[:, None] - c0) / ps).astype(np.int64) - 1
        cand = b0[:, :, None] + np.arange(k)[None, None, :]
        a_idx = np.broadcast_to(axis[None, :, None], cand.shape)
        ix = np.where(horiz[:, None, None], a_idx, cand)
        iy = np.where(horiz[:, None, None], cand, a_idx)
        valid = (ix >= 0) & (ix < n) & (iy >= 0) & (iy < n) & ok[:, None, None]
        ixc, iyc = np.clip(ix, 0, n - 1), np.clip(iy, 0, n - 1)
        cx = ll + ps / 2 + ixc * ps
        cy = ll + ps / 2 + iyc * ps
        dist = (cx - sx[:, None, None]) * (-uy)[:, None, None] + \
            (cy - sy[:, None, None]) * ux[:, None, None]
        sel = valid & (cx * cx + cy * cy < fov * fov) & \
            (np.abs(dist) < lor_width / 2)
        li = np.broadcast_to(np.arange(lo, lo + len(sg))[:, None, None],
                             cand.shape)[sel]
        t = (((cx - sx[:, None, None]) * ux[:, None, None] +
              (cy - sy[:, None, None]) * uy[:, None, None]) /
             length[:, None, None])[sel].astype(np.float32)
        out.append((li, iyc[sel] * n + ixc[sel], ixc[sel], iyc[sel], t))
    cat = [np.concatenate([o[i] for o in out]) for i in range(5)]
    lor, raster, x, y, t = cat
    order = np.lexsort((raster, t, lor))
    return lor[order], x[order], y[order], t[order]


def ring_geometry(n_detectors=48, ring_radius=0.4, lor_width=0.02,
                  n_pixels=80, pixel_size=0.005, n_planes=None,
                  fov_radius=None, sigma_z=0.015, sigma_dl=0.06, length=0.5,
                  n_planes_half=1, weight_seed=5):
    """The LOR tube geometry described in the module docstring.  Weights are
    positive (for the calculate_weight path they are recomputed)."""
    n_planes = n_pixels if n_planes is None else n_planes
    z_left = -n_planes * pixel_size / 2
    fov = fov_radius if fov_radius is not None else n_pixels * pixel_size / 2
    ang = 2 * np.pi * np.arange(n_detectors) / n_detectors
    centres = np.stack([ring_radius * np.cos(ang),
                        ring_radius * np.sin(ang)], 1).astype(np.float32)
    n_lors = (n_detectors + 1) * n_detectors // 2
    first = np.repeat(np.arange(n_detectors), np.arange(1, n_detectors + 1))
    second = np.concatenate([np.arange(f + 1) for f in range(n_detectors)])
    seg = np.concatenate([centres[second], centres[first]], 1)
    widths = np.full(n_lors, lor_width, np.float32)
    lor, x, y, t = _tube_infos(seg, lor_width, n_pixels, pixel_size, fov)
    counts = np.bincount(lor, minlength=n_lors)
    end = np.cumsum(counts).astype(np.int64)
    begin = end - counts
    pixels = np.concatenate([np.stack([x, y], 1),
                             [[n_pixels // 2, n_pixels // 2]]]).astype(np.int16)
    positions = np.concatenate([t, np.array([2.0], np.float32)])
    rng = np.random.default_rng(weight_seed)
    weights = rng.uniform(0.2, 1.0, size=(n_planes_half, len(positions))
                          ).astype(np.float32)
    weights[:, -1] = 0.0
    return HybridGeometry(n_detectors, ring_radius, n_pixels, n_pixels,
                          pixel_size, n_planes, z_left, sigma_z, sigma_dl,
                          length, seg, widths, pixels, positions, weights,
                          begin, end)


GEOMETRY_MAGIC = b"PETg"
_GRID_HEADER = np.dtype([("n_columns", "<i2"), ("n_rows", "<i2"),
                         ("pixel_size", "<f4"), ("llx", "<f4"),
                         ("lly", "<f4")])
_LOR_HEADER = np.dtype([("first", "<i2"), ("second", "<i2"),
                        ("seg", "<f4", (4,)), ("width", "<f4"),
                        ("n_pixels", "<u8")])
_PIXEL_INFO = np.dtype([("x", "<i2"), ("y", "<i2"), ("t", "<f4"),
                        ("distance", "<f4"), ("fill", "<f4"),
                        ("weight", "<f4")])


def read_geometry(path, n_planes, z_left=None, sigma_z=0.015, sigma_dl=0.06,
                  length=0.5) -> HybridGeometry:
    """A PET2D::Barrel::Geometry<float, short> binary file (the reference's
    2d_barrel_geometry output) flattened as GeometrySOA(geometry, 1) does.

    File (src/2d/barrel/geometry.h:56-105, native little-endian): the magic
    "PETg", int16 n_detectors, the grid (pixel_grid.h:82-101: int16
    n_columns, n_rows, float pixel_size, lower-left x, y), then LOR records
    to the end (lor_geometry.h:58-84: int16 first, second, the segment's
    start and end points (line_segment.h:38-56), float width, uint64
    n_pixels, n_pixels x {int16 x, y, float t, distance, fill, weight}); a
    later record for the same LOR (index first (first + 1) / 2 + second,
    lor.h:36-38) replaces the earlier one.  The flattening
    (geometry_soa.h:187-206) walks the LOR slots in index order and lays the
    infos out slot after slot; a slot no record filled is a zero LOR
    (first = second = 0), so it empties LOR 0, and keeps info range [0, 0)
    (its segment and width, which the reference leaves unset, are 0 here).
    The padding info of the module docstring (centre pixel, weight 0,
    position 2) follows the file's infos, so that the info one past the last
    LOR's stays inside the arrays.  The voxel grid's axis comes
    from the caller as the reference's command line does (--n-planes,
    --z-left; z_left None centres it).  The reconstruction's grid is centred
    on the scanner axis (PixelGrid(n_columns, n_rows, pixel_size)), so a file
    whose lower-left corner is not (-n_columns, -n_rows) pixel_size / 2 is
    refused.  ring_radius (used only to draw events) is the mean distance of
    the segments' end points from the axis."""
    data = np.fromfile(path, np.uint8)
    if data.size < 4 + 2 + _GRID_HEADER.itemsize or \
            data[:4].tobytes() != GEOMETRY_MAGIC:
        raise ValueError(f"{path}: not a binary float barrel geometry file")
    n_det = int(np.frombuffer(data, "<i2", 1, 4)[0])
    grid = np.frombuffer(data, _GRID_HEADER, 1, 6)[0]
    n_columns, n_rows = int(grid["n_columns"]), int(grid["n_rows"])
    ps = np.float32(grid["pixel_size"])
    ll = (np.float32(-ps * np.float32(n_columns) / np.float32(2)),
          np.float32(-ps * np.float32(n_rows) / np.float32(2)))
    if n_det <= 0 or n_columns <= 0 or n_rows <= 0 or not ps > 0:
        raise ValueError(f"{path}: bad header")
    for got, want in ((grid["llx"], ll[0]), (grid["lly"], ll[1])):
        if abs(float(got) - float(want)) > 1e-6 * abs(float(want)):
            raise ValueError(f"{path}: pixel grid not centred on the axis "
                             f"(lower-left {grid['llx']}, {grid['lly']})")
    n_lors = (n_det + 1) * n_det // 2
    slot = [None] * n_lors
    off = 6 + _GRID_HEADER.itemsize
    while off + _LOR_HEADER.itemsize <= data.size:
        h = np.frombuffer(data, _LOR_HEADER, 1, off)[0]
        off += _LOR_HEADER.itemsize
        m = int(h["n_pixels"])
        if m * _PIXEL_INFO.itemsize > data.size - off:
            break  # a short record ends the stream, as a failed read does
        infos = np.frombuffer(data, _PIXEL_INFO, m, off)
        off += m * _PIXEL_INFO.itemsize
        f, s = int(h["first"]), int(h["second"])
        li = f * (f + 1) // 2 + s
        if not (0 <= s <= f < n_det):
            raise ValueError(f"{path}: LOR ({f}, {s}) outside the "
                             f"{n_det}-detector scanner")
        slot[li] = (h["seg"], h["width"], infos)
    seg = np.zeros((n_lors, 4), np.float32)
    widths = np.zeros(n_lors, np.float32)
    begin = np.zeros(n_lors, np.int64)
    end = np.zeros(n_lors, np.int64)
    parts, size = [], 0
    for li, rec in enumerate(slot):
        target = li if rec is not None else 0
        sg, w, infos = rec if rec is not None else \
            (np.zeros(4, np.float32), np.float32(0), None)
        seg[target], widths[target] = sg, w
        begin[target] = size
        if infos is not None and len(infos):
            parts.append(infos)
            size += len(infos)
        end[target] = size
    pad = np.zeros(1, _PIXEL_INFO)
    pad["x"], pad["y"], pad["t"] = n_columns // 2, n_rows // 2, 2.0
    infos = np.concatenate(parts + [pad])
    pixels = np.stack([infos["x"], infos["y"]], 1).astype(np.int16)
    ends = np.concatenate([seg[:, :2], seg[:, 2:]])
    r = np.hypot(ends[:, 0].astype(np.float64), ends[:, 1])
    ring_radius = float(r[r > 0].mean()) if np.any(r > 0) else 0.0
    z_left = -n_planes * float(ps) / 2 if z_left is None else z_left
    return HybridGeometry(n_det, ring_radius, n_columns, n_rows, float(ps),
                          n_planes, z_left, sigma_z, sigma_dl, length, seg,
                          widths, pixels,
                          np.ascontiguousarray(infos["t"], np.float32),
                          np.ascontiguousarray(infos["weight"],
                                               np.float32)[None, :],
                          begin, end)


def write_geometry(geo: HybridGeometry, path):
    """geo written as a PET2D::Barrel::Geometry<float, short> binary file
    (the layout read_geometry describes, geometry.h:91-103): every LOR slot
    in index order with its own infos and first plane-half's weights;
    infos outside every LOR's range (ring_geometry's padding) are not
    written.  The file keeps no distance or fill (the reconstruction reads
    neither): they are written as 0 and 1."""
    n_det = geo.n_detectors
    ps = np.float32(geo.pixel_size)
    head = np.zeros(1, _GRID_HEADER)
    head["n_columns"], head["n_rows"], head["pixel_size"] = \
        geo.n_columns, geo.n_rows, ps
    head["llx"] = -ps * np.float32(geo.n_columns) / np.float32(2)
    head["lly"] = -ps * np.float32(geo.n_rows) / np.float32(2)
    first = np.repeat(np.arange(n_det), np.arange(1, n_det + 1))
    second = np.concatenate([np.arange(f + 1) for f in range(n_det)])
    with open(path, "wb") as out:
        out.write(GEOMETRY_MAGIC)
        out.write(np.int16(n_det).tobytes())
        out.write(head.tobytes())
        for li in range(geo.n_lors):
            b, e = int(geo.info_begin[li]), int(geo.info_end[li])
            h = np.zeros(1, _LOR_HEADER)
            h["first"], h["second"] = first[li], second[li]
            h["seg"], h["width"], h["n_pixels"] = geo.seg[li], \
                geo.widths[li], e - b
            out.write(h.tobytes())
            rec = np.zeros(e - b, _PIXEL_INFO)
            rec["x"], rec["y"] = geo.pixels[b:e, 0], geo.pixels[b:e, 1]
            rec["t"], rec["fill"] = geo.positions[b:e], 1.0
            rec["weight"] = geo.weights[0, b:e]
            out.write(rec.tobytes())


@dataclass(frozen=True)
class HybridWorkload:
    """A named hybrid reconstruction workload: ring geometry parameters,
    the emission phantom of hybrid_events and the event count."""
    name: str
    n_detectors: int
    ring_radius: float
    lor_width: float
    n_pixels: int
    pixel_size: float
    n_planes: int
    fov_radius: float
    sigma_z: float
    sigma_dl: float
    length: float
    radius: float
    half_z: float
    n_events: int

    def geometry(self):
        return ring_geometry(self.n_detectors, self.ring_radius,
                             self.lor_width, self.n_pixels, self.pixel_size,
                             self.n_planes, self.fov_radius, self.sigma_z,
                             self.sigma_dl, self.length)

    def events(self, geo, n=None, seed=0):
        return hybrid_events(geo, self.n_events if n is None else n, seed,
                             radius=self.radius, half_z=self.half_z)


WORKLOADS = {w.name: w for w in [
    # the test ring: 48 detectors, tubes 2 cm wide
    HybridWorkload("hybrid-ring48-80x80x80-5mm-5e4", 48, 0.4, 0.02, 80,
                   0.005, 80, 0.2, 0.015, 0.06, 0.5, 0.12, 0.1, 50_000),
    # a fully packed J-PET ring (config/full_ring.cfg: 384 detectors 7 mm
    # wide at R 0.4278 m, 0.5 m long, FOV radius 0.4 m), 160 x 160 x 100
    # voxels of 5 mm, sigma_z 1.5 cm, sigma_dl 6 cm (the command line's
    # defaults), a 25 cm cylinder phantom, 1e7 events
    HybridWorkload("hybrid-ring384-160x160x100-5mm-1e7", 384, 0.4278, 0.007,
                   160, 0.005, 100, 0.4, 0.015, 0.06, 0.5, 0.25, 0.2,
                   10_000_000),
    # the same ring and phantom on 320 x 320 x 200 voxels of 2.5 mm: 3.7e7
    # pixel infos, LORs of up to 960 infos, whose infos x planes exceed LDS
    # (the LOR-space iteration then works in chunks of planes)
    HybridWorkload("hybrid-ring384-320x320x200-2.5mm-1e7", 384, 0.4278,
                   0.007, 320, 0.0025, 200, 0.4, 0.015, 0.06, 0.5, 0.25, 0.2,
                   10_000_000),
]}


def hybrid_events(geo: HybridGeometry, n, seed=0, radius=0.12, half_z=0.1,
                  hot=(0.04, 0.03, 0.0, 0.03, 0.3), max_tan=None,
                  batch=1 << 18):
    """(n, 5) float32 events: first, second, z_up, z_dn, dl.  The phantom is
    a cylinder (radius, |z| < half_z) with a hot cylinder hot = (x, y, z,
    radius, share of emissions), both uniform."""
    rng = np.random.Generator(np.random.Philox(seed))
    N = geo.n_detectors
    Rr = geo.ring_radius
    half_l = geo.length / 2
    tmax = max_tan if max_tan is not None else 1.1 * half_l / Rr
    out = np.empty((n, 5), np.float32)
    filled = 0
    hx, hy, hz, hr, share = hot
    while filled < n:
        m = batch
        u = rng.random((m, 6))
        in_hot = u[:, 0] < share
        rad = np.where(in_hot, hr, radius) * np.sqrt(u[:, 1])
        th = 2 * np.pi * u[:, 2]
        px = np.where(in_hot, hx, 0.0) + rad * np.cos(th)
        py = np.where(in_hot, hy, 0.0) + rad * np.sin(th)
        pz = np.where(in_hot, hz + hr * (2 * u[:, 3] - 1),
                      half_z * (2 * u[:, 3] - 1))
        phi = np.pi * u[:, 4]
        dx, dy = np.cos(phi), np.sin(phi)
        # crossings of the line p + s d with the ring
        b = px * dx + py * dy
        c = px * px + py * py - Rr * Rr
        root = np.sqrt(b * b - c)
        s1, s2 = -b - root, -b + root
        a1 = np.arctan2(py + s1 * dy, px + s1 * dx)
        a2 = np.arctan2(py + s2 * dy, px + s2 * dx)
        d1 = np.round(a1 / (2 * np.pi / N)).astype(np.int64) % N
        d2 = np.round(a2 / (2 * np.pi / N)).astype(np.int64) % N
        first = np.maximum(d1, d2)
        second = np.minimum(d1, d2)
        li = first * (first + 1) // 2 + second
        sx, sy = geo.seg[li, 0], geo.seg[li, 1]
        ex, ey = geo.seg[li, 2], geo.seg[li, 3]
        lx, ly = ex - sx, ey - sy
        L = np.hypot(lx, ly)
        ok = (first != second) & (L > 0)
        L = np.where(ok, L, 1.0)
        ux, uy = lx / L, ly / L
        mx, my = (sx + ex) / 2, (sy + ey) / 2
        up = (px - mx) * ux + (py - my) * uy
        R = L / 2
        tan = tmax * (2 * u[:, 5] - 1)
        z_up = pz + (R - up) * tan
        z_dn = pz - (R + up) * tan
        dl = -2 * up * np.sqrt(tan * tan + 1)
        ok &= (np.abs(z_up) < half_l) & (np.abs(z_dn) < half_l)
        idx = np.nonzero(ok)[0]
        g = rng.standard_normal((m, 3))
        take = min(len(idx), n - filled)
        idx = idx[:take]
        sl = slice(filled, filled + take)
        out[sl, 0] = first[idx]
        out[sl, 1] = second[idx]
        out[sl, 2] = z_up[idx] + geo.sigma_z * g[idx, 0]
        out[sl, 3] = z_dn[idx] + geo.sigma_z * g[idx, 1]
        out[sl, 4] = dl[idx] + geo.sigma_dl * g[idx, 2]
        filled += take
    return out


# ---------------------------------------------------------------------------
# Binding of include/hybrid_mlem.h (lib/libstrip_mlem.so)
# ---------------------------------------------------------------------------

import ctypes  # noqa: E402

HYBRID_MLEM_CPU = 0
HYBRID_MLEM_LEGACY_GPU = 1

RESPONSE_DTYPE = np.dtype([("first", "<i2"), ("second", "<i2"),
                           ("z_up", "<f4"), ("z_dn", "<f4"), ("dl", "<f4")])
# hybrid_mlem_frame_event, with the C layout's padding
FRAME_EVENT_DTYPE = np.dtype({
    "names": ["first", "second", "up", "right", "tan", "sec", "info_begin",
              "info_end", "plane_begin", "plane_end"],
    "formats": ["<i2", "<i2", "<f4", "<f4", "<f4", "<f4", "<i8", "<i8",
                "<i2", "<i2"],
    "offsets": [0, 2, 4, 8, 12, 16, 24, 32, 40, 42],
    "itemsize": 48})


class GeometryC(ctypes.Structure):
    """hybrid_mlem_geometry."""
    _fields_ = [("n_detectors", ctypes.c_int32),
                ("n_pixel_infos", ctypes.c_int64),
                ("n_planes_half", ctypes.c_int32),
                ("lor_line_segments", ctypes.c_void_p),
                ("lor_widths", ctypes.c_void_p),
                ("pixels", ctypes.c_void_p),
                ("pixel_positions", ctypes.c_void_p),
                ("pixel_weights", ctypes.c_void_p),
                ("lor_pixel_info_begin", ctypes.c_void_p),
                ("lor_pixel_info_end", ctypes.c_void_p),
                ("n_columns", ctypes.c_int32), ("n_rows", ctypes.c_int32),
                ("n_planes", ctypes.c_int32),
                ("pixel_size", ctypes.c_double), ("z_left", ctypes.c_double),
                ("sigma_z", ctypes.c_double), ("sigma_dl", ctypes.c_double),
                ("length", ctypes.c_double)]


_bound = None


def exported_symbols():
    """Names of the hybrid C-ABI entry points this binding uses."""
    return ["hybrid_mlem_create", "hybrid_mlem_destroy",
            "hybrid_mlem_prepare_weights", "hybrid_mlem_add_responses",
            "hybrid_mlem_clear_events", "hybrid_mlem_n_events",
            "hybrid_mlem_get_frame_events", "hybrid_mlem_set_rho",
            "hybrid_mlem_get_rho", "hybrid_mlem_get_sensitivity",
            "hybrid_mlem_get_weights", "hybrid_mlem_iterate",
            "hybrid_mlem_last_iteration_ms", "hybrid_mlem_voxel_visits",
            "hybrid_mlem_set_algorithm", "hybrid_mlem_get_algorithm",
            "hybrid_mlem_set_frame_events", "hybrid_mlem_device_name",
            "hybrid_mlem_get_timing"]


def load_library():
    """lib/libstrip_mlem.so with the hybrid entry points' signatures set
    (raises if the library has not been built)."""
    global _bound
    if _bound is not None:
        return _bound
    from . import load_library as _load
    lib = _load()
    P = ctypes.c_void_p
    i64 = ctypes.c_int64
    fp = ctypes.POINTER(ctypes.c_float)
    sigs = {
        "hybrid_mlem_create": (ctypes.c_int, [ctypes.POINTER(GeometryC),
                                              ctypes.c_int, ctypes.c_int,
                                              ctypes.POINTER(P)]),
        "hybrid_mlem_destroy": (None, [P]),
        "hybrid_mlem_prepare_weights": (ctypes.c_int, [P, ctypes.c_int,
                                                       ctypes.c_int, fp]),
        "hybrid_mlem_add_responses": (ctypes.c_int, [P, P, i64]),
        "hybrid_mlem_clear_events": (ctypes.c_int, [P]),
        "hybrid_mlem_n_events": (i64, [P]),
        "hybrid_mlem_get_frame_events": (ctypes.c_int, [P, P, P]),
        "hybrid_mlem_set_rho": (ctypes.c_int, [P, fp]),
        "hybrid_mlem_get_rho": (ctypes.c_int, [P, fp]),
        "hybrid_mlem_get_sensitivity": (ctypes.c_int, [P, fp]),
        "hybrid_mlem_get_weights": (ctypes.c_int, [P, fp]),
        "hybrid_mlem_iterate": (ctypes.c_int, [P, ctypes.c_int]),
        "hybrid_mlem_last_iteration_ms": (ctypes.c_int, [P, fp]),
        "hybrid_mlem_voxel_visits": (ctypes.c_int, [P, ctypes.POINTER(i64)]),
        "hybrid_mlem_set_algorithm": (ctypes.c_int, [P, ctypes.c_int]),
        "hybrid_mlem_get_algorithm": (ctypes.c_int, [P]),
        "hybrid_mlem_set_frame_events": (ctypes.c_int, [P, P, i64]),
        "hybrid_mlem_device_name": (ctypes.c_int, [P, ctypes.c_char_p,
                                                   ctypes.c_int]),
        "hybrid_mlem_get_timing": (ctypes.c_int, [
            P, ctypes.POINTER(ctypes.c_double),
            ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _bound = lib
    return lib


def responses(events) -> np.ndarray:
    """(n, 5) events (first, second, z_up, z_dn, dl) as the C ABI's
    hybrid_mlem_response records (the binary event file's layout)."""
    ev = np.asarray(events)
    out = np.empty(ev.shape[0], RESPONSE_DTYPE)
    out["first"] = ev[:, 0]
    out["second"] = ev[:, 1]
    out["z_up"] = ev[:, 2]
    out["z_dn"] = ev[:, 3]
    out["dl"] = ev[:, 4]
    return out


class HybridReconstruction:
    """The GPU hybrid reconstruction of one geometry (include/hybrid_mlem.h),
    used like the reference's command line drives
    PET3D::Hybrid::Reconstruction (reconstruction_cmd.cpp:459-560)."""

    def __init__(self, geo: HybridGeometry, semantics=HYBRID_MLEM_CPU,
                 device=0):
        from . import StripMLEMError
        self._err = StripMLEMError
        self._lib = load_library()
        self.geo = geo
        keep = [np.ascontiguousarray(geo.seg, np.float32),
                np.ascontiguousarray(geo.widths, np.float32),
                np.ascontiguousarray(geo.pixels, np.int16),
                np.ascontiguousarray(geo.positions, np.float32),
                np.ascontiguousarray(geo.weights[0], np.float32),
                np.ascontiguousarray(geo.info_begin, np.int64),
                np.ascontiguousarray(geo.info_end, np.int64)]
        g = GeometryC(geo.n_detectors, geo.n_infos, geo.n_planes_half,
                      *[a.ctypes.data for a in keep], geo.n_columns,
                      geo.n_rows, geo.n_planes,
                      float(np.float32(geo.pixel_size)),
                      float(np.float32(geo.z_left)),
                      float(np.float32(geo.sigma_z)),
                      float(np.float32(geo.sigma_dl)),
                      float(np.float32(geo.length)))
        self._ctx = ctypes.c_void_p()
        self._check(self._lib.hybrid_mlem_create(ctypes.byref(g), semantics,
                                                 device,
                                                 ctypes.byref(self._ctx)))

    def _check(self, rc):
        if rc != 0:
            msg = self._lib.strip_mlem_last_error().decode(errors="replace")
            raise self._err(f"status {rc}: {msg}")

    def close(self):
        if self._ctx:
            self._lib.hybrid_mlem_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def prepare_weights(self, calculate_weight=True, sensitivity_mode=0,
                        sensitivity=None):
        s = None
        if sensitivity is not None:
            s = np.ascontiguousarray(sensitivity, np.float32)
            if s.size != self.geo.n_rows * self.geo.n_columns:
                raise ValueError("sensitivity must have n_rows x n_columns "
                                 "values")
        fp = ctypes.POINTER(ctypes.c_float)
        self._check(self._lib.hybrid_mlem_prepare_weights(
            self._ctx, int(bool(calculate_weight)), int(sensitivity_mode),
            s.ctypes.data_as(fp) if s is not None else None))

    def add_responses(self, events):
        r = responses(events)
        self._check(self._lib.hybrid_mlem_add_responses(
            self._ctx, r.ctypes.data if len(r) else None, len(r)))

    def clear_events(self):
        self._check(self._lib.hybrid_mlem_clear_events(self._ctx))

    def set_frame_events(self, frames):
        """Replaces the events held by frame events (a FRAME_EVENT_DTYPE
        array), with no filter."""
        f = np.ascontiguousarray(frames, FRAME_EVENT_DTYPE)
        self._check(self._lib.hybrid_mlem_set_frame_events(
            self._ctx, f.ctypes.data if len(f) else None, len(f)))

    def device_name(self):
        buf = ctypes.create_string_buffer(256)
        self._check(self._lib.hybrid_mlem_device_name(self._ctx, buf, 256))
        return buf.value.decode()

    @property
    def n_events(self):
        return int(self._lib.hybrid_mlem_n_events(self._ctx))

    def frame_events(self):
        n = self.n_events
        ints = np.empty((n, 4), np.int32)
        floats = np.empty((n, 4), np.float32)
        self._check(self._lib.hybrid_mlem_get_frame_events(
            self._ctx, ints.ctypes.data, floats.ctypes.data))
        return ints, floats

    @property
    def rho(self):
        out = np.empty(self.geo.shape, np.float32)
        self._check(self._lib.hybrid_mlem_get_rho(
            self._ctx, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    @rho.setter
    def rho(self, value):
        a = np.ascontiguousarray(value, np.float32)
        if a.shape != self.geo.shape:
            raise ValueError(f"rho must have shape {self.geo.shape}")
        self._check(self._lib.hybrid_mlem_set_rho(
            self._ctx, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))

    @property
    def sensitivity(self):
        out = np.empty((self.geo.n_rows, self.geo.n_columns), np.float32)
        self._check(self._lib.hybrid_mlem_get_sensitivity(
            self._ctx, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    @property
    def weights(self):
        out = np.empty(self.geo.n_infos, np.float32)
        self._check(self._lib.hybrid_mlem_get_weights(
            self._ctx, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def iterate(self, n=1):
        self._check(self._lib.hybrid_mlem_iterate(self._ctx, int(n)))

    def set_algorithm(self, algorithm):
        """0 automatic, 1 wavefront per event, 2 LOR space."""
        self._check(self._lib.hybrid_mlem_set_algorithm(self._ctx,
                                                        int(algorithm)))

    @property
    def algorithm(self):
        """The algorithm of the last iteration (1 or 2)."""
        a = self._lib.hybrid_mlem_get_algorithm(self._ctx)
        if a < 0:
            self._check(a)
        return a

    def last_iteration_ms(self):
        ms = ctypes.c_float()
        self._check(self._lib.hybrid_mlem_last_iteration_ms(
            self._ctx, ctypes.byref(ms)))
        return ms.value

    def timing(self):
        """(event kernels ms, gather ms, iterations) summed over the last
        iterate call."""
        a, b, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        self._check(self._lib.hybrid_mlem_get_timing(
            self._ctx, ctypes.byref(a), ctypes.byref(b), ctypes.byref(n)))
        return a.value, b.value, n.value

    def voxel_visits(self):
        v = ctypes.c_int64()
        self._check(self._lib.hybrid_mlem_voxel_visits(self._ctx,
                                                       ctypes.byref(v)))
        return v.value
