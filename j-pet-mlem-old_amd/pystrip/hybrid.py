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
(reconstruction.h:132-137), which for the last LOR with infos is one past
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
    seg = np.zeros((n_lors, 4), np.float32)
    widths = np.full(n_lors, lor_width, np.float32)
    # pixel centres of the grid (PixelGrid::center_at, pixel_grid.h:54-57)
    ll = -n_pixels * pixel_size / 2
    ix, iy = np.meshgrid(np.arange(n_pixels), np.arange(n_pixels))
    cx = (ll + pixel_size / 2 + ix * pixel_size).ravel()
    cy = (ll + pixel_size / 2 + iy * pixel_size).ravel()
    in_fov = cx * cx + cy * cy < fov * fov
    pix_all = np.stack([ix.ravel(), iy.ravel()], 1).astype(np.int16)
    begin = np.zeros(n_lors, np.int64)
    end = np.zeros(n_lors, np.int64)
    pixels, positions = [], []
    pos = 0
    for first in range(n_detectors):
        for second in range(first + 1):
            li = lor_index(first, second)
            s, e = centres[second], centres[first]
            seg[li] = (s[0], s[1], e[0], e[1])
            begin[li] = pos
            if first != second:
                d = (e - s).astype(np.float64)
                length_ = float(np.hypot(*d))
                u = d / length_
                nrm = np.array([-u[1], u[0]])
                dist = (cx - s[0]) * nrm[0] + (cy - s[1]) * nrm[1]
                sel = in_fov & (np.abs(dist) < lor_width / 2)
                if sel.any():
                    t = ((cx[sel] - s[0]) * u[0] + (cy[sel] - s[1]) * u[1]) \
                        / length_
                    t = t.astype(np.float32)
                    order = np.argsort(t, kind="stable")
                    pixels.append(pix_all[sel][order])
                    positions.append(t[order])
                    pos += int(sel.sum())
            end[li] = pos
    pixels.append(np.array([[n_pixels // 2, n_pixels // 2]], np.int16))
    positions.append(np.array([2.0], np.float32))
    pixels = np.concatenate(pixels)
    positions = np.concatenate(positions)
    rng = np.random.default_rng(weight_seed)
    weights = rng.uniform(0.2, 1.0, size=(n_planes_half, len(positions))
                          ).astype(np.float32)
    weights[:, -1] = 0.0
    return HybridGeometry(n_detectors, ring_radius, n_pixels, n_pixels,
                          pixel_size, n_planes, z_left, sigma_z, sigma_dl,
                          length, seg, widths, pixels, positions, weights,
                          begin, end)


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
