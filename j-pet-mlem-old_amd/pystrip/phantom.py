"""Seeded strip-PET event generator and the benchmark workloads.

Follows the distributional semantics of the reference's ``2d_strip_phantom``
(src/2d/strip/phantom_cmd.cpp:136-444):
  - region chosen with probability intensity x area
    (src/2d/geometry/phantom.h:120-133,183-191);
  - point uniform in the region: angle = 2 pi U, r = sqrt(U)
    (src/2d/geometry/ellipse.h:83-92);
  - non-additive phantoms reject a point lying in an earlier region and draw
    again (phantom.h:193-204), so earlier regions take precedence;
  - direction phi = pi/2 * U(-1, 1), tan = sin(phi)/cos(phi)
    (phantom.h:211-215, src/2d/strip/event.h:25-28);
  - exact detection if |z_u|, |z_d| < L/2 (scanner.h:196-214) on the exact
    projection (scanner.h:107-113);
  - Gaussian smearing of z_u, z_d (sigma_z) and dl (sigma_dl)
    (scanner.h:257-267).
The random stream is numpy's Philox4x32-10 counter-based generator (Salmon et
al., SC'11) with a fixed batch size, so a (seed, n) pair always yields the same
events.  It is this project's own stream: the reference's libstdc++
distributions re-seeded per emission through srand48 cannot be reproduced.
"""
from __future__ import annotations

import hashlib
import math
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Ellipse:
    """One phantom region: `ellipse x y a b angle intensity` (phantom.h:144-156);
    x is the z coordinate of the strip scanner, angle in degrees."""

    x: float
    y: float
    a: float
    b: float
    angle: float
    intensity: float

    @property
    def area(self):
        return math.pi * self.a * self.b

    def contains(self, px, py):
        # Ellipse::contains (src/2d/geometry/ellipse.h:29-33)
        t = math.radians(self.angle)
        s, c = math.sin(t), math.cos(t)
        A = c * c / (self.a * self.a) + s * s / (self.b * self.b)
        B = s * s / (self.a * self.a) + c * c / (self.b * self.b)
        C = s * c * (1 / (self.a * self.a) - 1 / (self.b * self.b))
        rx, ry = px - self.x, py - self.y
        return A * rx * rx + 2 * C * rx * ry + B * ry * ry <= 1

    def text(self):
        return (f"ellipse {self.x:g} {self.y:g} {self.a:g} {self.b:g} "
                f"{self.angle:g} {self.intensity:g}")


@dataclass
class Workload:
    name: str
    radius: float
    length: float
    n_z: int
    n_y: int
    pixel: float
    sigma_z: float
    sigma_dl: float
    phantom: list
    n_events: int
    iterations: int
    seed: int = 0
    additive: bool = False
    note: str = ""

    def geometry(self):
        from . import Geometry
        return Geometry.make(self.radius, self.length, self.n_z, self.n_y,
                             self.pixel, self.sigma_z, self.sigma_dl)

    def oracle_geometry_tuple(self):
        return (self.radius, self.length, self.n_z, self.n_y, self.pixel,
                self.pixel, self.sigma_z, self.sigma_dl, 0.0, 0.0)

    def scaled(self, n_events, iterations=None, name=None):
        return Workload(name or f"{self.name}-n{n_events}", self.radius,
                        self.length, self.n_z, self.n_y, self.pixel,
                        self.sigma_z, self.sigma_dl, self.phantom, n_events,
                        self.iterations if iterations is None else iterations,
                        self.seed, self.additive, self.note)

    def events(self, n=None, seed=None):
        return generate_events(self.phantom, self.radius, self.length,
                               self.sigma_z, self.sigma_dl,
                               self.n_events if n is None else n,
                               self.seed if seed is None else seed,
                               self.additive)


def generate_events(phantom, radius, length, sigma_z, sigma_dl, n, seed=0,
                    additive=False, batch=1 << 20):
    """Returns an (n, 3) float32 array of detected, smeared responses
    (z_u, z_d, dl)."""
    n = int(n)
    rng = np.random.Generator(np.random.Philox(seed))
    regions = list(phantom)
    weights = np.array([r.intensity * r.area for r in regions], dtype=np.float64)
    cdf = np.cumsum(weights)
    cdf /= cdf[-1]
    ang = np.radians([r.angle for r in regions])
    sin_a, cos_a = np.sin(ang), np.cos(ang)
    cx = np.array([r.x for r in regions])
    cy = np.array([r.y for r in regions])
    ra = np.array([r.a for r in regions])
    rb = np.array([r.b for r in regions])
    # ellipse equation coefficients for the precedence test
    eA = cos_a**2 / ra**2 + sin_a**2 / rb**2
    eB = sin_a**2 / ra**2 + cos_a**2 / rb**2
    eC = sin_a * cos_a * (1 / ra**2 - 1 / rb**2)
    half_l = length / 2.0
    out = np.empty((n, 3), dtype=np.float32)
    filled = 0
    while filled < n:
        m = batch
        u = rng.random((m, 5))
        k = np.searchsorted(cdf, u[:, 0], side="left")
        k = np.minimum(k, len(regions) - 1)
        theta = 2 * np.pi * u[:, 1]
        r = np.sqrt(u[:, 2])
        lx = ra[k] * r * np.cos(theta)
        ly = rb[k] * r * np.sin(theta)
        px = cos_a[k] * lx - sin_a[k] * ly + cx[k]
        py = sin_a[k] * lx + cos_a[k] * ly + cy[k]
        keep = np.ones(m, dtype=bool)
        if not additive:
            for j in range(len(regions) - 1):
                sel = k > j
                if not sel.any():
                    continue
                rx, ry = px - cx[j], py - cy[j]
                inside = (eA[j] * rx * rx + 2 * eC[j] * rx * ry +
                          eB[j] * ry * ry) <= 1
                keep &= ~(sel & inside)
        phi = 0.5 * np.pi * (2 * u[:, 3] - 1)
        tan = np.sin(phi) / np.cos(phi)
        # to_projection_space_tan (scanner.h:107-113): point = (z, y)
        z_u = px + (radius - py) * tan
        z_d = px - (radius + py) * tan
        dl = -2 * py * np.sqrt(tan * tan + 1)
        keep &= (np.abs(z_u) < half_l) & (np.abs(z_d) < half_l)
        idx = np.nonzero(keep)[0]
        g = rng.standard_normal((m, 3))
        take = min(len(idx), n - filled)
        idx = idx[:take]
        out[filled:filled + take, 0] = z_u[idx] + sigma_z * g[idx, 0]
        out[filled:filled + take, 1] = z_d[idx] + sigma_z * g[idx, 1]
        out[filled:filled + take, 2] = dl[idx] + sigma_dl * g[idx, 2]
        filled += take
    return out


def events_digest(events: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(events, np.float32).tobytes()
                          ).hexdigest()


def phantom_image(phantom, n_z, n_y, pixel, additive=False):
    """Piecewise-constant intensity of the phantom sampled at pixel centres
    (row-major y*n_z+z, rows from +y down, as Scanner::pixel_center)."""
    z = (np.arange(n_z) + 0.5) * pixel - n_z * pixel / 2
    y = n_y * pixel / 2 - (np.arange(n_y) + 0.5) * pixel
    Z, Y = np.meshgrid(z, y)
    img = np.zeros((n_y, n_z))
    assigned = np.zeros((n_y, n_z), dtype=bool)
    for r in phantom:
        inside = r.contains(Z, Y)
        if additive:
            img += r.intensity * inside
        else:
            new = inside & ~assigned
            img[new] = r.intensity
            assigned |= inside
    return img


# --------------------------------------------------------------------------
# The workloads (sizes in millimetres)
# --------------------------------------------------------------------------

PHANTOM_3 = [  # small regions first: earlier regions take precedence
    Ellipse(40, -15, 20, 25, -20, 4),
    Ellipse(-35, 10, 30, 20, 30, 2),
    Ellipse(0, 0, 80, 60, 0, 1),
]

PHANTOM_HEAD_10 = [
    Ellipse(-60, 40, 3, 3, 0, 3),
    Ellipse(50, 60, 4, 4, 0, 3),
    Ellipse(20, -70, 5, 5, 0, 3),
    Ellipse(-90, -30, 6, 6, 0, 3),
    Ellipse(70, 0, 40, 25, 15, 0.2),
    Ellipse(-70, 0, 40, 25, -15, 0.5),
    Ellipse(0, 80, 35, 20, 0, 1.5),
    Ellipse(0, -110, 30, 20, 0, 2),
    Ellipse(0, 0, 170, 130, 0, 0.8),
    Ellipse(0, 0, 180, 140, 0, 1),
]

PHANTOM_6 = [
    Ellipse(30, 20, 15, 10, 0, 3),
    Ellipse(-40, -20, 20, 12, 25, 2),
    Ellipse(60, -40, 10, 10, 0, 0.3),
    Ellipse(-20, 50, 25, 15, -10, 1.5),
    Ellipse(0, 0, 110, 80, 0, 0.8),
    Ellipse(0, 0, 120, 90, 0, 1),
]

PHANTOM_HEAD_12 = [
    Ellipse(-60, 40, 1.5, 1.5, 0, 3),
    Ellipse(50, 60, 2, 2, 0, 3),
    Ellipse(20, -70, 3, 3, 0, 3),
    Ellipse(-90, -30, 4, 4, 0, 3),
    Ellipse(100, 50, 2.5, 2.5, 0, 3),
    Ellipse(-120, 60, 3, 3, 0, 3),
    Ellipse(70, 0, 40, 25, 15, 0.2),
    Ellipse(-70, 0, 40, 25, -15, 0.5),
    Ellipse(0, 80, 35, 20, 0, 1.5),
    Ellipse(0, -110, 30, 20, 0, 2),
    Ellipse(0, 0, 170, 130, 0, 0.8),
    Ellipse(0, 0, 180, 140, 0, 1),
]

WORKLOADS = {
    # CPU-runnable plain case
    "strip-256x160-2mm-1e6": Workload(
        "strip-256x160-2mm-1e6", 350.0, 500.0, 256, 160, 2.0, 10.0, 42.0,
        PHANTOM_3, 1_000_000, 10),
    # plain method at scale (the headline single-GPU workload)
    "strip-1000x400-1mm-2e7": Workload(
        "strip-1000x400-1mm-2e7", 400.0, 1000.0, 1000, 400, 1.0, 10.0, 42.0,
        PHANTOM_HEAD_10, 20_000_000, 20),
    # differentiation workload (R, L chosen: 400 / 1000)
    "strip-512x256-1mm-5e6": Workload(
        "strip-512x256-1mm-5e6", 400.0, 1000.0, 512, 256, 1.0, 10.0, 42.0,
        PHANTOM_6, 5_000_000, 8),
    # largest grid (OSEM, streamed)
    "strip-2000x800-0.5mm-2e8": Workload(
        "strip-2000x800-0.5mm-2e8", 400.0, 1000.0, 2000, 800, 0.5, 10.0, 42.0,
        PHANTOM_HEAD_12, 200_000_000, 5),
}


def random_phantom(rng, half_z, half_y, n_regions):
    """A random 2-8 ellipse phantom inside the grid (batched workload)."""
    regions = []
    outer_a = 0.8 * half_z * rng.uniform(0.6, 1.0)
    outer_b = 0.8 * half_y * rng.uniform(0.6, 1.0)
    for _ in range(n_regions - 1):
        a = rng.uniform(0.05, 0.3) * outer_a
        b = rng.uniform(0.05, 0.3) * outer_b
        x = rng.uniform(-0.5, 0.5) * outer_a
        y = rng.uniform(-0.5, 0.5) * outer_b
        regions.append(Ellipse(x, y, a, b, rng.uniform(-90, 90),
                               rng.uniform(0.2, 4.0)))
    regions.append(Ellipse(0, 0, outer_a, outer_b, 0, 1.0))
    return regions


# --------------------------------------------------------------------------
# The batched workload: many independent reconstructions
# --------------------------------------------------------------------------

BATCH_GRIDS = [(128, 128), (256, 128), (512, 256), (512, 512)]


def batched_workload(n_instances=64, seed=0, min_events=10_000,
                     max_events=2_000_000, event_scale=1.0):
    """Members of the batched benchmark: grid drawn from BATCH_GRIDS, pixel
    size uniform in [1, 2] mm (0.25 mm steps), event count log-uniform in
    [min_events, max_events] (times event_scale), a random 2-8 ellipse
    phantom, sigma_z 10 mm, sigma_dl 42 mm.  The strips sit 50 mm beyond the
    grid (R = half grid height + 50 mm) and are as long as the grid is wide
    plus 100 mm.  Returns a list of Workload."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_instances):
        n_z, n_y = BATCH_GRIDS[rng.integers(len(BATCH_GRIDS))]
        pixel = 1.0 + 0.25 * rng.integers(5)
        half_z, half_y = n_z * pixel / 2, n_y * pixel / 2
        n = int(round(math.exp(rng.uniform(math.log(min_events),
                                           math.log(max_events)))))
        n = max(1, int(n * event_scale))
        ph = random_phantom(rng, half_z, half_y, int(rng.integers(2, 9)))
        out.append(Workload(f"batch-{i}-{n_z}x{n_y}-{pixel:g}mm-{n}",
                            half_y + 50.0, 2 * half_z + 100.0, n_z, n_y,
                            pixel, 10.0, 42.0, ph, n, 10, seed=seed * 1000 + i))
    return out
