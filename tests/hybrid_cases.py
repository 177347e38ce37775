"""Hybrid test geometries with a repeated voxel visit.

An event's pixel-info range ends one info past the first position above its
box (src/3d/hybrid/reconstruction.h:131-138).  When the box reaches past a
LOR's last info, that extra info is the next LOR's first; if its pixel is
also one of the event's own infos, the event visits that voxel twice.  The
reference's CPU path then adds the voxel's cached weight (the extra info's,
written last) on both visits; its CUDA kernel adds each visit's own weight.

repeated_pixel_geometry(geo) inserts, at the front of LOR l + 1, an info
with the pixel of LOR l's last info (position 0, so l + 1 stays sorted);
repeated_pixel_events(...) draws events on LOR l whose box covers that last
info and reaches past it.
"""
import numpy as np


def _pick_lor(geo):
    n = geo.n_detectors
    sizes = geo.info_end - geo.info_begin
    best = None
    for first in range(n - 1, 1, -1):
        for second in range(first - 2, -1, -1):
            li = first * (first + 1) // 2 + second
            if sizes[li] < 20 or sizes[li + 1] < 20:
                continue
            pj = geo.pixels[geo.info_end[li] - 1]
            nxt = geo.pixels[geo.info_begin[li + 1]:geo.info_end[li + 1]]
            if np.any(np.all(nxt == pj, axis=1)):
                continue
            best = (li, first, second)
            break
        if best:
            break
    assert best is not None
    return best


def repeated_pixel_geometry(geo, weight=0.7):
    """A copy of geo with the inserted info; returns (geometry, lor, first,
    second)."""
    from dataclasses import replace
    li, first, second = _pick_lor(geo)
    at = int(geo.info_end[li])
    pj = geo.pixels[at - 1]
    pixels = np.insert(geo.pixels, at, pj, axis=0)
    positions = np.insert(geo.positions, at, np.float32(0.0))
    weights = np.insert(geo.weights, at, np.float32(weight), axis=1)
    begin = geo.info_begin.copy()
    end = geo.info_end.copy()
    begin[li + 1:] += 1
    end[li + 1:] += 1
    g = replace(geo, pixels=pixels, positions=positions, weights=weights,
                info_begin=begin, info_end=end)
    return g, li, first, second


def repeated_pixel_events(geo, li, first, second, n, seed=0):
    """(n, 5) events on LOR li whose emission lies near the LOR's last info
    (tan 0 up to noise), so that the box reaches past it."""
    rng = np.random.default_rng(seed)
    s = geo.seg[li].astype(np.float64)
    R = np.hypot(s[2] - s[0], s[3] - s[1]) / 2
    t = float(geo.positions[geo.info_end[li] - 1])
    up = (t - 0.5) * 2 * R - rng.uniform(0.0, 0.02, n)
    z = rng.uniform(-0.05, 0.05, n)
    tan = rng.uniform(-0.01, 0.01, n)
    ev = np.empty((n, 5), np.float32)
    ev[:, 0] = first
    ev[:, 1] = second
    ev[:, 2] = z + (R - up) * tan
    ev[:, 3] = z - (R + up) * tan
    ev[:, 4] = -2 * up * np.sqrt(1 + tan * tan)
    return ev
