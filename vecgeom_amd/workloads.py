"""Seeded synthetic inputs shared by the tests and the benchmark.

Counter-based Philox streams (numpy's Philox4x32-10) keyed by a per-shape seed,
so the CPU baseline and the GPU kernels see identical points and directions.
Directions are isotropic unit vectors: z = 2u - 1, phi = 2 pi v.
"""
import numpy as np

SEED_BASE = 0x5EED0001


def rng(seed):
    return np.random.Generator(np.random.Philox(key=seed))


def isotropic_dirs(g, n):
    z = 2.0 * g.random(n) - 1.0
    phi = 2.0 * np.pi * g.random(n)
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    return np.ascontiguousarray(np.stack([r * np.cos(phi), r * np.sin(phi), z]))


def uniform_in_box(g, n, lo, hi, scale=1.2):
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    c, h = 0.5 * (lo + hi), 0.5 * (hi - lo) * scale
    u = g.random((3, n))
    return np.ascontiguousarray(c[:, None] + (2.0 * u - 1.0) * h[:, None])


# Plain FP64 CPU-runnable configuration: box (10, 20, 30) and full-phi tube (5, 10, 15).
CONFIG0_SHAPES = [("box", [10.0, 20.0, 30.0]), ("tube", [5.0, 10.0, 15.0, 0.0, 2.0 * np.pi])]


def shape_stream(extent, n, seed):
    """Points uniform in 1.2x the bounding box, isotropic directions."""
    g = rng(seed)
    lo, hi = extent
    return uniform_in_box(g, n, lo, hi), isotropic_dirs(g, n)


def near_surface(g, n, lo, hi, contains_fn, dist_out_fn, band=1e-3):
    """Points within `band` of the surface: inside points are pushed along a random
    direction to the boundary (DistanceToOut) and offset by U(-band, band) along a
    second random direction. contains_fn / dist_out_fn evaluate the shape on (3, m)
    arrays; the caller chooses the implementation that produces them."""
    out = np.empty((3, 0))
    tries = 0
    while out.shape[1] < n and tries < 64:
        tries += 1
        m = max(1024, 2 * (n - out.shape[1]))
        p = uniform_in_box(g, m, lo, hi, scale=1.0)
        inside = contains_fn(p).astype(bool)
        p = np.ascontiguousarray(p[:, inside])
        if p.shape[1] == 0:
            continue
        d = isotropic_dirs(g, p.shape[1])
        t = dist_out_fn(p, d)
        ok = np.isfinite(t) & (t >= 0) & (t < 1e30)
        s = p[:, ok] + t[ok] * d[:, ok]
        k = s.shape[1]
        e = isotropic_dirs(g, k)
        delta = (2.0 * g.random(k) - 1.0) * band
        out = np.concatenate([out, s + delta * e], axis=1)
    return np.ascontiguousarray(out[:, :n])
