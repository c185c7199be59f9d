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


def convex_hull_sphere(radius=20.0, nphi=32, nbands=17):
    """Tessellated convex hull of a UV sphere: nphi longitude segments and nbands latitude
    bands (two triangle caps, the other bands quadrilaterals split into two triangles), so
    nphi * (2 * nbands - 2) triangles -- 1024 for the defaults. Vertices are ordered
    anticlockwise seen from outside. Returns the tessellated parameter list
    [ntri, then 9 doubles per triangle]."""
    theta = np.pi * np.arange(nbands + 1) / nbands
    phi = 2.0 * np.pi * np.arange(nphi) / nphi

    def vert(i, j):
        if i == 0:
            return np.array([0.0, 0.0, radius])
        if i == nbands:
            return np.array([0.0, 0.0, -radius])
        t, f = theta[i], phi[j % nphi]
        return radius * np.array([np.sin(t) * np.cos(f), np.sin(t) * np.sin(f), np.cos(t)])

    tris = []

    def add(a, b, c):
        n = np.cross(b - a, c - b)
        if np.dot(n, a + b + c) < 0:  # make the normal point outwards
            b, c = c, b
        tris.append((a, b, c))

    for j in range(nphi):
        add(vert(0, 0), vert(1, j), vert(1, j + 1))
        add(vert(nbands, 0), vert(nbands - 1, j + 1), vert(nbands - 1, j))
        for i in range(1, nbands - 1):
            a, b, c, d = vert(i, j), vert(i + 1, j), vert(i + 1, j + 1), vert(i, j + 1)
            add(a, b, c)
            add(a, c, d)
    return [float(len(tris))] + [float(x) for t in tris for v in t for x in v]


# Batched shape set evaluated in one launch (the shapes implemented so far; each
# entry is (kind, constructor parameters)). Tube: rmin 2, rmax 8, dz 10, phi 30-270 deg;
# polycone with 10 z planes; general trapezoid; cone with an inner radius and slopes;
# 8-sided polyhedron with 6 z planes (inner radii in the middle planes); sphere shell
# with theta (30-130 deg) and phi (20-290 deg) cuts; Torus2 with rtor 50, rmax 10; orb;
# tessellated convex hull of 1024 triangles (a UV sphere of radius 20).
def mix_shapes():
    r = np.radians
    return [
        ("box", [10.0, 20.0, 30.0]),
        ("tube", [2.0, 8.0, 10.0, float(r(30)), float(r(240))]),
        ("cone", [5.0, 10.0, 7.0, 15.0, 10.0, float(r(20)), float(r(300))]),
        ("polycone", [float(r(30)), float(r(270)), 10,
                      -50, -40, -25, -10, 0, 5, 20, 30, 40, 50,
                      0, 5, 5, 8, 8, 3, 3, 10, 10, 0,
                      20, 25, 25, 30, 40, 40, 30, 35, 35, 20]),
        ("polyhedron", [0.0, float(2 * np.pi), 8, 6, -20, -10, -5, 5, 10, 20, 0, 5, 5, 3, 0, 0,
                        10, 15, 15, 12, 12, 8]),
        ("trapezoid", [60.0, float(r(20)), float(r(5)), 40.0, 30.0, 40.0, float(r(10)), 16.0, 10.0, 14.0,
                       float(r(10))]),
        ("sphere", [5.0, 10.0, float(r(20)), float(r(270)), float(r(30)), float(r(100))]),
        ("torus2", [0.0, 10.0, 50.0, 0.0, float(2 * np.pi)]),
        ("orb", [10.0]),
        ("tessellated", convex_hull_sphere(20.0)),
    ]


def mix_counts(nshapes, seed=SEED_BASE + 2, lo=14, hi=20):
    """Per-shape query counts drawn log-uniform in [2^lo, 2^hi]."""
    g = rng(seed)
    return [int(2.0 ** (lo + (hi - lo) * g.random())) for _ in range(nshapes)]


def mix_inputs(shapes, counts, contains_fns, dist_out_fns, seed=SEED_BASE + 3):
    """Concatenated SOA inputs: per shape, half uniform in 1.2x the bounding box and half
    within 1e-3 mm of the surface; isotropic directions. extents[i] = (lo, hi)."""
    ps, ds = [], []
    for i, (ext, n) in enumerate(zip(shapes, counts)):
        g = rng(seed + 7919 * i)
        lo, hi = ext
        n_near = n // 2
        p_uni = uniform_in_box(g, n - n_near, lo, hi)
        p_near = near_surface(g, n_near, lo, hi, contains_fns[i], dist_out_fns[i])
        p = np.concatenate([p_uni, p_near], axis=1)
        ps.append(p)
        ds.append(isotropic_dirs(g, p.shape[1]))
    return np.ascontiguousarray(np.concatenate(ps, axis=1)), np.ascontiguousarray(np.concatenate(ds, axis=1))
