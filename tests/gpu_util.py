"""Shared helpers of the GPU parity tests (test infrastructure)."""
import numpy as np
import torch

from oracle.pyoracle import OracleShape
from vecgeom_amd import SOA3D, workloads

INF_THRESHOLD = 1e300


def inputs_for(kind, params, n, seed, near_fraction=0.5):
    """Half uniform in 1.2x the bounding box, half within 1e-3 mm of the surface."""
    o = OracleShape(kind, params)
    lo, hi = o.extent()
    g = workloads.rng(seed)
    n_near = int(n * near_fraction)
    p_uni = workloads.uniform_in_box(g, n - n_near, lo, hi)
    p_near = workloads.near_surface(g, n_near, lo, hi, lambda p: o.query("Contains", p),
                                    lambda p, d: o.query("DistanceToOut", p, d))
    p = np.ascontiguousarray(np.concatenate([p_uni, p_near], axis=1))
    d = workloads.isotropic_dirs(g, p.shape[1])
    return o, p, d


def grazing_rays(kind, params, n, seed):
    """Rays that skim the surface far from their start: two nearby surface points (within 1e-6 mm of
    the surface, neighbours along a random projection) define a direction close to tangent or
    across an edge, and the start point lies 5-60 mm back along it. These are the rays whose hit or
    miss depends on rounding."""
    o = OracleShape(kind, params)
    lo, hi = o.extent()
    g = workloads.rng(seed)
    s = workloads.near_surface(g, n, lo, hi, lambda q: o.query("Contains", q),
                               lambda q, e: o.query("DistanceToOut", q, e), band=1e-6)
    key = np.argsort(s[0] * 0.37 + s[1] * 0.71 + s[2] * 0.6)
    s1 = s[:, key]
    e = np.roll(s1, 1, axis=1) - s1
    nrm = np.linalg.norm(e, axis=0)
    ok = nrm > 1e-9
    e, s1 = e[:, ok] / nrm[ok], s1[:, ok]
    p = s1 - e * g.uniform(5.0, 60.0, size=s1.shape[1])
    return o, np.ascontiguousarray(p), np.ascontiguousarray(e)


def to_dev(a):
    return SOA3D.from_array(a, device="cuda")


def compare_float(ref, got, tol):
    """Returns (n_bad, n_bitexact, max_abs_err over finite pairs)."""
    ref = np.asarray(ref, dtype=np.float64)
    got = np.asarray(got, dtype=np.float64)
    both_inf = (ref >= INF_THRESHOLD) & (got >= INF_THRESHOLD)
    diff = np.abs(ref - got)
    diff[both_inf] = 0.0
    bad = ~(both_inf | (diff <= tol))
    fin = ~both_inf
    maxerr = float(diff[fin].max()) if fin.any() else 0.0
    return int(bad.sum()), int((ref == got).sum()), maxerr
