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
