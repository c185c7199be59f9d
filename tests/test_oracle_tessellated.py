"""Pins of the tessellated oracle beyond the reference's known answers (which only cover
Trd-like solids probed along the axes):

- a tessellated box against the Box oracle (Contains, Inside, distances, SafetyToOut),
- the convex 1024-triangle hull of a UV sphere against convex-polytope formulas written
  independently of the facet kernels: inside = every plane offset negative; ray clipping
  against the plane set for DistanceToIn/Out; SafetyToOut = smallest plane distance; and
  SafetyToIn = Euclidean distance to the closest triangle (closest-point-on-triangle by
  region tests, Ericson, Real-Time Collision Detection 5.1.5).
Points within 1e-6 mm of a plane are left out, where the tolerance rules decide.
"""
import numpy as np
import pytest

from oracle.pyoracle import OracleShape
from tests.golden.extract_kats import trd_like_triangles
from vecgeom_amd import workloads

pytestmark = []


def planes(params):
    n = int(params[0])
    t = np.array(params[1:1 + 9 * n]).reshape(n, 3, 3)
    nrm = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 1])
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    off = -np.einsum("ij,ij->i", nrm, t[:, 0])
    return t, nrm, off


def sample(n, half, seed):
    g = np.random.default_rng(seed)
    p = g.uniform(-1.3 * half, 1.3 * half, size=(3, n))
    d = g.normal(size=(3, n))
    d /= np.linalg.norm(d, axis=0)
    return np.ascontiguousarray(p), np.ascontiguousarray(d)


def test_tessellated_box_matches_box_oracle():
    h = np.array([10.0, 20.0, 30.0])
    tess = OracleShape("tessellated", trd_like_triangles(h[0], h[0], h[1], h[1], h[2]))
    box = OracleShape("box", list(h))
    p, d = sample(4000, h[:, None], 1)
    near = np.min(np.abs(np.abs(p) - h[:, None]), axis=0) < 1e-6
    p, d = np.ascontiguousarray(p[:, ~near]), np.ascontiguousarray(d[:, ~near])
    inside = np.all(np.abs(p) < h[:, None], axis=0)
    assert np.array_equal(tess.query("Contains", p), box.query("Contains", p))
    assert np.array_equal(tess.query("Inside", p), box.query("Inside", p))
    di_t, di_b = tess.query("DistanceToIn", p, d), box.query("DistanceToIn", p, d)
    out = ~inside
    hit = di_b[out] < 1e300
    assert np.array_equal(di_t[out] < 1e300, hit)
    assert np.allclose(di_t[out][hit], di_b[out][hit], rtol=0, atol=1e-9)
    assert np.all(di_t[inside] < 0)
    do_t, do_b = tess.query("DistanceToOut", p, d), box.query("DistanceToOut", p, d)
    assert np.allclose(do_t[inside], do_b[inside], rtol=0, atol=1e-9)
    assert np.all(do_t[out] < 0)
    so_t, so_b = tess.query("SafetyToOut", p), box.query("SafetyToOut", p)
    assert np.allclose(so_t[inside], so_b[inside], rtol=0, atol=1e-9)
    # the box safety from outside is the largest axis distance, a lower bound of the exact one
    si_t, si_b = tess.query("SafetyToIn", p), box.query("SafetyToIn", p)
    assert np.all(si_t[out] >= si_b[out] - 1e-9)


def closest_on_triangle(p, a, b, c):
    ab, ac, ap = b - a, c - a, p - a
    d1, d2 = ab @ ap, ac @ ap
    if d1 <= 0 and d2 <= 0:
        return a
    bp = p - b
    d3, d4 = ab @ bp, ac @ bp
    if d3 >= 0 and d4 <= d3:
        return b
    vc = d1 * d4 - d3 * d2
    if vc <= 0 and d1 >= 0 and d3 <= 0:
        return a + d1 / (d1 - d3) * ab
    cp = p - c
    d5, d6 = ab @ cp, ac @ cp
    if d6 >= 0 and d5 <= d6:
        return c
    vb = d5 * d2 - d1 * d6
    if vb <= 0 and d2 >= 0 and d6 <= 0:
        return a + d2 / (d2 - d6) * ac
    va = d3 * d6 - d5 * d4
    if va <= 0 and (d4 - d3) >= 0 and (d5 - d6) >= 0:
        return b + (d4 - d3) / ((d4 - d3) + (d5 - d6)) * (c - b)
    den = 1.0 / (va + vb + vc)
    return a + ab * (vb * den) + ac * (vc * den)


def test_convex_hull_against_plane_set():
    prm = workloads.convex_hull_sphere(20.0)
    tri, nrm, off = planes(prm)
    s = OracleShape("tessellated", prm)
    p, d = sample(3000, 20.0, 2)
    saf = nrm @ p + off[:, None]                      # (nfacets, npoints)
    keep = np.min(np.abs(saf), axis=0) > 1e-6
    p, d, saf = np.ascontiguousarray(p[:, keep]), np.ascontiguousarray(d[:, keep]), saf[:, keep]
    inside = np.all(saf < 0, axis=0)
    assert inside.sum() > 300 and (~inside).sum() > 300
    assert np.array_equal(s.query("Contains", p).astype(bool), inside)
    assert np.array_equal(s.query("Inside", p), np.where(inside, 1, 3))
    ndd = nrm @ d
    with np.errstate(divide="ignore", invalid="ignore"):
        t = -saf / ndd
    t_enter = np.max(np.where(ndd < 0, t, -np.inf), axis=0)
    t_exit = np.min(np.where(ndd > 0, t, np.inf), axis=0)
    di = s.query("DistanceToIn", p, d)
    out = ~inside
    hit = out & (t_enter <= t_exit) & (t_exit > 0)
    assert np.allclose(di[hit], t_enter[hit], rtol=0, atol=1e-9)
    assert np.all(di[out & ~hit] > 1e300)
    do = s.query("DistanceToOut", p, d)
    assert np.allclose(do[inside], t_exit[inside], rtol=0, atol=1e-9)
    so = s.query("SafetyToOut", p)
    assert np.allclose(so[inside], np.min(-saf[:, inside], axis=0), rtol=0, atol=1e-9)
    si = s.query("SafetyToIn", p)
    idx = np.nonzero(out)[0][:300]
    ref = [min(np.linalg.norm(closest_on_triangle(p[:, i], *tri[k]) - p[:, i]) for k in range(len(tri)))
           for i in idx]
    assert np.allclose(si[idx], ref, rtol=0, atol=1e-9)
    assert np.all(s.query("SafetyToIn", np.ascontiguousarray(p[:, inside])) == -1)
    assert np.all(s.query("SafetyToOut", np.ascontiguousarray(p[:, out])) == -1)


def test_extent_and_degenerate_facets():
    prm = workloads.convex_hull_sphere(20.0)
    lo, hi = OracleShape("tessellated", prm).extent()
    v = np.array(prm[1:]).reshape(-1, 3)
    assert np.array_equal(lo, v.min(axis=0)) and np.array_equal(hi, v.max(axis=0))
    # a zero-area facet is dropped (Tile::AddVertex rejects it), the rest of the solid is kept
    bad = list(trd_like_triangles(10, 10, 10, 10, 10))
    bad[0] += 1
    bad += [0, 0, 0, 1e-6, 0, 0, 2e-6, 0, 0]
    s = OracleShape("tessellated", bad)
    assert s.query("Contains", np.zeros((3, 1)))[0] == 1
