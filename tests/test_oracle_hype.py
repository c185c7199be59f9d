"""Pins of the Hype oracle beyond the reference's known answers (test/unit_tests/TestHype.cpp,
which probes along the axes only):

- the closed form of TestHype.cpp:140 (DistanceToOut along +-z from a mid-radius point hits
  the inner hyperboloid at z = sqrt((rho^2 - rmin^2) / tan^2(stIn))), which the known-answer
  extractor cannot evaluate;
- Inside and the exact ray distances against an independent description of the solid:
  |z| <= dz and rmin^2 + tIn^2 z^2 <= rho^2 <= rmax^2 + tOut^2 z^2, with the first boundary
  crossing found from the roots of the four surfaces along the ray (numpy quadratics) and the
  inside test at the midpoints between them. Points within 1e-6 mm of a surface and rays whose
  crossing is within 1e-6 mm of another root are left out (tolerance rules decide there);
- the safety bounds the shape tester checks: SafetyToIn <= DistanceToIn and
  SafetyToOut <= DistanceToOut for every direction, safeties positive away from the surface.
  The one exception is the reference's own: beyond an end cap and inside the end radius of a
  hyperbolic inner surface, SafetyToIn is the distance to the cap's inner edge
  (HypeImplementation.h:266-270), which exceeds the distance to the flaring inner surface.
"""
import math

import numpy as np
import pytest

from oracle.pyoracle import OracleShape

SHAPES = [
    (10.0, 20.0, math.pi / 4, math.pi / 3, 50.0),  # TestHype.cpp:53 (b1)
    (0.0, 20.0, 0.0, math.pi / 3, 50.0),           # no inner surface
    (10.0, 20.0, 0.0, math.pi / 4, 50.0),          # cylindrical inner surface
]


def implicit(params, p):
    rmin, rmax, sti, sto, dz = params
    ti2, to2 = math.tan(sti) ** 2, math.tan(sto) ** 2
    rho2 = p[0] ** 2 + p[1] ** 2
    z2 = p[2] ** 2
    # signed margins (positive inside), in mm to first order
    m_z = dz - np.abs(p[2])
    m_o = (rmax ** 2 + to2 * z2 - rho2) / (2 * np.sqrt(rmax ** 2 + to2 * z2))
    has_inner = rmin > 0 or sti != 0
    m_i = (rho2 - rmin ** 2 - ti2 * z2) / (2 * np.sqrt(np.maximum(rmin ** 2 + ti2 * z2, 1e-30))) if has_inner \
        else np.full(p.shape[1], np.inf)
    return np.minimum(m_z, np.minimum(m_o, m_i))


def roots(params, p, d):
    """All positive parameters where the ray meets one of the four surfaces (n x 8, inf padded)."""
    rmin, rmax, sti, sto, dz = params
    out = []
    for z0 in (dz, -dz):
        with np.errstate(divide="ignore", invalid="ignore"):
            t = (z0 - p[2]) / d[2]
        out.append(np.where(np.isfinite(t) & (t > 0), t, np.inf))
    for r0, t2 in ((rmax, math.tan(sto) ** 2), (rmin, math.tan(sti) ** 2)):
        a = d[0] ** 2 + d[1] ** 2 - t2 * d[2] ** 2
        b = p[0] * d[0] + p[1] * d[1] - t2 * p[2] * d[2]
        c = p[0] ** 2 + p[1] ** 2 - t2 * p[2] ** 2 - r0 ** 2
        disc = b * b - a * c
        ok = disc >= 0
        sq = np.sqrt(np.where(ok, disc, 0.0))
        with np.errstate(divide="ignore", invalid="ignore"):
            for sgn in (-1.0, 1.0):
                t = np.where(np.abs(a) > 1e-300, (-b + sgn * sq) / a, -c / (2 * b))
                out.append(np.where(ok & np.isfinite(t) & (t > 0), t, np.inf))
    return np.sort(np.stack(out, axis=1), axis=1)


def first_crossing(params, p, d, want_inside_after):
    """Smallest root after which the inside state is `want_inside_after` (inf if none), and
    whether that root is isolated from its neighbours by more than 1e-6."""
    r = roots(params, p, d)
    n = p.shape[1]
    best = np.full(n, np.inf)
    isolated = np.ones(n, dtype=bool)
    found = np.zeros(n, dtype=bool)
    for k in range(r.shape[1]):
        t = r[:, k]
        nxt = r[:, k + 1] if k + 1 < r.shape[1] else np.full(n, np.inf)
        mid = np.where(np.isfinite(nxt), 0.5 * (t + nxt), t + 1.0)
        q = p + mid[None, :] * d
        state = implicit(params, q) > 0
        hit = ~found & np.isfinite(t) & (state == want_inside_after)
        prev = r[:, k - 1] if k > 0 else np.full(n, -np.inf)
        isolated = np.where(hit, (t - prev > 1e-6) & (nxt - t > 1e-6), isolated)
        best = np.where(hit, t, best)
        found |= hit
    return best, isolated


def sample(params, n, seed):
    rmin, rmax, sti, sto, dz = params
    re = math.sqrt(rmax ** 2 + math.tan(sto) ** 2 * dz ** 2)
    g = np.random.default_rng(seed)
    p = g.uniform([-1.2 * re, -1.2 * re, -1.2 * dz], [1.2 * re, 1.2 * re, 1.2 * dz], size=(n, 3)).T
    d = g.normal(size=(3, n))
    d /= np.linalg.norm(d, axis=0)
    return np.ascontiguousarray(p), np.ascontiguousarray(d)


def test_closed_form_inner_hyperboloid_hit():
    """TestHype.cpp:140-163: distZ = sqrt((|pmid|^2 - rmin^2) / tan^2(stIn))."""
    params = SHAPES[0]
    s = OracleShape("hype", list(params))
    dist_z = math.sqrt((15.0 ** 2 - 10.0 ** 2) / math.tan(math.pi / 4) ** 2)
    pts = [(15, 0, 0), (-15, 0, 0), (0, 15, 0), (0, -15, 0)]
    for p in pts:
        for dzs in (1.0, -1.0):
            got = s.query("DistanceToOut", np.array(p, float).reshape(3, 1), np.array([0, 0, dzs]).reshape(3, 1))[0]
            assert abs(got - dist_z) < 1e-9, (p, dzs, got, dist_z)


@pytest.mark.parametrize("params", SHAPES)
def test_inside_matches_implicit_form(params):
    s = OracleShape("hype", list(params))
    p, _ = sample(params, 20000, 1)
    m = implicit(params, p)
    keep = np.abs(m) > 1e-6
    ins = s.query("Inside", np.ascontiguousarray(p[:, keep]))
    exp = np.where(m[keep] > 0, 1, 3)
    assert np.array_equal(ins, exp)
    assert np.array_equal(s.query("Contains", np.ascontiguousarray(p[:, keep])).astype(bool), m[keep] > 0)


@pytest.mark.parametrize("params", SHAPES)
def test_distances_match_quadric_roots(params):
    s = OracleShape("hype", list(params))
    p, d = sample(params, 20000, 2)
    m = implicit(params, p)
    keep = np.abs(m) > 1e-6
    p, d, m = np.ascontiguousarray(p[:, keep]), np.ascontiguousarray(d[:, keep]), m[keep]
    inside = m > 0
    # from outside: first entry
    pi, di = np.ascontiguousarray(p[:, ~inside]), np.ascontiguousarray(d[:, ~inside])
    exp, iso = first_crossing(params, pi, di, True)
    got = s.query("DistanceToIn", pi, di)
    both_inf = (got >= 1e300) & ~np.isfinite(exp)
    ok = both_inf | (np.abs(got - exp) <= 1e-9 * np.maximum(1, np.abs(exp)))
    assert np.all(ok | ~iso), int((~ok & iso).sum())
    assert iso.mean() > 0.95
    # from inside: first exit
    po, do = np.ascontiguousarray(p[:, inside]), np.ascontiguousarray(d[:, inside])
    exp, iso = first_crossing(params, po, do, False)
    got = s.query("DistanceToOut", po, do)
    ok = np.abs(got - exp) <= 1e-9 * np.maximum(1, np.abs(exp))
    assert np.all(ok | ~iso), int((~ok & iso).sum())
    assert np.all(s.query("DistanceToIn", po, do) < 0)
    assert np.all(s.query("DistanceToOut", pi, di) < 0)


@pytest.mark.parametrize("params", SHAPES)
def test_safeties_bound_distances(params):
    s = OracleShape("hype", list(params))
    p, _ = sample(params, 2000, 3)
    m = implicit(params, p)
    keep = np.abs(m) > 1e-6
    p, m = np.ascontiguousarray(p[:, keep]), m[keep]
    inside = m > 0
    g = np.random.default_rng(4)
    sin_, sout = s.query("SafetyToIn", p), s.query("SafetyToOut", p)
    assert np.all(sin_[inside] == -1.0) and np.all(sout[~inside] == -1.0)
    assert np.all(sout[inside] > 0) and np.all(sin_[~inside] >= 0)
    rmin, rmax, sti, sto, dz = params
    end_in = math.sqrt(math.tan(sti) ** 2 * dz ** 2 + rmin ** 2)
    cap_hole = (np.abs(p[2]) > dz) & (np.hypot(p[0], p[1]) < end_in) & (sti != 0)
    for _ in range(16):
        d = g.normal(size=p.shape)
        d = np.ascontiguousarray(d / np.linalg.norm(d, axis=0))
        di = s.query("DistanceToIn", p, d)
        do = s.query("DistanceToOut", p, d)
        assert np.all((sin_ <= di + 1e-9)[~inside & ~cap_hole])
        assert np.all(sout[inside] <= do[inside] + 1e-9)
