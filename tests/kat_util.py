"""Helpers for the known-answer fixtures in tests/golden/kat_*.json.

ApproxEqual follows test/unit_tests/ApproxEqual.h of the reference: equal, or
absolute 1e-6 when one side is zero, or both beyond 1e100 with equal sign
(infinity / kInfLength), or relative 1e-6.
"""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
INSIDE = {"Inside": 1, "Surface": 2, "Outside": 3}


def approx_equal(x, y, tol=1e-6):
    if x == y:
        return True
    if x * y == 0.0:
        return abs(x - y) < tol
    if abs(x) > 1e100 or abs(y) > 1e100:
        return x * y > 0 and abs(x) > 1e100 and abs(y) > 1e100
    return abs(x - y) / (abs(x) + abs(y)) < tol


def gentrap_twisted(params):
    """True for a twisted GenTrap (VecGeom/volumes/GenTrapStruct.h:246-276: some lateral face whose bottom
    and top edges are not parallel); params = x[8] y[8] dz."""
    x, y = params[:8], params[8:16]
    for i in range(4):
        j = (i + 1) % 4
        dx1, dy1 = x[j] - x[i], y[j] - y[i]
        dx2, dy2 = x[j + 4] - x[i + 4], y[j + 4] - y[i + 4]
        if (dx1 == 0 and dy1 == 0) or (dx2 == 0 and dy2 == 0):
            continue
        if abs(dy1 * dx2 - dx1 * dy2) >= 1e-9:
            return True
    return False


def load(name):
    with open(os.path.join(GOLDEN, "kat_%s.json" % name)) as f:
        return json.load(f)["records"]


def normal_point(rec, distance):
    """The point of a Normal record: rec["p"], or base + Dist * dir for a chained record, with Dist from
    distance(shape_params, method, p, d) (the reference test feeds its own query result back)."""
    c = rec.get("chain")
    if c is None:
        return rec["p"]
    q = c["dist"]
    dist = float(distance(q["shape"], q["method"], q["p"], q["d"]))
    return [b + dist * v for b, v in zip(c["base"], c["dir"])]


def check_normal(rec, normal, valid):
    """ApproxEqual on each component (test/unit_tests/ApproxEqual.h:22-26), the validity flag, or the
    length / squared length of the normal."""
    e = rec["expect"]
    if e["kind"] == "normal":
        return all(approx_equal(float(a), float(b)) for a, b in zip(normal, e["value"]))
    if e["kind"] == "normal_valid":
        return bool(valid) == e["value"]
    m2 = sum(float(a) * float(a) for a in normal)
    if e["kind"] == "normal_mag":
        return approx_equal(m2 ** 0.5, e["value"])
    if e["kind"] == "normal_mag2":
        return approx_equal(m2, e["value"])
    raise ValueError(e)


def check(rec, got):
    e = rec["expect"]
    if e["kind"] == "approx":
        return approx_equal(got, e["value"])
    if e["kind"] == "anyof":
        return any(approx_equal(got, v) for v in e["values"])
    if e["kind"] == "inside":
        return int(got) == INSIDE[e["value"]]
    if e["kind"] == "negative":
        return got < 0
    if e["kind"] == "nonpositive":
        return got <= 0
    if e["kind"] == "positive":
        return got > 0
    if e["kind"] == "nonnegative":
        return got >= 0
    if e["kind"] == "exact":  # assert(s->M(..) == v): the reference's own value, bit for bit
        return got == e["value"]
    if e["kind"] == "absdiff":
        return abs(got - e["value"]) < e["tol"]
    if e["kind"] == "bool":
        return bool(got) == e["value"]
    if e["kind"] == "extent":  # got = (lo, hi); ApproxEqual on each component
        v = got[0] if e["which"] == "lo" else got[1]
        return all(approx_equal(float(a), float(b)) for a, b in zip(v, e["value"]))
    raise ValueError(e)


def soa(vecs):
    return np.ascontiguousarray(np.array(vecs, dtype=np.float64).T)
