"""Pins the oracle's Normal restatement by geometry rather than by its formulas.

For surface points reached by DistanceToOut from interior points (where a navigator asks for normals),
the normal must be (a) valid, (b) of unit length, (c) perpendicular to the surface: two exit points of
nearly parallel rays from the same interior point lie on the same face, and the chord between them is
tangent to within its length over the curvature radius, and (d) outward: an exiting ray has a
non-negative component along it. A dropped term, a wrong sign, a swapped component or the normal of
the wrong surface fails (c) or (d). The reference's Normal asserts are checked separately
(test_oracle_kat.py).

The paraboloid's parabolic surface is excluded from (c): the reference computes that surface's normal
into a shadowed local (ParaboloidImplementation.h:441-449) and returns (0, 0, Sign(a)) there, which the
restatement keeps; its z planes are checked like every other face.
"""
import math

import numpy as np
import pytest

from oracle.pyoracle import OracleShape
from tests.golden.extract_kats import trd_like_triangles
from vecgeom_amd import workloads

CASES = [
    ("box", [10.0, 20.0, 30.0]),
    ("tube", [2.0, 8.0, 10.0, math.radians(30), math.radians(240)]),
    ("tube", [0.0, 50.0, 50.0, 0.0, 2 * math.pi]),
    ("trapezoid", [60, math.radians(20), math.radians(5), 40, 30, 40, math.radians(10), 16, 10, 14,
                   math.radians(10)]),
    ("orb", [10.0]),
    ("cone", [5.0, 10.0, 7.0, 15.0, 10.0, math.radians(20), math.radians(300)]),
    ("cone", [20.0, 60.0, 80.0, 140.0, 100.0, math.radians(10), math.radians(300)]),
    ("polycone", [math.radians(30), math.radians(270), 10, -50, -40, -25, -10, 0, 5, 20, 30, 40, 50,
                  0, 5, 5, 8, 8, 3, 3, 10, 10, 0, 20, 25, 25, 30, 40, 40, 30, 35, 35, 20]),
    ("sphere", [0.0, 10.0, math.radians(20), math.radians(270), math.radians(30), math.radians(100)]),
    ("sphere", [6.0, 8.0, 0.0, 2 * math.pi, 0.0, math.pi]),
    ("polyhedron", [0.0, 2 * math.pi, 8, 6, -20, -10, -5, 5, 10, 20, 0, 5, 5, 3, 0, 0, 10, 15, 15, 12, 12, 8]),
    ("polyhedron", [0.5, 4.0, 8, 6, -20, -10, -5, 5, 10, 20, 0, 5, 5, 3, 0, 0, 10, 15, 15, 12, 12, 8]),
    ("torus2", [2.0, 10.0, 50.0, 0.3, 2.0]),
    ("tessellated", workloads.convex_hull_sphere(20.0)),
    ("tessellated", trd_like_triangles(10, 30, 20, 40, 40)),
    ("hype", [10.0, 20.0, math.pi / 4, math.pi / 3, 50.0]),
    ("trd", [10.0, 30.0, 20.0, 40.0, 40.0]),
    ("paraboloid", [6.0, 9.0, 10.0]),
    ("parallelepiped", [20.0, 30.0, 40.0, math.radians(30), math.radians(30), math.radians(45)]),
    ("tet", [6.0, 4.0, 5.0, -5.0, -6.0, 4.0, -4.0, 5.0, -6.0, 5.0, -5.0, -4.0]),
    ("ellipticaltube", [5.0, 10.0, 20.0]),
    ("ellipsoid", [3.0, 4.0, 5.0, -4.5, 3.5]),
    # a GenTrap given clockwise: given anticlockwise, the reference reverts the vertices but takes the face
    # normal by the reverted polygon's edge index from normals made in the order given
    # (GenTrapStruct.h:67-101, GenTrapImplementation.h:654-660), which is not the exit face's
    ("gentrap", [-4.0, -3.0, 6.0, 5.0, -1.4, -0.8, 4.6, 4.0, -5.0, 3.0, 4.0, -4.0, -2.5, 2.3, 2.9, -1.9, 8.0]),
    ("ellipticalcone", [0.5, 0.4, 10.0, 5.0]),
    ("ellipticalcone", [0.3, 0.9, 20.0, 8.0]),
    ("cuttube", [0.0, 10.0, 20.0, 0.0, 2 * math.pi, 0.0, -0.7, -0.71, 0.7, 0.0, 0.71]),
    ("cuttube", [3.0, 8.0, 15.0, math.radians(30), math.radians(240), 0.3, 0.2, -1.0, -0.4, 0.1, 1.0]),
]


def exit_pairs(o, n, seed):
    """Interior points, a random direction d and a direction d2 within 1e-5 rad of it; their exit points."""
    lo, hi = o.extent()
    g = workloads.rng(seed)
    p = workloads.uniform_in_box(g, n, lo, hi, scale=0.9)
    p = p[:, o.query("Contains", p).astype(bool)]
    d = workloads.isotropic_dirs(g, p.shape[1])
    d2 = d + 1e-5 * workloads.isotropic_dirs(g, p.shape[1])
    d2 /= np.linalg.norm(d2, axis=0)
    t1, t2 = o.query("DistanceToOut", p, d), o.query("DistanceToOut", p, d2)
    ok = (t1 < 1e290) & (t2 < 1e290) & (t1 > 1e-3)
    return p[:, ok] + t1[ok] * d[:, ok], p[:, ok] + t2[ok] * d2[:, ok], d[:, ok]


@pytest.mark.parametrize("kind,params", CASES, ids=["%s%d" % (c[0], i) for i, c in enumerate(CASES)])
def test_normal_is_the_outward_surface_normal(kind, params):
    o = OracleShape(kind, params)
    x1, x2, d = exit_pairs(o, 4000, seed=11)
    n1, v1 = o.normal(x1)
    n2, _ = o.normal(x2)
    assert x1.shape[1] > 500
    # (a) exit points are on the surface. The polycone finds its section at z - kTolerance, which on the
    # bottom cap lies below the first plane: the reference answers (0, 0, -1) with valid = false there
    # (UnplacedPolycone.cpp:523-528)
    if kind == "polycone":
        bottom = np.abs(x1[2] - params[3]) < 1e-9
        assert not v1[bottom].any() and np.all(n1[:, bottom] == np.array([[0.0], [0.0], [-1.0]]))
        v1 = np.where(bottom, 1, v1).astype(np.uint8)
    assert v1.mean() > 0.99, v1.mean()
    # (b) unit length wherever the shape normalises (paraboloid normals too: Normalized())
    mag = np.linalg.norm(n1[:, v1.astype(bool)], axis=0)
    assert np.all(np.abs(mag - 1) < 1e-12), np.abs(mag - 1).max()
    # (c) perpendicular to chords within one face (an edge between the two exits shows as a jump
    # of the normal; those pairs are left out)
    chord = x2 - x1
    L = np.linalg.norm(chord, axis=0)
    same = (np.linalg.norm(n1 - n2, axis=0) < 1e-3) & (L > 1e-9) & v1.astype(bool)
    if kind == "paraboloid":  # only the z planes: the parabolic surface's normal is the reference's (0,0,Sign(a))
        same &= np.abs(np.abs(x1[2]) - params[2]) < 1e-9
    assert same.sum() > 0.3 * x1.shape[1] if kind != "paraboloid" else same.sum() > 50, same.sum()
    cos = np.abs(np.sum(n1 * chord, axis=0))[same] / L[same]
    assert np.all(cos < 1e-3), (cos.max(), x1[:, same][:, np.argmax(cos)])
    # (d) outward: exiting rays point along the normal (or graze it)
    dot = np.sum(n1 * d, axis=0)[same]
    assert np.all(dot > -1e-9), dot.min()
    assert np.mean(dot > 0) > 0.99


def test_closed_forms():
    """Values the geometry fixes exactly: an orb's normal is p / |p|; a box face centre's normal is the
    axis; a tube barrel point's normal is (x, y, 0) / r; a box edge averages its two faces."""
    orb = OracleShape("orb", [10.0])
    p = np.array([[6.0, 0.0, 8.0], [0.0, 10.0, 0.0]]).T
    n, v = orb.normal(p)
    assert np.allclose(n, p / 10.0, atol=1e-15) and v.all()
    box = OracleShape("box", [10.0, 20.0, 30.0])
    n, v = box.normal(np.array([[10.0, 0, 0], [0, -20.0, 0], [10.0, 20.0, 0]]).T)
    assert np.array_equal(n[:, 0], [1, 0, 0]) and np.array_equal(n[:, 1], [0, -1, 0])
    assert np.allclose(n[:, 2], [math.sqrt(0.5), math.sqrt(0.5), 0], atol=1e-16) and v.all()
    tube = OracleShape("tube", [0.0, 5.0, 10.0, 0.0, 2 * math.pi])
    n, v = tube.normal(np.array([[3.0, 4.0, 2.0]]).T)
    assert np.allclose(n[:, 0], [0.6, 0.8, 0.0], atol=1e-15) and v.all()
