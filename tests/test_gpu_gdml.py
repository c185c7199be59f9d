"""Navigation on geometries read from GDML (vgmi_gdml_read), GPU against the oracle, as the
reference's TestVGDMLNavigation checks its navigators on a loaded file: located volumes identical,
20 steps with relocation identical (states and steps bit for bit), safeties identical, in FP64 and
mixed precision. Geometries: the reference's one-solid and two-volume files, a file with one
daughter of every supported solid under general rotations, and the calorimeter written to GDML."""
import os

import numpy as np
import pytest
import torch

from oracle.pyoracle import OracleGeometry, OracleShape
from tests import gdml_util
from vecgeom_amd import SOA3D, geometry, workloads
from vecgeom_amd.gdml import load_gdml
from vecgeom_amd.shapes import FP64, MIXED

pytestmark = pytest.mark.gpu

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "gdml")
FILES = ["allSolids.gdml", "aRotation.gdml", "aPosition.gdml", "twoBoxes.gdml", "oneBox.gdml", "oneConeSegment.gdml",
         "oneGeneralTrapezoid.gdml", "oneHyperbolicTube.gdml", "oneOrb.gdml", "oneParaboloid.gdml",
         "onePolycone.gdml", "onePolyhedron.gdml", "oneSphere.gdml", "oneTesselatedSolid.gdml", "oneTorus.gdml",
         "oneTrapezoid.gdml", "oneTube.gdml", "oneEllipsoid.gdml",
         "oneCutTube.gdml", "oneEllipticalCone.gdml"]
KIND = {v: k for k, v in __import__("oracle.pyoracle", fromlist=["SHAPE_TYPES"]).SHAPE_TYPES.items()}


def _points(g, n, seed):
    """Uniform points in 1.1 x the world's extent, isotropic directions."""
    t, p = g.shapes[g.logicals[g.world]]
    lo, hi = (np.array(v) for v in OracleShape(KIND[t], p).extent())
    c, h = 0.5 * (lo + hi), 0.55 * (hi - lo)
    rng = workloads.rng(seed)
    return workloads.uniform_in_box(rng, n, c - h, c + h, scale=1.0), workloads.isotropic_dirs(rng, n)


def _u32(a):
    return torch.as_tensor(np.asarray(a).astype(np.int32), device="cuda")


def _geometries():
    out = [(f, lambda f=f: load_gdml(os.path.join(FIX, f), 1.0)) for f in FILES]

    def calo():
        import tempfile
        g = geometry.small_calorimeter()
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "calo.gdml")
            gdml_util.write(g, path)
            return load_gdml(path, 1.0)
    out.append(("calorimeter", calo))
    out.append(("allSolids-specialized", lambda: load_gdml(os.path.join(FIX, "allSolids.gdml"), 1.0, specialized=True)))
    out.append(("parallelepipeds-and-tets", planar_solids))
    return out


def planar_solids():
    """Parallelepipeds, tets, elliptical tubes, ellipsoids and planar generic trapezoids placed with
    rotations in a box world, one parallelepiped holding a tet, one elliptical tube holding a cut
    ellipsoid, one generic trapezoid and one cut tube holding a tet."""
    from tests.test_nav_path import _rot
    from vecgeom_amd.shapes import BOX, CUTTUBE, ELLIPTICALCONE, ELLIPSOID, ELLIPTICALTUBE, GENTRAP, PARALLELEPIPED, TET
    g = geometry.GeometryDesc()
    w = g.add_logical(g.add_shape(BOX, [300.0, 300.0, 300.0]), "world")
    g.set_world(w)
    para = g.add_logical(g.add_shape(PARALLELEPIPED, [40.0, 30.0, 50.0, 0.3, 0.25, 0.7]), "para")
    tet = g.add_logical(g.add_shape(TET, [0.0, 0.0, 30.0, -20.0, -15.0, -10.0, 25.0, -10.0, -10.0, 0.0, 25.0, -10.0]),
                        "tet")
    small = g.add_logical(g.add_shape(TET, [0.0, 0.0, 8.0, -6.0, -5.0, -4.0, 7.0, -4.0, -4.0, 0.0, 7.0, -4.0]),
                          "small tet")
    g.place(para, small, (5.0, -3.0, 2.0) + _rot("x", 0.4))
    for k, (x, y, axis, ang) in enumerate([(-150, -150, "z", 0.3), (150, -140, "y", 0.8), (-140, 150, "x", -0.5)]):
        g.place(w, para, (float(x), float(y), 20.0 * k) + _rot(axis, ang))
    for k, (x, y, axis, ang) in enumerate([(150, 150, "z", 1.1), (0, 0, "x", 0.2), (0, -180, "y", -1.3)]):
        g.place(w, tet, (float(x), float(y), -30.0 * k) + _rot(axis, ang))
    eltube = g.add_logical(g.add_shape(ELLIPTICALTUBE, [45.0, 25.0, 60.0]), "elliptical tube")
    ell = g.add_logical(g.add_shape(ELLIPSOID, [20.0, 15.0, 40.0, -30.0, 25.0]), "ellipsoid")
    g.place(eltube, ell, (8.0, 2.0, -5.0) + _rot("z", 0.5))
    g.place(w, eltube, (150.0, 0.0, 100.0) + _rot("x", 0.6))
    g.place(w, eltube, (-150.0, 10.0, -120.0) + _rot("y", 1.2))
    g.place(w, ell, (80.0, 200.0, 0.0) + _rot("z", -0.4))
    gtra = g.add_logical(g.add_shape(GENTRAP, [-40.0, 50.0, 60.0, -30.0, -14.0, 40.0, 46.0, -8.0,
                                                -50.0, -40.0, 40.0, 30.0, -25.0, -19.0, 29.0, 23.0, 45.0]),
                         "generic trapezoid")
    g.place(gtra, small, (0.0, 0.0, -10.0) + _rot("z", 0.9))
    g.place(w, gtra, (-200.0, -20.0, 150.0) + _rot("y", 0.7))
    g.place(w, gtra, (220.0, 60.0, -180.0) + _rot("x", -1.0))
    ctub = g.add_logical(g.add_shape(CUTTUBE, [0.0, 30.0, 40.0, 0.0, 2 * np.pi, 0.0, -0.3, -1.0, 0.2, 0.0, 1.0]),
                         "cut tube")
    g.place(ctub, small, (0.0, 5.0, 0.0) + _rot("y", 0.3))
    g.place(w, ctub, (-20.0, 200.0, 180.0) + _rot("z", 0.8))
    elco = g.add_logical(g.add_shape(ELLIPTICALCONE, [0.5, 0.4, 60.0, 40.0]), "elliptical cone")
    g.place(w, elco, (-180.0, 180.0, -180.0) + _rot("x", 0.5))
    g.place(w, elco, (200.0, -200.0, 200.0) + _rot("y", -0.9))
    return g


@pytest.mark.parametrize("name,make", _geometries(), ids=[n for n, _ in _geometries()])
@pytest.mark.parametrize("prec", [FP64, MIXED], ids=["fp64", "mixed"])
def test_gdml_navigation(name, make, prec):
    from vecgeom_amd.navigation import Geometry
    print("gdml geometry", name, "precision", prec, flush=True)
    g = make()
    G, O = Geometry(g), OracleGeometry(g)
    has_torus = any(t == 7 for t, _ in g.shapes)
    n = 1 << 13
    p, d = _points(g, n, seed=31)
    nav = O.locate(p, threads=8)
    got = G.LocateGlobalPoint(SOA3D.from_array(p), precision=prec).cpu().numpy().astype(np.uint32)
    assert np.array_equal(got, nav), int((got != nav).sum())
    keep = nav != 0  # the steps start inside the world
    p, d, nav = np.ascontiguousarray(p[:, keep]), np.ascontiguousarray(d[:, keep]), nav[keep]
    saf = G.ComputeSafety(SOA3D.from_array(p), _u32(nav), precision=prec).cpu().numpy()
    assert np.array_equal(saf, O.safety(p, nav, threads=8))
    last = np.zeros_like(nav)
    D = SOA3D.from_array(d)
    for s in range(20):
        step, nav_o, last_o, bnd = G.ComputeStep(SOA3D.from_array(p), D, 25.0, _u32(nav), _u32(last),
                                                 relocate=True, precision=prec)
        rs, rn, rl, rb = O.step(p, d, 25.0, nav, last, relocate=True, threads=8)
        assert np.array_equal(nav_o.cpu().numpy().astype(np.uint32), rn), s
        assert np.array_equal(last_o.cpu().numpy().astype(np.uint32), rl), s
        assert np.array_equal(bnd.cpu().numpy(), rb), s
        st = step.cpu().numpy()
        diff = float(np.max(np.abs(st - rs), initial=0.0))
        if diff:
            print("step %d: %d steps differ, max %.3g mm" % (s, int((st != rs).sum()), diff), flush=True)
        # the torus's cubic uses the device math library (last-bit differences, held to 1e-9 mm as in the
        # shape tests); every other solid's step is bit-identical
        assert diff <= (1e-9 if has_torus else 0.0), (s, diff)
        alive = rn != 0
        if not alive.any():
            break
        p = np.ascontiguousarray((p + rs * d)[:, alive])
        d = np.ascontiguousarray(d[:, alive])
        D = SOA3D.from_array(d)
        nav, last = rn[alive], rl[alive]
