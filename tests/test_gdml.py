"""GDML ingestion (vgmi_gdml_read) on the CPU: the reference's one-solid GDML files (copied under
tests/golden/gdml/) give the solids the reference's middleware builds from them, with its units,
halvings and GeoManager::MakeInstance substitutions; rotations follow the passive convention,
checked against a product of elementary rotations; errors are reported for unsupported solids and
malformed documents; a written calorimeter comes back exactly."""
import math
import os

import numpy as np
import pytest

from tests import gdml_util
from vecgeom_amd import VgmiError, geometry
from vecgeom_amd import shapes as S
from vecgeom_amd.gdml import load_gdml

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "gdml")
MM = 0.1  # the reference's default internal value of one millimetre
CM = 10 * MM
DEG = math.pi / 180.


def _one(name, mm=MM, specialized=False):
    g = load_gdml(os.path.join(FIX, name), mm, specialized=specialized)
    assert g.world == 0 and len(g.logicals) == 1 and not g.placements
    return g.shapes[0]


# (file, expected (type, params)) worked out by hand from each file's solid element, as the reference's
# default build (VECGEOM_NO_SPECIALIZATION=ON) makes them: every solid as written
ONE_SOLID = [
    ("oneBox.gdml", (S.BOX, [100 * CM / 2] * 3)),
    ("oneOrb.gdml", (S.ORB, [100 * CM])),
    ("aConstant.gdml", (S.ORB, [100 * CM])),  # r="radius", <constant name="radius" value="100"/>
    ("aVariable.gdml", (S.ORB, [0.0])),  # <variable> is not read: an unknown name is 0, as in the reference
    ("oneConeSegment.gdml", (S.CONE, [15 * CM, 30 * CM, 70 * CM, 100 * CM, 50 * CM / 2, 1.0, 4.0])),
    ("oneGeneralTrapezoid.gdml", (S.TRAPEZOID, [100 * CM / 2, 1.0, 2.0, 150 * CM / 2, 100 * CM / 2, 10 * CM / 2,
                                                1.0, 150 * CM / 2, 10 * CM / 2, 10 * CM / 2, 1.0])),
    ("oneHyperbolicTube.gdml", (S.HYPE, [10 * CM, 20 * CM, 3.0, 4.0, 100 * CM / 2])),
    ("oneParaboloid.gdml", (S.PARABOLOID, [50 * CM, 75 * CM, 100 * CM])),
    ("onePolycone.gdml", (S.POLYCONE, [1.0, 4.0, 2, 50 * MM, 80 * MM, 15 * MM, 70 * MM, 30 * MM, 100 * MM])),
    ("onePolyhedron.gdml", (S.POLYHEDRON, [1.0, 4.0, 11, 2, 50 * MM, 80 * MM, 15 * MM, 70 * MM, 30 * MM, 100 * MM])),
    ("oneTorus.gdml", (S.TORUS2, [10 * CM, 100 * CM, 20 * CM, 1.0, 3.0])),
    ("oneTrapezoid.gdml", (S.TRD, [90 * CM / 2, 80 * CM / 2, 60 * CM / 2, 50 * CM / 2, 100 * CM / 2])),
    ("oneTube.gdml", (S.TUBE, [0.0, 100 * CM, 100 * CM / 2, 0.0, 360 * DEG])),
    # semi-axes and cuts as written (processEllipsoid halves nothing)
    ("oneEllipsoid.gdml", (S.ELLIPSOID, [30 * CM, 70 * CM, 100 * CM, 20 * CM, 40 * CM])),
    # z halved; the cut normals scaled by the length unit like lengths (they are made unit afterwards)
    ("oneEllipticalCone.gdml", (S.ELLIPTICALCONE, [70 * CM, 30 * CM, 100 * CM, 15 * CM])),  # all in lengths
    ("oneCutTube.gdml", (S.CUTTUBE, [1 * CM, 5 * CM, 20 * CM / 2, 1.0, 4.0, -15 * CM, -15 * CM, -15 * CM,
                                     10 * CM, 10 * CM, 10 * CM])),
]


# the same files through the substitutions of a build with VECGEOM_NO_SPECIALIZATION=OFF, where they differ
ONE_SOLID_SPECIALIZED = {
    # two z planes: the factory's cone of half-height |z1 - z0| / 2 (centred at z = 0)
    "onePolycone.gdml": (S.CONE, [15 * MM, 30 * MM, 70 * MM, 100 * MM, abs(80 * MM - 50 * MM) / 2, 1.0, 4.0]),
}


@pytest.mark.parametrize("name,exp", ONE_SOLID, ids=[c[0] for c in ONE_SOLID])
def test_one_solid_files(name, exp):
    t, p = _one(name)
    assert t == exp[0], (t, exp[0])
    assert p == [float(x) for x in exp[1]], (p, exp[1])
    exp = ONE_SOLID_SPECIALIZED.get(name, exp)
    t, p = _one(name, specialized=True)
    assert (t, p) == (exp[0], [float(x) for x in exp[1]]), (t, p)


def test_full_sphere_is_an_orb():
    # 0..360 deg and 0..180 deg: a sphere as written; with the substitutions MakeInstance<UnplacedSphere>
    # returns an orb when the ranges are full
    sphere = (S.SPHERE, [0.0, 100 * CM, 0.0, 360 * DEG, 0.0, 180 * DEG])
    assert _one("oneSphere.gdml") == sphere
    full = 360 * DEG >= 2 * math.pi and 180 * DEG == math.pi
    assert _one("oneSphere.gdml", specialized=True) == ((S.ORB, [100 * CM]) if full else sphere)


def test_tessellated_facets():
    t, p = _one("oneTesselatedSolid.gdml")
    assert t == S.TESSELLATED and p[0] == 8  # six triangles and one quadrilateral (two triangles)
    v = {"v1": (10, 10, 0), "v2": (-10, 10, 0), "v3": (-10, -10, 0), "v4": (10, -10, 0), "v5": (7, 3, 20),
         "v6": (-3, 7, 5)}
    tri = [("v1", "v2", "v6"), ("v2", "v3", "v6"), ("v3", "v4", "v5"), ("v4", "v1", "v5"), ("v1", "v6", "v5"),
           ("v6", "v3", "v5"), ("v4", "v3", "v2"), ("v4", "v2", "v1")]
    exp = [c * CM for a in tri for k in a for c in v[k]]
    assert p[1:] == exp


def test_positions_and_units():
    g = load_gdml(os.path.join(FIX, "aPosition.gdml"))
    assert g.shapes == [(S.BOX, [5 * CM] * 3), (S.BOX, [50 * CM] * 3)]
    assert g.logicals == [0, 1] and g.world == 1
    (m, d, t), = g.placements
    assert (m, d) == (1, 0) and t[:3] == (7 * CM, 17 * CM, 37 * CM) and t[3:] == (1., 0., 0., 0., 1., 0., 0., 0., 1.)
    # the same file at one millimetre per unit
    g = load_gdml(os.path.join(FIX, "twoBoxes.gdml"), 1.0)
    assert g.shapes == [(S.BOX, [50.0] * 3), (S.BOX, [500.0] * 3)]


def _rx(a):
    return np.array([[1, 0, 0], [0, math.cos(a), -math.sin(a)], [0, math.sin(a), math.cos(a)]])


def _ry(a):
    return np.array([[math.cos(a), 0, math.sin(a)], [0, 1, 0], [-math.sin(a), 0, math.cos(a)]])


def _rz(a):
    return np.array([[math.cos(a), -math.sin(a), 0], [math.sin(a), math.cos(a), 0], [0, 0, 1]])


def test_rotation_is_the_passive_xyz_rotation():
    g = load_gdml(os.path.join(FIX, "aRotation.gdml"))
    (m, d, t), = g.placements
    r = np.array(t[3:]).reshape(3, 3)
    x, y, z = 7 * DEG, 17 * DEG, 37 * DEG
    # the inverse of rotating about x, then y, then z: Rx(-x) Ry(-y) Rz(-z)
    assert np.allclose(r, _rx(-x) @ _ry(-y) @ _rz(-z), atol=1e-15, rtol=0)
    assert np.allclose(r @ r.T, np.eye(3), atol=1e-15) and abs(np.linalg.det(r) - 1) < 1e-15


def test_unsupported_solids_fail(tmp_path):
    text = BODY % ('<twistedtubs name="c" twistedangle="30" endinnerrad="10" endouterrad="15" zlen="40" '
                   'phi="90" aunit="deg"/>', "<volume name='W'><solidref ref='c'/></volume>", "W")
    with pytest.raises(VgmiError, match="twistedtubs"):
        load_gdml(_write(tmp_path, text))


def test_arb8_vertices_carry_no_unit(tmp_path):
    # processGenTrap reads v1x..v8y without the length unit (DECLAREANDGETPLAINVAR) and dz with it; the
    # reference's oneArbitraryTrapezoid.gdml names two solids "world", which the reader refuses
    with pytest.raises(VgmiError, match="defined twice"):
        load_gdml(os.path.join(FIX, "oneArbitraryTrapezoid.gdml"))
    arb8 = ('<arb8 name="a" v1x="-3" v1y="-3" v2x="-3" v2y="3" v3x="3" v3y="3" v4x="3" v4y="-3" v5x="-2" '
            'v5y="-2" v6x="-2" v6y="2" v7x="2" v7y="2" v8x="2" v8y="-2" dz="5" lunit="cm"/>')
    g = load_gdml(_write(tmp_path, BODY % (arb8, "<volume name='W'><solidref ref='a'/></volume>", "W")))
    assert g.shapes[0] == (S.GENTRAP, [-3.0, -3.0, 3.0, 3.0, -2.0, -2.0, 2.0, 2.0, -3.0, 3.0, 3.0, -3.0, -2.0, 2.0,
                                       2.0, -2.0, 5 * CM])


def _write(tmp_path, text, name="t.gdml"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


BODY = """<gdml>
 <solids>%s</solids>
 <structure>%s</structure>
 <setup name="default" version="1.0"><world ref="%s"/></setup>
</gdml>"""


@pytest.mark.parametrize("text,word", [
    ("<gdml><solids><box name='b' x='1' y='1' z='1'></solids></gdml>", "closes"),
    ("<gdml><solids>", "missing"),
    ("<gdml a=1/>", "quoted"),
    ("<notgdml/>", "root element"),
    (BODY % ("<box name='b' x='1' y='1' z='1'/><box name='b' x='2' y='2' z='2'/>",
             "<volume name='W'><solidref ref='b'/></volume>", "W"), "twice"),
    (BODY % ("<box name='b' x='1' y='1' z='1'/>", "<volume name='W'><solidref ref='c'/></volume>", "W"),
     "unknown solid"),
    (BODY % ("<box name='b' x='1' y='1' z='1'/>", "<volume name='W'><solidref ref='b'/></volume>", "X"),
     "unknown volume"),
    (BODY % ("<box name='b' x='1' y='1' z='1'/><box name='w' x='9' y='9' z='9'/>",
             "<volume name='D'><solidref ref='b'/></volume><volume name='W'><solidref ref='w'/>"
             "<physvol><volumeref ref='D'/><scale name='s' x='1' y='1' z='-1'/></physvol></volume>", "W"),
     "reflected"),
])
def test_malformed_documents_fail(tmp_path, text, word):
    with pytest.raises(VgmiError, match=word):
        load_gdml(_write(tmp_path, text))


def test_parser_details(tmp_path):
    text = """<?xml version="1.0" encoding="UTF-8"?>
<!DOCTYPE gdml [ <!ENTITY x "y"> ]>
<!-- a comment -->
<gdml xmlns:xsi="http://www.w3.org/2001/XMLSchema-instance">
  <define>
    <constant name="half" value="0.5e1"/>
    <variable name="ignored" value="3"/>
    <position name='p' unit="m" x=' 0.25' y="2*cm" z="half"/>
  </define>
  <![CDATA[ <box name="not" x="1"/> ]]>
  <solids>
    <box name="a&amp;b" x="1.5" y="2" z="3" lunit="cm"/>
    <box name="w" x="10" y="10" z="10" lunit="m"/>
  </solids>
  <structure>
    <volume name="D"><solidref ref="a&amp;b"/></volume>
    <volume name="W"><solidref ref="w"/><physvol name="pv"><volumeref ref="D"/><positionref ref="p"/></physvol></volume>
  </structure>
  <setup name="default" version="1.0"><world ref="W"/></setup>
</gdml>"""
    g = load_gdml(_write(tmp_path, text), 1.0)
    assert g.shapes == [(S.BOX, [7.5, 10.0, 15.0]), (S.BOX, [5000.0] * 3)]
    # x: 0.25 m; y: "2*cm" reads as its leading number 2 (in the position's unit, m); z: the constant 5
    assert g.placements[0][2][:3] == (250.0, 2000.0, 5000.0)
    assert g.names == ["D", "W"] and g.placement_names == ["pv"]


def test_make_instance_substitutions(tmp_path):
    solids = """
      <cone name="c1" rmin1="0" rmax1="5" rmin2="0" rmax2="5" z="20" startphi="0" deltaphi="7" aunit="rad"/>
      <cone name="c2" rmin1="2" rmax1="5" rmin2="2" rmax2="5" z="20" startphi="0" deltaphi="1" aunit="rad"/>
      <cone name="c3" rmin1="0" rmax1="5" rmin2="0" rmax2="6" z="20" startphi="0" deltaphi="8" aunit="rad"/>
      <hype name="h1" rmin="0" rmax="5" inst="0" outst="0" z="20"/>
      <hype name="h2" rmin="1" rmax="5" inst="0" outst="0" z="20"/>
      <hype name="h3" rmin="0" rmax="5" inst="0" outst="0.5" z="20"/>
      <trap name="t1" z="20" theta="0" phi="0" y1="4" x1="6" x2="6" alpha1="0" y2="4" x3="6" x4="6" alpha2="0"/>
      <trap name="t2" z="20" theta="0" phi="0" y1="4" x1="6" x2="6" alpha1="0" y2="4" x3="8" x4="8" alpha2="0"/>
      <trap name="t3" z="20" theta="0" phi="0" y1="4" x1="6" x2="6" alpha1="0" y2="5" x3="8" x4="8" alpha2="0"/>
      <trd name="d1" x1="6" x2="6" y1="4" y2="4" z="20"/>
      <trd name="d2" x1="6" x2="8" y1="4" y2="4" z="20"/>
      <sphere name="s1" rmin="0" rmax="5" startphi="0" deltaphi="360" starttheta="0" deltatheta="90" aunit="deg"/>
      <trap name="t4" z="20" theta="0.2" phi="0.3" y1="4" x1="6" x2="6" alpha1="0.1" y2="4" x3="6" x4="6" alpha2="0.1"/>
      <para name="p1" x="6" y="4" z="20" alpha="0.1" theta="0.2" phi="0.3"/>
      <box name="w" x="100" y="100" z="100"/>"""
    text = BODY % (solids, "<volume name='W'><solidref ref='w'/></volume>", "W")
    # the reference's default build: every solid as written
    g = load_gdml(_write(tmp_path, text), 1.0)
    assert [t for t, _ in g.shapes] == [S.CONE] * 3 + [S.HYPE] * 3 + [S.TRAPEZOID] * 3 + [S.TRD] * 2 + [
        S.SPHERE, S.TRAPEZOID, S.PARALLELEPIPED, S.BOX]
    assert g.shapes[10] == (S.TRD, [3.0, 4.0, 2.0, 2.0, 10.0]) and g.shapes[0][1][6] == 7.0
    assert g.shapes[13] == (S.PARALLELEPIPED, [3.0, 2.0, 10.0, 0.1, 0.2, 0.3])
    # with VECGEOM_NO_SPECIALIZATION=OFF's substitutions
    g = load_gdml(_write(tmp_path, text), 1.0, specialized=True)
    two_pi = 2 * math.pi
    exp = [(S.TUBE, [0.0, 5.0, 10.0, 0.0, two_pi]), (S.TUBE, [2.0, 5.0, 10.0, 0.0, 1.0]),
           (S.CONE, [0.0, 5.0, 0.0, 6.0, 10.0, 0.0, two_pi]),
           (S.TUBE, [0.0, 5.0, 10.0, 0.0, two_pi]), (S.TUBE, [1.0, 5.0, 10.0, 0.0, two_pi]),
           (S.HYPE, [0.0, 5.0, 0.0, 0.5, 10.0]),
           (S.BOX, [3.0, 2.0, 10.0]), (S.TRD, [3.0, 4.0, 2.0, 2.0, 10.0, 1.0]), (S.TRD, [3.0, 4.0, 2.0, 2.5, 10.0]),
           (S.BOX, [3.0, 2.0, 10.0]), (S.TRD, [3.0, 4.0, 2.0, 2.0, 10.0, 1.0]),
           (S.SPHERE, [0.0, 5.0, 0.0, 360 * DEG, 0.0, 90 * DEG]), (S.PARALLELEPIPED, [3.0, 2.0, 10.0, 0.1, 0.2, 0.3]),
           (S.PARALLELEPIPED, [3.0, 2.0, 10.0, 0.1, 0.2, 0.3]), (S.BOX, [50.0] * 3)]
    assert g.shapes == exp


@pytest.mark.parametrize("which", ["small", "full"])
def test_calorimeter_round_trip(tmp_path, which):
    g = geometry.small_calorimeter() if which == "small" else geometry.calorimeter()
    path = str(tmp_path / "calo.gdml")
    order = gdml_util.write(g, path)
    got = load_gdml(path, 1.0)
    shapes, logicals, placements, world = gdml_util.expected(g, order)
    assert got.shapes == shapes and got.logicals == logicals and got.world == world
    assert len(got.placements) == len(placements)
    for (m, d, t), (em, ed, et) in zip(got.placements, placements):
        assert (m, d) == (em, ed) and list(t) == et
    # the written rotation is the description's to rounding
    for (m, d, t) in g.placements:
        a = gdml_util.z_angle(t)
        assert np.allclose([math.cos(a), math.sin(a)], [t[3], t[4]], atol=1e-15, rtol=0)
    # the same file in the reference's default unit (0.1 per mm)
    got = load_gdml(path)
    assert got.shapes[0][1] == [v * MM for v in shapes[0][1]]


def test_all_solids_file_bvh_matches_brute_force():
    """The file with every supported solid: the oracle's BVH navigation equals its brute-force
    navigation (every daughter tested), and every located volume contains its point."""
    from oracle.pyoracle import OracleGeometry
    g = load_gdml(os.path.join(FIX, "allSolids.gdml"), 1.0)
    assert [t for t, _ in g.shapes] == [S.BOX, S.BOX, S.TRD, S.TRD, S.PARABOLOID, S.HYPE, S.CONE, S.POLYCONE,
                                         S.POLYHEDRON, S.SPHERE, S.TORUS2, S.TUBE, S.ORB, S.TRAPEZOID, S.TESSELLATED]
    assert g.shapes[3][1] == [20.0, 30.0, 15.0, 15.0, 25.0]  # y1 == y2, a UniversalTrd as written
    assert load_gdml(os.path.join(FIX, "allSolids.gdml"), 1.0, specialized=True).shapes[3][1][5] == 1.0  # Trd1
    o, ob = OracleGeometry(g), OracleGeometry(g, use_bvh=False)
    rng = np.random.default_rng(7)
    centres = np.array([t[:3] for _, _, t in g.placements]).T  # the daughters' origins (in their cores)
    p = np.concatenate([rng.uniform(-300, 300, size=(3, 20000)), centres], axis=1)
    d = rng.normal(size=p.shape)
    d /= np.linalg.norm(d, axis=0)
    nav = o.locate(p)
    assert np.array_equal(nav, ob.locate(p))
    assert len(set(nav.tolist())) >= 17  # the world, the 13 daughters and cores
    last = np.zeros_like(nav)
    for s in range(15):
        a = o.step(p, d, 25.0, nav, last)
        b = ob.step(p, d, 25.0, nav, last)
        for x, y in zip(a, b):
            assert np.array_equal(x, y), s
        step, nav, last, _ = a
        p = p + step * d


def test_tet_corners_are_named_positions(tmp_path):
    """<tet vertex1..4="name"> refers to <define><position> elements (Middleware.cpp:1155-1172) in their own
    length unit; an unknown name reads (0, 0, 0) as the middleware's position map does."""
    text = """<gdml>
 <define>
  <position name="a" x="0" y="0" z="2" unit="cm"/>
  <position name="b" x="0" y="0" z="0"/>
  <position name="c" x="20" y="0" z="0"/>
 </define>
 <solids>
  <tet name="t1" vertex1="a" vertex2="b" vertex3="c" vertex4="nowhere"/>
  <box name="w" x="100" y="100" z="100"/>
 </solids>
 <structure><volume name="W"><solidref ref="w"/></volume></structure>
 <setup name="default" version="1.0"><world ref="W"/></setup>
</gdml>"""
    g = load_gdml(_write(tmp_path, text), 1.0)
    assert g.shapes[0] == (S.TET, [0.0, 0.0, 20.0, 0.0, 0.0, 0.0, 20.0, 0.0, 0.0, 0.0, 0.0, 0.0])
