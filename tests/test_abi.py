"""CPU checks of the C ABI boundary: the library loads and exports every
function that include/vgmi.h declares (no compute calls: no GPU here)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"VGMI_API\s+[\w\s\*]+?\b(vgmi_\w+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    from vecgeom_amd import _lib
    assert os.path.exists(_lib.LIB_PATH), "libvgmi.so not built"
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 7
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_version_and_error_paths_without_gpu():
    from vecgeom_amd import VgmiError, lib, make_shape
    assert b"gfx950" in lib().vgmi_version()
    # invalid parameters are rejected on the host, before any device work
    import pytest
    with pytest.raises(VgmiError):
        make_shape("box", [1.0, -2.0, 3.0])
    with pytest.raises(VgmiError):
        make_shape("tube", [1.0, 2.0])
    s = make_shape("tube", [5.0, 10.0, 15.0, 0.0, 6.283185307179586])
    assert s.Extent() == ([-10.0, -10.0, -15.0], [10.0, 10.0, 15.0])


def test_extent_matches_oracle():
    """The product's host-side extent (used for ABBoxes and sampling) equals the oracle's."""
    import math
    from oracle.pyoracle import OracleShape
    from vecgeom_amd import make_shape
    cases = [("box", [10.0, 20.0, 30.0]), ("tube", [2.0, 8.0, 10.0, math.radians(30), math.radians(240)]),
             ("orb", [10.0]), ("cone", [5.0, 10.0, 7.0, 15.0, 10.0, math.radians(20), math.radians(300)]),
             ("trapezoid", [60, math.radians(20), math.radians(5), 40, 30, 40, math.radians(10), 16, 10, 14,
                            math.radians(10)]),
             ("sphere", [6.0, 8.0, 0.0, 2 * math.pi, 0.0, math.pi]),
             ("sphere", [0.0, 10.0, math.radians(20), math.radians(270), math.radians(30), math.radians(100)]),
             ("sphere", [45.0, 50.0, -math.pi / 4, 0.5 * math.pi, math.pi / 4, 0.5 * math.pi]),
             ("sphere", [0.0, 50.0, 0.0, 2 * math.pi, 0.75 * math.pi, 0.25 * math.pi]),
             ("sphere", [6.0, 8.0, math.pi / 6, math.pi / 3, 0.0, math.pi]),
             ("trd", [10.0, 30.0, 20.0, 40.0, 40.0]), ("trd", [30.0, 10.0, 25.0, 40.0]), ("trd", [5.0, 6.0, 7.0]),
             ("paraboloid", [6.0, 9.0, 10.0]), ("paraboloid", [0.0, 8.0, 10.0]),
             ("parallelepiped", [20.0, 30.0, 40.0, math.radians(30), math.radians(30), math.radians(45)]),
             ("parallelepiped", [8.0, 5.0, 12.0, math.radians(-20), math.radians(15), math.radians(200)]),
             ("tet", [0.0, 0.0, 2.0, 0.0, 0.0, 0.0, 2.0, 0.0, 0.0, 0.0, 2.0, 0.0]),
             ("tet", [6.0, 4.0, 5.0, -5.0, -6.0, 4.0, -4.0, 5.0, -6.0, 5.0, -5.0, -4.0]),
             ("ellipticaltube", [5.0, 10.0, 20.0]), ("ellipsoid", [3.0, 4.0, 5.0, -4.5, 3.5]),
             ("ellipsoid", [6.0, 9.0, 4.0, 1.0, 4.0]), ("ellipsoid", [6.0, 9.0, 4.0, -4.0, -1.0]),
             ("ellipsoid", [10.0, 7.0, 12.0]),
             ("gentrap", [-3.0, -3.0, 3.0, 3.0, -2.0, -2.0, 2.0, 2.0, -3.0, 3.0, 3.0, -3.0, -2.0, 2.0, 2.0, -2.0, 5.0]),
             ("gentrap", [-4.0, 5.0, 6.0, -3.0, -1.4, 4.0, 4.6, -0.8, -5.0, -4.0, 4.0, 3.0, -2.5, -1.9, 2.9, 2.3, 8.0]),
             # cut tubes: full circle, and phi ranges that reduce x, y and both z limits
             ("cuttube", [0.0, 10.0, 20.0, 0.0, 2 * math.pi, 0.0, -0.7, -0.71, 0.7, 0.0, 0.71]),
             ("cuttube", [3.0, 8.0, 15.0, math.radians(30), math.radians(240), 0.3, 0.2, -1.0, -0.4, 0.1, 1.0]),
             ("cuttube", [1.0, 5.0, 10.0, 1.0, 4.0, -1.0, -1.0, -1.0, 1.0, 1.0, 1.0]),
             ("cuttube", [2.0, 6.0, 8.0, math.radians(-40), math.radians(80), 0.5, 0.0, -1.0, -0.5, 0.0, 1.0]),
             ("ellipticalcone", [0.5, 0.4, 10.0, 5.0]), ("ellipticalcone", [1.0, 2.0, 3.0, 4.0]),
             ("ellipticalcone", [0.0, 0.3, 5.0, 1.0])]
    # shapes with a variable-length block need a device for creation: extent from the host record
    cases = [c for c in cases if c[0] != "polyhedron"]
    for kind, params in cases:
        lo, hi = make_shape(kind, params).Extent()
        olo, ohi = OracleShape(kind, params).extent()
        assert list(lo) == list(olo) and list(hi) == list(ohi), (kind, params, lo, hi, olo, ohi)


def test_batch_rejects_mismatched_counts_before_native_call():
    import pytest
    from vecgeom_amd import make_shape
    from vecgeom_amd.batch import ShapeBatch
    box = make_shape("box", [1.0, 2.0, 3.0])
    with pytest.raises(ValueError):
        ShapeBatch([box], [10, 20])


def test_product_extent_matches_reference_kats():
    """Host-side extents of the product (shapes without a variable-length block) against the
    Extent asserts of the reference's unit tests."""
    from tests import kat_util
    from vecgeom_amd import make_shape
    fails, n = [], 0
    for kind in ["box", "tube", "trapezoid", "cone", "orb", "sphere", "torus2"]:
        for r in kat_util.load(kind):
            if r["method"] == "Extent":
                n += 1
                got = make_shape(kind, r["shape"]).Extent()
                if not kat_util.check(r, got):
                    fails.append((kind, r["line"], got))
    assert n >= 10 and not fails, fails


def test_server_rejects_bad_arguments_without_gpu():
    """vgmi_server_* validate their arguments on the host before touching a device."""
    import ctypes as C
    from vecgeom_amd import lib, serving
    L = serving._declare()
    h = C.c_void_p()
    assert L.vgmi_server_create(None, 0, 10.0, 1024, 2, 4096, C.byref(h)) == -1
    assert b"geometry" in lib().vgmi_last_error()
    assert L.vgmi_server_create(None, 0, 10.0, 1024, 2, 4096, None) == -1
    t = C.c_uint64()
    assert L.vgmi_server_submit(None, *([None] * 7), 1, *([None] * 5), C.byref(t)) == -1
    assert L.vgmi_server_wait(None, 1, None) == -1
    done = C.c_int()
    assert L.vgmi_server_poll(None, 1, C.byref(done)) == -1
    assert L.vgmi_server_cancel(None, 1) == -1
    assert L.vgmi_server_set_hbm_reserve(None, 0) == -1
    assert L.vgmi_server_destroy(None) == 0


def test_tessellated_facet_count_is_bounded():
    """The facet count is checked before any size arithmetic (at most 2^23 facets)."""
    import math
    import pytest
    from vecgeom_amd import VgmiError, make_shape
    for bad in (float(1 << 24), 1e300, math.nan, 3.0, -5.0):
        with pytest.raises(VgmiError, match="triangles"):
            make_shape("tessellated", [bad] + [0.0] * 36)


def test_device_array_check():
    """The wrappers' tensor check: device, dtype, contiguity and length (host tensors here)."""
    import pytest
    import torch
    from vecgeom_amd.soa import device_array
    with pytest.raises(ValueError, match="device tensor"):
        device_array(torch.zeros(4, dtype=torch.float64), torch.float64, 4, "x")
    with pytest.raises(ValueError, match="torch tensor"):
        device_array([0.0] * 4, torch.float64, 4, "x")


def test_cut_tube_and_elliptical_cone_parameter_rules():
    """A cut tube whose bottom normal points up or top normal points down is refused
    (UnplacedCutTube.h:54-66); an elliptical cone's values below 2 kTolerance are raised to it and
    zcut is held to the height (UnplacedEllipticalCone.cpp:28-42), as the oracle does."""
    import pytest
    from oracle.pyoracle import OracleShape
    from vecgeom_amd import VgmiError, make_shape
    for bad in ([0.0, 5.0, 10.0, 0.0, 6.0, 0.0, 0.0, 1.0, 0.0, 0.0, 1.0],
                [0.0, 5.0, 10.0, 0.0, 6.0, 0.0, 0.0, -1.0, 0.0, 0.0, -1.0]):
        with pytest.raises(VgmiError, match="normal"):
            make_shape("cuttube", bad)
    lo, hi = make_shape("ellipticalcone", [0.5, 0.4, 10.0, 50.0]).Extent()
    olo, ohi = OracleShape("ellipticalcone", [0.5, 0.4, 10.0, 50.0]).extent()
    assert hi[2] == 10.0 and list(lo) == list(olo) and list(hi) == list(ohi)
