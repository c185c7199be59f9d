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
