"""The header-only C++ adapter (include/vgmi.hpp) compiled with g++ into a host program, linked
against libvgmi.so and run without a GPU: a host-only geometry closed through vgmi::Geometry
gives the same NavIndex table as the Python wrapper and the oracle, vgmi::UnplacedVolume gives
the shape's extent, and failures arrive as vgmi::Error with the ABI's status codes."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle.pyoracle import OracleGeometry
from vecgeom_amd import _lib, geometry

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _write_desc(g, path):
    with open(path, "w") as f:
        f.write("%d\n" % len(g.shapes))
        for t, params in g.shapes:
            f.write("%d %d %s\n" % (t, len(params), " ".join(repr(float(x)) for x in params)))
        f.write("%d %s\n" % (len(g.logicals), " ".join(str(x) for x in g.logicals)))
        f.write("%d\n" % len(g.placements))
        for m, d, t in g.placements:
            f.write("%d %d %s\n" % (m, d, " ".join(repr(float(x)) for x in t)))
        f.write("%d\n" % g.world)


@pytest.fixture(scope="module")
def adapter(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    assert os.path.exists(_lib.LIB_PATH), "libvgmi.so not built"
    out = tmp_path_factory.mktemp("adapter") / "adapter_main"
    libdir = os.path.dirname(_lib.LIB_PATH)
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "adapter_main.cpp"), "-o", str(out), "-L", libdir, "-lvgmi",
           "-Wl,-rpath," + libdir]
    subprocess.run(cmd, check=True)
    return out


def _run(exe, g, tmp_path):
    desc = tmp_path / "desc.txt"
    _write_desc(g, desc)
    r = subprocess.run([str(exe), str(desc)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = dict(line.split(" ", 1) if " " in line else (line, "") for line in r.stdout.splitlines())
    return lines


@pytest.mark.parametrize("which", ["small", "full"])
def test_adapter_closes_host_only_geometry(adapter, tmp_path, which):
    g = geometry.small_calorimeter() if which == "small" else geometry.calorimeter()
    out = _run(adapter, g, tmp_path)
    from vecgeom_amd.navigation import Geometry
    G = Geometry(g, host_only=True)
    words = np.array([int(x) for x in out["navindex"].split()], dtype=np.uint32)
    assert np.array_equal(words, G.navindex())
    assert np.array_equal(words, OracleGeometry(g).navindex())
    info = [int(x) for x in out["info"].split()]
    ref = G.info()
    assert info[:5] == [ref["placed"], ref["logical"], ref["navindex_words"], ref["max_depth"], ref["world_placed"]]
    assert info[5] == 0  # host-only: nothing on the device
    assert int(out["world"]) == len(g.placements)
    # extent of the first shape, against the Python wrapper of the same ABI call
    from vecgeom_amd.shapes import UnplacedShape
    t, params = g.shapes[0]
    lo, hi = UnplacedShape(t, params).Extent()
    assert [float(x) for x in out["extent"].split()] == list(lo) + list(hi)
    assert int(out["errors"]) == 3
