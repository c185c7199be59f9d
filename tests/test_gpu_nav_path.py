"""GPU parity of the NavStateIndex <-> path interop (vgmi_nav_to_path / from_path / distance /
global_to_local) with the oracle, on every touchable of the full calorimeter and of a four-level
geometry with general rotations. Bar: integer results identical; GlobalToLocal bit-identical
(both apply the same cached FP64 matrix in the same operation order)."""
import numpy as np
import pytest
import torch

from oracle.pyoracle import OracleGeometry
from tests.test_nav_path import nested, touchables
from vecgeom_amd import SOA3D, geometry
from vecgeom_amd.navigation import Geometry

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=["calorimeter", "nested"])
def geo(request):
    g = geometry.calorimeter() if request.param == "calorimeter" else nested()
    o = OracleGeometry(g)
    return g, o, Geometry(g), touchables(g, o.navindex())


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_to_path_and_back_on_every_touchable(geo):
    g, o, G, ts = geo
    L = o.max_depth()
    nav = np.array([t[0] for t in ts] + [0], np.uint32)
    rl, rpv, rix = o.nav_to_path(nav, L)
    levels, pv, ix = G.NavToPath(_t(nav.view(np.int32)))
    assert np.array_equal(levels.cpu().numpy(), rl)
    assert np.array_equal(pv.cpu().numpy(), rpv) and np.array_equal(ix.cpu().numpy(), rix)
    back_pv = G.PathToNav(levels, pv, G.PATH_PLACED).cpu().numpy().view(np.uint32)
    back_ix = G.PathToNav(levels, ix, G.PATH_INDICES).cpu().numpy().view(np.uint32)
    assert np.array_equal(back_pv, nav) and np.array_equal(back_ix, nav)
    # wider arrays than the depth are accepted; narrower are refused
    l2, pv2, _ = G.NavToPath(_t(nav.view(np.int32)), max_levels=L + 2)
    assert np.array_equal(pv2[:L].cpu().numpy(), rpv) and (pv2[L:] == -1).all()
    with pytest.raises(Exception, match="depth"):
        G.NavToPath(_t(nav.view(np.int32)), max_levels=L - 1)


def test_invalid_paths_match_the_oracle(geo):
    g, o, G, ts = geo
    L = o.max_depth()
    rng = np.random.default_rng(9)
    sel = rng.integers(0, len(ts), 3000)
    nav = np.array([ts[i][0] for i in sel], np.uint32)
    levels, pv, ix = o.nav_to_path(nav, L)
    # corrupt a third of the paths at a random level with a random (often invalid) value
    bad = rng.random(len(sel)) < 0.35
    lev = rng.integers(0, L, len(sel))
    for path in (pv, ix):
        path = path.copy()
        cols = np.nonzero(bad)[0]
        path[lev[cols], cols] = rng.integers(-3, len(g.placements) + 3, len(cols))
        for kind in (0, 1):
            want = o.nav_from_path(levels, path, kind)
            got = G.PathToNav(_t(levels), _t(path), kind).cpu().numpy().view(np.uint32)
            assert np.array_equal(got, want), kind


def test_distance_matches_the_oracle(geo):
    g, o, G, ts = geo
    rng = np.random.default_rng(10)
    a = np.array([ts[i][0] for i in rng.integers(0, len(ts), 50000)] + [0, 0, 1], np.uint32)
    b = np.array([ts[i][0] for i in rng.integers(0, len(ts), 50000)] + [0, 1, 0], np.uint32)
    b[:1000] = a[:1000]
    want = o.nav_distance(a, b)
    got = G.Distance(_t(a.view(np.int32)), _t(b.view(np.int32))).cpu().numpy()
    assert np.array_equal(got, want)


def test_global_to_local_matches_the_oracle(geo):
    g, o, G, ts = geo
    rng = np.random.default_rng(11)
    n = 20000
    nav = np.array([ts[i][0] for i in rng.integers(0, len(ts), n)], np.uint32)
    nav[:5] = 0
    p = rng.uniform(-2000, 2000, size=(3, n))
    for tolevel in (-1, 0, 1, 2, 3, 5):
        want = o.global_to_local(nav, p, tolevel)
        got = G.GlobalToLocal(_t(nav.view(np.int32)), SOA3D.from_array(p), tolevel).cpu()
        assert np.array_equal(got, want), tolevel
