"""CPU checks of the navigation oracle and of the product's host geometry builder.

Pins the oracle with the reference's own cross-checks:
 - BVH navigation against the brute-force NewSimpleNavigator loop, step by step
   (persistency/gdml/test/TestVGDMLNavigation/TestVGDMLNavigation.cpp:72-90);
 - the NavIndex table against a depth-first walk of the placement tree
   (test/benchmark/TestNavIndex.cu:30-138: push/pop/top/level/daughters/matrix);
 - located volumes contain the point, none of their daughters does (closed form).
The product builder (vecgeom_amd, host-only build: no GPU calls) must produce
the same NavIndex words and BVH arrays as the oracle.
"""
import numpy as np
import pytest

from oracle.pyoracle import OracleGeometry, OracleShape
from vecgeom_amd import geometry

SHAPE_NAMES = {0: "box", 1: "tube", 5: "trapezoid"}


@pytest.fixture(scope="module")
def small():
    g = geometry.small_calorimeter()
    return g, OracleGeometry(g), OracleGeometry(g, use_bvh=False)


@pytest.fixture(scope="module")
def full():
    g = geometry.calorimeter()
    return g, OracleGeometry(g)


def _walk(g):
    """(pvol, level, parent index, child slot, global matrix) in depth-first order."""
    daughters = {}
    for k, (m, d, t) in enumerate(g.placements):
        daughters.setdefault(m, []).append(k)
    world = len(g.placements)
    out = []

    def rec(pv, lv, level, mat):
        out.append((pv, level, mat))
        for slot, k in enumerate(daughters.get(lv, [])):
            t = np.array(g.placements[k][2])
            M = np.eye(4)
            M[:3, :3] = t[3:].reshape(3, 3)
            M[:3, 3] = t[:3]
            rec(k, g.placements[k][1], level + 1, mat @ M)

    rec(world, g.world, 0, np.eye(4))
    return out, daughters


def test_navindex_table_matches_placement_tree(small):
    g, o, _ = small
    t = o.navindex()
    walk, daughters = _walk(g)
    # reconstruct node indices by following push/pop from the world
    idx = 1
    nodes = []
    stack = [(1, len(g.placements), g.world)]
    while stack:
        ni, pv, lv = stack.pop()
        nodes.append(ni)
        assert t[ni + 1] == pv
        nd = t[ni + 2] >> 16
        assert nd == len(daughters.get(lv, []))
        for slot in reversed(range(nd)):
            child = t[ni + 3 + slot]
            assert child > ni and t[child] == ni  # pop(child) == node
            k = daughters[lv][slot]
            stack.append((child, k, g.placements[k][1]))
    assert len(nodes) == g.touchable_count()
    # levels and cached global matrices in depth-first order
    for ni, (pv, level, mat) in zip(sorted(nodes), walk):
        assert t[ni + 1] == pv and (t[ni + 2] & 0xFF) == level
        nd = t[ni + 2] >> 16
        ti = ni + 3 + nd + ((nd + 1) & 1)
        ti += ti & 1
        m = t[ti:ti + 24].view(np.float64)
        R = m[3:].reshape(3, 3)
        assert np.allclose(R, mat[:3, :3], atol=1e-12) and np.allclose(m[:3], mat[:3, 3], atol=1e-9)


def test_builder_tables_match_oracle(small, full):
    from vecgeom_amd.navigation import Geometry
    for g, o in ((small[0], small[1]), full):
        G = Geometry(g, host_only=True)
        assert np.array_equal(G.navindex(), o.navindex())
        for lv in range(len(g.logicals)):
            a, b = G.bvh(lv), o.bvh(lv)
            assert (a is None) == (b is None)
            if a is not None:
                assert a["depth"] == b["depth"]
                for k in ("nchild", "offset", "primid", "aabbs"):
                    assert np.array_equal(a[k], b[k]), (lv, k)


def test_bvh_navigation_matches_brute_force(small, full):
    for g, o, ob in (small, (full[0], full[1], OracleGeometry(full[0], use_bvh=False))):
        p, d = geometry.world_points(g, 4000, seed=11)
        nav = o.locate(p)
        assert np.array_equal(nav, ob.locate(p))
        last = np.zeros_like(nav)
        for s in range(30):
            alive = nav != 0
            if not alive.any():
                break
            a = o.step(p, d, 10.0, nav, last)
            b = ob.step(p, d, 10.0, nav, last)
            for x, y in zip(a, b):
                assert np.array_equal(x, y), s
            step, nav, last, _ = a
            p = p + step * d


def test_located_volume_contains_point(small):
    g, o, _ = small
    p, _ = geometry.world_points(g, 3000, seed=5)
    nav = o.locate(p)
    t = o.navindex()
    walk, daughters = _walk(g)
    mats = {}
    # global matrix per node from the table
    for i in range(p.shape[1]):
        ni = nav[i]
        if ni == 0:
            continue
        nd = t[ni + 2] >> 16
        ti = ni + 3 + nd + ((nd + 1) & 1)
        ti += ti & 1
        m = t[ti:ti + 24].view(np.float64)
        R, tr = m[3:].reshape(3, 3), m[:3]
        local = (p[:, i] - tr) @ R
        pv = t[ni + 1]
        lv = g.world if pv == len(g.placements) else g.placements[pv][1]
        st, prm = g.shapes[g.logicals[lv]]
        assert OracleShape(SHAPE_NAMES[st], prm).query("Contains", local.reshape(3, 1))[0] == 1
        for k in daughters.get(lv, []):
            tt = np.array(g.placements[k][2])
            dl = (local - tt[:3]) @ tt[3:].reshape(3, 3)
            st2, prm2 = g.shapes[g.logicals[g.placements[k][1]]]
            assert OracleShape(SHAPE_NAMES[st2], prm2).query("Contains", dl.reshape(3, 1))[0] == 0


def test_safety_is_a_lower_bound_of_step(small):
    """BVH safety <= brute-force daughter safety is not required, but the isotropic
    safety must never exceed the distance travelled to the next boundary."""
    g, o, ob = small
    p, d = geometry.world_points(g, 3000, seed=9)
    nav = o.locate(p)
    s_bvh = o.safety(p, nav)
    s_brute = ob.safety(p, nav)
    step, _, _, bnd = o.step(p, d, 1e9, nav)
    alive = nav != 0
    assert np.all(s_bvh[alive] >= s_brute[alive] - 1e-12)
    assert np.all(s_bvh[alive] <= step[alive] + 1e-9)
