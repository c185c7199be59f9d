"""NavStateIndex <-> path interop (NavStatePath-style states) in the oracle.

Pins: every touchable of the geometry is enumerated by an independent depth-first walk of the
placement tree (its NavIndex found by following the table's daughter links, its placed volumes
and daughter positions along the way, its global matrix composed from the placements); the
oracle's conversions must give exactly those paths, the paths must give back the NavIndex, the
NavStateIndex::Distance of random pairs must equal the brute-force count of up and down moves
over the two paths, and GlobalToLocal must agree with the composed matrices.
"""
import numpy as np
import pytest

from oracle.pyoracle import OracleGeometry
from vecgeom_amd import geometry


def touchables(g, table):
    """[(nav, [pvols], [daughter indices], 4x4 global matrix)] by a depth-first walk."""
    daughters = {}
    for k, (m, d, t) in enumerate(g.placements):
        daughters.setdefault(m, []).append(k)
    world = len(g.placements)
    out = []
    stack = [(1, [world], [0], g.world, np.eye(4))]
    while stack:
        nav, pv, ix, lv, mat = stack.pop()
        out.append((nav, pv, ix, mat))
        for slot, k in enumerate(daughters.get(lv, [])):
            t = np.array(g.placements[k][2])
            M = np.eye(4)
            M[:3, :3] = t[3:].reshape(3, 3)
            M[:3, 3] = t[:3]
            stack.append((int(table[nav + 3 + slot]), pv + [k], ix + [slot], g.placements[k][1], mat @ M))
    return out


def _rot(axis, a):
    c, s = np.cos(a), np.sin(a)
    R = {"x": [[1, 0, 0], [0, c, -s], [0, s, c]], "y": [[c, 0, s], [0, 1, 0], [-s, 0, c]],
         "z": [[c, -s, 0], [s, c, 0], [0, 0, 1]]}[axis]
    return tuple(np.array(R, dtype=float).ravel())


def nested():
    """Four levels, general rotations at every level, repeated logical volumes."""
    from vecgeom_amd.shapes import BOX, TUBE
    g = geometry.GeometryDesc()
    w = g.add_logical(g.add_shape(BOX, [500, 500, 500]), "world")
    g.set_world(w)
    a = g.add_logical(g.add_shape(BOX, [120, 120, 120]), "a")
    b = g.add_logical(g.add_shape(TUBE, [0, 40, 30, 0, 2 * np.pi]), "b")
    c = g.add_logical(g.add_shape(BOX, [5, 6, 7]), "c")
    for k, (x, axis, ang) in enumerate([(-250, "z", 0.3), (250, "x", -0.7)]):
        g.place(w, a, (x, 10.0 * k, -20.0) + _rot(axis, ang))
    for k, (y, ang) in enumerate([(-60, 0.2), (0, 1.1), (60, -0.5)]):
        g.place(a, b, (5.0, y, 3.0 * k) + _rot("y", ang))
    for k, z in enumerate([-12.0, 12.0]):
        g.place(b, c, (3.0, -2.0, z) + _rot("z", 0.4 * k))
    return g


@pytest.fixture(scope="module", params=["small", "nested"])
def geo(request):
    g = geometry.small_calorimeter() if request.param == "small" else nested()
    o = OracleGeometry(g)
    return g, o, touchables(g, o.navindex())


def _as_arrays(ts, L):
    n = len(ts)
    nav = np.array([t[0] for t in ts], np.uint32)
    levels = np.array([len(t[1]) for t in ts], np.int32)
    pv = np.full((L, n), -1, np.int32)
    ix = np.full((L, n), -1, np.int32)
    for i, t in enumerate(ts):
        pv[:len(t[1]), i] = t[1]
        ix[:len(t[2]), i] = t[2]
    return nav, levels, pv, ix


def test_paths_of_every_touchable(geo):
    g, o, ts = geo
    assert len(ts) == g.touchable_count()
    L = o.max_depth()
    nav, levels, pv, ix = _as_arrays(ts, L)
    got_l, got_pv, got_ix = o.nav_to_path(nav, L)
    assert np.array_equal(got_l, levels) and np.array_equal(got_pv, pv) and np.array_equal(got_ix, ix)
    assert np.array_equal(o.nav_from_path(levels, pv, 1), nav)
    assert np.array_equal(o.nav_from_path(levels, ix, 0), nav)
    # the outside state
    l0, p0, i0 = o.nav_to_path(np.zeros(1, np.uint32), L)
    assert l0[0] == 0 and (p0 == -1).all() and (i0 == -1).all()


def test_invalid_paths_give_the_outside_state(geo):
    g, o, ts = geo
    L = o.max_depth()
    deep = [t for t in ts if len(t[1]) >= 2][:50]
    nav, levels, pv, ix = _as_arrays(deep, L)
    bad_ix = ix.copy()
    bad_ix[1] = 10 ** 6                      # no such daughter
    assert (o.nav_from_path(levels, bad_ix, 0) == 0).all()
    bad_pv = pv.copy()
    bad_pv[0] = pv[1]                         # first entry not the world
    assert (o.nav_from_path(levels, bad_pv, 1) == 0).all()
    bad_pv = pv.copy()
    bad_pv[1] = len(g.placements) + 5         # not a placed volume
    assert (o.nav_from_path(levels, bad_pv, 1) == 0).all()
    assert (o.nav_from_path(np.zeros_like(levels), pv, 1) == 0).all()
    # a first daughter index other than 0 is not read (ResetPathFromListOfIndices pushes the world)
    ix2 = ix.copy()
    ix2[0] = 7
    assert np.array_equal(o.nav_from_path(levels, ix2, 0), nav)


def test_distance_is_the_number_of_moves(geo):
    g, o, ts = geo
    rng = np.random.default_rng(3)
    a = rng.integers(0, len(ts), 4000)
    b = rng.integers(0, len(ts), 4000)
    b[:200] = a[:200]                          # same state
    want = []
    for i, j in zip(a, b):
        pa, pb = ts[i][1], ts[j][1]
        common = 0
        while common < min(len(pa), len(pb)) and pa[:common + 1] == pb[:common + 1]:
            common += 1
        want.append((len(pa) - common) + (len(pb) - common))
    nav_a = np.array([ts[i][0] for i in a], np.uint32)
    nav_b = np.array([ts[j][0] for j in b], np.uint32)
    got = o.nav_distance(nav_a, nav_b)
    assert np.array_equal(got, np.array(want, np.int32))
    assert (got[:200] == 0).all()


def test_global_to_local_matches_composed_placements(geo):
    g, o, ts = geo
    rng = np.random.default_rng(4)
    sel = rng.integers(0, len(ts), 2000)
    nav = np.array([ts[i][0] for i in sel], np.uint32)
    p = rng.uniform(-150, 150, size=(3, len(sel)))
    by_path = {tuple(t[1]): t for t in ts}
    for tolevel in (-1, 0, 1, 2, 3):
        got = o.global_to_local(nav, p, tolevel)
        for c, i in enumerate(sel):
            depth = len(ts[i][1]) - 1
            if tolevel > depth:
                want = p[:, c]                    # outside state: identity
            else:
                k = depth if tolevel < 0 else tolevel
                # the ancestor at level k is the prefix of the path
                anc = by_path[tuple(ts[i][1][:k + 1])]
                M = anc[3]
                want = M[:3, :3].T @ (p[:, c] - M[:3, 3])
            assert np.allclose(got[:, c], want, rtol=0, atol=1e-9), (i, tolevel)
