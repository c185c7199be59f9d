"""GPU parity of the BVH navigation kernels (FP64 and mixed precision) against the CPU oracle.

Bar (from the north star): next-volume indices / navigation states identical, FP64
steps and safeties within kTolerance (expected bit-identical: the kernels evaluate
the same FP64 operations as the oracle, without contraction). The mixed-precision
kernels use FP32 box tests only as conservative filters in front of the same FP64
decisions, so they are held to the same bit-identical bar.
"""
import numpy as np
import pytest
import torch

from oracle.pyoracle import OracleGeometry
from vecgeom_amd import SOA3D, geometry
from vecgeom_amd.shapes import FP64, MIXED

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=["small", "calorimeter"])
def geo(request):
    from vecgeom_amd.navigation import Geometry
    g = geometry.small_calorimeter() if request.param == "small" else geometry.calorimeter()
    print("geometry", request.param, flush=True)
    return g, Geometry(g), OracleGeometry(g)


def _dev_u32(a):
    return torch.as_tensor(np.asarray(a).astype(np.int32), device="cuda")


PRECISIONS = pytest.mark.parametrize("prec", [FP64, MIXED], ids=["fp64", "mixed"])


@PRECISIONS
def test_locate(geo, prec):
    g, G, O = geo
    p, _ = geometry.world_points(g, 1 << 16, seed=21)
    got = G.LocateGlobalPoint(SOA3D.from_array(p), precision=prec).cpu().numpy().astype(np.uint32)
    ref = O.locate(p, threads=8)
    assert np.array_equal(got, ref), int((got != ref).sum())


@PRECISIONS
def test_step_by_step(geo, prec):
    g, G, O = geo
    n = 1 << 14
    p, d = geometry.world_points(g, n, seed=22)
    nav = O.locate(p, threads=8)
    last = np.zeros_like(nav)
    D = SOA3D.from_array(d)
    for relocate in (False, True):
        pp, nv, ls = p.copy(), nav.copy(), last.copy()
        for s in range(20):
            step, nav_o, last_o, bnd = G.ComputeStep(SOA3D.from_array(pp), D, 10.0, _dev_u32(nv), _dev_u32(ls),
                                                     relocate=relocate, precision=prec)
            rs, rn, rl, rb = O.step(pp, d, 10.0, nv, ls, relocate=relocate, threads=8)
            assert np.array_equal(nav_o.cpu().numpy().astype(np.uint32), rn), (relocate, s)
            assert np.array_equal(last_o.cpu().numpy().astype(np.uint32), rl), (relocate, s)
            assert np.array_equal(bnd.cpu().numpy(), rb), (relocate, s)
            st = step.cpu().numpy()
            assert np.max(np.abs(st - rs)) <= 1e-9, (relocate, s)
            assert np.array_equal(st, rs), (relocate, s, int((st != rs).sum()))
            pp = pp + rs * d
            nv, ls = rn, rl


@PRECISIONS
def test_step_and_safety(geo, prec):
    """ComputeStepAndSafetyAndPropagatedState: the step and out-state of ComputeStepAndPropagatedState
    and the safety of the start point (VNavigator.h:1014-1064), where requested."""
    g, G, O = geo
    n = 1 << 14
    p, d = geometry.world_points(g, n, seed=25)
    nav = O.locate(p, threads=8)
    last = np.zeros_like(nav)
    calc = np.random.default_rng(5).random(n) < 0.7
    D = SOA3D.from_array(d)
    pp, nv, ls = p.copy(), nav.copy(), last.copy()
    for s in range(10):
        step, nav_o, last_o, bnd, saf = G.ComputeStepAndSafety(
            SOA3D.from_array(pp), D, 10.0, _dev_u32(nv), _dev_u32(ls),
            calc_safety=torch.as_tensor(calc, device="cuda"), precision=prec)
        rs, rn, rl, rb = O.step(pp, d, 10.0, nv, ls, relocate=True, threads=8)
        rsaf = np.where(calc & (nv != 0), O.safety(pp, nv, threads=8), 0.0)
        assert np.array_equal(nav_o.cpu().numpy().astype(np.uint32), rn), s
        assert np.array_equal(last_o.cpu().numpy().astype(np.uint32), rl), s
        assert np.array_equal(bnd.cpu().numpy(), rb), s
        assert np.array_equal(step.cpu().numpy(), rs), s
        assert np.array_equal(saf.cpu().numpy(), rsaf), (s, float(np.max(np.abs(saf.cpu().numpy() - rsaf))))
        pp = pp + rs * d
        nv, ls = rn, rl


@PRECISIONS
def test_axis_aligned_directions_and_outside_points(geo, prec):
    """Directions with exactly zero components (the NonZero reciprocal paths of the slab tests and
    the shapes) and points outside the world (state 0), step by step against the oracle."""
    g, G, O = geo
    n = 6 * 2048
    p, _ = geometry.world_points(g, n, seed=26)
    axes = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], dtype=np.float64)
    d = np.ascontiguousarray(np.repeat(axes, n // 6, axis=0).T)
    p[:, :64] *= 1.5  # some points outside the world box
    p = np.ascontiguousarray(p)
    nav = O.locate(p, threads=8)
    got = G.LocateGlobalPoint(SOA3D.from_array(p), precision=prec).cpu().numpy().astype(np.uint32)
    assert np.array_equal(got, nav)
    assert (nav == 0).sum() > 0
    last = np.zeros_like(nav)
    D = SOA3D.from_array(d)
    pp, nv, ls = p.copy(), nav.copy(), last.copy()
    for s in range(15):
        step, nav_o, last_o, bnd = G.ComputeStep(SOA3D.from_array(pp), D, 10.0, _dev_u32(nv), _dev_u32(ls),
                                                 relocate=True, precision=prec)
        rs, rn, rl, rb = O.step(pp, d, 10.0, nv, ls, relocate=True, threads=8)
        assert np.array_equal(nav_o.cpu().numpy().astype(np.uint32), rn), s
        assert np.array_equal(last_o.cpu().numpy().astype(np.uint32), rl), s
        assert np.array_equal(step.cpu().numpy(), rs), s
        pp = pp + rs * d
        nv, ls = rn, rl


@PRECISIONS
def test_safety(geo, prec):
    g, G, O = geo
    p, _ = geometry.world_points(g, 1 << 15, seed=23)
    nav = O.locate(p, threads=8)
    got = G.ComputeSafety(SOA3D.from_array(p), _dev_u32(nav), precision=prec).cpu().numpy()
    ref = O.safety(p, nav, threads=8)
    assert np.array_equal(got, ref), float(np.max(np.abs(got - ref)))


@PRECISIONS
@pytest.mark.parametrize("lanes", [0, 1, 8, 64])
@pytest.mark.parametrize("regroup", [0, 7, 10])
def test_propagate_matches_oracle(geo, regroup, lanes, prec):
    """The fused loop, unsorted (0) and regrouped by volume every 7 / 10 steps, with automatic,
    1, 8 and 64 tracks per wavefront: same results."""
    g, G, O = geo
    n, nsteps = (1 << 16 if regroup else 1 << 14), 50  # regrouping starts at 65536 tracks
    p, d = geometry.world_points(g, n, seed=24)
    nav = O.locate(p, threads=8)
    last = np.zeros_like(nav)
    P, D = SOA3D.from_array(p), SOA3D.from_array(d)
    nv, ls = _dev_u32(nav), _dev_u32(last)
    taken = torch.zeros(n, dtype=torch.int32, device="cuda")
    G.set_regroup(regroup)
    G.set_lanes_per_wave(lanes)
    try:
        G.Propagate(P, D, 10.0, nv, ls, nsteps, taken, precision=prec)
    finally:
        G.set_regroup(10)
        G.set_lanes_per_wave(0)
    rp, rn, rl, rtaken = O.propagate(p, d, 10.0, nav, last, nsteps, threads=8)
    assert np.array_equal(nv.cpu().numpy().astype(np.uint32), rn)
    assert np.array_equal(ls.cpu().numpy().astype(np.uint32), rl)
    assert int(taken.sum()) == rtaken
    assert np.array_equal(P.cpu(), rp), float(np.max(np.abs(P.cpu() - rp)))


def _placed_cases(g):
    """Placement ids of a layer tube, a rotated trapezoid cell and a box plate, and the world."""
    from vecgeom_amd.shapes import BOX, TRAPEZOID, TUBE
    want, out = {TUBE: None, TRAPEZOID: None, BOX: None}, []
    for k, (m, d, t) in enumerate(g.placements):
        st = g.shapes[g.logicals[d]][0]
        rotated = any(abs(a - b) > 1e-12 for a, b in zip(t[3:], geometry.IDENTITY[3:]))
        if st in want and want[st] is None and (st != TRAPEZOID or rotated):
            want[st] = k
    return [k for k in want.values() if k is not None] + [len(g.placements)]


def _placed_points(g, pid, n, seed):
    """Points in the placed volume's frame (half in 1.3x its extent, half near its surface) and the
    same points in its mother's frame (master = R local + t)."""
    from oracle.pyoracle import OracleShape, SHAPE_TYPES
    from vecgeom_amd import workloads
    names = {v: k for k, v in SHAPE_TYPES.items()}
    lv = g.world if pid == len(g.placements) else g.placements[pid][1]
    st, params = g.shapes[g.logicals[lv]]
    o = OracleShape(names[st], params)
    lo, hi = o.extent()
    rg = workloads.rng(seed)
    half = n // 2
    loc = np.concatenate([workloads.uniform_in_box(rg, half, lo, hi, scale=1.3),
                          workloads.near_surface(rg, n - half, lo, hi, lambda q: o.query("Contains", q),
                                                 lambda q, e: o.query("DistanceToOut", q, e))], axis=1)
    d = workloads.isotropic_dirs(rg, loc.shape[1])
    t = np.array(geometry.IDENTITY if pid == len(g.placements) else g.placements[pid][2])
    R = t[3:].reshape(3, 3)
    master = t[:3, None] + R @ loc
    dm = R @ d
    return np.ascontiguousarray(loc), np.ascontiguousarray(d), np.ascontiguousarray(master), np.ascontiguousarray(dm)


@PRECISIONS
def test_placed_queries(geo, prec):
    """VPlacedVolume SOA queries (SpecializedPlacedVolImplHelper.h:98-179) against the oracle's placed
    dispatch: FP64 bit for bit; mixed: decisions identical, certified distances within 2.5e-4 mm."""
    g, G, O = geo
    for pid in _placed_cases(g):
        loc, dl, mas, dm = _placed_points(g, pid, 1 << 13, seed=40 + pid % 97)
        print("placed volume", pid, flush=True)
        for op in ("Contains", "Inside", "DistanceToIn", "SafetyToIn", "PlacedDistanceToOut", "DistanceToOut",
                   "SafetyToOut"):
            local = op in ("DistanceToOut", "SafetyToOut")
            p, d = (loc, dl) if local else (mas, dm)
            directed = "Distance" in op
            got = G.PlacedQuery(pid, op, SOA3D.from_array(p), SOA3D.from_array(d) if directed else None,
                                precision=prec).cpu().numpy()
            ref = O.placed_query(pid, 6 if op == "PlacedDistanceToOut" else op, p, d if directed else None)
            if prec == FP64 or op in ("Contains", "Inside"):
                assert np.array_equal(got, ref), (pid, op, int((got != ref).sum()))
            else:
                assert np.array_equal(got < 0, ref < 0) and np.array_equal(got >= 1e300, ref >= 1e300), (pid, op)
                fin = (ref < 1e300) & (ref >= 0)
                assert np.max(np.abs(got[fin] - ref[fin]), initial=0) <= 2.5e-4, (pid, op)
        # the placed dispatch equals the unplaced shape on the local points (pins the transform)
        from oracle.pyoracle import OracleShape, SHAPE_TYPES
        lv = g.world if pid == len(g.placements) else g.placements[pid][1]
        st, params = g.shapes[g.logicals[lv]]
        un = OracleShape({v: k for k, v in SHAPE_TYPES.items()}[st], params)
        ins_loc = un.query("Inside", loc)
        ins_mas = O.placed_query(pid, "Inside", mas)
        assert (ins_loc == ins_mas).mean() > 0.999, pid  # equal up to the rounding of the transform


def test_placed_normals(geo):
    """VPlacedVolume::Normal (source/PlacedVolume.cpp:126-137) against the oracle's, bit for bit, at points
    near and on the placed volume's surface given in its mother's frame; and the normal is the rotated
    normal of the unplaced shape at the local point (pins the frame handling)."""
    from oracle.pyoracle import OracleShape, SHAPE_TYPES
    g, G, O = geo
    names = {v: k for k, v in SHAPE_TYPES.items()}
    for pid in _placed_cases(g):
        loc, dl, mas, _ = _placed_points(g, pid, 1 << 13, seed=70 + pid % 89)
        lv = g.world if pid == len(g.placements) else g.placements[pid][1]
        st, params = g.shapes[g.logicals[lv]]
        un = OracleShape(names[st], params)
        inside = un.query("Contains", loc).astype(bool)
        t_out = un.query("DistanceToOut", loc[:, inside], dl[:, inside])
        ok = t_out < 1e290
        exits = loc[:, inside][:, ok] + t_out[ok] * dl[:, inside][:, ok]
        t = np.array(geometry.IDENTITY if pid == len(g.placements) else g.placements[pid][2])
        R = t[3:].reshape(3, 3)
        loc_all = np.concatenate([loc, exits], axis=1)
        mas_all = np.ascontiguousarray(np.concatenate([mas, t[:3, None] + R @ exits], axis=1))
        print("placed volume", pid, "points", mas_all.shape[1], flush=True)
        got_n, got_v = G.PlacedNormal(pid, SOA3D.from_array(mas_all))
        got_n, got_v = got_n.cpu(), got_v.cpu().numpy()
        ref_n, ref_v = O.placed_normal(pid, mas_all)
        assert np.array_equal(got_n, ref_n) and np.array_equal(got_v, ref_v), pid
        ln, lvld = un.normal(loc_all, init=np.zeros_like(loc_all))
        close = np.all(np.abs(R @ ln - ref_n) < 1e-9, axis=0) & (lvld == ref_v)
        assert close.mean() > 0.999, (pid, close.mean())


@PRECISIONS
def test_level_locate(geo, prec):
    """BVHLevelLocator::LevelLocate and LevelLocateExclVol per point against the oracle: the daughter
    found and the point in its frame, bit for bit; with the found daughter excluded, the next one."""
    from oracle.pyoracle import OracleShape, SHAPE_TYPES
    from vecgeom_amd import workloads
    g, G, O = geo
    names = {v: k for k, v in SHAPE_TYPES.items()}
    mothers = sorted({m for m, _, _ in g.placements})
    lvs = [g.world] + [m for m in mothers if m != g.world][:3]
    for k, lv in enumerate(lvs):
        st, params = g.shapes[g.logicals[lv]]
        lo, hi = OracleShape(names[st], params).extent()
        p = workloads.uniform_in_box(workloads.rng(60 + k), 1 << 14, lo, hi, scale=1.0)
        P = SOA3D.from_array(p)
        got, dl = G.LevelLocate(lv, P, precision=prec)
        ref, rdl = O.level_locate(lv, p)
        got = got.cpu().numpy()
        assert np.array_equal(got, ref), (lv, int((got != ref).sum()))
        assert (ref >= 0).mean() > 0.01, lv
        assert np.array_equal(dl.cpu(), rdl), lv
        excl = torch.as_tensor(ref, dtype=torch.int32, device="cuda")
        got2, dl2 = G.LevelLocate(lv, P, excl=excl, precision=prec)
        ref2, rdl2 = O.level_locate(lv, p, excl=ref)
        assert np.array_equal(got2.cpu().numpy(), ref2), lv
        assert np.array_equal(dl2.cpu(), rdl2), lv
    # an out-of-range logical volume finds nothing
    bad, _ = G.LevelLocate(len(g.logicals) + 5, SOA3D.from_array(np.zeros((3, 4))), precision=prec)
    assert (bad.cpu().numpy() == -1).all()
