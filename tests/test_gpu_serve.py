"""GPU tests of the continuous-batching server (vgmi_server_*).

Bar: every request's result is bit-identical whether it runs alone or packed with
others into shared launches (tracks are independent and the FP64 kernels are
deterministic), and matches the oracle's transport loop (points within kTolerance,
expected bit-identical; NavStateIndex values and step counts exact); the mixed-precision
server returns the FP64 server's results.
"""
import threading
import time

import numpy as np
import pytest

from oracle.pyoracle import OracleGeometry
from vecgeom_amd import FP64, MIXED, VgmiError, geometry
from vecgeom_amd.serving import Server, poisson_load, summarize

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def geo():
    from vecgeom_amd.navigation import Geometry
    g = geometry.small_calorimeter()
    return g, Geometry(g)


def _requests(g, sizes, seed):
    p, d = geometry.world_points(g, sum(sizes), seed=seed)
    rng = np.random.default_rng(seed)
    out, o = [], 0
    for n in sizes:
        out.append((p[:, o:o + n], d[:, o:o + n], rng.integers(1, 101, size=n, dtype=np.int32)))
        o += n
    return out


def test_alone_equals_in_the_mix(geo):
    g, G = geo
    reqs = _requests(g, [1, 63, 1000, 4097, 20000, 300, 65536, 17], seed=31)
    with Server(G, lane_tracks=1 << 17, lanes=3) as S:
        alone = [S.wait(S.submit(*r)) for r in reqs]
        # the same requests from 8 threads at once, twice over, so launches carry several requests
        mixed = [None] * (2 * len(reqs))

        def run(k):
            rr = [S.submit(*reqs[i % len(reqs)]) for i in range(k, 2 * len(reqs), 8)]
            for i, r in zip(range(k, 2 * len(reqs), 8), rr):
                mixed[i] = S.wait(r)

        th = [threading.Thread(target=run, args=(k,)) for k in range(8)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        st = S.stats()
    assert st["launches"] < len(reqs) + 2 * len(reqs), st
    assert st["tracks"] == 3 * sum(len(r[2]) for r in reqs)
    # step counts are at most 100, so every launch replays a captured graph
    assert st["graphs_per_lane"] >= 5 and st["graph_launches"] == st["launches"], st
    for i, m in enumerate(mixed):
        a = alone[i % len(reqs)]
        assert np.array_equal(a.p_out, m.p_out), i
        assert np.array_equal(a.nav, m.nav), i
        assert np.array_equal(a.taken, m.taken), i
        assert m.latency_us > 0


def test_mixed_server_equals_fp64(geo):
    """The mixed-precision server runs the mixed navigation kernels (FP32 box filters in front
    of the FP64 decisions): its results are the FP64 server's, bit for bit."""
    g, G = geo
    reqs = _requests(g, [700, 5000, 30000], seed=35)
    with Server(G, precision=FP64) as S:
        ref = [S.wait(S.submit(*r)) for r in reqs]
    with Server(G, precision=MIXED) as S:
        got = [S.wait(S.submit(*r)) for r in reqs]
    for a, b in zip(ref, got):
        assert np.array_equal(a.p_out, b.p_out)
        assert np.array_equal(a.nav, b.nav)
        assert np.array_equal(a.taken, b.taken)


def test_graph_and_direct_launches_agree(geo):
    """A request whose tracks ask for more steps than the captured graphs cover is launched
    directly; the tracks it shares with a graph-launched request give the same results."""
    g, G = geo
    (p, d, nsteps), = _requests(g, [3000], seed=36)
    longer = nsteps.copy()
    longer[0] = 150  # beyond the graphs' step count: direct launch
    with Server(G, lane_tracks=1 << 14, lanes=2) as S:
        a = S.wait(S.submit(p, d, nsteps))
        st0 = S.stats()
        b = S.wait(S.submit(p, d, longer))
        st1 = S.stats()
    assert st0["graph_launches"] == 1 and st1["graph_launches"] == 1 and st1["launches"] == 2
    assert np.array_equal(a.p_out[:, 1:], b.p_out[:, 1:])
    assert np.array_equal(a.nav[1:], b.nav[1:])
    assert np.array_equal(a.taken[1:], b.taken[1:])


def test_matches_oracle_transport(geo):
    g, G = geo
    O = OracleGeometry(g)
    (p, d, nsteps), = _requests(g, [70000], seed=32)  # above the regrouping threshold (65536)
    with Server(G) as S:
        r = S.wait(S.submit(p, d, nsteps))
    nav0 = O.locate(p, threads=8)
    for k in np.unique(nsteps):
        sel = np.nonzero(nsteps == k)[0]
        rp, rn, rl, rtaken = O.propagate(np.ascontiguousarray(p[:, sel]), np.ascontiguousarray(d[:, sel]), 10.0,
                                         nav0[sel], np.zeros(len(sel), np.uint32), int(k), threads=8)
        assert np.array_equal(r.nav[sel], rn), k
        assert int(r.taken[sel].sum()) == rtaken, k
        assert np.max(np.abs(r.p_out[:, sel] - rp)) <= 1e-9, k
        assert np.array_equal(r.p_out[:, sel], rp), k


def test_admission(geo):
    g, G = geo
    (p, d, nsteps), = _requests(g, [5000], seed=33)
    with pytest.raises(VgmiError, match="unknown precision"):
        Server(G, precision=7)
    with Server(G, precision=FP64, lane_tracks=4096, lanes=1, queue_tracks=4096) as S:
        with pytest.raises(VgmiError, match="not admitted"):
            S.submit(p, d, nsteps)  # larger than a lane
        r = S.wait(S.submit(p[:, :4096], d[:, :4096], nsteps[:4096]))
        assert r.taken.min() >= 1
        assert S.stats()["refused"] == 1
    with pytest.raises(VgmiError, match="free HBM"):
        Server(G, lane_tracks=1 << 40, lanes=16, queue_tracks=1 << 41)


def test_poisson_load_completes(geo):
    g, G = geo
    pool = geometry.world_points(g, 1 << 17, seed=34)
    with Server(G, lane_tracks=1 << 17, lanes=4) as S:
        rec, refused, wall = poisson_load(S, pool, rate_hz=100.0, duration_s=1.0, producers=16, seed=1,
                                          log2_max=16.0)
        st = S.stats()
    s = summarize(rec, refused, wall)
    print(s, st)
    assert s["requests"] > 20 and refused == 0
    assert st["tracks"] == s["tracks"] and st["steps"] == s["nav_steps"]
    assert s["p99_ms"] > 0


@pytest.fixture(scope="module")
def calo():
    from vecgeom_amd.navigation import Geometry
    g = geometry.calorimeter()
    return g, Geometry(g)


@pytest.mark.parametrize("precision", [FP64, MIXED])
def test_full_calorimeter_requests(calo, precision):
    """The serving workload's own geometry (the 23 081-touchable calorimeter) and request sizes
    (1k, 64k and 200k tracks of U{1..100} steps): each request's result is the same alone and in
    the mix, and a sample of its tracks matches the oracle's transport loop step for step."""
    g, G = calo
    print("calorimeter requests, precision", precision, flush=True)
    reqs = _requests(g, [1000, 65536, 200000], seed=41)
    with Server(G, precision=precision, lane_tracks=1 << 18, lanes=3) as S:
        alone = [S.wait(S.submit(*r)) for r in reqs]
        mixed = [None] * len(reqs)

        def run(i):
            mixed[i] = S.wait(S.submit(*reqs[i]))

        th = [threading.Thread(target=run, args=(i,)) for i in range(len(reqs))]
        for t in th:
            t.start()
        for t in th:
            t.join()
    for a, m in zip(alone, mixed):
        assert np.array_equal(a.p_out, m.p_out)
        assert np.array_equal(a.nav, m.nav)
        assert np.array_equal(a.taken, m.taken)
    O = OracleGeometry(g)
    rng = np.random.default_rng(5)
    for (p, d, nsteps), r in zip(reqs, alone):
        sel = np.sort(rng.choice(len(nsteps), size=min(256, len(nsteps)), replace=False))
        nav0 = O.locate(np.ascontiguousarray(p[:, sel]), threads=8)
        for j, i in enumerate(sel):
            rp, rn, rl, rt = O.propagate(np.ascontiguousarray(p[:, i:i + 1]), np.ascontiguousarray(d[:, i:i + 1]),
                                         10.0, nav0[j:j + 1], np.zeros(1, np.uint32), int(nsteps[i]))
            assert r.nav[i] == rn[0] and r.taken[i] == rt, (i, int(nsteps[i]))
            assert np.max(np.abs(r.p_out[:, i] - rp[:, 0])) <= 1e-9, i


def test_destroy_with_queued_requests_and_a_blocked_waiter(geo):
    """Destroying a server whose queue still holds requests: their waiters return an error (the
    request was never launched) instead of blocking, and destroy returns once they have left."""
    g, G = geo
    reqs = _requests(g, [30000] * 6, seed=37)
    S = Server(G, lane_tracks=1 << 15, lanes=1)
    tickets = [S.submit(*r) for r in reqs]
    outcome = []

    def wait_last():
        try:
            S.wait(tickets[-1])
            outcome.append("done")
        except VgmiError as e:
            outcome.append(str(e))

    t = threading.Thread(target=wait_last)
    t.start()
    S.close()
    t.join(timeout=60)
    assert not t.is_alive()
    assert outcome and (outcome[0] == "done" or "never launched" in outcome[0]), outcome


def test_poll_and_cancel(geo):
    """vgmi_server_poll never blocks and turns true once a request is done; a request still in
    the queue can be cancelled (its wait then reports the cancellation), one that is done cannot;
    the other requests' results are unaffected."""
    g, G = geo
    reqs = _requests(g, [30000] * 5, seed=41)
    with Server(G, lane_tracks=1 << 15, lanes=1) as S:
        tickets = [S.submit(*r) for r in reqs]
        cancelled = []
        for t in tickets[::-1]:
            try:
                S.cancel(t)
                cancelled.append(t)
            except VgmiError as e:
                assert "on the GPU" in str(e) or "already done" in str(e), str(e)
        assert cancelled, "with one lane the last requests are still queued"
        for t in tickets:
            if t in cancelled:
                assert S.poll(t)
                with pytest.raises(VgmiError, match="cancelled"):
                    S.wait(t)
            else:
                S.wait(t)
        kept = [(r, t) for r, t in zip(reqs, tickets) if t not in cancelled]
        assert kept
        for r, t in kept:
            alone = S.wait(S.submit(*r))
            assert np.array_equal(alone.nav, t.nav) and np.array_equal(alone.p_out, t.p_out)
        t = S.submit(*reqs[0])
        while not S.poll(t):
            time.sleep(1e-4)
        S.wait(t)
        with pytest.raises(VgmiError, match="unknown ticket"):
            S.poll(t)


def test_admission_by_free_hbm(geo):
    """With a free-HBM reserve above what the GPU has, submissions are refused; a reserve the GPU
    meets admits them again and the results are those of an unrestricted server."""
    g, G = geo
    import torch
    free, total = torch.cuda.mem_get_info()
    (p, d, nsteps), = _requests(g, [4096], seed=43)
    with Server(G, lane_tracks=1 << 15, lanes=2) as S:
        ref = S.wait(S.submit(p, d, nsteps))
        S.set_hbm_reserve(total + (1 << 30))
        before = S.stats()["refused"]
        with pytest.raises(VgmiError, match="free HBM"):
            S.submit(p, d, nsteps)
        assert S.stats()["refused"] == before + 1
        S.set_hbm_reserve(1 << 20)
        r = S.wait(S.submit(p, d, nsteps))
        assert np.array_equal(r.nav, ref.nav) and np.array_equal(r.p_out, ref.p_out)
        S.set_hbm_reserve(0)
