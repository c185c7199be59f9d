"""Continuous-batching track-transport server over the C ABI (vgmi_server_*).

A request is a set of tracks (global points, directions, per-track step counts) in host
memory. The library locates every track from the world, runs up to nsteps[i]
ComputeStepAndPropagatedState steps per track on the device (the loop a transport engine
runs around VNavigator::ComputeStepAndPropagatedState, VecGeom/navigation/VNavigator.h:436-540),
and writes final points, NavStateIndex values and steps taken back to host arrays.
Requests from many threads are packed into shared launches as lanes free up.

`poisson_load` is the load generator: producer threads with Poisson arrivals, batch
sizes log-uniform in [2^10, 2^17.6] tracks and per-track step counts uniform in 1..100.
"""
import ctypes
import queue
import threading
import time

import numpy as np

from ._lib import check, lib
from .shapes import FP64


def _declare():
    L = lib()
    if getattr(L, "_serve_declared", False):
        return L
    vp, i64, ci, d = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double
    sigs = {
        "vgmi_server_create": [vp, ci, d, i64, ci, i64, ctypes.POINTER(vp)],
        "vgmi_server_submit": [vp] + [vp] * 7 + [i64] + [vp] * 5 + [ctypes.POINTER(ctypes.c_uint64)],
        "vgmi_server_wait": [vp, ctypes.c_uint64, ctypes.POINTER(d)],
        "vgmi_server_poll": [vp, ctypes.c_uint64, ctypes.POINTER(ci)],
        "vgmi_server_cancel": [vp, ctypes.c_uint64],
        "vgmi_server_set_hbm_reserve": [vp, i64],
        "vgmi_server_stats": [vp, vp],
        "vgmi_server_graph_stats": [vp, vp],
        "vgmi_server_destroy": [vp],
    }
    for name, args in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    L._serve_declared = True
    return L


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Request:
    """One submitted request; its arrays stay referenced until `Server.wait` returns."""

    def __init__(self, p, d, nsteps):
        self.p = np.ascontiguousarray(p, dtype=np.float64)
        self.d = np.ascontiguousarray(d, dtype=np.float64)
        self.nsteps = np.ascontiguousarray(nsteps, dtype=np.int32)
        n = self.nsteps.shape[0]
        if self.p.shape != (3, n) or self.d.shape != (3, n):
            raise ValueError("points and directions must be (3, n) with n = len(nsteps)")
        self.p_out = np.empty((3, n), dtype=np.float64)
        self.nav = np.empty(n, dtype=np.uint32)
        self.taken = np.empty(n, dtype=np.int32)
        self.ticket = None
        self.latency_us = None

    @property
    def n(self):
        return self.nsteps.shape[0]


class Server:
    """vgmi_server: lanes of `lane_tracks` tracks, `lanes` launches in flight."""

    def __init__(self, geom, precision=FP64, step_limit=10.0, lane_tracks=1 << 18, lanes=4,
                 queue_tracks=1 << 21):
        L = _declare()
        h = ctypes.c_void_p()
        check(L.vgmi_server_create(geom._h, precision, float(step_limit), int(lane_tracks), int(lanes),
                                   int(queue_tracks), ctypes.byref(h)), "vgmi_server_create")
        self._h = h
        self.geom = geom  # keeps the geometry alive while the server runs

    def submit(self, p, d, nsteps):
        r = Request(p, d, nsteps)
        t = ctypes.c_uint64()
        check(lib().vgmi_server_submit(self._h, _ptr(r.p[0]), _ptr(r.p[1]), _ptr(r.p[2]), _ptr(r.d[0]),
                                       _ptr(r.d[1]), _ptr(r.d[2]), _ptr(r.nsteps), r.n, _ptr(r.p_out[0]),
                                       _ptr(r.p_out[1]), _ptr(r.p_out[2]), _ptr(r.nav), _ptr(r.taken),
                                       ctypes.byref(t)), "vgmi_server_submit")
        r.ticket = t.value
        return r

    def wait(self, r):
        lat = ctypes.c_double()
        check(lib().vgmi_server_wait(self._h, r.ticket, ctypes.byref(lat)), "vgmi_server_wait")
        r.latency_us = lat.value
        return r

    def poll(self, r):
        """True once the request has finished (wait() then returns at once); never blocks."""
        done = ctypes.c_int()
        check(lib().vgmi_server_poll(self._h, r.ticket, ctypes.byref(done)), "vgmi_server_poll")
        return bool(done.value)

    def cancel(self, r):
        """Cancels a request that is still queued; wait() then raises for it. Raises when the
        request is already on the GPU or done."""
        check(lib().vgmi_server_cancel(self._h, r.ticket), "vgmi_server_cancel")

    def set_hbm_reserve(self, nbytes):
        """Admission by free HBM: submissions and batches are refused while less than `nbytes`
        of device memory is free (0 turns the check off)."""
        check(lib().vgmi_server_set_hbm_reserve(self._h, int(nbytes)), "vgmi_server_set_hbm_reserve")

    def stats(self):
        s = np.zeros(4, dtype=np.int64)
        check(lib().vgmi_server_stats(self._h, _ptr(s)), "vgmi_server_stats")
        g = np.zeros(2, dtype=np.int64)
        check(lib().vgmi_server_graph_stats(self._h, _ptr(g)), "vgmi_server_graph_stats")
        return {"launches": int(s[0]), "tracks": int(s[1]), "steps": int(s[2]), "refused": int(s[3]),
                "graph_launches": int(g[0]), "graphs_per_lane": int(g[1])}

    def close(self):
        if getattr(self, "_h", None) is not None:
            check(lib().vgmi_server_destroy(self._h), "vgmi_server_destroy")
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def track_pool(desc, n, seed):
    """Host pool of world points and isotropic directions that requests slice from."""
    from .geometry import world_points
    return world_points(desc, n, seed)


def poisson_load(server, pool, rate_hz, duration_s, producers=16, seed=0, log2_min=10.0, log2_max=17.6,
                 max_steps=100):
    """Open-loop load: `producers` threads, each with Poisson arrivals at rate_hz / producers.

    Returns per-request (tracks, steps done, latency in us) and the wall time from the first
    submission to the last completion. Refused submissions (queue full) are counted.
    """
    p_all, d_all = pool
    npool = p_all.shape[1]
    records, refused = [], [0]
    lock = threading.Lock()
    t_start = time.perf_counter() + 0.05
    t_end = t_start + duration_s

    def producer(k):
        rng = np.random.default_rng([seed, k])
        pending = queue.Queue()

        def waiter():
            while True:
                r = pending.get()
                if r is None:
                    return
                server.wait(r)
                with lock:
                    records.append((r.n, int(r.taken.sum()), r.latency_us))

        w = threading.Thread(target=waiter)
        w.start()
        t = t_start + rng.exponential(producers / rate_hz)
        while t < t_end:
            n = int(2.0 ** rng.uniform(log2_min, log2_max))
            off = int(rng.integers(0, npool - n + 1))
            nsteps = rng.integers(1, max_steps + 1, size=n, dtype=np.int32)
            now = time.perf_counter()
            if t > now:
                time.sleep(t - now)
            try:
                pending.put(server.submit(p_all[:, off:off + n], d_all[:, off:off + n], nsteps))
            except RuntimeError:
                with lock:
                    refused[0] += 1
            t += rng.exponential(producers / rate_hz)
        pending.put(None)
        w.join()

    threads = [threading.Thread(target=producer, args=(k,)) for k in range(producers)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    wall = time.perf_counter() - t_start
    return records, refused[0], wall


def summarize(records, refused, wall, small_max=1 << 16):
    """p50/p99 latency (all requests and those of at most `small_max` tracks), nav-steps/s."""
    if not records:
        return {"requests": 0, "refused": refused}
    a = np.array(records, dtype=np.float64)
    lat_ms = a[:, 2] / 1e3
    small = a[:, 0] <= small_max
    out = {"requests": int(len(a)), "refused": int(refused), "tracks": int(a[:, 0].sum()),
           "nav_steps": int(a[:, 1].sum()), "wall_s": round(wall, 3),
           "nav_steps_per_s": round(a[:, 1].sum() / wall, 1),
           "p50_ms": round(float(np.percentile(lat_ms, 50)), 3), "p99_ms": round(float(np.percentile(lat_ms, 99)), 3)}
    if small.any():
        out["p50_ms_le64k"] = round(float(np.percentile(lat_ms[small], 50)), 3)
        out["p99_ms_le64k"] = round(float(np.percentile(lat_ms[small], 99)), 3)
    tiny = a[:, 0] <= (1 << 14)
    if tiny.any():
        out["p99_ms_le16k"] = round(float(np.percentile(lat_ms[tiny], 99)), 3)
    return out
