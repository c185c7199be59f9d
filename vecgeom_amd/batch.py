"""One-launch batched shape queries over a mix of shapes, with HIP graph replay and
an autotuner for the launch shape (points per tile, block size, resident blocks per CU).

The tuned launch shape is persisted as JSON keyed by the GPU name and the mix
signature (shape types and counts), so repeated runs reuse it.
"""
import ctypes
import json
import os
import time

import torch

from ._lib import check, lib
from .shapes import FP64, _stream
from .soa import device_array

METHODS = ["Contains", "Inside", "DistanceToIn", "DistanceToOut", "SafetyToIn", "SafetyToOut"]
DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "vecgeom_amd", "batch_tuning.json")


def _declare():
    L = lib()
    if getattr(L, "_batch_declared", False):
        return L
    vp, i, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t
    sigs = {
        "vgmi_batch_create": [vp, vp, i, ctypes.POINTER(vp)],
        "vgmi_batch_destroy": [vp],
        "vgmi_batch_set_tuning": [vp, i, i, i],
        "vgmi_batch_info": [vp, vp],
        "vgmi_batch_query_all": [vp, i] + [vp] * 12 + [vp],
        "vgmi_batch_graph_capture": [vp, i] + [vp] * 12 + [vp],
        "vgmi_batch_graph_launch": [vp, vp],
    }
    for name, args in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    L._batch_declared = True
    return L


class ShapeBatch:
    """A batch of (shape, count) pairs; points of shape k follow those of shape k-1."""

    def __init__(self, shapes, counts):
        L = _declare()
        self.shapes = list(shapes)  # keep the shapes alive: the batch references their device data
        self.counts = [int(c) for c in counts]
        if len(self.shapes) != len(self.counts) or not self.shapes:
            raise ValueError("ShapeBatch needs one count per shape (%d shapes, %d counts)"
                             % (len(self.shapes), len(self.counts)))
        arr = (ctypes.c_void_p * len(self.shapes))(*[s.handle.value for s in self.shapes])
        cnt = (ctypes.c_int64 * len(self.counts))(*self.counts)
        h = ctypes.c_void_p()
        check(L.vgmi_batch_create(arr, cnt, len(self.shapes), ctypes.byref(h)), "vgmi_batch_create")
        self._h = h
        self.total = sum(self.counts)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().vgmi_batch_destroy(h)
            except Exception:
                pass
            self._h = None

    def signature(self):
        return ",".join("%d:%d" % (s.type_id, c) for s, c in zip(self.shapes, self.counts))

    def set_tuning(self, tile, block, blocks_per_cu):
        check(lib().vgmi_batch_set_tuning(self._h, int(tile), int(block), int(blocks_per_cu)), "vgmi_batch_set_tuning")

    def info(self):
        a = (ctypes.c_int64 * 6)()
        check(lib().vgmi_batch_info(self._h, a), "vgmi_batch_info")
        return {"shapes": a[0], "points": a[1], "tiles": a[2], "tile": a[3], "block": a[4], "blocks_per_cu": a[5]}

    def alloc_outputs(self, device="cuda"):
        n = self.total
        return {"Contains": torch.empty(n, dtype=torch.uint8, device=device),
                "Inside": torch.empty(n, dtype=torch.int32, device=device),
                **{m: torch.empty(n, dtype=torch.float64, device=device) for m in METHODS[2:]}}

    def _args(self, points, dirs, outs):
        n = self.total
        if len(points) != n or len(dirs) != n:
            raise ValueError("points (%d) and directions (%d) must hold the batch's %d points"
                             % (len(points), len(dirs), n))
        dt = {"Contains": torch.uint8, "Inside": torch.int32}
        for m in METHODS:
            if outs.get(m) is not None:
                device_array(outs[m], dt.get(m, torch.float64), n, m)
        ptr = lambda m: outs[m].data_ptr() if outs.get(m) is not None else None  # noqa: E731
        return [*points.ptrs(), *dirs.ptrs()] + [ptr(m) for m in METHODS]

    def query_all(self, points, dirs, outs=None, precision=FP64):
        outs = outs if outs is not None else self.alloc_outputs(points.x.device)
        check(lib().vgmi_batch_query_all(self._h, precision, *self._args(points, dirs, outs), _stream()),
              "vgmi_batch_query_all")
        return outs

    def capture(self, points, dirs, outs, precision=FP64):
        """Capture the launch for these buffers into a HIP graph."""
        self._captured = (points, dirs, outs)
        check(lib().vgmi_batch_graph_capture(self._h, precision, *self._args(points, dirs, outs), _stream()),
              "vgmi_batch_graph_capture")

    def replay(self):
        check(lib().vgmi_batch_graph_launch(self._h, _stream()), "vgmi_batch_graph_launch")


def _time_launch(fn, reps):
    s = torch.cuda.current_stream()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    fn()
    a.record(s)
    for _ in range(reps):
        fn()
    b.record(s)
    torch.cuda.synchronize()
    return a.elapsed_time(b) * 1e3 / reps


def autotune(batch, points, dirs, outs, tiles=(128, 256, 512, 1024, 2048, 4096), blocks=(64, 128, 256, 512),
             blocks_per_cu=(2, 4, 8, 16), reps=5, cache=DEFAULT_CACHE, use_cache=True):
    """Pick the fastest launch shape for this mix on this GPU; returns (config, timings)."""
    gpu = torch.cuda.get_device_name(0)
    key = "%s|%s" % (gpu, batch.signature())
    db = {}
    if cache and os.path.exists(cache):
        try:
            with open(cache) as f:
                db = json.load(f)
        except (OSError, ValueError):
            db = {}
    if use_cache and key in db:
        c = db[key]
        batch.set_tuning(c["tile"], c["block"], c["blocks_per_cu"])
        return c, []
    results = []
    t0 = time.time()
    for tile in tiles:
        for block in blocks:
            for bpc in blocks_per_cu:
                if block * bpc > 2048:  # more resident threads per CU than the hardware holds
                    continue
                batch.set_tuning(tile, block, bpc)
                us = _time_launch(lambda: batch.query_all(points, dirs, outs), reps)
                results.append({"tile": tile, "block": block, "blocks_per_cu": bpc, "us": round(us, 2)})
    best = min(results, key=lambda r: r["us"])
    batch.set_tuning(best["tile"], best["block"], best["blocks_per_cu"])
    best = dict(best, tuned_s=round(time.time() - t0, 1))
    if cache:
        try:
            os.makedirs(os.path.dirname(cache), exist_ok=True)
            db[key] = best
            with open(cache, "w") as f:
                json.dump(db, f, indent=1)
        except OSError:
            pass
    return best, results
