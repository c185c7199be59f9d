#!/usr/bin/env python3
"""Benchmark of the batched shape-query and BVH-navigation hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Headline (value): shape queries/s of the one-launch batched FP64 shape set — every
implemented shape of the 10-shape mix, per-shape query counts log-uniform in
2^14..2^20, half the points uniform in 1.2x the bounding box and half within
1e-3 mm of the surface, isotropic directions, float64 SOA resident in HBM; a step
is one replay of the captured HIP graph evaluating all six methods
(Contains, Inside, DistanceToIn/Out, SafetyToIn/Out) for every point.

Also reported in the same JSON line:
  box_tube: Box (10,20,30) and full-phi Tube (5,10,15), 2^20 points each, six methods;
  nav:      calorimeter (23k touchables), 2^20 tracks x 50 ComputeStepAndPropagatedState
            steps (10 mm limit), fused transport loop; nav-steps/s; LevelLocate and
            LocateGlobalPoint of the same 2^20 world points (points/s).
  serve:    the continuous-batching server on the calorimeter: 16 producer threads with
            Poisson arrivals (rates swept), 2^U(10,17.6) tracks per request, U{1..100}
            steps per track; p50/p99 latency and nav-steps/s per rate.
Each throughput figure carries a roofline entry (HIP events on the launch stream) and a
CPU baseline (the oracle restatement of the scalar backend, 1 thread, bounded sample).
Algorithmic bytes: 56 B per directed query, 32 B per safety, 28 B per Inside, 25 B per
Contains, 85 B per point of the six-method batch, 88 B per nav-step (p, d, limit and state
in, step and state out), 56 B per LevelLocate (point and volume in; daughter and its local
point out), 28 B per LocateGlobalPoint. The VALU-bound kernels (batched mix, transport loop)
also carry a `valu` entry from the committed rocprofv3 counters of the same launch shape:
issued VALU cycles over the SIMDs' cycles, and FP64 throughput over the vector peak.

The batched launch runs at a fixed launch shape (--tuning, default the autotuner's choice on
one MI355X, the shape the committed profiles were taken with); --autotune picks it afresh.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_VECTOR_PEAK = 78.6e12  # FLOP/s, MI355X
NUM_SIMDS = 256 * 4
VALU_CYCLES_PER_INST = 4     # a wave64 instruction occupies a 16-lane SIMD for four cycles
DEFAULT_TUNING = "256,256,4"  # tile, block, blocks per CU of the batched launch
NAV_BYTES_PER_STEP = 88
LEVEL_LOCATE_BYTES = 56
LOCATE_BYTES = 28
METHODS = ["Contains", "Inside", "DistanceToIn", "DistanceToOut", "SafetyToIn", "SafetyToOut"]
# algorithmic bytes per query (float64 SOA in, result out)
BYTES = {"Contains": 25, "Inside": 28, "DistanceToIn": 56, "DistanceToOut": 56, "SafetyToIn": 32,
         "SafetyToOut": 32}
# all six methods in one pass: p and d read once, six results written
BYTES_ALL = 24 + 24 + 1 + 4 + 4 * 8


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1 << 20, help="points per shape of the box/tube workload")
    ap.add_argument("--tracks", type=int, default=1 << 20)
    ap.add_argument("--nav-steps", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=6.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mix", action="store_true", help="skip the batched mix (the line then has value null)")
    ap.add_argument("--no-nav", action="store_true")
    ap.add_argument("--no-box-tube", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-rates", default="50,100,200",
                    help="Poisson request rates (requests/s, all producers together) of the serving sweep")
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--copy-check", action="store_true",
                    help="only run a plain 1 GiB device copy 5 times (calibrates HBM byte counters) and exit")
    ap.add_argument("--autotune", action="store_true", help="pick the batched launch shape with the autotuner")
    ap.add_argument("--tuning", default=DEFAULT_TUNING,
                    help="tile,block,blocks_per_cu for the batched launch (default %(default)s)")
    return ap.parse_args()


def cpu_info():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count()


def host_cores():
    """Threads for the all-cores CPU baseline: the host's CPUs, at most 16 (a GPU's share of the host)."""
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        return max(1, min(16, os.cpu_count() or 1))


def measured_traffic(kernel, prefix=None):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary
    (profiles/traffic.json, scaled by a plain-copy calibration), or None. With `prefix`, an entry of
    another instantiation of the same kernel (another block size) stands in when `kernel` has none."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            kernels = json.load(f)["kernels"]
        name = kernel if kernel in kernels or prefix is None else next(
            (n for n in kernels if n.startswith(prefix)), kernel)
        k = kernels.get(name)
        if k is None:
            return None
        out = {"bytes": k["traffic_bytes"], "workload": k["workload"], "kernel": name,
               "source": "profiles/pmc_hbm_traffic.md"}
        if "nav_steps" in k:
            out["nav_steps"] = k["nav_steps"]
        return out
    except (OSError, ValueError, KeyError):
        return None


def valu_entry(kernel):
    """VALU issue fraction and FP64 throughput of `kernel` from the committed rocprofv3 counters
    (profiles/counters.json: SQ_INSTS_VALU, SQ_INSTS_VALU_FLOPS_FP64, GRBM_GUI_ACTIVE summed over the
    8 XCDs, and the average dispatch time, per dispatch), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "counters.json")) as f:
            c = json.load(f)["kernels"][kernel]
    except (OSError, ValueError, KeyError):
        return None
    cycles = c["GRBM_GUI_ACTIVE"] / 8.0
    # SQ_INSTS_VALU_FLOPS_FP64 counts per wavefront instruction: x64 lanes is an upper bound
    flops = 64 * c["SQ_INSTS_VALU_FLOPS_FP64"] / (c["dispatch_ns"] * 1e-9)
    return {"issue_frac": round(c["SQ_INSTS_VALU"] * VALU_CYCLES_PER_INST / (NUM_SIMDS * cycles), 4),
            "fp64_tflops_upper": round(flops / 1e12, 2), "fp64_frac_upper": round(flops / FP64_VECTOR_PEAK, 4),
            "kernel": kernel, "workload": c["workload"], "source": c["source"]}


def event_time_us(fn, reps):
    import torch
    s = torch.cuda.current_stream()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(s)
    for _ in range(reps):
        fn()
    b.record(s)
    torch.cuda.synchronize()
    return a.elapsed_time(b) * 1e3 / reps


# ------------------------------------------------------------------ headline: batched mix
def run_mix(args):
    import numpy as np
    import torch

    from vecgeom_amd import SOA3D, make_shape, workloads
    from vecgeom_amd.batch import ShapeBatch, autotune

    shapes = workloads.mix_shapes()
    gpu = [make_shape(k, p) for k, p in shapes]
    counts = workloads.mix_counts(len(shapes))

    def dev_fn(s, m):
        def f(p, d=None):
            P = SOA3D.from_array(p)
            if d is None:
                return getattr(s, m)(P).cpu().numpy()
            return getattr(s, m)(P, SOA3D.from_array(d)).cpu().numpy()
        return f

    exts = [tuple(np.array(v) for v in s.Extent()) for s in gpu]
    p, d = workloads.mix_inputs(exts, counts, [dev_fn(s, "Contains") for s in gpu],
                                [dev_fn(s, "DistanceToOut") for s in gpu])
    P, D = SOA3D.from_array(p), SOA3D.from_array(d)
    batch = ShapeBatch(gpu, counts)
    outs = batch.alloc_outputs()
    if args.autotune:
        tuning, _ = autotune(batch, P, D, outs)
        tuning = dict(tuning, source="autotuner")
    else:
        tile, block, bpc = (int(x) for x in args.tuning.split(","))
        batch.set_tuning(tile, block, bpc)
        tuning = {"tile": tile, "block": block, "blocks_per_cu": bpc, "source": "--tuning"}
    batch.capture(P, D, outs)
    for _ in range(args.warmup):
        batch.replay()
    torch.cuda.synchronize()
    us = event_time_us(batch.replay, args.steps)
    total = batch.total
    queries = total * len(METHODS)
    gbs = BYTES_ALL * total / (us * 1e-6) / 1e9
    res = {"value": round(queries / (us * 1e-6), 1), "ms_per_step": round(us / 1e3, 5),
           "config": {"workload": "batched shape mix, one launch (HIP graph replay), all six methods",
                      "shapes": [[k, c] for (k, _), c in zip(shapes, counts)], "points": total,
                      "queries_per_step": queries, "near_surface_fraction": 0.5,
                      "not_yet_in_mix": [k for k in ("box", "tube", "cone", "polycone", "polyhedron", "trapezoid",
                                                      "sphere", "torus2", "orb", "tessellated")
                                         if k not in {s for s, _ in shapes}],
                      "tuning": tuning},
           "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": measured_traffic("vgmi::k_batch<%d, false>" % batch.info()["block"], "vgmi::k_batch<"),
                        "kernel": "k_batch",
                        "bytes_per_launch": BYTES_ALL * total, "avg_launch_us": round(us, 2)},
           "valu": valu_entry("vgmi::k_batch<%d, false>" % batch.info()["block"])}
    # the same mix and launch shape in mixed precision (FP32 bulk, FP64 near the surface)
    from vecgeom_amd.shapes import MIXED
    mouts = batch.alloc_outputs()
    batch.capture(P, D, mouts, precision=MIXED)
    for _ in range(args.warmup):
        batch.replay()
    torch.cuda.synchronize()
    mus = event_time_us(batch.replay, args.steps)
    res["mixed"] = {"value": round(queries / (mus * 1e-6), 1), "unit": "queries/s", "ms_per_step": round(mus / 1e3, 5),
                    "vs_fp64": round(us / mus, 3),
                    "decisions_equal_fp64": bool(torch.equal(mouts["Contains"], outs["Contains"])
                                                 and torch.equal(mouts["Inside"], outs["Inside"]))}
    if not args.no_cpu:
        from oracle.pyoracle import OracleShape
        orc = [OracleShape(k, prm) for k, prm in shapes]

        def mix_rate(sample, threads, seconds):
            done, used = 0, 0.0
            while used < seconds:
                off = 0
                for o, c in zip(orc, counts):
                    m = min(c, sample)
                    idx = off + (np.arange(m, dtype=np.int64) * c) // m  # evenly spaced over both halves
                    ps = np.ascontiguousarray(p[:, idx])
                    dsm = np.ascontiguousarray(d[:, idx])
                    t0 = time.perf_counter()
                    for meth in METHODS:
                        o.query(meth, ps, dsm if meth.startswith("Distance") else None, threads=threads)
                    used += time.perf_counter() - t0
                    done += m * len(METHODS)
                    off += c
            return done, used

        sample = 1 << 13
        done, used = mix_rate(sample, 1, args.cpu_seconds)
        model, ncpu = cpu_info()
        res["cpu_baseline"] = {"value": round(done / used, 1), "unit": "queries/s", "cores": 1, "kind": "port",
                               "sample": "%d queries: %d points evenly spaced over every shape's range of the same "
                                         "mix (uniform and near-surface halves alike) x 6 methods, repeated for "
                                         "~%.0f s; host %s, %d logical CPUs"
                                         % (done, sample, args.cpu_seconds, model, ncpu)}
        cores = host_cores()
        done, used = mix_rate(1 << 15, cores, args.cpu_seconds / 2)
        res["cpu_baseline"]["all_cores"] = {"value": round(done / used, 1), "cores": cores,
                                            "sample": "%d queries, %d evenly spaced points of every shape, %d threads"
                                                      % (done, 1 << 15, cores)}
    return res


# ------------------------------------------------------------------ box / tube
def run_box_tube(args):
    import numpy as np
    import torch

    from vecgeom_amd import SOA3D, make_shape, workloads

    n = args.n
    shapes = []
    for i, (kind, params) in enumerate(workloads.CONFIG0_SHAPES):
        s = make_shape(kind, params)
        p, d = workloads.shape_stream(s.Extent(), n, workloads.SEED_BASE + i)
        P, D = SOA3D.from_array(p), SOA3D.from_array(d)
        outs = {m: torch.empty(n, dtype=torch.uint8 if m == "Contains" else
                               (torch.int32 if m == "Inside" else torch.float64), device="cuda") for m in METHODS}
        shapes.append((kind, params, s, P, D, outs, p, d))
    per_kernel = []
    total_us = 0.0
    for kind, params, s, P, D, outs, _, _ in shapes:
        for m in METHODS:
            f = (lambda s=s, m=m, P=P, D=D, o=outs[m]: getattr(s, m)(P, D, out=o)) if m.startswith("Distance") \
                else (lambda s=s, m=m, P=P, o=outs[m]: getattr(s, m)(P, out=o))
            for _ in range(args.warmup):
                f()
            us = event_time_us(f, max(10, args.steps))
            total_us += us
            per_kernel.append({"kernel": "%s.%s" % (kind, m), "avg_us": round(us, 3),
                               "GBps": round(BYTES[m] * n / (us * 1e-6) / 1e9, 1)})
    dom = max(per_kernel, key=lambda k: k["avg_us"])
    kind, m = dom["kernel"].split(".")
    queries = n * len(METHODS) * len(shapes)
    res = {"queries_per_s": round(queries / (total_us * 1e-6), 1), "points_per_shape": n,
           "shapes": [[k, p] for k, p, *_ in shapes],
           "roofline": {"bound": "hbm", "achieved": dom["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(dom["GBps"] / HBM_PEAK_GBS, 4),
                        "traffic": measured_traffic("vgmi::k_query_fp64<vgmi::%sData, %d>" % (
                            {"box": "Box", "tube": "Tube"}[kind], METHODS.index(m))),
                        "kernel": "k_query_fp64 (%s)" % dom["kernel"], "bytes_per_launch": BYTES[m] * n},
           "kernels": per_kernel}
    # mixed precision on the same shapes and points: throughput, and the error of every method
    # against the FP64 kernels (which reproduce the oracle bit for bit, tests/test_gpu_shapes.py)
    from vecgeom_amd.shapes import MIXED
    mixed_us, errors = 0.0, {}
    for kind, params, s, P, D, outs, _, _ in shapes:
        errors[kind] = {}
        for m in METHODS:
            om = torch.empty_like(outs[m])
            f = (lambda s=s, m=m, P=P, D=D, o=om: getattr(s, m)(P, D, out=o, precision=MIXED)) \
                if m.startswith("Distance") else (lambda s=s, m=m, P=P, o=om: getattr(s, m)(P, out=o, precision=MIXED))
            for _ in range(args.warmup):
                f()
            mixed_us += event_time_us(f, max(10, args.steps))
            ref = outs[m]
            if m in ("Contains", "Inside"):
                errors[kind][m] = {"agreement": float((om == ref).double().mean())}
                continue
            r, g = ref.double(), om.double()
            fin = (r < 1e300) & (g < 1e300) & (r >= 0) & (g >= 0)
            err = (g[fin] - r[fin]).abs()
            both = ((r >= 1e300) == (g >= 1e300)) & ((r < 0) == (g < 0))
            errors[kind][m] = {"max_abs_mm": float(err.max()) if err.numel() else 0.0,
                               "p99_abs_mm": float(torch.quantile(err[:1 << 20], 0.99)) if err.numel() else 0.0,
                               "inf_and_sign_agreement": float(both.double().mean())}
    res["mixed"] = {"queries_per_s": round(queries / (mixed_us * 1e-6), 1),
                    "over_fp64": round(total_us / mixed_us, 3), "errors_vs_fp64": errors}
    if not args.no_cpu:
        from oracle.pyoracle import OracleShape
        sample = 1 << 17
        done, used = 0, 0.0
        while used < args.cpu_seconds / 2:
            for kind, params, s, P, D, outs, p, d in shapes:
                o = OracleShape(kind, params)
                ps, dsm = np.ascontiguousarray(p[:, :sample]), np.ascontiguousarray(d[:, :sample])
                t0 = time.perf_counter()
                for meth in METHODS:
                    o.query(meth, ps, dsm if meth.startswith("Distance") else None, threads=1)
                used += time.perf_counter() - t0
                done += sample * len(METHODS)
        model, ncpu = cpu_info()
        res["cpu_baseline"] = {"value": round(done / used, 1), "unit": "queries/s", "cores": 1, "kind": "port",
                               "sample": "%d queries (first 2^17 points per shape x 6 methods); host %s, %d CPUs"
                                         % (done, model, ncpu)}
    return res


# ------------------------------------------------------------------ navigation
def run_nav(args):
    import numpy as np
    import torch

    from vecgeom_amd import SOA3D, geometry
    from vecgeom_amd.navigation import Geometry

    g = geometry.calorimeter()
    G = Geometry(g)
    n, nsteps, limit = args.tracks, args.nav_steps, 10.0
    p, d = geometry.world_points(g, n, seed=0x5EED0004)
    P0, D = SOA3D.from_array(p), SOA3D.from_array(d)
    nav0 = G.LocateGlobalPoint(P0)
    torch.cuda.synchronize()
    P = SOA3D.empty(n)
    nav, last = torch.empty_like(nav0), torch.zeros_like(nav0)
    taken = torch.zeros(n, dtype=torch.int32, device="cuda")

    def reset():
        P.x.copy_(P0.x), P.y.copy_(P0.y), P.z.copy_(P0.z)
        nav.copy_(nav0)
        last.zero_()

    from vecgeom_amd.shapes import FP64, MIXED
    s = torch.cuda.current_stream()
    res = {"tracks": n, "nsteps": nsteps, "step_limit_mm": limit}
    kname = {"fp64": "vgmi::nav::k_propagate<295u>", "mixed": "vgmi::nav::k_propagate<2147483943u>"}
    for name, prec in (("fp64", FP64), ("mixed", MIXED)):
        for _ in range(max(1, args.warmup)):
            reset()
            G.Propagate(P, D, limit, nav, last, nsteps, taken, precision=prec)
        torch.cuda.synchronize()
        steps_per_run = int(taken.sum())
        total_us = 0.0
        reps = max(3, args.steps // 4)
        for _ in range(reps):
            reset()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(s)
            G.Propagate(P, D, limit, nav, last, nsteps, taken, precision=prec)
            b.record(s)
            torch.cuda.synchronize()
            total_us += a.elapsed_time(b) * 1e3
        avg_us = total_us / reps
        bytes_run = NAV_BYTES_PER_STEP * steps_per_run
        gbs = bytes_run / (avg_us * 1e-6) / 1e9
        traffic = measured_traffic(kname[name])
        if traffic is not None and traffic.get("nav_steps"):
            # HBM bytes per nav-step of the profiled dispatches over the 88 algorithmic ones
            traffic["over_algorithmic"] = round(traffic["bytes"] / (NAV_BYTES_PER_STEP * traffic["nav_steps"]), 2)
        res[name] = {"nav_steps_per_s": round(steps_per_run / (avg_us * 1e-6), 1), "avg_launch_us": round(avg_us, 2),
                     "steps_per_launch": steps_per_run,
                     "roofline": {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                                  "kernel": "k_propagate (the transport loop: regrouping sorts and %d-step "
                                            "dispatches)" % 10,
                                  "bytes_per_launch": bytes_run,
                                  "note": "88 B per nav-step; the fused loop keeps the track state in registers "
                                          "between steps, so its HBM traffic is geometry and spills"},
                     "valu": valu_entry(kname[name])}
        if name == "mixed":
            res[name]["final_state_equal_fp64"] = bool(torch.equal(nav, nav_fp64) and torch.equal(P.x, px_fp64))
        else:
            nav_fp64, px_fp64 = nav.clone(), P.x.clone()
    res["nav_steps_per_s"] = res["fp64"]["nav_steps_per_s"]
    res["mixed_over_fp64"] = round(res["mixed"]["nav_steps_per_s"] / res["fp64"]["nav_steps_per_s"], 3)
    # one LevelLocate per track in the world volume (the world's placement is the identity, so the
    # global points are the world's local points), and the full descent of LocateGlobalPoint
    world_lv = torch.full((n,), g.world, dtype=torch.int32, device="cuda")
    for name, prec in (("fp64", FP64), ("mixed", MIXED)):
        for label, fn, nbytes in (("level_locate", lambda: G.LevelLocate(world_lv, P0, precision=prec),
                                   LEVEL_LOCATE_BYTES),
                                  ("locate_global_point", lambda: G.LocateGlobalPoint(P0, out=nav, precision=prec),
                                   LOCATE_BYTES)):
            for _ in range(max(1, args.warmup)):
                fn()
            us = event_time_us(fn, max(5, args.steps // 2))
            gbs = nbytes * n / (us * 1e-6) / 1e9
            res[label + "_" + name] = {"points_per_s": round(n / (us * 1e-6), 1), "avg_launch_us": round(us, 2),
                                       "roofline": {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS,
                                                    "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                                                    "bytes_per_launch": nbytes * n}}
    if not args.no_cpu:
        from oracle.pyoracle import OracleGeometry
        O = OracleGeometry(g)
        m = min(n, 1 << 14)

        def nav_rate(m, threads, seconds):
            pn = nav0[:m].cpu().numpy().astype(np.uint32)
            t0 = time.perf_counter()
            done, reps_c = 0, 0
            while time.perf_counter() - t0 < seconds:
                _, _, _, k = O.propagate(p[:, :m], d[:, :m], limit, pn, np.zeros_like(pn), nsteps, threads=threads)
                done += k
                reps_c += 1
            return done, reps_c, time.perf_counter() - t0

        done, reps_c, dt = nav_rate(m, 1, args.cpu_seconds)
        model, ncpu = cpu_info()
        res["cpu_baseline"] = {"value": round(done / dt, 1), "unit": "nav-steps/s", "cores": 1, "kind": "port",
                               "sample": "%d x (first %d tracks x %d steps) = %d nav-steps; host %s, %d CPUs"
                                         % (reps_c, m, nsteps, done, model, ncpu)}
        cores = host_cores()
        ma = min(n, 1 << 16)
        done, reps_c, dt = nav_rate(ma, cores, args.cpu_seconds / 2)
        res["cpu_baseline"]["all_cores"] = {"value": round(done / dt, 1), "cores": cores,
                                            "sample": "%d x (first %d tracks x %d steps), %d threads"
                                                      % (reps_c, ma, nsteps, cores)}
    return res


# ------------------------------------------------------------------ serving loop
def run_serve(args):
    import numpy as np

    from vecgeom_amd import geometry
    from vecgeom_amd.navigation import Geometry
    from vecgeom_amd.serving import Server, poisson_load, summarize

    g = geometry.calorimeter()
    G = Geometry(g)
    pool = geometry.world_points(g, 1 << 20, seed=0x5EED0005)
    from vecgeom_amd.shapes import FP64, MIXED
    out = {"workload": "calorimeter; 16 producer threads, Poisson arrivals; batch 2^U(10,17.6) tracks; "
                       "U{1..100} steps per track; 10 mm step limit; latency = submit to results in host memory",
           "lanes": 4, "lane_tracks": 1 << 18}
    for name, prec in (("fp64", FP64), ("mixed", MIXED)):
        res = {}
        with Server(G, precision=prec, lane_tracks=1 << 18, lanes=4, queue_tracks=1 << 23) as S:
            rng = np.random.default_rng(0)
            alone = {}
            for n in (1024, 16384, 65536):
                lat = []
                for _ in range(5):
                    r = S.wait(S.submit(pool[0][:, :n], pool[1][:, :n], rng.integers(1, 101, size=n, dtype=np.int32)))
                    lat.append(r.latency_us / 1e3)
                alone[str(n)] = round(float(np.median(lat)), 3)
            res["alone_latency_ms"] = alone
            sweep = []
            for rate in [float(x) for x in args.serve_rates.split(",") if x]:
                rec, refused, wall = poisson_load(S, pool, rate_hz=rate, duration_s=args.serve_seconds, producers=16,
                                                  seed=int(rate))
                sm = summarize(rec, refused, wall)
                sm["rate_hz"] = rate
                if sm.get("nav_steps_per_s"):
                    # at the offered load, not at saturation: the rate is set by the arrivals
                    gbs = NAV_BYTES_PER_STEP * sm["nav_steps_per_s"] / 1e9
                    sm["roofline"] = {"bound": "hbm", "achieved": round(gbs, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                      "frac": round(gbs / HBM_PEAK_GBS, 5), "basis": "88 B per nav-step at the "
                                      "offered load"}
                sweep.append(sm)
            res["sweep"] = sweep
        out[name] = res
    return out


def run_copy_check():
    import torch
    n = 1 << 30
    a = torch.empty(n, dtype=torch.uint8, device="cuda").fill_(1)
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_check": {"bytes_read": n, "bytes_written": n, "launches": 5}}), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only; --gpus %d is not supported "
              "(shape queries and navigation tracks shard trivially across GPUs)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.copy_check:
        run_copy_check()
        return
    mix = {"value": None, "ms_per_step": None, "config": {"workload": "batched shape mix skipped (--no-mix)"},
           "roofline": None}
    if not args.no_mix:
        print("bench: batched shape mix", file=sys.stderr, flush=True)
        mix = run_mix(args)
    extra = {}
    if not args.no_box_tube:
        print("bench: box/tube", file=sys.stderr, flush=True)
        extra["box_tube"] = run_box_tube(args)
    if not args.no_nav:
        print("bench: calorimeter navigation", file=sys.stderr, flush=True)
        extra["nav"] = run_nav(args)
    if not args.no_serve:
        print("bench: serving loop", file=sys.stderr, flush=True)
        extra["serve"] = run_serve(args)
    line = {"metric": "shape queries/s (FP64 GPU, batched one-launch mix)", "value": mix["value"],
            "unit": "queries/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": mix["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64", "data": "synthetic", "config": mix["config"], "roofline": mix["roofline"],
            "cpu_baseline": mix.get("cpu_baseline"), "valu": mix.get("valu"), "mixed": mix.get("mixed")}
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
