"""One-launch batched queries: results identical to per-shape launches and to the
oracle, under several launch shapes, and through HIP graph replay."""
import numpy as np
import pytest
import torch

from oracle.pyoracle import OracleShape
from vecgeom_amd import SOA3D, make_shape, workloads
from vecgeom_amd.batch import METHODS, ShapeBatch
from tests import gpu_util
from tests.test_gpu_shapes import TOL_BY_KIND

pytestmark = pytest.mark.gpu


# uneven per-shape counts (one per shape of the mix), including a single-point shape
COUNTS = [3000, 5000, 777, 4096, 1, 2500, 1500, 333, 2048, 640, 100, 4000]


def _mix(counts):
    shapes = workloads.mix_shapes()
    gpu = [make_shape(k, p) for k, p in shapes]
    orc = [OracleShape(k, p) for k, p in shapes]
    exts = [o.extent() for o in orc]
    p, d = workloads.mix_inputs(exts, counts, [lambda x, o=o: o.query("Contains", x) for o in orc],
                                [lambda x, y, o=o: o.query("DistanceToOut", x, y) for o in orc])
    return shapes, gpu, orc, p, d


def test_batch_matches_oracle_and_single_launches():
    counts = COUNTS[:len(workloads.mix_shapes())]
    shapes, gpu, orc, p, d = _mix(counts)
    P, D = SOA3D.from_array(p), SOA3D.from_array(d)
    b = ShapeBatch(gpu, counts)
    for tuning in [(64, 64, 1), (4096, 256, 8), (1000, 512, 2)]:
        print("tuning", tuning, flush=True)
        b.set_tuning(*tuning)
        out = b.query_all(P, D)
        off = 0
        for (kind, _), s, o, n in zip(shapes, gpu, orc, counts):
            sl = slice(off, off + n)
            ps, dsl = np.ascontiguousarray(p[:, sl]), np.ascontiguousarray(d[:, sl])
            for m in METHODS:
                ref = o.query(m, ps, dsl if m.startswith("Distance") else None)
                got = out[m][off:off + n].cpu().numpy()
                if kind in TOL_BY_KIND and m not in ("Contains", "Inside"):
                    bad, _, maxerr = gpu_util.compare_float(ref, got, TOL_BY_KIND[kind])
                    assert bad == 0, (kind, m, tuning, maxerr)
                else:
                    assert np.array_equal(ref, got), (kind, m, tuning)
            off += n


def test_batch_graph_replay():
    counts = COUNTS[::-1][:len(workloads.mix_shapes())]
    shapes, gpu, orc, p, d = _mix(counts)
    P, D = SOA3D.from_array(p), SOA3D.from_array(d)
    b = ShapeBatch(gpu, counts)
    eager = b.query_all(P, D)
    torch.cuda.synchronize()
    outs = b.alloc_outputs()
    b.capture(P, D, outs)
    for rep in range(3):
        for t in outs.values():
            t.zero_()
        torch.cuda.synchronize()
        b.replay()
        torch.cuda.synchronize()
        for m in METHODS:
            assert torch.equal(outs[m], eager[m]), (rep, m)


def test_batch_tessellated_lds_and_pool_walks():
    """Tessellated solids whose nodes and planes fit the launch's LDS (the 1024- and 12-facet
    solids) next to one that does not (2304 facets, walked in the pool), over tile sizes that
    make one workgroup take several consecutive tiles of one solid; every method equals the
    oracle bit for bit, and a partial output set equals the full one."""
    from tests.golden.extract_kats import trd_like_triangles
    shapes = [("tessellated", workloads.convex_hull_sphere(20.0)),
              ("box", [10.0, 20.0, 30.0]),
              ("tessellated", workloads.convex_hull_sphere(20.0, nphi=48, nbands=25)),
              ("tessellated", trd_like_triangles(10, 30, 20, 40, 40))]
    counts = [1500, 200, 700, 900]
    gpu = [make_shape(k, p) for k, p in shapes]
    orc = [OracleShape(k, p) for k, p in shapes]
    exts = [o.extent() for o in orc]
    p, d = workloads.mix_inputs(exts, counts, [lambda x, o=o: o.query("Contains", x) for o in orc],
                                [lambda x, y, o=o: o.query("DistanceToOut", x, y) for o in orc])
    P, D = SOA3D.from_array(p), SOA3D.from_array(d)
    refs, off = [], 0
    for o, n in zip(orc, counts):
        ps, dsl = np.ascontiguousarray(p[:, off:off + n]), np.ascontiguousarray(d[:, off:off + n])
        refs.append({m: o.query(m, ps, dsl if m.startswith("Distance") else None) for m in METHODS})
        off += n
    b = ShapeBatch(gpu, counts)
    for tuning in [(64, 64, 1), (128, 128, 2), (256, 256, 4)]:
        print("tuning", tuning, flush=True)
        b.set_tuning(*tuning)
        out = b.query_all(P, D)
        off = 0
        for (kind, _), ref, n in zip(shapes, refs, counts):
            for m in METHODS:
                assert np.array_equal(ref[m], out[m][off:off + n].cpu().numpy()), (kind, m, tuning)
            off += n
        part = b.alloc_outputs()
        for m in ("Contains", "SafetyToOut", "DistanceToIn"):
            part[m] = None
        b.query_all(P, D, part)
        for m in METHODS:
            if part[m] is not None:
                assert torch.equal(part[m], out[m]), (m, tuning)


def test_batch_tessellated_deep_points():
    """Points deep inside the 1024-facet hull (their safety searches are queued for whole
    wavefronts, and with 512-point tiles the queue overflows so lanes also search themselves)
    and far outside it: every method equals the oracle bit for bit."""
    params = workloads.convex_hull_sphere(20.0)
    s = make_shape("tessellated", params)
    o = OracleShape("tessellated", params)
    g = np.random.default_rng(11)
    n_in, n_out = 3000, 1000
    r = 6.0 * g.random(n_in) ** (1.0 / 3.0)
    u = workloads.isotropic_dirs(g, n_in)
    far = workloads.isotropic_dirs(g, n_out) * (25.0 + 30.0 * g.random(n_out))
    p = np.ascontiguousarray(np.concatenate([u * r, far], axis=1))
    d = np.ascontiguousarray(workloads.isotropic_dirs(g, n_in + n_out))
    P, D = SOA3D.from_array(p), SOA3D.from_array(d)
    ref = {m: o.query(m, p, d if m.startswith("Distance") else None) for m in METHODS}
    b = ShapeBatch([s], [n_in + n_out])
    for tuning in [(512, 512, 1), (256, 256, 4), (64, 64, 2)]:
        b.set_tuning(*tuning)
        out = b.query_all(P, D)
        for m in METHODS:
            assert np.array_equal(ref[m], out[m].cpu().numpy()), (m, tuning)


def l_prism_triangles(a=10.0, h=5.0):
    """Tessellated L-shaped prism (non-convex): the polygon (0,0) (2a,0) (2a,a) (a,a) (a,2a) (0,2a),
    star-shaped from (0,0) so its faces are fans from there, extruded over z in [-h, h]; vertices
    anticlockwise seen from outside."""
    poly = [(0.0, 0.0), (2 * a, 0.0), (2 * a, a), (a, a), (a, 2 * a), (0.0, 2 * a)]
    tris = []
    for i in range(1, len(poly) - 1):
        (x0, y0), (x1, y1), (x2, y2) = poly[0], poly[i], poly[i + 1]
        tris.append([(x0, y0, h), (x1, y1, h), (x2, y2, h)])
        tris.append([(x0, y0, -h), (x2, y2, -h), (x1, y1, -h)])
    for i in range(len(poly)):
        (ax, ay), (bx, by) = poly[i], poly[(i + 1) % len(poly)]
        tris.append([(ax, ay, -h), (bx, by, -h), (bx, by, h)])
        tris.append([(ax, ay, -h), (bx, by, h), (ax, ay, h)])
    return [float(len(tris))] + [float(c) for t in tris for v in t for c in v]


def test_batch_tessellated_non_convex():
    """A non-convex tessellated solid (L-shaped prism, walked in LDS, deep points queued for
    wavefronts) next to the convex hull: every method equals the oracle bit for bit."""
    shapes = [("tessellated", l_prism_triangles()), ("tessellated", workloads.convex_hull_sphere(20.0))]
    counts = [6000, 2000]
    gpu = [make_shape(k, p) for k, p in shapes]
    orc = [OracleShape(k, p) for k, p in shapes]
    exts = [o.extent() for o in orc]
    p, d = workloads.mix_inputs(exts, counts, [lambda x, o=o: o.query("Contains", x) for o in orc],
                                [lambda x, y, o=o: o.query("DistanceToOut", x, y) for o in orc])
    P, D = SOA3D.from_array(p), SOA3D.from_array(d)
    refs, off = [], 0
    for o, n in zip(orc, counts):
        ps, dsl = np.ascontiguousarray(p[:, off:off + n]), np.ascontiguousarray(d[:, off:off + n])
        refs.append({m: o.query(m, ps, dsl if m.startswith("Distance") else None) for m in METHODS})
        off += n
    assert refs[0]["Contains"].any() and not refs[0]["Contains"].all()
    b = ShapeBatch(gpu, counts)
    for tuning in [(512, 512, 2), (256, 256, 4), (64, 64, 1)]:
        b.set_tuning(*tuning)
        out = b.query_all(P, D)
        off = 0
        for ref, n in zip(refs, counts):
            for m in METHODS:
                assert np.array_equal(ref[m], out[m][off:off + n].cpu().numpy()), (m, tuning, off)
            off += n


def test_batch_mixed_matches_single_shape_mixed():
    """precision=MIXED: every point's six answers equal the single-shape mixed launches bit for bit
    (one shape of every type, uneven counts, several launch shapes, a partial output set and a
    graph replay); decisions equal the FP64 batch's."""
    from vecgeom_amd import MIXED
    from tests.test_gpu_shapes import CASES
    seen, shapes = set(), []
    for kind, params in CASES:
        if kind not in seen:
            seen.add(kind)
            shapes.append((kind, params))
    counts = [(1237 * (k + 3)) % 4000 + 1 for k in range(len(shapes))]
    gpu = [make_shape(k, p) for k, p in shapes]
    orc = [OracleShape(k, p) for k, p in shapes]
    exts = [o.extent() for o in orc]
    p, d = workloads.mix_inputs(exts, counts, [lambda x, o=o: o.query("Contains", x) for o in orc],
                                [lambda x, y, o=o: o.query("DistanceToOut", x, y) for o in orc])
    P, D = SOA3D.from_array(p), SOA3D.from_array(d)
    refs, off = [], 0
    for s, n in zip(gpu, counts):
        Ps = SOA3D.from_array(np.ascontiguousarray(p[:, off:off + n]))
        Ds = SOA3D.from_array(np.ascontiguousarray(d[:, off:off + n]))
        refs.append({m: (getattr(s, m)(Ps, Ds, precision=MIXED) if m.startswith("Distance")
                         else getattr(s, m)(Ps, precision=MIXED)) for m in METHODS})
        off += n
    b = ShapeBatch(gpu, counts)
    fp64 = b.query_all(P, D)
    for tuning in [(64, 64, 1), (256, 256, 4), (1000, 512, 2)]:
        print("tuning", tuning, flush=True)
        b.set_tuning(*tuning)
        out = b.query_all(P, D, precision=MIXED)
        off = 0
        for (kind, _), ref, n in zip(shapes, refs, counts):
            for m in METHODS:
                assert torch.equal(ref[m], out[m][off:off + n]), (kind, m, tuning)
            off += n
        for m in ("Contains", "Inside"):
            assert torch.equal(out[m], fp64[m]), (m, tuning)
    part = b.alloc_outputs()
    for m in ("Inside", "SafetyToIn", "DistanceToOut"):
        part[m] = None
    b.query_all(P, D, part, precision=MIXED)
    for m in METHODS:
        if part[m] is not None:
            assert torch.equal(part[m], out[m]), m
    outs = b.alloc_outputs()
    b.capture(P, D, outs, precision=MIXED)
    for rep in range(2):
        for t in outs.values():
            t.zero_()
        torch.cuda.synchronize()
        b.replay()
        torch.cuda.synchronize()
        for m in METHODS:
            assert torch.equal(outs[m], out[m]), (rep, m)
