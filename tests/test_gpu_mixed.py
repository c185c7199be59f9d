"""Mixed precision (FP32 bulk, FP64 near-surface re-evaluation) against the FP64 oracle.

Bar: Inside and Contains identical to the FP64 oracle on every point (decisions
exact); distances and safeties carry FP32 rounding only, reported per shape as
max and p99 absolute error and bounded here by 2e-5 relative to the point's
scale plus 1e-6 mm. Wrong-side results (negative values: DistanceToOut of an
outside point, DistanceToIn of an inside point, negative safeties) carry no
magnitude in the reference's convention and are compared by sign. A few
grazing rays (at most 1 in 5000) may exceed the bound: their FP32 input
rounding is amplified by 1/cos of the incidence angle.
"""
import numpy as np
import pytest

from tests import gpu_util
from tests.test_gpu_shapes import CASES

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind,params", CASES, ids=["%s%d" % (c[0], i) for i, c in enumerate(CASES)])
def test_mixed_decisions_exact_and_errors_bounded(kind, params):
    from vecgeom_amd import MIXED, make_shape
    print("case", kind, params[:12], flush=True)
    o, p, d = gpu_util.inputs_for(kind, params, 1 << 15, seed=99)
    s = make_shape(kind, params)
    P, D = gpu_util.to_dev(p), gpu_util.to_dev(d)
    scale = np.max(np.abs(p), axis=0) + 1.0
    for m in ["Contains", "Inside", "DistanceToIn", "DistanceToOut", "SafetyToIn", "SafetyToOut"]:
        dirs = d if m.startswith("Distance") else None
        ref = o.query(m, p, dirs)
        got = (getattr(s, m)(P, D, precision=MIXED) if dirs is not None else getattr(s, m)(P, precision=MIXED))
        got = got.cpu().numpy()
        if m in ("Contains", "Inside"):
            assert np.array_equal(ref, got), (m, int((ref != got).sum()))
            continue
        wrong_r, wrong_g = ref < 0, got < 0
        assert np.count_nonzero(wrong_r != wrong_g) <= max(2, len(ref) // 5000), (m, "wrong-side flips")
        inf_r, inf_g = ref >= 1e300, got >= 1e300
        flips = int((inf_r != inf_g).sum())
        both = ~inf_r & ~inf_g & ~wrong_r & ~wrong_g
        err = np.abs(ref[both] - got[both])
        bound = 2e-5 * (scale[both] + np.abs(ref[both])) + 1e-6
        print("%-10s %-13s max %.3g p99 %.3g hit/miss flips %d" % (kind, m, err.max(initial=0),
                                                                 np.percentile(err, 99) if err.size else 0, flips))
        assert flips <= max(2, len(ref) // 5000), (m, flips)
        assert np.count_nonzero(err > bound) <= max(2, len(ref) // 5000), (m, float(err.max(initial=0)))


def test_mixed_tessellated_is_fp64():
    """The tessellated solid has no FP32 pass: a MIXED query returns the FP64 results bit for bit."""
    from vecgeom_amd import FP64, MIXED, make_shape
    from vecgeom_amd import workloads
    params = workloads.convex_hull_sphere(20.0)
    o, p, d = gpu_util.inputs_for("tessellated", params, 1 << 13, seed=5)
    s = make_shape("tessellated", params)
    P, D = gpu_util.to_dev(p), gpu_util.to_dev(d)
    for m in ("Contains", "Inside", "DistanceToIn", "DistanceToOut", "SafetyToIn", "SafetyToOut"):
        args = (P, D) if m.startswith("Distance") else (P,)
        a = getattr(s, m)(*args, precision=FP64)
        b = getattr(s, m)(*args, precision=MIXED)
        assert a.equal(b), m
