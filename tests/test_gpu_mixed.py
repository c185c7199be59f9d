"""Mixed precision (FP32 bulk, FP64 near-surface and undecided re-evaluation) against the FP64 oracle.

Bar: Inside and Contains identical to the FP64 oracle on every point; every boundary decision of
a directed query identical too: the sign of every result (wrong-side results are negative and
carry no magnitude in the reference's convention, so they are compared by sign) and every hit or
miss. Distances of the shapes with certified FP32 kernels (box, tube, trapezoid, orb) are within
kMixDistErr = 2.5e-4 mm of the FP64 ones; the other shapes' directed queries are the FP64 ones
(within the FP64 parity bar of 1e-9 mm).
Safeties carry FP32 rounding, bounded by 2e-5 relative to the point's scale plus 1e-6 mm. The
inputs are the parity workload (half uniform, half within 1e-3 mm of the surface) and rays that
skim the surface far from their start point, where a hit or miss depends on rounding. Per-shape
max and p99 errors are printed.
"""
import numpy as np
import pytest

from tests import gpu_util
from tests.test_gpu_shapes import CASES

pytestmark = pytest.mark.gpu

CERTIFIED = {"box", "tube", "trapezoid", "orb"}
DIST_ERR = 2.5e-4


def _check(kind, params, o, p, d, label):
    from vecgeom_amd import MIXED, make_shape
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
        wrong = int(((ref < 0) != (got < 0)).sum())
        inf_r, inf_g = ref >= 1e300, got >= 1e300
        flips = int((inf_r != inf_g).sum())
        both = ~inf_r & ~inf_g & (ref >= 0) & (got >= 0)
        err = np.abs(ref[both] - got[both])
        print("%-10s %-6s %-13s max %.3g p99 %.3g wrong-side flips %d hit/miss flips %d" % (
            kind, label, m, err.max(initial=0), np.percentile(err, 99) if err.size else 0, wrong, flips))
        assert wrong == 0 and flips == 0, (m, wrong, flips)
        if m.startswith("Distance"):
            bound = DIST_ERR if kind in CERTIFIED else 1e-9  # the others run the FP64 kernels
            assert err.max(initial=0) <= bound, (m, float(err.max(initial=0)))
        else:
            bound = 2e-5 * (scale[both] + np.abs(ref[both])) + 1e-6
            assert np.all(err <= bound), (m, float(err.max(initial=0)))


@pytest.mark.parametrize("kind,params", CASES, ids=["%s%d" % (c[0], i) for i, c in enumerate(CASES)])
def test_mixed_decisions_exact_and_errors_bounded(kind, params):
    print("case", kind, params[:12], flush=True)
    o, p, d = gpu_util.inputs_for(kind, params, 1 << 15, seed=99)
    _check(kind, params, o, p, d, "mix")


@pytest.mark.parametrize("kind,params", [c for c in CASES if c[0] != "tessellated"],
                         ids=["%s%d" % (c[0], i) for i, c in enumerate(CASES) if c[0] != "tessellated"])
def test_mixed_grazing_rays_exact(kind, params):
    print("case", kind, params[:12], flush=True)
    o, p, d = gpu_util.grazing_rays(kind, params, 1 << 14, seed=7)
    _check(kind, params, o, p, d, "graze")


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
