"""vgmi_nav_propagate_noalloc (the transport loop whose regrouping workspace is a fixed device array
of the library, for callers that cannot allocate) against vgmi_nav_propagate, through the C ABI:
identical positions, states and step counts on two calls in a row, the second with fewer tracks
(both large enough to regroup, so both use the fixed workspace), in both precisions."""
import ctypes

import numpy as np
import pytest
import torch

from vecgeom_amd import FP64, MIXED, SOA3D, geometry
from vecgeom_amd._lib import lib

pytestmark = pytest.mark.gpu

VGMI_ERR_CAPACITY = -4


def _fn():
    f = lib().vgmi_nav_propagate_noalloc
    vp = ctypes.c_void_p
    f.argtypes = [vp, ctypes.c_int] + [vp] * 6 + [ctypes.c_double, vp, vp, vp, ctypes.c_size_t, ctypes.c_int, vp]
    f.restype = ctypes.c_int
    return f


def _state(G, p, d, n):
    P, D = SOA3D.from_array(p), SOA3D.from_array(d)
    nav = G.LocateGlobalPoint(P)
    last = torch.zeros(n, dtype=torch.int32, device="cuda")
    taken = torch.zeros(n, dtype=torch.int32, device="cuda")
    return P, D, nav, last, taken


def test_noalloc_equals_propagate_twice():
    from vecgeom_amd.navigation import Geometry
    g = geometry.small_calorimeter()
    G = Geometry(g)
    f = _fn()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for n in (120000, 70000):  # >= 65536 tracks: regrouped, so the fixed workspace is used
        p, d = geometry.world_points(g, n, seed=n)
        for prec in (FP64, MIXED):
            print("noalloc vs propagate: %d tracks, precision %d" % (n, prec), flush=True)
            P1, D1, nav1, last1, taken1 = _state(G, p, d, n)
            G.Propagate(P1, D1, 10.0, nav1, last1, 30, taken=taken1, precision=prec)
            P2, D2, nav2, last2, taken2 = _state(G, p, d, n)
            rc = f(G._h, prec, *P2.ptrs(), *D2.ptrs(), 10.0, nav2.data_ptr(), last2.data_ptr(), taken2.data_ptr(),
                   n, 30, stream)
            assert rc == 0, lib().vgmi_last_error()
            torch.cuda.synchronize()
            for a, b in ((P1.x, P2.x), (P1.y, P2.y), (P1.z, P2.z), (nav1, nav2), (last1, last2), (taken1, taken2)):
                assert torch.equal(a, b)
            assert int(taken2.sum()) > n  # the tracks moved


def test_noalloc_refuses_more_tracks_than_its_workspace():
    from vecgeom_amd.navigation import Geometry
    G = Geometry(geometry.small_calorimeter())
    f = _fn()
    one = torch.zeros(1, dtype=torch.float64, device="cuda")
    u = torch.zeros(1, dtype=torch.int32, device="cuda")
    ptr = one.data_ptr()
    rc = f(G._h, FP64, ptr, ptr, ptr, ptr, ptr, ptr, 10.0, u.data_ptr(), u.data_ptr(), None, (1 << 21) + 1, 5,
           None)
    assert rc == VGMI_ERR_CAPACITY
    assert b"at most" in lib().vgmi_last_error()
