"""Closed geometry on the device and the BVH navigator interface
(VNavigator::ComputeStep / ComputeStepAndPropagatedState, VSafetyEstimator::ComputeSafety,
GlobalLocator::LocateGlobalPoint; VecGeom/navigation/VNavigator.h:51-68,
VecGeom/navigation/VSafetyEstimator.h:45,98, VecGeom/navigation/GlobalLocator.h:35-82).

States are NavStateIndex values held in uint32 device tensors (0 = outside the world).
"""
import ctypes

import numpy as np
import torch

from ._lib import check, lib
from .shapes import FP64, _stream
from .soa import SOA3D, device_array, same_length

_U32 = (torch.int32, torch.uint32) if hasattr(torch, "uint32") else (torch.int32,)


def _limit(step_limit, n, dev):
    if not torch.is_tensor(step_limit):
        return torch.full((n,), float(step_limit), dtype=torch.float64, device=dev)
    return device_array(step_limit, torch.float64, n, "step_limit")

HOST_ONLY = 1


def _declare():
    L = lib()
    if getattr(L, "_nav_declared", False):
        return L
    vp, i64, ci, sz = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_size_t
    sigs = {
        "vgmi_geom_build": [vp, ci, ctypes.POINTER(vp)],
        "vgmi_geom_destroy": [vp],
        "vgmi_geom_info": [vp, vp],
        "vgmi_geom_navindex": [vp, vp, i64, vp],
        "vgmi_geom_bvh": [vp, ci] + [vp] * 6,
        "vgmi_nav_locate": [vp, ci] + [vp] * 4 + [sz, vp],
        "vgmi_nav_step": [vp, ci] + [vp] * 13 + [sz, ci, vp],
        "vgmi_nav_safety": [vp, ci] + [vp] * 5 + [sz, vp],
        "vgmi_nav_step_safety": [vp, ci] + [vp] * 15 + [sz, vp],
        "vgmi_nav_propagate": [vp, ci] + [vp] * 6 + [ctypes.c_double] + [vp] * 3 + [sz, ci, vp],
        "vgmi_nav_level_locate": [vp, ci] + [vp] * 9 + [sz, vp],
        "vgmi_placed_query": [vp, ci, ci, ci] + [vp] * 8 + [sz, vp],
        "vgmi_placed_normal": [vp, ci] + [vp] * 7 + [sz, vp],
        "vgmi_nav_propagate_noalloc": [vp, ci] + [vp] * 6 + [ctypes.c_double] + [vp] * 3 + [sz, ci, vp],
        "vgmi_nav_to_path": [vp, vp, sz, ci, vp, vp, vp, vp],
        "vgmi_nav_from_path": [vp, ci, vp, vp, sz, ci, vp, vp],
        "vgmi_nav_distance": [vp, vp, vp, sz, vp, vp],
        "vgmi_nav_global_to_local": [vp, vp, ci] + [vp] * 6 + [sz, vp],
        "vgmi_geom_set_regroup": [vp, ci],
        "vgmi_geom_set_lanes_per_wave": [vp, ci],
    }
    for name, args in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    L._nav_declared = True
    return L


class _Desc(ctypes.Structure):
    _fields_ = [("nshapes", ctypes.c_int), ("shape_type", ctypes.c_void_p), ("shape_param_offset", ctypes.c_void_p),
                ("shape_params", ctypes.c_void_p), ("nlogical", ctypes.c_int), ("logical_shape", ctypes.c_void_p),
                ("nplaced", ctypes.c_int), ("placed_mother", ctypes.c_void_p), ("placed_logical", ctypes.c_void_p),
                ("placed_transform", ctypes.c_void_p), ("world_logical", ctypes.c_int)]


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Geometry:
    """A closed geometry: host tables plus their device copy (unless host_only)."""

    def __init__(self, desc, host_only=False):
        L = _declare()
        self.desc = desc
        arrs = desc.arrays()
        stype, soff, sparams, lvshape, plm, pll, plt = arrs
        d = _Desc(len(stype), _p(stype), _p(soff), _p(sparams), len(lvshape), _p(lvshape), len(plm), _p(plm),
                  _p(pll), _p(plt), desc.world)
        h = ctypes.c_void_p()
        check(L.vgmi_geom_build(ctypes.byref(d), HOST_ONLY if host_only else 0, ctypes.byref(h)), "vgmi_geom_build")
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().vgmi_geom_destroy(h)
            except Exception:
                pass
            self._h = None

    def info(self):
        c = np.zeros(6, np.int64)
        check(lib().vgmi_geom_info(self._h, _p(c)), "vgmi_geom_info")
        return {"placed": int(c[0]), "logical": int(c[1]), "navindex_words": int(c[2]), "max_depth": int(c[3]),
                "world_placed": int(c[4]), "device_bytes": int(c[5])}

    def navindex(self):
        n = ctypes.c_int64()
        check(lib().vgmi_geom_navindex(self._h, None, 0, ctypes.byref(n)), "vgmi_geom_navindex")
        out = np.zeros(n.value, np.uint32)
        check(lib().vgmi_geom_navindex(self._h, _p(out), n.value, None), "vgmi_geom_navindex")
        return out

    def bvh(self, lv):
        counts = np.zeros(3, np.int32)
        if lib().vgmi_geom_bvh(self._h, lv, _p(counts), None, None, None, None, None) != 0:
            return None
        depth, nn, nd = (int(x) for x in counts)
        nchild, offset, prim = np.zeros(nn, np.int32), np.zeros(nn, np.int32), np.zeros(nd, np.int32)
        nodes, aabbs = np.zeros(6 * nn), np.zeros(6 * nd)
        check(lib().vgmi_geom_bvh(self._h, lv, _p(counts), _p(nchild), _p(offset), _p(prim), _p(nodes), _p(aabbs)),
              "vgmi_geom_bvh")
        return {"depth": depth, "nchild": nchild, "offset": offset, "primid": prim, "nodes": nodes.reshape(-1, 6),
                "aabbs": aabbs.reshape(-1, 6)}

    # ---- device navigation
    def set_regroup(self, chunk_steps):
        """Regroup tracks by volume every chunk_steps steps in the fused transport loops (0: never)."""
        check(lib().vgmi_geom_set_regroup(self._h, int(chunk_steps)), "vgmi_geom_set_regroup")

    def set_lanes_per_wave(self, lanes):
        """Tracks per wavefront of the fused loops (power of two up to 64; 0 = automatic)."""
        check(lib().vgmi_geom_set_lanes_per_wave(self._h, int(lanes)), "vgmi_geom_set_lanes_per_wave")

    def LocateGlobalPoint(self, points, out=None, precision=FP64):
        n = len(points)
        if out is None:
            out = torch.empty(n, dtype=torch.int32, device=points.x.device)
        device_array(out, _U32, n, "out")
        check(lib().vgmi_nav_locate(self._h, precision, *points.ptrs(), out.data_ptr(), n, _stream()),
              "vgmi_nav_locate")
        return out

    # ---- NavStateIndex <-> path interop (NavStatePath-style states: one placed volume per level)
    PATH_INDICES, PATH_PLACED = 0, 1

    def NavToPath(self, nav, max_levels=None):
        """Per state: (levels = GetCurrentLevel, placed volumes per level, daughter indices per level), the
        last two as (max_levels, n) int32 tensors with -1 beyond each state's levels
        (NavStateIndex::ValueAt, GetPathAsListOfIndices)."""
        n = len(nav)
        dev = nav.device
        device_array(nav, _U32, n, "nav")
        L = int(max_levels if max_levels is not None else self.info()["max_depth"])
        levels = torch.empty(n, dtype=torch.int32, device=dev)
        pvols = torch.empty((L, n), dtype=torch.int32, device=dev)
        idx = torch.empty((L, n), dtype=torch.int32, device=dev)
        check(lib().vgmi_nav_to_path(self._h, nav.data_ptr(), n, L, levels.data_ptr(), pvols.data_ptr(),
                                     idx.data_ptr(), _stream()), "vgmi_nav_to_path")
        return levels, pvols, idx

    def PathToNav(self, levels, path, kind=PATH_PLACED):
        """NavIndex per path: path is (max_levels, n) int32, placed volumes world first (kind PATH_PLACED) or
        daughter indices (PATH_INDICES, NavStateIndex::ResetPathFromListOfIndices). Paths the geometry
        does not hold give 0."""
        n = len(levels)
        device_array(levels, torch.int32, n, "levels")
        if not (torch.is_tensor(path) and path.dim() == 2 and path.shape[1] == n and path.is_contiguous()
                and path.dtype == torch.int32 and path.device == levels.device):
            raise ValueError("path must be a contiguous (max_levels, n) int32 tensor on the levels' device")
        out = torch.empty(n, dtype=torch.int32, device=levels.device)
        check(lib().vgmi_nav_from_path(self._h, int(kind), levels.data_ptr(), path.data_ptr(), n, path.shape[0],
                                       out.data_ptr(), _stream()), "vgmi_nav_from_path")
        return out

    def Distance(self, nav_a, nav_b):
        """NavStateIndex::Distance per pair (moves up to the deepest shared level, then down)."""
        n = len(nav_a)
        device_array(nav_a, _U32, n, "nav_a")
        device_array(nav_b, _U32, n, "nav_b")
        out = torch.empty(n, dtype=torch.int32, device=nav_a.device)
        check(lib().vgmi_nav_distance(self._h, nav_a.data_ptr(), nav_b.data_ptr(), n, out.data_ptr(), _stream()),
              "vgmi_nav_distance")
        return out

    def GlobalToLocal(self, nav, points, tolevel=-1):
        """NavStateIndex::GlobalToLocal(point[, tolevel]): points in the frame of each state's ancestor at
        tolevel (-1: the top volume)."""
        n = len(points)
        device_array(nav, _U32, n, "nav")
        out = SOA3D.empty(n, device=points.x.device)
        check(lib().vgmi_nav_global_to_local(self._h, nav.data_ptr(), int(tolevel), *points.ptrs(), *out.ptrs(), n,
                                             _stream()), "vgmi_nav_global_to_local")
        return out

    def LevelLocate(self, logical, points, excl=None, precision=FP64):
        """VLevelLocator::LevelLocate / LevelLocateExclVol (VecGeom/navigation/BVHLevelLocator.h:44-77) per
        point: logical (int32 tensor or int) is the logical volume the point (in its frame) is in, excl
        (int32 tensor, int or None) the placed volume to skip. Returns (daughter placed ids, -1 for none;
        the points in the daughters' frames as an SOA3D)."""
        n = len(points)
        dev = points.x.device
        if not torch.is_tensor(logical):
            logical = torch.full((n,), int(logical), dtype=torch.int32, device=dev)
        device_array(logical, torch.int32, n, "logical")
        if excl is not None and not torch.is_tensor(excl):
            excl = torch.full((n,), int(excl), dtype=torch.int32, device=dev)
        if excl is not None:
            device_array(excl, torch.int32, n, "excl")
        out = torch.empty(n, dtype=torch.int32, device=dev)
        dl = SOA3D.empty(n, device=dev)
        check(lib().vgmi_nav_level_locate(self._h, precision, logical.data_ptr(),
                                          excl.data_ptr() if excl is not None else None, *points.ptrs(),
                                          out.data_ptr(), *dl.ptrs(), n, _stream()), "vgmi_nav_level_locate")
        return out, dl

    def PlacedQuery(self, placed_id, op, points, dirs=None, step_max=None, precision=FP64):
        """VPlacedVolume SOA query of placement placed_id (the world: len(placements)); op is a method name
        of OPS or "PlacedDistanceToOut". Contains, Inside, DistanceToIn, SafetyToIn and PlacedDistanceToOut
        take points in the frame the volume is placed in; DistanceToOut and SafetyToOut in its own frame."""
        from .shapes import OPS
        code = 6 if op == "PlacedDistanceToOut" else OPS[op]
        n = len(points)
        same_length(points, dirs)
        if code in (2, 3, 6) and dirs is None:
            raise ValueError("%s needs directions" % op)
        out = torch.empty(n, dtype={0: torch.uint8, 1: torch.int32}.get(code, torch.float64), device=points.x.device)
        if step_max is not None:
            device_array(step_max, torch.float64, n, "step_max")
        d = dirs.ptrs() if dirs is not None else (None, None, None)
        check(lib().vgmi_placed_query(self._h, int(placed_id), code, precision, *points.ptrs(), *d,
                                      step_max.data_ptr() if step_max is not None else None, out.data_ptr(), n,
                                      _stream()), "vgmi_placed_query(%s)" % op)
        return out

    def PlacedNormal(self, placed_id, points):
        """VPlacedVolume::Normal of placement placed_id at points of the frame it is placed in: returns
        (normals in that frame as an SOA3D, valid uint8 tensor)."""
        n, dev = len(points), points.x.device
        out = SOA3D.empty(n, device=dev)
        valid = torch.empty(n, dtype=torch.uint8, device=dev)
        check(lib().vgmi_placed_normal(self._h, int(placed_id), *points.ptrs(), *out.ptrs(), valid.data_ptr(), n,
                                       _stream()), "vgmi_placed_normal")
        return out, valid

    def ComputeStep(self, points, dirs, step_limit, nav_in, last_in=None, relocate=True, precision=FP64):
        """Returns (step, nav_out, last_out, boundary) device tensors."""
        n = len(points)
        dev = points.x.device
        same_length(points, dirs)
        step_limit = _limit(step_limit, n, dev)
        device_array(nav_in, _U32, n, "nav_in")
        if last_in is not None:
            device_array(last_in, _U32, n, "last_in")
        step = torch.empty(n, dtype=torch.float64, device=dev)
        nav_out = torch.empty(n, dtype=torch.int32, device=dev)
        last_out = torch.empty(n, dtype=torch.int32, device=dev)
        bnd = torch.empty(n, dtype=torch.uint8, device=dev)
        check(lib().vgmi_nav_step(self._h, precision, *points.ptrs(), *dirs.ptrs(), step_limit.data_ptr(),
                                  nav_in.data_ptr(), last_in.data_ptr() if last_in is not None else None,
                                  nav_out.data_ptr(), last_out.data_ptr(), bnd.data_ptr(), step.data_ptr(), n,
                                  1 if relocate else 0, _stream()), "vgmi_nav_step")
        return step, nav_out, last_out, bnd

    def ComputeStepAndSafety(self, points, dirs, step_limit, nav_in, last_in=None, calc_safety=None,
                             precision=FP64):
        """ComputeStepAndSafetyAndPropagatedState for every track: returns (step, nav_out, last_out,
        boundary, safety); safety is that of the start point where calc_safety (default: all)."""
        n = len(points)
        dev = points.x.device
        same_length(points, dirs)
        step_limit = _limit(step_limit, n, dev)
        device_array(nav_in, _U32, n, "nav_in")
        if last_in is not None:
            device_array(last_in, _U32, n, "last_in")
        step = torch.empty(n, dtype=torch.float64, device=dev)
        safety = torch.empty(n, dtype=torch.float64, device=dev)
        nav_out = torch.empty(n, dtype=torch.int32, device=dev)
        last_out = torch.empty(n, dtype=torch.int32, device=dev)
        bnd = torch.empty(n, dtype=torch.uint8, device=dev)
        calc = None
        if calc_safety is not None:
            device_array(calc_safety, (torch.uint8, torch.bool), n, "calc_safety")
            calc = calc_safety.to(torch.uint8)
        check(lib().vgmi_nav_step_safety(self._h, precision, *points.ptrs(), *dirs.ptrs(), step_limit.data_ptr(),
                                         nav_in.data_ptr(), last_in.data_ptr() if last_in is not None else None,
                                         calc.data_ptr() if calc is not None else None, nav_out.data_ptr(),
                                         last_out.data_ptr(), bnd.data_ptr(), step.data_ptr(), safety.data_ptr(), n,
                                         _stream()), "vgmi_nav_step_safety")
        return step, nav_out, last_out, bnd, safety

    def ComputeSafety(self, points, nav, precision=FP64):
        n = len(points)
        device_array(nav, _U32, n, "nav")
        out = torch.empty(n, dtype=torch.float64, device=points.x.device)
        check(lib().vgmi_nav_safety(self._h, precision, *points.ptrs(), nav.data_ptr(), out.data_ptr(), n,
                                    _stream()), "vgmi_nav_safety")
        return out

    def Propagate(self, points, dirs, step_limit, nav, last, nsteps, taken=None, precision=FP64):
        """In-place transport loop on the device (points, nav, last are updated)."""
        n = len(points)
        same_length(points, dirs)
        device_array(nav, _U32, n, "nav")
        device_array(last, _U32, n, "last")
        if taken is not None:
            device_array(taken, torch.int32, n, "taken")
        check(lib().vgmi_nav_propagate(self._h, precision, *points.ptrs(), *dirs.ptrs(), float(step_limit),
                                       nav.data_ptr(), last.data_ptr(),
                                       taken.data_ptr() if taken is not None else None, n, int(nsteps), _stream()),
              "vgmi_nav_propagate")
        return taken
