"""ctypes wrapper around oracle/liboracle.so (the CPU scalar FP64 restatement).

Test infrastructure only: imported by tests/, by __graft_entry__.smoke() and by
bench.py's CPU-baseline leg. The product package never imports it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

OPS = {"Contains": 0, "Inside": 1, "DistanceToIn": 2, "DistanceToOut": 3, "SafetyToIn": 4, "SafetyToOut": 5}
SHAPE_TYPES = {"box": 0, "tube": 1, "cone": 2, "polycone": 3, "polyhedron": 4, "trapezoid": 5,
               "sphere": 6, "torus2": 7, "orb": 8, "tessellated": 9, "hype": 10,
               "trd": 11, "paraboloid": 12, "parallelepiped": 13, "tet": 14,
               "ellipticaltube": 15, "ellipsoid": 16, "gentrap": 17, "cuttube": 18, "ellipticalcone": 19}


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = ctypes.CDLL(path)
        vp, dp, i64 = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int64
        L.vgo_shape_new.restype = vp
        L.vgo_shape_new.argtypes = [ctypes.c_int, dp, ctypes.c_int]
        L.vgo_shape_free.argtypes = [vp]
        L.vgo_shape_extent.argtypes = [vp, dp, dp]
        L.vgo_shape_query.restype = ctypes.c_int
        L.vgo_shape_query.argtypes = [vp, ctypes.c_int] + [vp] * 7 + [vp, i64, ctypes.c_int]
        L.vgo_shape_normal.restype = ctypes.c_int
        L.vgo_shape_normal.argtypes = [vp] + [vp] * 7 + [i64]
        _LIB = L
    return _LIB


def _dptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class OracleShape:
    """One unplaced shape of the oracle, parameters as in the reference constructor."""

    def __init__(self, kind, params):
        self.kind = kind
        self.params = [float(x) for x in params]
        arr = (ctypes.c_double * len(self.params))(*self.params)
        self._h = lib().vgo_shape_new(SHAPE_TYPES[kind], arr, len(self.params))
        if not self._h:
            raise ValueError("oracle rejected %s%r" % (kind, tuple(params)))

    def __del__(self):
        if getattr(self, "_h", None) and _LIB is not None:
            _LIB.vgo_shape_free(self._h)
            self._h = None

    def extent(self):
        lo = (ctypes.c_double * 3)()
        hi = (ctypes.c_double * 3)()
        lib().vgo_shape_extent(self._h, lo, hi)
        return np.array(lo[:]), np.array(hi[:])

    def query(self, op, p, d=None, step=None, threads=1):
        """p, d: float64 arrays of shape (3, n) (SOA). Returns the output array."""
        p = np.ascontiguousarray(p, dtype=np.float64)
        n = p.shape[1]
        if d is not None:
            d = np.ascontiguousarray(d, dtype=np.float64)
        if op == "Contains":
            out = np.empty(n, np.uint8)
        elif op == "Inside":
            out = np.empty(n, np.int32)
        else:
            out = np.empty(n, np.float64)
        px, py, pz = (np.ascontiguousarray(p[i]) for i in range(3))
        dx = dy = dz = None
        if d is not None:
            dx, dy, dz = (np.ascontiguousarray(d[i]) for i in range(3))
        st = None if step is None else np.ascontiguousarray(step, dtype=np.float64)
        rc = lib().vgo_shape_query(self._h, OPS[op], _dptr(px), _dptr(py), _dptr(pz), _dptr(dx), _dptr(dy),
                                   _dptr(dz), _dptr(st), _dptr(out), n, int(threads))
        if rc != 0:
            raise RuntimeError("vgo_shape_query failed: %d" % rc)
        return out

    def normal(self, p, init=None):
        """Normal per point of p (3, n). init (3, n): the caller's vectors, kept where the reference
        leaves its normal argument unassigned (default NaN). Returns (normals (3, n), valid uint8 (n,))."""
        p = np.ascontiguousarray(p, dtype=np.float64)
        n = p.shape[1]
        nv = np.full((3, n), np.nan) if init is None else np.array(init, dtype=np.float64, copy=True)
        nv = np.ascontiguousarray(nv)
        valid = np.zeros(n, np.uint8)
        px, py, pz = (np.ascontiguousarray(p[i]) for i in range(3))
        nx, ny, nz = (np.ascontiguousarray(nv[i]) for i in range(3))
        rc = lib().vgo_shape_normal(self._h, _dptr(px), _dptr(py), _dptr(pz), _dptr(nx), _dptr(ny), _dptr(nz),
                                    _dptr(valid), n)
        if rc != 0:
            raise RuntimeError("vgo_shape_normal failed: %d" % rc)
        return np.stack([nx, ny, nz]), valid


def _np_ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class OracleGeometry:
    """Oracle geometry + BVH navigator (use_bvh=False: brute-force daughter loops)."""

    def __init__(self, desc, use_bvh=True):
        L = lib()
        if not getattr(L, "_geom_declared", False):
            vp, i64, ci = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
            L.vgo_geom_new.restype = vp
            L.vgo_geom_new.argtypes = [ci, vp, vp, vp, ci, vp, ci, vp, vp, vp, ci, ci]
            L.vgo_geom_free.argtypes = [vp]
            L.vgo_geom_navindex.restype = i64
            L.vgo_geom_navindex.argtypes = [vp, vp, i64]
            L.vgo_geom_maxdepth.restype = ci
            L.vgo_geom_maxdepth.argtypes = [vp]
            L.vgo_geom_bvh.restype = ci
            L.vgo_geom_bvh.argtypes = [vp, ci] + [vp] * 6
            L.vgo_nav_locate.restype = ci
            L.vgo_nav_locate.argtypes = [vp] + [vp] * 4 + [i64, ci]
            L.vgo_nav_step.restype = ci
            L.vgo_nav_step.argtypes = [vp] + [vp] * 13 + [i64, ci, ci]
            L.vgo_nav_safety.restype = ci
            L.vgo_nav_safety.argtypes = [vp] + [vp] * 5 + [i64, ci]
            L.vgo_nav_level_locate.restype = ci
            L.vgo_nav_level_locate.argtypes = [vp] + [vp] * 9 + [i64, ci]
            L.vgo_placed_query.restype = ci
            L.vgo_placed_query.argtypes = [vp, ci, ci] + [vp] * 8 + [i64, ci]
            L.vgo_placed_normal.restype = ci
            L.vgo_placed_normal.argtypes = [vp, ci] + [vp] * 7 + [i64]
            L.vgo_nav_to_path.restype = ci
            L.vgo_nav_to_path.argtypes = [vp, vp, i64, ci, vp, vp, vp]
            L.vgo_nav_from_path.restype = ci
            L.vgo_nav_from_path.argtypes = [vp, ci, vp, vp, i64, ci, vp]
            L.vgo_nav_distance.restype = ci
            L.vgo_nav_distance.argtypes = [vp, vp, vp, i64, vp]
            L.vgo_nav_global_to_local.restype = ci
            L.vgo_nav_global_to_local.argtypes = [vp, vp, ci] + [vp] * 6 + [i64]
            L.vgo_nav_propagate.restype = i64
            L.vgo_nav_propagate.argtypes = [vp] + [vp] * 6 + [ctypes.c_double, vp, vp, i64, ci, ci]
            L._geom_declared = True
        self.desc = desc
        self._arrs = desc.arrays()
        stype, soff, sparams, lvshape, plm, pll, plt = self._arrs
        self._h = L.vgo_geom_new(len(stype), _np_ptr(stype), _np_ptr(soff), _np_ptr(sparams), len(lvshape),
                                 _np_ptr(lvshape), len(plm), _np_ptr(plm), _np_ptr(pll), _np_ptr(plt), desc.world,
                                 1 if use_bvh else 0)
        if not self._h:
            raise ValueError("oracle rejected the geometry description")

    def __del__(self):
        if getattr(self, "_h", None) and _LIB is not None:
            _LIB.vgo_geom_free(self._h)
            self._h = None

    def navindex(self):
        n = lib().vgo_geom_navindex(self._h, None, 0)
        out = np.zeros(n, dtype=np.uint32)
        lib().vgo_geom_navindex(self._h, _np_ptr(out), n)
        return out

    def max_depth(self):
        return lib().vgo_geom_maxdepth(self._h)

    def bvh(self, lv):
        counts = np.zeros(3, dtype=np.int32)
        if lib().vgo_geom_bvh(self._h, lv, _np_ptr(counts), None, None, None, None, None) != 0:
            return None
        depth, nn, nd = (int(x) for x in counts)
        nchild = np.zeros(nn, np.int32)
        offset = np.zeros(nn, np.int32)
        prim = np.zeros(nd, np.int32)
        nodes = np.zeros(6 * nn)
        aabbs = np.zeros(6 * nd)
        lib().vgo_geom_bvh(self._h, lv, _np_ptr(counts), _np_ptr(nchild), _np_ptr(offset), _np_ptr(prim),
                           _np_ptr(nodes), _np_ptr(aabbs))
        return {"depth": depth, "nchild": nchild, "offset": offset, "primid": prim, "nodes": nodes.reshape(-1, 6),
                "aabbs": aabbs.reshape(-1, 6)}

    def locate(self, p, threads=1):
        p = np.ascontiguousarray(p, dtype=np.float64)
        n = p.shape[1]
        out = np.zeros(n, np.uint32)
        px, py, pz = (np.ascontiguousarray(p[i]) for i in range(3))
        lib().vgo_nav_locate(self._h, _np_ptr(px), _np_ptr(py), _np_ptr(pz), _np_ptr(out), n, threads)
        return out

    def step(self, p, d, limit, nav, last=None, relocate=True, threads=1):
        p = np.ascontiguousarray(p, dtype=np.float64)
        d = np.ascontiguousarray(d, dtype=np.float64)
        n = p.shape[1]
        lim = np.ascontiguousarray(np.broadcast_to(np.asarray(limit, dtype=np.float64), (n,)))
        nav = np.ascontiguousarray(nav, dtype=np.uint32)
        last = np.zeros(n, np.uint32) if last is None else np.ascontiguousarray(last, dtype=np.uint32)
        nav_out = np.zeros(n, np.uint32)
        last_out = np.zeros(n, np.uint32)
        bnd = np.zeros(n, np.uint8)
        step = np.zeros(n)
        comps = [np.ascontiguousarray(a[i]) for a in (p, d) for i in range(3)]
        lib().vgo_nav_step(self._h, *[_np_ptr(c) for c in comps], _np_ptr(lim), _np_ptr(nav), _np_ptr(last),
                           _np_ptr(nav_out), _np_ptr(last_out), _np_ptr(bnd), _np_ptr(step), n,
                           1 if relocate else 0, threads)
        return step, nav_out, last_out, bnd

    def safety(self, p, nav, threads=1):
        p = np.ascontiguousarray(p, dtype=np.float64)
        n = p.shape[1]
        nav = np.ascontiguousarray(nav, dtype=np.uint32)
        out = np.zeros(n)
        px, py, pz = (np.ascontiguousarray(p[i]) for i in range(3))
        lib().vgo_nav_safety(self._h, _np_ptr(px), _np_ptr(py), _np_ptr(pz), _np_ptr(nav), _np_ptr(out), n, threads)
        return out

    def level_locate(self, lv, p, excl=None, threads=1):
        """BVH LevelLocate per point in logical volume lv[i]: (placed daughter or -1, daughter-frame point)."""
        p = np.ascontiguousarray(p, dtype=np.float64)
        n = p.shape[1]
        lv = np.ascontiguousarray(np.broadcast_to(np.asarray(lv, dtype=np.int32), (n,)))
        ex = None if excl is None else np.ascontiguousarray(np.broadcast_to(np.asarray(excl, dtype=np.int32), (n,)))
        out = np.zeros(n, np.int32)
        dl = np.zeros((3, n))
        px, py, pz = (np.ascontiguousarray(p[i]) for i in range(3))
        dx, dy, dz = (np.ascontiguousarray(dl[i]) for i in range(3))
        lib().vgo_nav_level_locate(self._h, _np_ptr(lv), _np_ptr(ex) if ex is not None else None, _np_ptr(px),
                                   _np_ptr(py), _np_ptr(pz), _np_ptr(out), _np_ptr(dx), _np_ptr(dy), _np_ptr(dz), n,
                                   threads)
        return out, np.stack([dx, dy, dz])

    def placed_query(self, pvol, op, p, d=None, step=None, threads=1):
        """Placed-volume query (op as OPS, or 6 = PlacedDistanceToOut) on SOA (3, n) arrays."""
        code = OPS[op] if isinstance(op, str) else int(op)
        p = np.ascontiguousarray(p, dtype=np.float64)
        n = p.shape[1]
        out = np.empty(n, {0: np.uint8, 1: np.int32}.get(code, np.float64))
        px, py, pz = (np.ascontiguousarray(p[i]) for i in range(3))
        dx = dy = dz = None
        if d is not None:
            d = np.ascontiguousarray(d, dtype=np.float64)
            dx, dy, dz = (np.ascontiguousarray(d[i]) for i in range(3))
        st = None if step is None else np.ascontiguousarray(step, dtype=np.float64)
        rc = lib().vgo_placed_query(self._h, int(pvol), code, _np_ptr(px), _np_ptr(py), _np_ptr(pz),
                                    _dptr(dx), _dptr(dy), _dptr(dz), _dptr(st), _np_ptr(out), n, threads)
        if rc != 0:
            raise ValueError("vgo_placed_query rejected placed volume %d / op %d" % (pvol, code))
        return out

    def placed_normal(self, pvol, p):
        """VPlacedVolume::Normal of placement pvol at points p (3, n) of the mother's frame: (normals, valid)."""
        p = np.ascontiguousarray(p, dtype=np.float64)
        n = p.shape[1]
        px, py, pz = (np.ascontiguousarray(p[i]) for i in range(3))
        nx, ny, nz = (np.empty(n) for _ in range(3))
        valid = np.empty(n, np.uint8)
        if lib().vgo_placed_normal(self._h, int(pvol), _np_ptr(px), _np_ptr(py), _np_ptr(pz), _np_ptr(nx),
                                   _np_ptr(ny), _np_ptr(nz), _np_ptr(valid), n) != 0:
            raise ValueError("vgo_placed_normal rejected placed volume %d" % pvol)
        return np.stack([nx, ny, nz]), valid

    def nav_to_path(self, nav, max_levels):
        """(levels, placed volumes (max_levels, n), daughter indices (max_levels, n)) per state."""
        nav = np.ascontiguousarray(nav, dtype=np.uint32)
        n = nav.shape[0]
        levels = np.empty(n, np.int32)
        pv = np.empty((max_levels, n), np.int32)
        ix = np.empty((max_levels, n), np.int32)
        if lib().vgo_nav_to_path(self._h, _np_ptr(nav), n, int(max_levels), _np_ptr(levels), _np_ptr(pv),
                                 _np_ptr(ix)) != 0:
            raise ValueError("a state has more than max_levels levels")
        return levels, pv, ix

    def nav_from_path(self, levels, path, kind):
        """NavIndex per path; kind 0: daughter indices, 1: placed volumes (world first)."""
        levels = np.ascontiguousarray(levels, dtype=np.int32)
        path = np.ascontiguousarray(path, dtype=np.int32)
        n = levels.shape[0]
        out = np.empty(n, np.uint32)
        lib().vgo_nav_from_path(self._h, int(kind), _np_ptr(levels), _np_ptr(path), n, path.shape[0], _np_ptr(out))
        return out

    def nav_distance(self, a, b):
        a = np.ascontiguousarray(a, dtype=np.uint32)
        b = np.ascontiguousarray(b, dtype=np.uint32)
        out = np.empty(a.shape[0], np.int32)
        lib().vgo_nav_distance(self._h, _np_ptr(a), _np_ptr(b), a.shape[0], _np_ptr(out))
        return out

    def global_to_local(self, nav, p, tolevel=-1):
        nav = np.ascontiguousarray(nav, dtype=np.uint32)
        p = np.ascontiguousarray(p, dtype=np.float64)
        n = nav.shape[0]
        comps = [np.ascontiguousarray(p[i]) for i in range(3)]
        out = [np.empty(n) for _ in range(3)]
        lib().vgo_nav_global_to_local(self._h, _np_ptr(nav), int(tolevel), *[_np_ptr(c) for c in comps],
                                      *[_np_ptr(c) for c in out], n)
        return np.stack(out)

    def propagate(self, p, d, limit, nav, last, nsteps, threads=1):
        """In-place transport loop; returns (p, nav, last, steps_taken)."""
        p = np.array(p, dtype=np.float64, order="C", copy=True)
        nav = np.array(nav, dtype=np.uint32, copy=True)
        last = np.array(last, dtype=np.uint32, copy=True)
        n = p.shape[1]
        px, py, pz = (np.ascontiguousarray(p[i]) for i in range(3))
        dx, dy, dz = (np.ascontiguousarray(d[i], dtype=np.float64) for i in range(3))
        taken = lib().vgo_nav_propagate(self._h, _np_ptr(px), _np_ptr(py), _np_ptr(pz), _np_ptr(dx), _np_ptr(dy),
                                        _np_ptr(dz), float(limit), _np_ptr(nav), _np_ptr(last), n, nsteps, threads)
        return np.stack([px, py, pz]), nav, last, int(taken)
