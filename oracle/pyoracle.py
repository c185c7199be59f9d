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
               "sphere": 6, "torus2": 7, "orb": 8, "tessellated": 9, "hype": 10}


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
