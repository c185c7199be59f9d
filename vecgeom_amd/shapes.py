"""Unplaced shapes with the VUnplacedVolume SOA query interface
(VecGeom/volumes/UnplacedVolume.h:88-223, SOA overloads :144,179 and the
VPlacedVolume SOA overloads VecGeom/volumes/PlacedVolume.h:273-421).

Every method runs the HIP kernels of libvgmi.so on the current torch stream
and returns a device tensor; there is no host fallback.
"""
import ctypes

import torch

from ._lib import check, lib
from .soa import SOA3D, device_array, same_length

(BOX, TUBE, CONE, POLYCONE, POLYHEDRON, TRAPEZOID, SPHERE, TORUS2, ORB, TESSELLATED, HYPE, TRD, PARABOLOID,
 PARALLELEPIPED, TET, ELLIPTICALTUBE, ELLIPSOID, GENTRAP, CUTTUBE, ELLIPTICALCONE) = range(20)
OPS = {"Contains": 0, "Inside": 1, "DistanceToIn": 2, "DistanceToOut": 3, "SafetyToIn": 4, "SafetyToOut": 5}
FP64, MIXED = 0, 1
_OUT_DTYPE = {0: torch.uint8, 1: torch.int32}


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class UnplacedShape:
    type_id = None

    def __init__(self, type_id, params):
        self.type_id = type_id
        self.params = [float(x) for x in params]
        arr = (ctypes.c_double * max(1, len(self.params)))(*self.params)
        h = ctypes.c_void_p()
        check(lib().vgmi_shape_create(type_id, arr, len(self.params), ctypes.byref(h)), "vgmi_shape_create")
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().vgmi_shape_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def handle(self):
        return self._h

    def Extent(self):
        lo, hi = (ctypes.c_double * 3)(), (ctypes.c_double * 3)()
        check(lib().vgmi_shape_extent(self._h, lo, hi), "vgmi_shape_extent")
        return list(lo), list(hi)

    # ---- batched SOA interface
    def _query(self, op, points, dirs=None, step_max=None, out=None, precision=FP64):
        n = len(points)
        code = OPS[op]
        if code in (2, 3) and dirs is None:
            raise ValueError("%s needs directions" % op)
        same_length(points, dirs)
        odt = _OUT_DTYPE.get(code, torch.float64)
        if out is None:
            out = torch.empty(n, dtype=odt, device=points.x.device)
        device_array(out, odt, n, "out")
        d = dirs.ptrs() if dirs is not None else (None, None, None)
        if step_max is not None:
            device_array(step_max, torch.float64, n, "step_max")
        sm = step_max.data_ptr() if step_max is not None else None
        check(lib().vgmi_shape_query(self._h, code, precision, *points.ptrs(), *d, sm, out.data_ptr(), n,
                                     _stream()), "vgmi_shape_query(%s)" % op)
        return out

    def Contains(self, points, out=None, precision=FP64):
        return self._query("Contains", points, out=out, precision=precision)

    def Inside(self, points, out=None, precision=FP64):
        return self._query("Inside", points, out=out, precision=precision)

    def DistanceToIn(self, points, dirs, step_max=None, out=None, precision=FP64):
        return self._query("DistanceToIn", points, dirs, step_max, out, precision)

    def DistanceToOut(self, points, dirs, step_max=None, out=None, precision=FP64):
        return self._query("DistanceToOut", points, dirs, step_max, out, precision)

    def SafetyToIn(self, points, out=None, precision=FP64):
        return self._query("SafetyToIn", points, out=out, precision=precision)

    def SafetyToOut(self, points, out=None, precision=FP64):
        return self._query("SafetyToOut", points, out=out, precision=precision)

    def Normal(self, points, out=None, valid=None):
        """Surface normals (VUnplacedVolume::Normal, VecGeom/volumes/UnplacedVolume.h:252): returns
        (normals as an SOA3D, valid uint8 tensor). Where the reference leaves the normal unassigned,
        the vector of `out` is kept (a fresh `out` starts at zero)."""
        n, dev = len(points), points.x.device
        if out is None:
            out = SOA3D.zeros(n, dev)
        same_length(points, out)
        if valid is None:
            valid = torch.empty(n, dtype=torch.uint8, device=dev)
        device_array(valid, torch.uint8, n, "valid")
        check(lib().vgmi_shape_normal(self._h, *points.ptrs(), *out.ptrs(), valid.data_ptr(), n, _stream()),
              "vgmi_shape_normal")
        return out, valid

    def QueryAll(self, points, dirs, precision=FP64):
        """All six methods in one pass; returns a dict of device tensors."""
        n, dev = len(points), points.x.device
        same_length(points, dirs)
        res = {"Contains": torch.empty(n, dtype=torch.uint8, device=dev),
               "Inside": torch.empty(n, dtype=torch.int32, device=dev)}
        for k in ("DistanceToIn", "DistanceToOut", "SafetyToIn", "SafetyToOut"):
            res[k] = torch.empty(n, dtype=torch.float64, device=dev)
        check(lib().vgmi_shape_query_all(self._h, precision, *points.ptrs(), *dirs.ptrs(),
                                         *(res[k].data_ptr() for k in ("Contains", "Inside", "DistanceToIn",
                                                                       "DistanceToOut", "SafetyToIn",
                                                                       "SafetyToOut")),
                                         n, _stream()), "vgmi_shape_query_all")
        return res


class UnplacedBox(UnplacedShape):
    def __init__(self, dx, dy, dz):
        super().__init__(BOX, [dx, dy, dz])


class UnplacedTube(UnplacedShape):
    def __init__(self, rmin, rmax, dz, sphi, dphi):
        super().__init__(TUBE, [rmin, rmax, dz, sphi, dphi])


_BY_NAME = {"box": BOX, "tube": TUBE, "cone": CONE, "polycone": POLYCONE, "polyhedron": POLYHEDRON,
            "trapezoid": TRAPEZOID, "sphere": SPHERE, "torus2": TORUS2, "orb": ORB, "tessellated": TESSELLATED,
            "hype": HYPE, "trd": TRD, "paraboloid": PARABOLOID, "parallelepiped": PARALLELEPIPED, "tet": TET,
            "ellipticaltube": ELLIPTICALTUBE, "ellipsoid": ELLIPSOID, "gentrap": GENTRAP, "cuttube": CUTTUBE, "ellipticalcone": ELLIPTICALCONE}


def make_shape(kind, params):
    """Create an unplaced shape by name with the reference constructor's parameter list."""
    return UnplacedShape(_BY_NAME[kind], params)


__all__ = ["UnplacedShape", "UnplacedBox", "UnplacedTube", "make_shape", "SOA3D", "FP64", "MIXED", "OPS"]
