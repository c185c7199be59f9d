"""GDML ingestion: vgdml::Frontend::Load (persistency/gdml/source/include/Frontend.h:12-21) over the
library's own reader (vgmi_gdml_read). The result is a GeometryDesc, with the logical volumes' and
placements' names, ready for navigation.Geometry."""
import ctypes
import os

import numpy as np

from ._lib import check, lib
from .geometry import GeometryDesc
from .navigation import _Desc

MM_UNIT_DEFAULT = 0.1  # GeoManager's default: one millimetre is 0.1 internal units (centimetres)


def _declare():
    L = lib()
    if getattr(L, "_gdml_declared", False):
        return L
    vp, ci = ctypes.c_void_p, ctypes.c_int
    for name, args in (("vgmi_gdml_read", [ctypes.c_char_p, ctypes.c_double, ctypes.c_int, ctypes.POINTER(vp)]),
                       ("vgmi_gdml_desc", [vp, ctypes.POINTER(_Desc)]),
                       ("vgmi_gdml_name", [vp, ci, ci, ctypes.POINTER(ctypes.c_char_p)]),
                       ("vgmi_gdml_destroy", [vp])):
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    L._gdml_declared = True
    return L


def _array(ptr, n, ctype):
    if n == 0 or not ptr:
        return np.zeros(0, dtype=np.dtype(ctype))
    return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctype)), shape=(n,)).copy()


def load_gdml(path, mm_unit=MM_UNIT_DEFAULT, specialized=False):
    """Read a GDML file into a GeometryDesc (lengths in units of 1 / mm_unit millimetres). Solids are
    those of the reference's default build (as written); specialized=True applies the
    GeoManager::MakeInstance substitutions of a build with VECGEOM_NO_SPECIALIZATION=OFF. Raises
    VgmiError for an unsupported solid, a reflected placement or a malformed document."""
    L = _declare()
    h = ctypes.c_void_p()
    check(L.vgmi_gdml_read(os.fsencode(path), float(mm_unit), 1 if specialized else 0, ctypes.byref(h)),
          "vgmi_gdml_read")
    try:
        d = _Desc()
        check(L.vgmi_gdml_desc(h, ctypes.byref(d)), "vgmi_gdml_desc")
        stype = _array(d.shape_type, d.nshapes, ctypes.c_int)
        soff = _array(d.shape_param_offset, d.nshapes + 1, ctypes.c_int)
        sparams = _array(d.shape_params, int(soff[-1]), ctypes.c_double)
        lvshape = _array(d.logical_shape, d.nlogical, ctypes.c_int)
        plm = _array(d.placed_mother, d.nplaced, ctypes.c_int)
        pll = _array(d.placed_logical, d.nplaced, ctypes.c_int)
        plt = _array(d.placed_transform, 12 * d.nplaced, ctypes.c_double)
        g = GeometryDesc()
        for i in range(d.nshapes):
            g.add_shape(int(stype[i]), sparams[soff[i]:soff[i + 1]].tolist())
        name = ctypes.c_char_p()
        for i in range(d.nlogical):
            check(L.vgmi_gdml_name(h, 0, i, ctypes.byref(name)), "vgmi_gdml_name")
            g.add_logical(int(lvshape[i]), name.value.decode())
        g.placement_names = []
        for k in range(d.nplaced):
            g.place(int(plm[k]), int(pll[k]), plt[12 * k:12 * k + 12].tolist())
            check(L.vgmi_gdml_name(h, 1, k, ctypes.byref(name)), "vgmi_gdml_name")
            g.placement_names.append(name.value.decode())
        g.set_world(int(d.world_logical))
        return g
    finally:
        L.vgmi_gdml_destroy(h)
