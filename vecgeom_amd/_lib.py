"""Loader for the native library libvgmi.so (HIP kernels for gfx950 + C ABI).

The package has no CPU fallback: if the library is missing or no GPU is
visible, every query raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvgmi.so")
_LIB = None


class VgmiError(RuntimeError):
    pass


def lib():
    """Load libvgmi.so once (torch is imported first so both share one HIP runtime)."""
    global _LIB
    if _LIB is None:
        import torch  # noqa: F401  (loads the HIP runtime torch was built with)
        if not os.path.exists(LIB_PATH):
            raise VgmiError("libvgmi.so is not built: run __graft_entry__.build() (make -C vecgeom_amd/csrc)")
        L = ctypes.CDLL(LIB_PATH)
        L.vgmi_version.restype = ctypes.c_char_p
        L.vgmi_last_error.restype = ctypes.c_char_p
        _declare(L)
        _LIB = L
    return _LIB


def check(rc, what=""):
    if rc != 0:
        msg = lib().vgmi_last_error().decode()
        raise VgmiError("%s failed (%d): %s" % (what, rc, msg))


def _declare(L):
    vp, i, d, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_size_t
    dp = ctypes.POINTER(ctypes.c_double)
    sigs = {
        "vgmi_shape_create": [i, dp, i, ctypes.POINTER(vp)],
        "vgmi_shape_destroy": [vp],
        "vgmi_shape_extent": [vp, dp, dp],
        "vgmi_shape_query": [vp, i, i] + [vp] * 8 + [sz, vp],
        "vgmi_shape_query_all": [vp, i] + [vp] * 12 + [sz, vp],
        "vgmi_shape_normal": [vp] + [vp] * 7 + [sz, vp],
    }
    for name, args in sigs.items():
        if hasattr(L, name):
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int
