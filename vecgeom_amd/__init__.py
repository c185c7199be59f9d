"""vecgeom_amd: MI355X-native batched shape queries and BVH navigation.

Host-side mirror of VecGeom's query interface over the C ABI of libvgmi.so
(include/vgmi.h). Device memory and streams come from torch; all arithmetic
runs in the library's HIP kernels.
"""
from ._lib import LIB_PATH, VgmiError, check, lib
from .soa import SOA3D
from .shapes import FP64, MIXED, OPS, UnplacedBox, UnplacedShape, UnplacedTube, make_shape

__all__ = ["LIB_PATH", "VgmiError", "check", "lib", "SOA3D", "FP64", "MIXED", "OPS", "UnplacedBox",
           "UnplacedShape", "UnplacedTube", "make_shape"]
