// Flat device geometry: the tables the navigation kernels read.
//
// Everything a track needs is reachable from a NavIndex (the uint32 state) in a
// few dependent loads: the NavIndex table gives the placed volume and the cached
// global matrix, the placed-volume table gives the logical volume and the local
// placement, the logical-volume table gives the shape record and the BVH arrays.
// The whole calorimeter (23k touchables, 2.9k placed volumes) is ~3 MB and stays
// resident in the L2/MALL caches; tracks stream through HBM.
#pragma once
#include "shape_data.h"
#include <stdint.h>

namespace vgmi {

// One placed volume (VPlacedVolume): logical volume, position in the mother's
// daughter list, and its transformation (translation, row-major rotation;
// VecGeom/base/Transformation3D.h:59-63).
struct PVolRec {
  int32_t lv;
  int32_t childId;
  int32_t hasT, hasR; // SetProperties flags (source/Transformation3D.cpp:118-126)
  double t[3];
  double r[9];
};

// One logical volume: shape, daughter list and BVH (source/BVH.cpp:48-87).
struct LVRec {
  int32_t shape;
  int32_t ndaughters;
  int32_t dOffset;    // daughters[dOffset .. dOffset + ndaughters) and their AABBs
  int32_t nodeBase;   // BVH nodes [nodeBase, nodeBase + nNodes)
  int32_t nNodes;     // 0 when the volume has no daughters
  int32_t primBase;   // BVH primitive ids (indices into this volume's daughter list)
  int32_t pad_[2];
};

struct AABBRec {
  double lo[3], hi[3];
};

// ---- Navigation records (device layout derived from the tables above at upload).
// Each record holds everything one traversal step needs from it, so a step issues its
// loads as a few independent 16-byte reads instead of a chain of dependent ones.

// Placed volume with its logical volume's fields folded in (128 B).
struct alignas(16) NavVol {
  int32_t lv, childId, shape, ndaughters; // shape: the logical volume's shape record
  int32_t nodeBase, slotBase, pad0, pad1; // BVH nodes; daughter slots (== LVRec::dOffset)
  double t[3], r[9];                      // placement (as PVolRec)
};

// BVH node (64 B): AABB, child count (-1: internal) and primitive offset (LVRec::primBase-relative).
struct alignas(16) NavNode {
  double lo[3], hi[3];
  int32_t nchild, offset, pad0, pad1;
};

// Daughter slot in BVH primitive order (64 B): slot slotBase + k holds the daughter of
// primitive prim[primBase + k], i.e. its ABBox, placed volume and shape record.
struct alignas(16) NavSlot {
  double lo[3], hi[3];
  int32_t pvol, shape, pad0, pad1;
};

// FP32 copies for the mixed-precision traversal (32 B each): the boxes are widened by
// kNavBox32Pad plus 2e-6 of the volume's coordinate range and rounded outwards, so every box
// test the FP64 traversal passes is passed in FP32 as well (the FP32 tests only filter;
// FP64 decides).
constexpr double kNavBox32Pad = 1e-3; // mm
struct alignas(16) NavNode32 {
  float lo[3], hi[3];
  int32_t nchild, offset;
};
struct alignas(16) NavSlot32 {
  float lo[3], hi[3];
  int32_t pvol, shape;
};

// Kernel-argument view of the device tables.
struct GeomView {
  const ShapeRecord *shapes;
  const PVolRec *pvols;
  const LVRec *lvs;
  const int32_t *daughters;  // placed-volume ids
  const AABBRec *dAABB;      // per daughter entry (same index as daughters)
  const int32_t *nchild;     // per BVH node; -1 marks an internal node
  const int32_t *offset;     // per BVH node
  const AABBRec *nodes;      // per BVH node
  const int32_t *prim;       // per BVH primitive slot
  const uint32_t *navind;    // NavIndex table (source/NavIndexTable.cpp:9-99 layout)
  const NavVol *vols;        // per placed volume
  const NavNode *bnodes;     // per BVH node (same index as nchild/offset/nodes)
  const NavSlot *slots;      // per daughter entry, in primitive order
  const NavNode32 *bnodes32; // FP32 copies of bnodes / slots
  const NavSlot32 *slots32;
  int32_t worldPvol;
  int32_t maxDepth;
  int32_t nlogical;
};

} // namespace vgmi
