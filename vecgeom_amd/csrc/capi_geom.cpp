// C ABI: geometry build (host tables + one device block) and table access.
#include "geom_internal.h"
#include "vgmi_internal.h"
#include "shape_host.h"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>

using namespace vgmi;

namespace {

size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

// FP32 box widened by `pad` and rounded outwards.
void widen32(const double lo[3], const double hi[3], double pad, float flo[3], float fhi[3])
{
  for (int k = 0; k < 3; ++k) {
    float a = float(lo[k] - pad), b = float(hi[k] + pad);
    if (double(a) > lo[k] - pad) a = std::nextafter(a, -HUGE_VALF);
    if (double(b) < hi[k] + pad) b = std::nextafter(b, HUGE_VALF);
    flo[k] = a, fhi[k] = b;
  }
}

// Widening of the FP32 boxes of one logical volume: kNavBox32Pad plus 2e-6 of the largest
// coordinate a point inside the volume or on its daughters' boxes can have, which exceeds
// the FP32 rounding of the point, of the box and of the test arithmetic.
double pad32(HostGeometry const &h, LVRec const &L)
{
  double lo[3], hi[3], R = 0.;
  ShapeExtent(h.shapes[L.shape], lo, hi);
  for (int k = 0; k < 3; ++k) R = std::max(R, std::max(std::fabs(lo[k]), std::fabs(hi[k])));
  for (int i = 0; i < L.nNodes; ++i)
    for (int k = 0; k < 3; ++k)
      R = std::max(R, std::max(std::fabs(h.nodes[L.nodeBase + i].lo[k]), std::fabs(h.nodes[L.nodeBase + i].hi[k])));
  return kNavBox32Pad + 2e-6 * R;
}

// Copies every host table into one device allocation and fills the view.
int upload(vgmi_geom &G)
{
  HostGeometry const &h = G.host;
  // navigation records: the placed-volume, BVH-node and daughter tables re-packed
  std::vector<NavVol> vols(h.pvols.size());
  for (size_t i = 0; i < h.pvols.size(); ++i) {
    PVolRec const &P = h.pvols[i];
    LVRec const &L   = h.lvs[P.lv];
    NavVol &V        = vols[i];
    V                = NavVol{};
    V.lv = P.lv, V.childId = P.childId, V.shape = L.shape, V.ndaughters = L.ndaughters;
    V.nodeBase = L.nodeBase, V.slotBase = L.dOffset;
    std::memcpy(V.t, P.t, sizeof(V.t));
    std::memcpy(V.r, P.r, sizeof(V.r));
  }
  std::vector<NavNode> bnodes(h.nchild.size());
  for (size_t i = 0; i < bnodes.size(); ++i) {
    bnodes[i] = NavNode{};
    std::memcpy(bnodes[i].lo, h.nodes[i].lo, sizeof(bnodes[i].lo));
    std::memcpy(bnodes[i].hi, h.nodes[i].hi, sizeof(bnodes[i].hi));
    bnodes[i].nchild = h.nchild[i];
    bnodes[i].offset = h.offset[i];
  }
  std::vector<NavSlot> slots(h.daughters.size());
  for (LVRec const &L : h.lvs)
    for (int k = 0; k < L.ndaughters; ++k) {
      const int prim = h.prim[L.primBase + k];
      NavSlot &S     = slots[L.dOffset + k];
      S              = NavSlot{};
      std::memcpy(S.lo, h.dAABB[L.dOffset + prim].lo, sizeof(S.lo));
      std::memcpy(S.hi, h.dAABB[L.dOffset + prim].hi, sizeof(S.hi));
      S.pvol  = h.daughters[L.dOffset + prim];
      S.shape = h.lvs[h.pvols[S.pvol].lv].shape;
    }
  struct Part {
    const void *src;
    size_t bytes;
    size_t off;
  };
  Part parts[] = {{h.shapes.data(), h.shapes.size() * sizeof(ShapeRecord), 0},
                  {h.pvols.data(), h.pvols.size() * sizeof(PVolRec), 0},
                  {h.lvs.data(), h.lvs.size() * sizeof(LVRec), 0},
                  {h.daughters.data(), h.daughters.size() * sizeof(int32_t), 0},
                  {h.dAABB.data(), h.dAABB.size() * sizeof(AABBRec), 0},
                  {h.nchild.data(), h.nchild.size() * sizeof(int32_t), 0},
                  {h.offset.data(), h.offset.size() * sizeof(int32_t), 0},
                  {h.nodes.data(), h.nodes.size() * sizeof(AABBRec), 0},
                  {h.prim.data(), h.prim.size() * sizeof(int32_t), 0},
                  {h.navind.data(), h.navind.size() * sizeof(uint32_t), 0},
                  {h.extraPool.data(), h.extraPool.size() * sizeof(double), 0},
                  {vols.data(), vols.size() * sizeof(NavVol), 0},
                  {bnodes.data(), bnodes.size() * sizeof(NavNode), 0},
                  {slots.data(), slots.size() * sizeof(NavSlot), 0},
                  {nullptr, 0, 0},
                  {nullptr, 0, 0}};
  std::vector<NavNode32> bnodes32(bnodes.size());
  std::vector<NavSlot32> slots32(slots.size());
  for (LVRec const &L : h.lvs) {
    if (L.ndaughters == 0) continue;
    const double pad = pad32(h, L);
    for (int i = L.nodeBase; i < L.nodeBase + L.nNodes; ++i) {
      widen32(bnodes[i].lo, bnodes[i].hi, pad, bnodes32[i].lo, bnodes32[i].hi);
      bnodes32[i].nchild = bnodes[i].nchild;
      bnodes32[i].offset = bnodes[i].offset;
    }
    for (int i = L.dOffset; i < L.dOffset + L.ndaughters; ++i) {
      widen32(slots[i].lo, slots[i].hi, pad, slots32[i].lo, slots32[i].hi);
      slots32[i].pvol  = slots[i].pvol;
      slots32[i].shape = slots[i].shape;
    }
  }
  parts[14] = Part{bnodes32.data(), bnodes32.size() * sizeof(NavNode32), 0};
  parts[15] = Part{slots32.data(), slots32.size() * sizeof(NavSlot32), 0};
  size_t total = 0;
  for (auto &p : parts) {
    p.off = total;
    total = align_up(total + p.bytes);
  }
  if (hipGetDevice(&G.device) != hipSuccess) return set_error(VGMI_ERR_DEVICE, "vgmi_geom_build: no HIP device");
  void *blk      = nullptr;
  hipError_t e   = hipMalloc(&blk, total ? total : 256);
  if (e != hipSuccess) return set_error(VGMI_ERR_DEVICE, "vgmi_geom_build: hipMalloc(%zu): %s", total, hipGetErrorString(e));
  std::vector<unsigned char> staging(total, 0);
  for (auto &p : parts)
    if (p.bytes) std::memcpy(staging.data() + p.off, p.src, p.bytes);
  // shape records that own variable-length data point into the uploaded pool
  {
    const double *pool = reinterpret_cast<const double *>(static_cast<unsigned char *>(blk) + parts[10].off);
    ShapeRecord *recs  = reinterpret_cast<ShapeRecord *>(staging.data() + parts[0].off);
    for (size_t i = 0; i < h.shapes.size(); ++i)
      if (h.extraOffset[i] >= 0) SetShapeExtraBase(recs[i], pool + h.extraOffset[i]);
    G.devShapes.assign(recs, recs + h.shapes.size());
  }
  e = hipMemcpy(blk, staging.data(), total, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(blk);
    return set_error(VGMI_ERR_DEVICE, "vgmi_geom_build: hipMemcpy: %s", hipGetErrorString(e));
  }
  auto at = [&](int i) { return reinterpret_cast<const void *>(static_cast<unsigned char *>(blk) + parts[i].off); };
  GeomView &v      = G.view;
  v.shapes         = static_cast<const ShapeRecord *>(at(0));
  v.pvols          = static_cast<const PVolRec *>(at(1));
  v.lvs            = static_cast<const LVRec *>(at(2));
  v.daughters      = static_cast<const int32_t *>(at(3));
  v.dAABB          = static_cast<const AABBRec *>(at(4));
  v.nchild         = static_cast<const int32_t *>(at(5));
  v.offset         = static_cast<const int32_t *>(at(6));
  v.nodes          = static_cast<const AABBRec *>(at(7));
  v.prim           = static_cast<const int32_t *>(at(8));
  v.navind         = static_cast<const uint32_t *>(at(9));
  v.vols           = static_cast<const NavVol *>(at(11));
  v.bnodes         = static_cast<const NavNode *>(at(12));
  v.slots          = static_cast<const NavSlot *>(at(13));
  v.bnodes32       = static_cast<const NavNode32 *>(at(14));
  v.slots32        = static_cast<const NavSlot32 *>(at(15));
  v.worldPvol      = h.worldPvol;
  v.maxDepth       = h.maxDepth;
  v.nlogical       = int32_t(h.lvs.size());
  G.device_block   = blk;
  G.device_bytes   = total;
  return VGMI_OK;
}

} // namespace

extern "C" {

vgmi_status vgmi_geom_build(const vgmi_geom_desc *desc, int flags, vgmi_geom **out)
{
  if (!out) return set_error(VGMI_ERR_INVALID, "vgmi_geom_build: null output handle");
  *out = nullptr;
  if (!desc || !desc->shape_type || !desc->shape_param_offset || !desc->logical_shape ||
      (desc->nplaced > 0 && (!desc->placed_mother || !desc->placed_logical || !desc->placed_transform)))
    return set_error(VGMI_ERR_INVALID, "vgmi_geom_build: incomplete description");
  vgmi_geom *G = new (std::nothrow) vgmi_geom();
  if (!G) return set_error(VGMI_ERR_INVALID, "vgmi_geom_build: out of host memory");
  std::string err;
  if (build_host_geometry(*desc, G->host, err) != 0) {
    delete G;
    return set_error(VGMI_ERR_INVALID, "vgmi_geom_build: %s", err.c_str());
  }
  for (auto const &r : G->host.shapes) G->typeMask |= 1u << unsigned(r.type);
  G->nlogical = int(G->host.lvs.size());
  if (!(flags & VGMI_GEOM_HOST_ONLY)) {
    if (int rc = upload(*G)) {
      delete G;
      return rc;
    }
  }
  *out = G;
  return VGMI_OK;
}

vgmi_status vgmi_geom_destroy(vgmi_geom *geom)
{
  if (!geom) return VGMI_OK;
  if (geom->device_block) (void)hipFree(geom->device_block);
  delete geom;
  return VGMI_OK;
}

vgmi_status vgmi_geom_info(const vgmi_geom *g, int64_t counts[6])
{
  if (!g || !counts) return set_error(VGMI_ERR_INVALID, "vgmi_geom_info: null argument");
  counts[0] = int64_t(g->host.pvols.size());
  counts[1] = int64_t(g->host.lvs.size());
  counts[2] = int64_t(g->host.navind.size());
  counts[3] = g->host.maxDepth;
  counts[4] = g->host.worldPvol;
  counts[5] = int64_t(g->device_bytes);
  return VGMI_OK;
}

vgmi_status vgmi_geom_navindex(const vgmi_geom *g, uint32_t *out, int64_t cap, int64_t *nwords)
{
  if (!g) return set_error(VGMI_ERR_INVALID, "vgmi_geom_navindex: null geometry");
  const int64_t n = int64_t(g->host.navind.size());
  if (nwords) *nwords = n;
  if (out) std::memcpy(out, g->host.navind.data(), size_t(std::min(cap, n)) * sizeof(uint32_t));
  return VGMI_OK;
}

vgmi_status vgmi_geom_bvh(const vgmi_geom *g, int logical, int counts[3], int *nchild, int *offset, int *primid,
                          double *nodes6, double *aabbs6)
{
  if (!g || !counts) return set_error(VGMI_ERR_INVALID, "vgmi_geom_bvh: null argument");
  HostGeometry const &h = g->host;
  if (logical < 0 || logical >= int(h.lvs.size())) return set_error(VGMI_ERR_INVALID, "vgmi_geom_bvh: bad logical");
  LVRec const &L = h.lvs[logical];
  if (L.ndaughters == 0) return set_error(VGMI_ERR_INVALID, "vgmi_geom_bvh: volume has no daughters");
  int bi = 0;
  for (int i = 0; i < logical; ++i) bi += h.lvs[i].ndaughters > 0;
  counts[0] = h.bvhDepth[bi];
  counts[1] = L.nNodes;
  counts[2] = L.ndaughters;
  for (int i = 0; i < L.nNodes; ++i) {
    if (nchild) nchild[i] = h.nchild[L.nodeBase + i];
    if (offset) offset[i] = h.offset[L.nodeBase + i];
    if (nodes6)
      for (int k = 0; k < 3; ++k) nodes6[6 * i + k] = h.nodes[L.nodeBase + i].lo[k], nodes6[6 * i + 3 + k] = h.nodes[L.nodeBase + i].hi[k];
  }
  for (int i = 0; i < L.ndaughters; ++i) {
    if (primid) primid[i] = h.prim[L.primBase + i];
    if (aabbs6)
      for (int k = 0; k < 3; ++k) aabbs6[6 * i + k] = h.dAABB[L.dOffset + i].lo[k], aabbs6[6 * i + 3 + k] = h.dAABB[L.dOffset + i].hi[k];
  }
  return VGMI_OK;
}

} // extern "C"
