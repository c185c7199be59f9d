// C ABI: geometry build (host tables + one device block) and table access.
#include "geom_internal.h"
#include "vgmi_internal.h"
#include "shape_host.h"
#include <cstring>
#include <new>

using namespace vgmi;

namespace {

size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

// Copies every host table into one device allocation and fills the view.
int upload(vgmi_geom &G)
{
  HostGeometry const &h = G.host;
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
                  {h.extraPool.data(), h.extraPool.size() * sizeof(double), 0}};
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
  v.worldPvol      = h.worldPvol;
  v.maxDepth       = h.maxDepth;
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
