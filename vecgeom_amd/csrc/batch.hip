// One-launch batched queries over many shapes with uneven point counts.
//
// The points of shape s occupy [offset_s, offset_s + count_s) of the SOA arrays.
// The range is cut into tiles of `tile` points, one tile per method group of the
// shape type (method_groups); tiles are sorted by the per-point cost of their group
// (most expensive first, longest-processing-time order) and
// consumed by a persistent grid through one atomic work counter. A tile belongs to
// exactly one shape, so every wavefront of a workgroup evaluates one shape type:
// the type switch is wave-uniform, the shape record is read through scalar loads,
// and no lane waits on another lane's more expensive shape.
//
// The launch sequence (counter reset kernel + persistent kernel) can be captured
// once into a hipGraph and replayed; the reset is a kernel, not a memset node.
#include "dev_mixed.h"
#include "dev_tess_tile.h"
#include "vgmi_internal.h"
#include <algorithm>
#include <vector>

namespace vgmi {

struct Tile {
  int32_t shape; // index into the batch's record table
  int32_t mask;  // methods this tile evaluates (kM* bits)
  int64_t begin, end;
};

// Method bits of a tile. A shape whose methods are long serial chains per point (polycone,
// polyhedron, tessellated) has each point range cut into several tiles, one per method group, so
// that different wavefronts run the groups side by side: the range then takes as long as its
// slowest group instead of the sum of all six.
constexpr int kMContains = 1, kMInside = 2, kMDistIn = 4, kMDistOut = 8, kMSafeIn = 16, kMSafeOut = 32;
constexpr int kMAll      = 63;
constexpr int kMPoint    = kMContains | kMInside | kMSafeIn | kMSafeOut;

VGMI_DEV BatchOut mask_outputs(BatchOut o, int m)
{
  if (!(m & kMContains)) o.contains = nullptr;
  if (!(m & kMInside)) o.inside = nullptr;
  if (!(m & kMDistIn)) o.din = nullptr;
  if (!(m & kMDistOut)) o.dout = nullptr;
  if (!(m & kMSafeIn)) o.sin = nullptr;
  if (!(m & kMSafeOut)) o.sout = nullptr;
  return o;
}

template <class Data>
VGMI_DEV void run_tile(Data const &s, int64_t b, int64_t e, BatchOut const &o, int nthreads)
{
  using S = ShapeOps<Data>;
  for (int64_t i = b + threadIdx.x; i < e; i += nthreads) {
    Vec3<double> p(__builtin_nontemporal_load(o.px + i), __builtin_nontemporal_load(o.py + i),
                   __builtin_nontemporal_load(o.pz + i));
    if (o.contains) __builtin_nontemporal_store(uint8_t(S::template Contains<double>(s, p)), o.contains + i);
    if (o.inside) __builtin_nontemporal_store(int32_t(S::template Inside<double>(s, p)), o.inside + i);
    if (o.sin) __builtin_nontemporal_store(S::template SafeIn<double>(s, p), o.sin + i);
    if (o.sout) __builtin_nontemporal_store(S::template SafeOut<double>(s, p), o.sout + i);
    if (o.din || o.dout) {
      Vec3<double> d(__builtin_nontemporal_load(o.dx + i), __builtin_nontemporal_load(o.dy + i),
                     __builtin_nontemporal_load(o.dz + i));
      if (o.din) __builtin_nontemporal_store(S::template DistIn<double>(s, p, d), o.din + i);
      if (o.dout) __builtin_nontemporal_store(S::template DistOut<double>(s, p, d), o.dout + i);
    }
  }
}

// The mixed-precision tile: the per-point answers of the single-shape mixed kernels, with the
// point's FP32 classification computed once for all methods of the tile.
template <class Data>
VGMI_DEV void run_tile_mixed(Data const &s, int64_t b, int64_t e, BatchOut const &o, int nthreads)
{
  const bool point = o.contains || o.inside || o.sin || o.sout;
  const bool dirs  = o.din || o.dout;
  for (int64_t i = b + threadIdx.x; i < e; i += nthreads) {
    Vec3<double> p(__builtin_nontemporal_load(o.px + i), __builtin_nontemporal_load(o.py + i),
                   __builtin_nontemporal_load(o.pz + i));
    Vec3<double> d(0., 0., 0.);
    if (dirs)
      d = Vec3<double>(__builtin_nontemporal_load(o.dx + i), __builtin_nontemporal_load(o.dy + i),
                       __builtin_nontemporal_load(o.dz + i));
    MixedState m{};
    if (point || (dirs && CertifiedF32Dist<Data>::value)) m = mixed_state(s, p);
    if (o.contains)
      __builtin_nontemporal_store(uint8_t(mixed_answer<Data, kOpContains>(s, m, p, d) != 0.), o.contains + i);
    if (o.inside) __builtin_nontemporal_store(int32_t(mixed_answer<Data, kOpInside>(s, m, p, d)), o.inside + i);
    if (o.sin) __builtin_nontemporal_store(mixed_answer<Data, kOpSafetyIn>(s, m, p, d), o.sin + i);
    if (o.sout) __builtin_nontemporal_store(mixed_answer<Data, kOpSafetyOut>(s, m, p, d), o.sout + i);
    if (o.din) __builtin_nontemporal_store(mixed_answer<Data, kOpDistIn>(s, m, p, d), o.din + i);
    if (o.dout) __builtin_nontemporal_store(mixed_answer<Data, kOpDistOut>(s, m, p, d), o.dout + i);
  }
}

// A tessellated tile. When the solid's image (nodes, planes, candidate boxes) fits in the
// launch's LDS (ldsDoubles), the workgroup copies it there once per solid (consecutive tiles of
// one solid reuse the copy; `staged` names the solid held) and the walks read it from LDS. The
// results are the pool walk's bit for bit.
template <int BLOCK>
VGMI_DEV void tess_tile(TessellatedData const &s, int si, Tile const &tile, BatchOut const &o, double *lds,
                        int ldsDoubles, int *staged, TessQueue *q)
{
  const int n = s.imageDoubles;
  if (n == 0 || n > ldsDoubles) {
    run_tile_tess(s, tess::pool_walk(s), tile.begin, tile.end, o, BLOCK, q);
    return;
  }
  if (*staged != si) { // block-uniform: written after the barrier that ends the copy
    stage_tess_image(s, lds, BLOCK);
    if (threadIdx.x == 0) *staged = si;
  }
  run_tile_tess(s, tess::lds_walk(s, lds), tile.begin, tile.end, o, BLOCK, q);
}

__global__ void k_batch_reset(uint32_t *counter) { *counter = 0u; }

// At least two wavefronts per SIMD: the register budget then counts accumulation registers too
// (without it the compiler may take 256 vector registers plus one accumulation register, which
// leaves room for a single wavefront).
// MIX selects the mixed-precision answers; the tessellated solid answers in FP64 either way (its
// single-shape mixed query does the same).
template <int BLOCK, bool MIX>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(2))) void k_batch(const ShapeRecord *__restrict__ recs, const Tile *__restrict__ tiles,
                                                 int ntiles, uint32_t *counter, BatchOut o, int ldsDoubles)
{
  extern __shared__ __align__(16) double lds[];
  __shared__ int s_tile, s_staged;
  __shared__ TessQueue s_queue;
  if (threadIdx.x == 0) s_staged = -1;
  while (true) {
    if (threadIdx.x == 0) {
      s_tile       = int(atomicAdd(counter, 1u));
      s_queue.n    = 0;
    }
    __syncthreads();
    const int t = __builtin_amdgcn_readfirstlane(s_tile);
    __syncthreads();
    if (t >= ntiles) break; // every wave of every block reaches this exit
    const Tile tile        = tiles[t];
    const int si           = __builtin_amdgcn_readfirstlane(tile.shape);
    const BatchOut mo      = mask_outputs(o, __builtin_amdgcn_readfirstlane(tile.mask));
    if (!mo.contains && !mo.inside && !mo.din && !mo.dout && !mo.sin && !mo.sout) continue;
    ShapeRecord const &rec = recs[si];
    if (rec.type == kTessellated) {
      tess_tile<BLOCK>(rec.u.tess, si, tile, mo, lds, ldsDoubles, &s_staged, &s_queue);
      continue;
    }
    switch (rec.type) {
#define X(E, T, M, NS)                                                                                               \
  case E:                                                                                                            \
    if constexpr (MIX)                                                                                               \
      run_tile_mixed<T>(rec.u.M, tile.begin, tile.end, mo, BLOCK);                                                  \
    else                                                                                                             \
      run_tile<T>(rec.u.M, tile.begin, tile.end, mo, BLOCK);                                                        \
    break;
      VGMI_FOR_EACH_SHAPE_BUT_TESSELLATED(X)
#undef X
    default:
      break;
    }
  }
}

// Method groups of a shape type and the relative cost of one point of each (sort key of the tile
// queue: the most expensive groups start first, longest-processing-time order). The costs are the
// per-point times of the groups measured on one MI355X, in units of a box point.
struct MethodGroup {
  int mask, cost;
};

static int method_groups(int type, MethodGroup g[3])
{
  switch (type) {
  case kBox:
  case kOrb:
    g[0] = {kMAll, 1};
    return 1;
  case kTrapezoid:
  case kTube:
  case kTrd:
  case kParaboloid:
  case kParallelepiped:
  case kTet:
  case kEllipticalTube:
  case kEllipsoid:
  case kGenTrap:
  case kCutTube:
  case kEllipticalCone:
    g[0] = {kMAll, 2};
    return 1;
  case kCone:
  case kHype:
  case kSphere:
    g[0] = {kMAll, 4};
    return 1;
  case kTorus2:
    g[0] = {kMAll, 8};
    return 1;
  case kPolycone:
  case kPolyhedron:
    g[0] = {kMDistIn, 12};
    g[1] = {kMDistOut, 10};
    g[2] = {kMPoint, 10};
    return 3;
  case kTessellated:
    g[0] = {kMPoint, 64};
    g[1] = {kMDistIn | kMDistOut, 24};
    return 2;
  default:
    g[0] = {kMAll, 16};
    return 1;
  }
}

} // namespace vgmi

using namespace vgmi;

struct vgmi_batch {
  std::vector<ShapeRecord> recs;
  std::vector<int64_t> counts, offsets;
  int64_t total = 0;
  int tile = 4096, block = 256, blocksPerCU = 8, numCU = 256;
  ShapeRecord *d_recs = nullptr;
  Tile *d_tiles = nullptr;
  uint32_t *d_counter = nullptr;
  int ntiles = 0;
  int ldsBytes = 0; // dynamic LDS per workgroup: the largest tessellated solid that fits kTessLdsMax
  hipGraph_t graph = nullptr;
  hipGraphExec_t graphExec = nullptr;
};

static int tess_lds_bytes(TessellatedData const &t) { return int(sizeof(double)) * t.imageDoubles; }

// Default points per tile: at least four tiles per resident block so the atomic queue can
// balance the expensive shapes, within [256, 4096] and a multiple of 64.
static int default_tile(vgmi_batch const *b)
{
  const int64_t blocks = int64_t(b->numCU) * b->blocksPerCU;
  int64_t t            = b->total / std::max<int64_t>(1, 4 * blocks);
  t                    = std::max<int64_t>(256, std::min<int64_t>(4096, t));
  return int(t / 64 * 64);
}

static vgmi_status batch_build_tiles(vgmi_batch *b)
{
  std::vector<Tile> tiles;
  std::vector<int> cost;
  for (size_t s = 0; s < b->recs.size(); ++s) {
    MethodGroup g[3];
    const int ng = method_groups(b->recs[s].type, g);
    for (int k = 0; k < ng; ++k)
      for (int64_t beg = b->offsets[s]; beg < b->offsets[s] + b->counts[s]; beg += b->tile) {
        tiles.push_back(Tile{int32_t(s), g[k].mask, beg, std::min(beg + b->tile, b->offsets[s] + b->counts[s])});
        cost.push_back(g[k].cost);
      }
  }
  std::vector<int> order(tiles.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = int(i);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[x] > cost[y]; });
  std::vector<Tile> sorted(tiles.size());
  for (size_t i = 0; i < order.size(); ++i) sorted[i] = tiles[order[i]];
  tiles.swap(sorted);
  if (b->d_tiles) (void)hipFree(b->d_tiles);
  b->d_tiles = nullptr;
  b->ntiles  = int(tiles.size());
  if (tiles.empty()) return VGMI_OK;
  hipError_t e = hipMalloc(&b->d_tiles, tiles.size() * sizeof(Tile));
  if (e == hipSuccess) e = hipMemcpy(b->d_tiles, tiles.data(), tiles.size() * sizeof(Tile), hipMemcpyHostToDevice);
  if (e != hipSuccess) return set_error(VGMI_ERR_DEVICE, "vgmi_batch: tile table upload: %s", hipGetErrorString(e));
  return VGMI_OK;
}

static void batch_drop_graph(vgmi_batch *b)
{
  if (b->graphExec) (void)hipGraphExecDestroy(b->graphExec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->graphExec = nullptr;
  b->graph     = nullptr;
}

template <bool MIX>
static void launch_k_batch(vgmi_batch const *b, BatchOut const &o, hipStream_t st, int grid, size_t lds, int ldsD)
{
  switch (b->block) {
  case 64:
    hipLaunchKernelGGL((k_batch<64, MIX>), dim3(grid), dim3(64), lds, st, b->d_recs, b->d_tiles, b->ntiles,
                       b->d_counter, o, ldsD);
    break;
  case 128:
    hipLaunchKernelGGL((k_batch<128, MIX>), dim3(grid), dim3(128), lds, st, b->d_recs, b->d_tiles, b->ntiles,
                       b->d_counter, o, ldsD);
    break;
  case 512:
    hipLaunchKernelGGL((k_batch<512, MIX>), dim3(grid), dim3(512), lds, st, b->d_recs, b->d_tiles, b->ntiles,
                       b->d_counter, o, ldsD);
    break;
  default:
    hipLaunchKernelGGL((k_batch<256, MIX>), dim3(grid), dim3(256), lds, st, b->d_recs, b->d_tiles, b->ntiles,
                       b->d_counter, o, ldsD);
    break;
  }
}

static vgmi_status batch_launch(vgmi_batch *b, BatchOut const &o, bool mixed, hipStream_t st)
{
  if (b->ntiles == 0) return VGMI_OK;
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(1), 0, st, b->d_counter);
  // resident workgroups per CU are bounded by the dynamic LDS too
  const int perCU =
      b->ldsBytes ? std::max(1, std::min(b->blocksPerCU, kLdsPerCU / (b->ldsBytes + kStaticLds))) : b->blocksPerCU;
  const int grid  = std::max(1, std::min(b->ntiles, b->numCU * perCU));
  const size_t lds = size_t(b->ldsBytes);
  const int ldsD   = b->ldsBytes / int(sizeof(double));
  if (mixed)
    launch_k_batch<true>(b, o, st, grid, lds, ldsD);
  else
    launch_k_batch<false>(b, o, st, grid, lds, ldsD);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(VGMI_ERR_DEVICE, "vgmi_batch launch: %s", hipGetErrorString(e));
  return VGMI_OK;
}

static vgmi_status batch_outputs(vgmi_batch *b, int precision, const double *px, const double *py, const double *pz,
                                 const double *dx, const double *dy, const double *dz, uint8_t *contains,
                                 int32_t *inside, double *din, double *dout, double *sin, double *sout, BatchOut &o)
{
  if (!b) return set_error(VGMI_ERR_INVALID, "vgmi_batch: null batch");
  if (precision != VGMI_FP64 && precision != VGMI_MIXED)
    return set_error(VGMI_ERR_INVALID, "vgmi_batch: precision must be VGMI_FP64 or VGMI_MIXED");
  if (b->total > 0 && (!px || !py || !pz)) return set_error(VGMI_ERR_INVALID, "vgmi_batch: null point pointer");
  if ((din || dout) && (!dx || !dy || !dz)) return set_error(VGMI_ERR_INVALID, "vgmi_batch: distances need directions");
  o = BatchOut{px, py, pz, dx, dy, dz, contains, inside, din, dout, sin, sout};
  return VGMI_OK;
}

extern "C" {

vgmi_status vgmi_batch_create(const vgmi_shape *const *shapes, const int64_t *counts, int nshapes, vgmi_batch **out)
{
  if (!out) return set_error(VGMI_ERR_INVALID, "vgmi_batch_create: null output handle");
  *out = nullptr;
  if (nshapes <= 0 || !shapes || !counts) return set_error(VGMI_ERR_INVALID, "vgmi_batch_create: empty batch");
  auto *b = new vgmi_batch();
  for (int s = 0; s < nshapes; ++s) {
    if (!shapes[s] || counts[s] < 0) {
      delete b;
      return set_error(VGMI_ERR_INVALID, "vgmi_batch_create: bad shape or count at %d", s);
    }
    b->recs.push_back(shapes[s]->rec);
    if (b->recs.back().type == kTessellated && counts[s] > 0) {
      const int need = tess_lds_bytes(b->recs.back().u.tess);
      if (need > 0 && need <= tess_lds_limit()) b->ldsBytes = std::max(b->ldsBytes, need);
    }
    b->offsets.push_back(b->total);
    b->counts.push_back(counts[s]);
    b->total += counts[s];
  }
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
    b->numCU = prop.multiProcessorCount;
  hipError_t e = hipMalloc(&b->d_recs, b->recs.size() * sizeof(ShapeRecord));
  if (e == hipSuccess)
    e = hipMemcpy(b->d_recs, b->recs.data(), b->recs.size() * sizeof(ShapeRecord), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc(&b->d_counter, 256);
  if (e != hipSuccess) {
    vgmi_batch_destroy(b);
    return set_error(VGMI_ERR_DEVICE, "vgmi_batch_create: %s", hipGetErrorString(e));
  }
  b->tile = default_tile(b);
  if (vgmi_status s = batch_build_tiles(b)) {
    vgmi_batch_destroy(b);
    return s;
  }
  *out = b;
  return VGMI_OK;
}

vgmi_status vgmi_batch_destroy(vgmi_batch *b)
{
  if (!b) return VGMI_OK;
  batch_drop_graph(b);
  if (b->d_recs) (void)hipFree(b->d_recs);
  if (b->d_tiles) (void)hipFree(b->d_tiles);
  if (b->d_counter) (void)hipFree(b->d_counter);
  delete b;
  return VGMI_OK;
}

vgmi_status vgmi_batch_set_tuning(vgmi_batch *b, int tile_points, int block, int blocks_per_cu)
{
  if (!b) return set_error(VGMI_ERR_INVALID, "vgmi_batch_set_tuning: null batch");
  if (tile_points < 64 || tile_points > (1 << 24) || (block != 64 && block != 128 && block != 256 && block != 512) ||
      blocks_per_cu < 1 || blocks_per_cu > 64)
    return set_error(VGMI_ERR_INVALID,
                     "vgmi_batch_set_tuning: tile in [64, 2^24], block in {64,128,256,512}, 1..64 blocks/CU");
  b->tile        = tile_points;
  b->block       = block;
  b->blocksPerCU = blocks_per_cu;
  batch_drop_graph(b);
  return batch_build_tiles(b);
}

vgmi_status vgmi_batch_info(const vgmi_batch *b, int64_t info[6])
{
  if (!b || !info) return set_error(VGMI_ERR_INVALID, "vgmi_batch_info: null argument");
  info[0] = int64_t(b->recs.size());
  info[1] = b->total;
  info[2] = b->ntiles;
  info[3] = b->tile;
  info[4] = b->block;
  info[5] = b->blocksPerCU;
  return VGMI_OK;
}

vgmi_status vgmi_batch_query_all(vgmi_batch *b, int precision, const double *px, const double *py, const double *pz,
                                 const double *dx, const double *dy, const double *dz, uint8_t *contains,
                                 int32_t *inside, double *dist_in, double *dist_out, double *safety_in,
                                 double *safety_out, void *stream)
{
  BatchOut o;
  if (vgmi_status s = batch_outputs(b, precision, px, py, pz, dx, dy, dz, contains, inside, dist_in, dist_out,
                                    safety_in, safety_out, o))
    return s;
  return batch_launch(b, o, precision == VGMI_MIXED, (hipStream_t)stream);
}

vgmi_status vgmi_batch_graph_capture(vgmi_batch *b, int precision, const double *px, const double *py,
                                     const double *pz, const double *dx, const double *dy, const double *dz,
                                     uint8_t *contains, int32_t *inside, double *dist_in, double *dist_out,
                                     double *safety_in, double *safety_out, void *stream)
{
  BatchOut o;
  if (vgmi_status s = batch_outputs(b, precision, px, py, pz, dx, dy, dz, contains, inside, dist_in, dist_out,
                                    safety_in, safety_out, o))
    return s;
  batch_drop_graph(b);
  hipStream_t cap = nullptr;
  hipError_t e    = hipStreamCreateWithFlags(&cap, hipStreamNonBlocking);
  if (e != hipSuccess) return set_error(VGMI_ERR_DEVICE, "vgmi_batch_graph_capture: stream: %s", hipGetErrorString(e));
  e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
  vgmi_status ls = VGMI_OK;
  if (e == hipSuccess) {
    ls = batch_launch(b, o, precision == VGMI_MIXED, cap);
    hipError_t e2 = hipStreamEndCapture(cap, &b->graph);
    if (e == hipSuccess) e = e2;
  }
  if (e == hipSuccess && ls == VGMI_OK) e = hipGraphInstantiate(&b->graphExec, b->graph, nullptr, nullptr, 0);
  (void)hipStreamDestroy(cap);
  (void)stream;
  if (ls != VGMI_OK) return ls;
  if (e != hipSuccess) {
    batch_drop_graph(b);
    return set_error(VGMI_ERR_DEVICE, "vgmi_batch_graph_capture: %s", hipGetErrorString(e));
  }
  return VGMI_OK;
}

vgmi_status vgmi_batch_graph_launch(vgmi_batch *b, void *stream)
{
  if (!b) return set_error(VGMI_ERR_INVALID, "vgmi_batch_graph_launch: null batch");
  if (!b->graphExec) return set_error(VGMI_ERR_INVALID, "vgmi_batch_graph_launch: no captured graph");
  hipError_t e = hipGraphLaunch(b->graphExec, (hipStream_t)stream);
  if (e != hipSuccess) return set_error(VGMI_ERR_DEVICE, "vgmi_batch_graph_launch: %s", hipGetErrorString(e));
  return VGMI_OK;
}

} // extern "C"
