// Single-shape SOA query kernels and the C ABI entry points for them.
//
// One launch streams one shape over n points: each lane reads its point (and
// direction) from the SOA arrays with unit-stride, fully coalesced 8-byte loads,
// the shape record arrives by value in the kernel arguments (scalar registers,
// no per-lane parameter traffic), and results are written back coalesced.
// The FP64 path is the reference's scalar-backend arithmetic; the MIXED path
// runs the same kernel in FP32 and re-evaluates every point whose FP32 result
// is within the near-surface band in FP64, so decisions stay exact.
#include "vgmi_internal.h"
#include "dev_tess_tile.h"
#include <algorithm>
#include <atomic>

namespace vgmi {

constexpr int kBlock = 256;

template <class R>
VGMI_DEV Vec3<R> load3(const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ z,
                       size_t i)
{
  return Vec3<R>(R(__builtin_nontemporal_load(x + i)), R(__builtin_nontemporal_load(y + i)),
                 R(__builtin_nontemporal_load(z + i)));
}

struct QueryArgs {
  const double *px, *py, *pz, *dx, *dy, *dz;
  void *out;
  size_t n;
};

template <class Data, int OP>
__global__ __launch_bounds__(kBlock) void k_query_fp64(Data s, QueryArgs a)
{
  using S      = ShapeOps<Data>;
  size_t i     = size_t(blockIdx.x) * kBlock + threadIdx.x;
  size_t strid = size_t(gridDim.x) * kBlock;
  for (; i < a.n; i += strid) {
    Vec3<double> p = load3<double>(a.px, a.py, a.pz, i);
    if constexpr (OP == kOpContains) {
      __builtin_nontemporal_store(uint8_t(S::template Contains<double>(s, p)), (uint8_t *)a.out + i);
    } else if constexpr (OP == kOpInside) {
      __builtin_nontemporal_store(int32_t(S::template Inside<double>(s, p)), (int32_t *)a.out + i);
    } else if constexpr (OP == kOpDistIn || OP == kOpDistOut) {
      Vec3<double> d = load3<double>(a.dx, a.dy, a.dz, i);
      double r       = OP == kOpDistIn ? S::template DistIn<double>(s, p, d) : S::template DistOut<double>(s, p, d);
      __builtin_nontemporal_store(r, (double *)a.out + i);
    } else if constexpr (OP == kOpSafetyIn) {
      __builtin_nontemporal_store(S::template SafeIn<double>(s, p), (double *)a.out + i);
    } else {
      __builtin_nontemporal_store(S::template SafeOut<double>(s, p), (double *)a.out + i);
    }
  }
}

struct QueryAllArgs {
  const double *px, *py, *pz, *dx, *dy, *dz;
  uint8_t *contains;
  int32_t *inside;
  double *din, *dout, *sin, *sout;
  size_t n;
};

template <class Data>
__global__ __launch_bounds__(kBlock) void k_query_all_fp64(Data s, QueryAllArgs a)
{
  using S      = ShapeOps<Data>;
  size_t i     = size_t(blockIdx.x) * kBlock + threadIdx.x;
  size_t strid = size_t(gridDim.x) * kBlock;
  for (; i < a.n; i += strid) {
    Vec3<double> p = load3<double>(a.px, a.py, a.pz, i);
    if (a.contains) __builtin_nontemporal_store(uint8_t(S::template Contains<double>(s, p)), a.contains + i);
    if (a.inside) __builtin_nontemporal_store(int32_t(S::template Inside<double>(s, p)), a.inside + i);
    if (a.sin) __builtin_nontemporal_store(S::template SafeIn<double>(s, p), a.sin + i);
    if (a.sout) __builtin_nontemporal_store(S::template SafeOut<double>(s, p), a.sout + i);
    if (a.din || a.dout) {
      Vec3<double> d = load3<double>(a.dx, a.dy, a.dz, i);
      if (a.din) __builtin_nontemporal_store(S::template DistIn<double>(s, p, d), a.din + i);
      if (a.dout) __builtin_nontemporal_store(S::template DistOut<double>(s, p, d), a.dout + i);
    }
  }
}

// A tessellated solid: tiles of kBlock points, the solid's image in LDS when it fits (copied once
// per workgroup), one test-ray walk, one safety walk and one walk along d shared by the requested
// methods, deep points' safety searches run by whole wavefronts (the batched launch's tile code).
__global__ __launch_bounds__(kBlock) void k_tess_query(TessellatedData s, BatchOut o, int64_t n, int useLds)
{
  extern __shared__ __align__(16) double lds[];
  __shared__ TessQueue q;
  if (useLds) stage_tess_image(s, lds, kBlock);
  for (int64_t t = blockIdx.x; t * kBlock < n; t += gridDim.x) {
    if (threadIdx.x == 0) q.n = 0;
    __syncthreads();
    const int64_t b = t * kBlock, e = b + kBlock < n ? b + kBlock : n;
    if (useLds)
      run_tile_tess(s, tess::lds_walk(s, lds), b, e, o, kBlock, &q);
    else
      run_tile_tess(s, tess::pool_walk(s), b, e, o, kBlock, &q);
    __syncthreads();
  }
}

// A tessellated solid's DistanceToIn / DistanceToOut with a step limit per point
// (TessellatedImplementation.h:191-199, 213-216, 250): one lane per point over the pool.
__global__ __launch_bounds__(kBlock) void k_tess_dist_step(TessellatedData s, QueryArgs a,
                                                           const double *__restrict__ step_max, int toIn)
{
  size_t i     = size_t(blockIdx.x) * kBlock + threadIdx.x;
  size_t strid = size_t(gridDim.x) * kBlock;
  for (; i < a.n; i += strid) {
    const Vec3<double> p = load3<double>(a.px, a.py, a.pz, i), d = load3<double>(a.dx, a.dy, a.dz, i);
    const tess::V3d pp(p.x, p.y, p.z), dd(d.x, d.y, d.z);
    const double sm = __builtin_nontemporal_load(step_max + i);
    const double r  = toIn ? tess::dist_in_step64(s, tess::pool_walk(s), pp, dd, sm)
                           : tess::dist_out_step64(s, tess::pool_walk(s), pp, dd, sm);
    __builtin_nontemporal_store(r, (double *)a.out + i);
  }
}

// A planar GenTrap's DistanceToIn with a step limit per point (TessellatedSection::DistanceToIn<false>
// bounds its search by min(z exit, stepMax)).
__global__ __launch_bounds__(kBlock) void k_gentrap_dist_in_step(GenTrapData s, QueryArgs a,
                                                                 const double *__restrict__ step_max)
{
  size_t i     = size_t(blockIdx.x) * kBlock + threadIdx.x;
  size_t strid = size_t(gridDim.x) * kBlock;
  for (; i < a.n; i += strid) {
    const Vec3<double> p = load3<double>(a.px, a.py, a.pz, i), d = load3<double>(a.dx, a.dy, a.dz, i);
    const double r = gtra::DistanceToInStep(s, p, d, __builtin_nontemporal_load(step_max + i));
    __builtin_nontemporal_store(r, (double *)a.out + i);
  }
}

// A CutTube's DistanceToIn with a step limit per point (CutTubeImplementation.h:176-243 compare the
// plane and tube crossings with stepMax).
__global__ __launch_bounds__(kBlock) void k_cuttube_dist_in_step(CutTubeData s, QueryArgs a,
                                                                 const double *__restrict__ step_max)
{
  size_t i     = size_t(blockIdx.x) * kBlock + threadIdx.x;
  size_t strid = size_t(gridDim.x) * kBlock;
  for (; i < a.n; i += strid) {
    const Vec3<double> p = load3<double>(a.px, a.py, a.pz, i), d = load3<double>(a.dx, a.dy, a.dz, i);
    const double r = ctub::DistanceToInStep(s, p, d, __builtin_nontemporal_load(step_max + i));
    __builtin_nontemporal_store(r, (double *)a.out + i);
  }
}

int tess_lds_limit()
{
  static std::atomic<int> cache[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return kTessLdsMax;
  int v = cache[dev].load(std::memory_order_relaxed);
  if (v > 0) return v;
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess || lds <= 0)
    lds = kLdsPerCU;
  v = std::min(kTessLdsMax, lds / 2 - kStaticLds);
  v = v > 1 ? v : 1;
  cache[dev].store(v, std::memory_order_relaxed);
  return v;
}

static hipError_t launch_tess(TessellatedData const &s, BatchOut const &o, size_t n, hipStream_t st)
{
  const bool useLds = s.imageDoubles > 0 && s.imageDoubles * int64_t(sizeof(double)) <= tess_lds_limit();
  int dev = 0, ncu = 256;
  if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  const size_t tiles = (n + kBlock - 1) / kBlock, cap = size_t(ncu) * (useLds ? 2 : 8);
  const size_t lds   = useLds ? size_t(s.imageDoubles) * sizeof(double) : 0;
  hipLaunchKernelGGL(k_tess_query, dim3(unsigned(tiles < cap ? tiles : cap)), dim3(kBlock), lds, st, s, o, int64_t(n),
                     int(useLds));
  return hipGetLastError();
}

static unsigned grid_for(size_t n)
{
  size_t blocks = (n + kBlock - 1) / kBlock;
  // grid-stride beyond 64 waves per CU: enough to cover latency on 256 CUs
  const size_t cap = size_t(256) * 64;
  return unsigned(blocks < cap ? blocks : cap);
}

template <class Data>
static hipError_t launch_query(Data const &s, int op, QueryArgs const &a, hipStream_t st)
{
  dim3 g(grid_for(a.n)), b(kBlock);
  switch (op) {
  case kOpContains:
    hipLaunchKernelGGL((k_query_fp64<Data, kOpContains>), g, b, 0, st, s, a);
    break;
  case kOpInside:
    hipLaunchKernelGGL((k_query_fp64<Data, kOpInside>), g, b, 0, st, s, a);
    break;
  case kOpDistIn:
    hipLaunchKernelGGL((k_query_fp64<Data, kOpDistIn>), g, b, 0, st, s, a);
    break;
  case kOpDistOut:
    hipLaunchKernelGGL((k_query_fp64<Data, kOpDistOut>), g, b, 0, st, s, a);
    break;
  case kOpSafetyIn:
    hipLaunchKernelGGL((k_query_fp64<Data, kOpSafetyIn>), g, b, 0, st, s, a);
    break;
  default:
    hipLaunchKernelGGL((k_query_fp64<Data, kOpSafetyOut>), g, b, 0, st, s, a);
    break;
  }
  return hipGetLastError();
}

} // namespace vgmi

using namespace vgmi;

extern "C" {

vgmi_status vgmi_shape_query(const vgmi_shape *shape, int op, int precision, const double *px, const double *py,
                             const double *pz, const double *dx, const double *dy, const double *dz,
                             const double *step_max, void *out, size_t n, void *stream)
{
  if (!shape) return set_error(VGMI_ERR_INVALID, "vgmi_shape_query: null shape");
  if (op < 0 || op >= kNumOps) return set_error(VGMI_ERR_INVALID, "vgmi_shape_query: unknown op %d", op);
  if (precision != VGMI_FP64 && precision != VGMI_MIXED)
    return set_error(VGMI_ERR_INVALID, "vgmi_shape_query: unknown precision %d", precision);
  if (n == 0) return VGMI_OK;
  if (!px || !py || !pz || !out) return set_error(VGMI_ERR_INVALID, "vgmi_shape_query: null point/output pointer");
  if ((op == kOpDistIn || op == kOpDistOut) && (!dx || !dy || !dz))
    return set_error(VGMI_ERR_INVALID, "vgmi_shape_query: DistanceToIn/Out need directions");
  // step_max: of the supported shapes only the tessellated solid and the GenTrap's and CutTube's
  // DistanceToIn use it
  // (the reference's other kernels take the argument and ignore it); they answer every point in FP64
  if (step_max && shape->rec.type == kTessellated && (op == kOpDistIn || op == kOpDistOut)) {
    QueryArgs a{px, py, pz, dx, dy, dz, out, n};
    hipLaunchKernelGGL(k_tess_dist_step, dim3(grid_for(n)), dim3(kBlock), 0, (hipStream_t)stream,
                       shape->rec.u.tess, a, step_max, int(op == kOpDistIn));
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(VGMI_ERR_DEVICE, "vgmi_shape_query launch: %s", hipGetErrorString(e));
    return VGMI_OK;
  }
  if (step_max && shape->rec.type == kGenTrap && op == kOpDistIn) {
    QueryArgs a{px, py, pz, dx, dy, dz, out, n};
    hipLaunchKernelGGL(k_gentrap_dist_in_step, dim3(grid_for(n)), dim3(kBlock), 0, (hipStream_t)stream,
                       shape->rec.u.gtra, a, step_max);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(VGMI_ERR_DEVICE, "vgmi_shape_query launch: %s", hipGetErrorString(e));
    return VGMI_OK;
  }
  if (step_max && shape->rec.type == kCutTube && op == kOpDistIn) {
    QueryArgs a{px, py, pz, dx, dy, dz, out, n};
    hipLaunchKernelGGL(k_cuttube_dist_in_step, dim3(grid_for(n)), dim3(kBlock), 0, (hipStream_t)stream,
                       shape->rec.u.ctub, a, step_max);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(VGMI_ERR_DEVICE, "vgmi_shape_query launch: %s", hipGetErrorString(e));
    return VGMI_OK;
  }
  // the tessellated solid has no FP32 pass (it answers every point in FP64), so MIXED takes the FP64 launch
  if (precision == VGMI_MIXED && shape->rec.type != kTessellated)
    return mixed_shape_query(shape, op, px, py, pz, dx, dy, dz, out, n, stream);
  QueryArgs a{px, py, pz, dx, dy, dz, out, n};
  hipStream_t st = (hipStream_t)stream;
  hipError_t e;
  switch (shape->rec.type) {
  case kTessellated: {
    BatchOut o{px, py, pz, dx, dy, dz, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    switch (op) {
    case kOpContains: o.contains = (uint8_t *)out; break;
    case kOpInside: o.inside = (int32_t *)out; break;
    case kOpDistIn: o.din = (double *)out; break;
    case kOpDistOut: o.dout = (double *)out; break;
    case kOpSafetyIn: o.sin = (double *)out; break;
    default: o.sout = (double *)out; break;
    }
    e = launch_tess(shape->rec.u.tess, o, n, st);
    break;
  }
#define X(E, T, M, NS)                                                                                               \
  case E:                                                                                                            \
    e = launch_query(shape->rec.u.M, op, a, st);                                                                     \
    break;
    VGMI_FOR_EACH_SHAPE_BUT_TESSELLATED(X)
#undef X
  default:
    return set_error(VGMI_ERR_UNSUPPORTED, "vgmi_shape_query: shape type %d has no kernel", shape->rec.type);
  }
  if (e != hipSuccess) return set_error(VGMI_ERR_DEVICE, "vgmi_shape_query launch: %s", hipGetErrorString(e));
  return VGMI_OK;
}

vgmi_status vgmi_shape_query_all(const vgmi_shape *shape, int precision, const double *px, const double *py,
                                 const double *pz, const double *dx, const double *dy, const double *dz,
                                 uint8_t *contains, int32_t *inside, double *dist_in, double *dist_out,
                                 double *safety_in, double *safety_out, size_t n, void *stream)
{
  if (!shape) return set_error(VGMI_ERR_INVALID, "vgmi_shape_query_all: null shape");
  if (precision != VGMI_FP64)
    return set_error(VGMI_ERR_UNSUPPORTED, "vgmi_shape_query_all: only VGMI_FP64 is implemented");
  if (n == 0) return VGMI_OK;
  if (!px || !py || !pz) return set_error(VGMI_ERR_INVALID, "vgmi_shape_query_all: null point pointer");
  if ((dist_in || dist_out) && (!dx || !dy || !dz))
    return set_error(VGMI_ERR_INVALID, "vgmi_shape_query_all: distances need directions");
  QueryAllArgs a{px, py, pz, dx, dy, dz, contains, inside, dist_in, dist_out, safety_in, safety_out, n};
  hipStream_t st = (hipStream_t)stream;
  dim3 g(grid_for(n)), b(kBlock);
  switch (shape->rec.type) {
  case kTessellated: {
    const BatchOut o{px, py, pz, dx, dy, dz, contains, inside, dist_in, dist_out, safety_in, safety_out};
    const hipError_t e = launch_tess(shape->rec.u.tess, o, n, st);
    if (e != hipSuccess)
      return set_error(VGMI_ERR_DEVICE, "vgmi_shape_query_all launch: %s", hipGetErrorString(e));
    return VGMI_OK;
  }
#define X(E, T, M, NS)                                                                                               \
  case E:                                                                                                            \
    hipLaunchKernelGGL((k_query_all_fp64<T>), g, b, 0, st, shape->rec.u.M, a);                                      \
    break;
    VGMI_FOR_EACH_SHAPE_BUT_TESSELLATED(X)
#undef X
  default:
    return set_error(VGMI_ERR_UNSUPPORTED, "vgmi_shape_query_all: shape type %d has no kernel", shape->rec.type);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(VGMI_ERR_DEVICE, "vgmi_shape_query_all launch: %s", hipGetErrorString(e));
  return VGMI_OK;
}

} // extern "C"
