// Mixed-precision single-shape queries: FP32 bulk arithmetic with FP64 re-evaluation
// of every point within the near-surface band.
//
// Each lane converts its point (and direction) to FP32 once and evaluates the
// method and the two safeties in FP32. A point whose FP32 safeties put it within
// kBand of a surface is evaluated again by the FP64 kernel, so every decision
// that depends on the tolerance bands (Inside/Contains, on-surface conventions,
// wrong-side -1) is the FP64 one; elsewhere the FP32 rounding error (about 1e-7
// of the local coordinates) is far below the distance to any surface, so
// Inside/Contains cannot flip and distances carry only the FP32 error. An FP32
// kernel that cannot decide within FP32 rounding (a polycone section hand-off)
// answers NaN, and that point is evaluated in FP64 as well.
#include "dev_shapes.h"
#include "vgmi_internal.h"

namespace vgmi {

constexpr int kMixBlock  = 256;
constexpr float kBand    = float(kMixedBand);
constexpr float kFltBig  = 1e30f;

template <class Data, int OP>
__global__ __launch_bounds__(kMixBlock) void k_query_mixed(Data s, const double *__restrict__ px,
                                                           const double *__restrict__ py,
                                                           const double *__restrict__ pz,
                                                           const double *__restrict__ dx,
                                                           const double *__restrict__ dy,
                                                           const double *__restrict__ dz, void *out, size_t n)
{
  using S        = ShapeOps<Data>;
  const size_t i = size_t(blockIdx.x) * kMixBlock + threadIdx.x;
  if (i >= n) return;
  const Vec3<double> p(__builtin_nontemporal_load(px + i), __builtin_nontemporal_load(py + i),
                       __builtin_nontemporal_load(pz + i));
  const Vec3<float> pf(float(p.x), float(p.y), float(p.z));
  const float sin32  = S::template SafeIn<float>(s, pf);
  const float sout32 = S::template SafeOut<float>(s, pf);
  // A consistent FP32 state has safeties of opposite sign (inside: SafetyToIn < 0 < SafetyToOut,
  // outside: the reverse). Equal signs mean the FP32 tolerance bands collapsed onto the surface
  // (a point within FP32 rounding of it is then both inside and outside): treat as near.
  const bool near = vabs(sin32) < kBand || vabs(sout32) < kBand || ((sin32 < 0.f) == (sout32 < 0.f));
  if constexpr (OP == kOpContains) {
    const bool r = near ? S::template Contains<double>(s, p) : S::template Contains<float>(s, pf);
    __builtin_nontemporal_store(uint8_t(r), (uint8_t *)out + i);
  } else if constexpr (OP == kOpInside) {
    const int r = near ? S::template Inside<double>(s, p) : S::template Inside<float>(s, pf);
    __builtin_nontemporal_store(int32_t(r), (int32_t *)out + i);
  } else if constexpr (OP == kOpSafetyIn) {
    const double r = near ? S::template SafeIn<double>(s, p) : double(sin32);
    __builtin_nontemporal_store(r, (double *)out + i);
  } else if constexpr (OP == kOpSafetyOut) {
    const double r = near ? S::template SafeOut<double>(s, p) : double(sout32);
    __builtin_nontemporal_store(r, (double *)out + i);
  } else {
    const Vec3<double> d(__builtin_nontemporal_load(dx + i), __builtin_nontemporal_load(dy + i),
                         __builtin_nontemporal_load(dz + i));
    double r;
    bool fp64 = near;
    if (!near) {
      const Vec3<float> df(float(d.x), float(d.y), float(d.z));
      const float f = OP == kOpDistIn ? S::template DistIn<float>(s, pf, df) : S::template DistOut<float>(s, pf, df);
      r             = f >= kFltBig ? kInfLength : double(f);
      fp64          = f != f;
    }
    if (fp64) r = OP == kOpDistIn ? S::template DistIn<double>(s, p, d) : S::template DistOut<double>(s, p, d);
    __builtin_nontemporal_store(r, (double *)out + i);
  }
}

template <class Data>
static hipError_t launch_mixed(Data const &s, int op, const double *px, const double *py, const double *pz,
                               const double *dx, const double *dy, const double *dz, void *out, size_t n,
                               hipStream_t st)
{
  dim3 g(unsigned((n + kMixBlock - 1) / kMixBlock)), b(kMixBlock);
  switch (op) {
  case kOpContains:
    hipLaunchKernelGGL((k_query_mixed<Data, kOpContains>), g, b, 0, st, s, px, py, pz, dx, dy, dz, out, n);
    break;
  case kOpInside:
    hipLaunchKernelGGL((k_query_mixed<Data, kOpInside>), g, b, 0, st, s, px, py, pz, dx, dy, dz, out, n);
    break;
  case kOpDistIn:
    hipLaunchKernelGGL((k_query_mixed<Data, kOpDistIn>), g, b, 0, st, s, px, py, pz, dx, dy, dz, out, n);
    break;
  case kOpDistOut:
    hipLaunchKernelGGL((k_query_mixed<Data, kOpDistOut>), g, b, 0, st, s, px, py, pz, dx, dy, dz, out, n);
    break;
  case kOpSafetyIn:
    hipLaunchKernelGGL((k_query_mixed<Data, kOpSafetyIn>), g, b, 0, st, s, px, py, pz, dx, dy, dz, out, n);
    break;
  default:
    hipLaunchKernelGGL((k_query_mixed<Data, kOpSafetyOut>), g, b, 0, st, s, px, py, pz, dx, dy, dz, out, n);
    break;
  }
  return hipGetLastError();
}

vgmi_status mixed_shape_query(const vgmi_shape *shape, int op, const double *px, const double *py, const double *pz,
                              const double *dx, const double *dy, const double *dz, void *out, size_t n, void *stream)
{
  hipStream_t st = (hipStream_t)stream;
  hipError_t e;
  switch (shape->rec.type) {
#define X(E, T, M, NS)                                                                                               \
  case E:                                                                                                            \
    e = launch_mixed(shape->rec.u.M, op, px, py, pz, dx, dy, dz, out, n, st);                                       \
    break;
    VGMI_FOR_EACH_SHAPE(X)
#undef X
  default:
    return set_error(VGMI_ERR_UNSUPPORTED, "vgmi_shape_query: no mixed kernel for shape type %d", shape->rec.type);
  }
  if (e != hipSuccess) return set_error(VGMI_ERR_DEVICE, "vgmi_shape_query (mixed) launch: %s", hipGetErrorString(e));
  return VGMI_OK;
}

} // namespace vgmi
