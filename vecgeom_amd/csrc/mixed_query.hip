// Mixed-precision single-shape queries: one lane per point, the per-point logic of
// dev_mixed.h (FP32 bulk, FP64 near-surface and undecided re-evaluation).
#include "dev_mixed.h"
#include "vgmi_internal.h"

namespace vgmi {

constexpr int kMixBlock = 256;

template <class Data, int OP>
__global__ __launch_bounds__(kMixBlock) void k_query_mixed(Data s, const double *__restrict__ px,
                                                           const double *__restrict__ py,
                                                           const double *__restrict__ pz,
                                                           const double *__restrict__ dx,
                                                           const double *__restrict__ dy,
                                                           const double *__restrict__ dz, void *out, size_t n)
{
  const size_t i = size_t(blockIdx.x) * kMixBlock + threadIdx.x;
  if (i >= n) return;
  const Vec3<double> p(__builtin_nontemporal_load(px + i), __builtin_nontemporal_load(py + i),
                       __builtin_nontemporal_load(pz + i));
  Vec3<double> d(0., 0., 0.);
  if constexpr (OP == kOpDistIn || OP == kOpDistOut)
    d = Vec3<double>(__builtin_nontemporal_load(dx + i), __builtin_nontemporal_load(dy + i),
                     __builtin_nontemporal_load(dz + i));
  const double r = mixed_point<Data, OP>(s, p, d);
  if constexpr (OP == kOpContains) {
    __builtin_nontemporal_store(uint8_t(r != 0.), (uint8_t *)out + i);
  } else if constexpr (OP == kOpInside) {
    __builtin_nontemporal_store(int32_t(r), (int32_t *)out + i);
  } else {
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
