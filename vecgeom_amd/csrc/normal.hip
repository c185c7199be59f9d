// Surface normals over SOA points: VUnplacedVolume::Normal (VecGeom/volumes/UnplacedVolume.h:252)
// as a batched launch. One lane per point: coalesced loads of the point, the shape record in the
// kernel arguments, coalesced stores of the normal and the validity flag. Lanes whose shape leaves
// the normal unassigned keep the caller's vector (no store).
#include "vgmi_internal.h"
#include "dev_normal.h"

namespace vgmi {

constexpr int kNormalBlock = 256;

// one lane per point, grid-strided beyond 64 workgroups per CU
static unsigned normal_grid(size_t n)
{
  const size_t blocks = (n + kNormalBlock - 1) / kNormalBlock, cap = size_t(256) * 64;
  return unsigned(blocks < cap ? blocks : cap);
}

struct NormalArgs {
  const double *px, *py, *pz;
  double *nx, *ny, *nz;
  uint8_t *valid;
  size_t n;
};

template <class Data>
__global__ __launch_bounds__(kNormalBlock) void k_normal(Data s, NormalArgs a)
{
  const size_t stride = size_t(gridDim.x) * kNormalBlock;
  for (size_t i = size_t(blockIdx.x) * kNormalBlock + threadIdx.x; i < a.n; i += stride) {
    const Vec3<double> p(a.px[i], a.py[i], a.pz[i]);
    Vec3<double> nv(0., 0., 0.);
    bool set         = false;
    const bool valid = nrm::Normal(s, p, nv, set);
    if (set) {
      a.nx[i] = nv.x;
      a.ny[i] = nv.y;
      a.nz[i] = nv.z;
    }
    a.valid[i] = uint8_t(valid);
  }
}

} // namespace vgmi

using namespace vgmi;

extern "C" {

vgmi_status vgmi_shape_normal(const vgmi_shape *shape, const double *px, const double *py, const double *pz,
                              double *nx, double *ny, double *nz, uint8_t *valid, size_t n, void *stream)
{
  if (!shape) return set_error(VGMI_ERR_INVALID, "vgmi_shape_normal: null shape");
  if (n == 0) return VGMI_OK;
  if (!px || !py || !pz || !nx || !ny || !nz || !valid)
    return set_error(VGMI_ERR_INVALID, "vgmi_shape_normal: null point/normal/valid pointer");
  const NormalArgs a{px, py, pz, nx, ny, nz, valid, n};
  const dim3 g(normal_grid(n)), b(kNormalBlock);
  hipStream_t st = (hipStream_t)stream;
  switch (shape->rec.type) {
#define X(E, T, M, NS)                                                                                               \
  case E:                                                                                                            \
    hipLaunchKernelGGL((k_normal<T>), g, b, 0, st, shape->rec.u.M, a);                                               \
    break;
    VGMI_FOR_EACH_SHAPE(X)
#undef X
  default:
    return set_error(VGMI_ERR_UNSUPPORTED, "vgmi_shape_normal: shape type %d has no kernel", shape->rec.type);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(VGMI_ERR_DEVICE, "vgmi_shape_normal launch: %s", hipGetErrorString(e));
  return VGMI_OK;
}

} // extern "C"
