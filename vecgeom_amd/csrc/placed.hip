// Placed-volume SOA queries: VPlacedVolume::{Contains, Inside, DistanceToIn, DistanceToOut, SafetyToIn,
// SafetyToOut} over SOA3D (VecGeom/volumes/PlacedVolume.h:273,299,332,372,405,421) and
// PlacedDistanceToOut, dispatched as CommonSpecializedVolImplHelper does
// (VecGeom/volumes/SpecializedPlacedVolImplHelper.h:98-179): the point (and direction) are taken from
// the mother's frame into the volume's frame with the placement's Transformation3D (general case,
// Transformation3D.h:734-758 and DoRotation :661-669), then the unplaced kernel runs. DistanceToOut and
// SafetyToOut take points already in the volume's frame (PlacedVolume.h:335-345, 408-416).
//
// One lane per point; the shape record and the 12 numbers of the transformation arrive by value in the
// kernel arguments. FP64 runs the reference's scalar arithmetic; MIXED keeps the transformation in
// FP64 and runs the mixed-precision logic of a single-shape query on the local point.
#include "dev_mixed.h"
#include "dev_nav.h"
#include "dev_normal.h"
#include "geom_internal.h"
#include "vgmi_internal.h"

namespace vgmi {

constexpr int kPlacedBlock   = 256;
constexpr int kOpPlacedDistOut = 6; // VGMI_OP_PLACED_DIST_OUT

struct Placement {
  double t[3], r[9];
};

struct PlacedArgs {
  const double *px, *py, *pz, *dx, *dy, *dz, *step_max;
  void *out;
  size_t n;
};

template <class Data>
VGMI_DEV double placed_dist(Data const &s, int op, bool mixed, Vec3<double> const &p, Vec3<double> const &d,
                            double stepMax)
{
  using S    = ShapeOps<Data>;
  const bool in = op == kOpDistIn;
  if constexpr (S::kType == kTessellated) {
    const tess::V3d pp(p.x, p.y, p.z), dd(d.x, d.y, d.z);
    return in ? tess::dist_in_step64(s, tess::pool_walk(s), pp, dd, stepMax)
              : tess::dist_out_step64(s, tess::pool_walk(s), pp, dd, stepMax);
  } else if constexpr (S::kType == kGenTrap) {
    if (in) return gtra::DistanceToInStep(s, p, d, stepMax); // the step limit bounds the entry search
    return S::template DistOut<double>(s, p, d);
  } else if constexpr (S::kType == kCutTube) {
    if (in) return ctub::DistanceToInStep(s, p, d, stepMax); // planes hit beyond the limit are not taken
    return S::template DistOut<double>(s, p, d);
  } else {
    (void)stepMax; // the reference's other kernels ignore the step limit
    if (mixed) return in ? mixed_point<Data, kOpDistIn>(s, p, d) : mixed_point<Data, kOpDistOut>(s, p, d);
    return in ? S::template DistIn<double>(s, p, d) : S::template DistOut<double>(s, p, d);
  }
}

template <class Data, int OP, bool MIXED>
__global__ __launch_bounds__(kPlacedBlock) void k_placed(Data s, Placement T, PlacedArgs a)
{
  using S      = ShapeOps<Data>;
  size_t i     = size_t(blockIdx.x) * kPlacedBlock + threadIdx.x;
  size_t strid = size_t(gridDim.x) * kPlacedBlock;
  for (; i < a.n; i += strid) {
    Vec3<double> p(a.px[i], a.py[i], a.pz[i]);
    // Contains, Inside, DistanceToIn, SafetyToIn and PlacedDistanceToOut work in the mother's frame
    constexpr bool kTransform = OP != kOpDistOut && OP != kOpSafetyOut;
    if constexpr (kTransform) p = nav::to_local(T.t, T.r, p);
    if constexpr (OP == kOpContains) {
      const bool r = MIXED && S::kType != kTessellated ? mixed_point<Data, kOpContains>(s, p, p) != 0.
                                                        : S::template Contains<double>(s, p);
      static_cast<uint8_t *>(a.out)[i] = uint8_t(r);
    } else if constexpr (OP == kOpInside) {
      const int r = MIXED && S::kType != kTessellated ? int(mixed_point<Data, kOpInside>(s, p, p))
                                                       : S::template Inside<double>(s, p);
      static_cast<int32_t *>(a.out)[i] = r;
    } else if constexpr (OP == kOpSafetyIn || OP == kOpSafetyOut) {
      double r;
      if (MIXED && S::kType != kTessellated)
        r = OP == kOpSafetyIn ? mixed_point<Data, kOpSafetyIn>(s, p, p) : mixed_point<Data, kOpSafetyOut>(s, p, p);
      else
        r = OP == kOpSafetyIn ? S::template SafeIn<double>(s, p) : S::template SafeOut<double>(s, p);
      static_cast<double *>(a.out)[i] = r;
    } else {
      Vec3<double> d(a.dx[i], a.dy[i], a.dz[i]);
      if constexpr (kTransform) d = nav::rotate(T.r, d);
      const double sm = a.step_max ? a.step_max[i] : kInfLength;
      const int op    = OP == kOpDistIn ? kOpDistIn : kOpDistOut;
      static_cast<double *>(a.out)[i] = placed_dist(s, op, MIXED, p, d, sm);
    }
  }
}

template <class Data, bool MIXED>
static hipError_t launch_placed(Data const &s, int op, Placement const &T, PlacedArgs const &a, hipStream_t st)
{
  size_t blocks = (a.n + kPlacedBlock - 1) / kPlacedBlock;
  blocks        = blocks < size_t(256) * 64 ? blocks : size_t(256) * 64;
  const dim3 g(static_cast<unsigned>(blocks)), b(kPlacedBlock);
  switch (op) {
  case kOpContains: hipLaunchKernelGGL((k_placed<Data, kOpContains, MIXED>), g, b, 0, st, s, T, a); break;
  case kOpInside: hipLaunchKernelGGL((k_placed<Data, kOpInside, MIXED>), g, b, 0, st, s, T, a); break;
  case kOpDistIn: hipLaunchKernelGGL((k_placed<Data, kOpDistIn, MIXED>), g, b, 0, st, s, T, a); break;
  case kOpDistOut: hipLaunchKernelGGL((k_placed<Data, kOpDistOut, MIXED>), g, b, 0, st, s, T, a); break;
  case kOpSafetyIn: hipLaunchKernelGGL((k_placed<Data, kOpSafetyIn, MIXED>), g, b, 0, st, s, T, a); break;
  case kOpSafetyOut: hipLaunchKernelGGL((k_placed<Data, kOpSafetyOut, MIXED>), g, b, 0, st, s, T, a); break;
  default: hipLaunchKernelGGL((k_placed<Data, kOpPlacedDistOut, MIXED>), g, b, 0, st, s, T, a); break;
  }
  return hipGetLastError();
}

struct PlacedNormalArgs {
  const double *px, *py, *pz;
  double *nx, *ny, *nz;
  uint8_t *valid;
  size_t n;
};

// VPlacedVolume::Normal (source/PlacedVolume.cpp:126-137): Transform, the unplaced Normal into a zero
// vector, InverseTransformDirection (Transformation3D.h:808-823 with transform_direction).
template <class Data>
__global__ __launch_bounds__(kPlacedBlock) void k_placed_normal(Data s, Placement T, PlacedNormalArgs a)
{
  const size_t stride = size_t(gridDim.x) * kPlacedBlock;
  for (size_t i = size_t(blockIdx.x) * kPlacedBlock + threadIdx.x; i < a.n; i += stride) {
    const Vec3<double> lp = nav::to_local(T.t, T.r, Vec3<double>(a.px[i], a.py[i], a.pz[i]));
    Vec3<double> ln(0., 0., 0.);
    bool set         = false;
    const bool valid = nrm::Normal(s, lp, ln, set);
    double x = ln.x * T.r[0], y = ln.x * T.r[3], z = ln.x * T.r[6];
    x += ln.y * T.r[1];
    x += ln.z * T.r[2];
    y += ln.y * T.r[4];
    y += ln.z * T.r[5];
    z += ln.y * T.r[7];
    z += ln.z * T.r[8];
    a.nx[i]    = x;
    a.ny[i]    = y;
    a.nz[i]    = z;
    a.valid[i] = uint8_t(valid);
  }
}

} // namespace vgmi

using namespace vgmi;

extern "C" {

vgmi_status vgmi_placed_query(const vgmi_geom *geom, int placed_id, int op, int precision, const double *px,
                              const double *py, const double *pz, const double *dx, const double *dy,
                              const double *dz, const double *step_max, void *out, size_t n, void *stream)
{
  if (!geom) return set_error(VGMI_ERR_INVALID, "vgmi_placed_query: null geometry");
  if (!geom->device_block) return set_error(VGMI_ERR_INVALID, "vgmi_placed_query: geometry was built host-only");
  const int npv = int(geom->host.pvols.size());
  if (placed_id < 0 || placed_id >= npv)
    return set_error(VGMI_ERR_INVALID, "vgmi_placed_query: placed volume %d not in [0, %d)", placed_id, npv);
  if (op < 0 || op > kOpPlacedDistOut) return set_error(VGMI_ERR_INVALID, "vgmi_placed_query: unknown op %d", op);
  if (precision != VGMI_FP64 && precision != VGMI_MIXED)
    return set_error(VGMI_ERR_INVALID, "vgmi_placed_query: unknown precision %d", precision);
  if (n == 0) return VGMI_OK;
  if (!px || !py || !pz || !out) return set_error(VGMI_ERR_INVALID, "vgmi_placed_query: null point/output pointer");
  const bool directed = op == kOpDistIn || op == kOpDistOut || op == kOpPlacedDistOut;
  if (directed && (!dx || !dy || !dz))
    return set_error(VGMI_ERR_INVALID, "vgmi_placed_query: distances need directions");
  PVolRec const &pv     = geom->host.pvols[placed_id];
  ShapeRecord const &sr = geom->devShapes[geom->host.lvs[pv.lv].shape];
  Placement T;
  for (int k = 0; k < 3; ++k) T.t[k] = pv.t[k];
  for (int k = 0; k < 9; ++k) T.r[k] = pv.r[k];
  const PlacedArgs a{px, py, pz, dx, dy, dz, directed ? step_max : nullptr, out, n};
  hipStream_t st = (hipStream_t)stream;
  hipError_t e;
  switch (sr.type) {
#define X(E, TT, M, NS)                                                                                              \
  case E:                                                                                                            \
    e = precision == VGMI_MIXED ? launch_placed<TT, true>(sr.u.M, op, T, a, st)                                     \
                                : launch_placed<TT, false>(sr.u.M, op, T, a, st);                                   \
    break;
    VGMI_FOR_EACH_SHAPE(X)
#undef X
  default:
    return set_error(VGMI_ERR_UNSUPPORTED, "vgmi_placed_query: shape type %d has no kernel", sr.type);
  }
  if (e != hipSuccess) return set_error(VGMI_ERR_DEVICE, "vgmi_placed_query launch: %s", hipGetErrorString(e));
  return VGMI_OK;
}

vgmi_status vgmi_placed_normal(const vgmi_geom *geom, int placed_id, const double *px, const double *py,
                               const double *pz, double *nx, double *ny, double *nz, uint8_t *valid, size_t n,
                               void *stream)
{
  if (!geom) return set_error(VGMI_ERR_INVALID, "vgmi_placed_normal: null geometry");
  if (!geom->device_block) return set_error(VGMI_ERR_INVALID, "vgmi_placed_normal: geometry was built host-only");
  const int npv = int(geom->host.pvols.size());
  if (placed_id < 0 || placed_id >= npv)
    return set_error(VGMI_ERR_INVALID, "vgmi_placed_normal: placed volume %d not in [0, %d)", placed_id, npv);
  if (n == 0) return VGMI_OK;
  if (!px || !py || !pz || !nx || !ny || !nz || !valid)
    return set_error(VGMI_ERR_INVALID, "vgmi_placed_normal: null point/normal/valid pointer");
  PVolRec const &pv     = geom->host.pvols[placed_id];
  ShapeRecord const &sr = geom->devShapes[geom->host.lvs[pv.lv].shape];
  Placement T;
  for (int k = 0; k < 3; ++k) T.t[k] = pv.t[k];
  for (int k = 0; k < 9; ++k) T.r[k] = pv.r[k];
  const PlacedNormalArgs a{px, py, pz, nx, ny, nz, valid, n};
  size_t blocks = (n + kPlacedBlock - 1) / kPlacedBlock;
  blocks        = blocks < size_t(256) * 64 ? blocks : size_t(256) * 64;
  const dim3 g(static_cast<unsigned>(blocks)), b(kPlacedBlock);
  hipStream_t st = (hipStream_t)stream;
  switch (sr.type) {
#define X(E, TT, M, NS)                                                                                              \
  case E:                                                                                                            \
    hipLaunchKernelGGL((k_placed_normal<TT>), g, b, 0, st, sr.u.M, T, a);                                             \
    break;
    VGMI_FOR_EACH_SHAPE(X)
#undef X
  default:
    return set_error(VGMI_ERR_UNSUPPORTED, "vgmi_placed_normal: shape type %d has no kernel", sr.type);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(VGMI_ERR_DEVICE, "vgmi_placed_normal launch: %s", hipGetErrorString(e));
  return VGMI_OK;
}

} // extern "C"
