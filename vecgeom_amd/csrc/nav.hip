// Navigation kernels and their C ABI entry points (one lane per track).
#include "dev_nav.h"
#include "geom_internal.h"
#include "vgmi_internal.h"

namespace vgmi {
namespace nav {

// Shape-type sets the kernels are instantiated for: a geometry built only from the simple
// solids gets kernels without the polycone/polyhedron/sphere/torus solvers, whose register
// demand would otherwise set the occupancy of every navigation kernel.
constexpr unsigned kSimpleTypes = (1u << kBox) | (1u << kTube) | (1u << kCone) | (1u << kTrapezoid) | (1u << kOrb);
constexpr unsigned kAllTypes    = ~0u;
// Waves per SIMD the transport-loop kernels are compiled for: the simple-solid variants are
// held to 128 VGPRs (4 waves, a few bytes of spill); the all-solid variants keep every register.
template <unsigned TM>
constexpr unsigned kLoopWaves = TM == kAllTypes ? 1u : 4u;

#define VGMI_NAV_STACK                                                                                               \
  __shared__ uint32_t lds_stack[kStackDepth * kNavBlock];                                                           \
  Stack stk{lds_stack + threadIdx.x}

template <unsigned TM>
__global__ __launch_bounds__(kNavBlock) void k_locate(GeomView g, const double *__restrict__ px,
                                                      const double *__restrict__ py, const double *__restrict__ pz,
                                                      uint32_t *__restrict__ out, size_t n)
{
  VGMI_NAV_STACK;
  const size_t i = size_t(blockIdx.x) * kNavBlock + threadIdx.x;
  if (i >= n) return;
  out[i] = locate_global<TM>(g, V3(px[i], py[i], pz[i]), stk);
}

struct StepArgs {
  const double *px, *py, *pz, *dx, *dy, *dz, *limit;
  const uint32_t *nav_in, *last_in;
  uint32_t *nav_out, *last_out;
  uint8_t *bnd_out;
  double *step_out;
  size_t n;
  int relocate;
};

template <unsigned TM>
__global__ __launch_bounds__(kNavBlock) void k_step(GeomView g, StepArgs a)
{
  VGMI_NAV_STACK;
  const size_t i = size_t(blockIdx.x) * kNavBlock + threadIdx.x;
  if (i >= a.n) return;
  State in{a.nav_in[i], a.last_in ? a.last_in[i] : 0u, false}, out = in;
  double step = 0.;
  if (in.nav != 0) {
    step = compute_step<TM>(g, V3(a.px[i], a.py[i], a.pz[i]), V3(a.dx[i], a.dy[i], a.dz[i]), a.limit[i], in, out,
                        a.relocate != 0, stk);
  } else {
    out.nav = 0;
  }
  a.step_out[i] = step;
  a.nav_out[i]  = out.nav;
  if (a.last_out) a.last_out[i] = out.last;
  if (a.bnd_out) a.bnd_out[i] = out.boundary;
}

template <unsigned TM>
__global__ __launch_bounds__(kNavBlock) void k_safety(GeomView g, const double *__restrict__ px,
                                                      const double *__restrict__ py, const double *__restrict__ pz,
                                                      const uint32_t *__restrict__ nav, double *__restrict__ out,
                                                      size_t n)
{
  VGMI_NAV_STACK;
  const size_t i = size_t(blockIdx.x) * kNavBlock + threadIdx.x;
  if (i >= n) return;
  const uint32_t s = nav[i];
  out[i]           = s ? compute_safety<TM>(g, V3(px[i], py[i], pz[i]), s, stk) : 0.;
}

struct PropArgs {
  double *px, *py, *pz;
  const double *dx, *dy, *dz;
  double limit;
  uint32_t *nav, *last;
  int32_t *taken;
  size_t n;
  int nsteps;
};

// The transport loop of one track stays in registers for all its steps; only the
// final state goes back to HBM.
template <unsigned TM>
__global__ __launch_bounds__(kNavBlock) __attribute__((amdgpu_waves_per_eu(kLoopWaves<TM>))) void k_propagate(GeomView g, PropArgs a)
{
  VGMI_NAV_STACK;
  const size_t i = size_t(blockIdx.x) * kNavBlock + threadIdx.x;
  if (i >= a.n) return;
  V3 p(a.px[i], a.py[i], a.pz[i]);
  const V3 d(a.dx[i], a.dy[i], a.dz[i]);
  State st{a.nav[i], a.last[i], false};
  int s = 0;
  for (; s < a.nsteps && st.nav != 0; ++s) {
    State out;
    const double step = compute_step<TM>(g, p, d, a.limit, st, out, true, stk);
    p                 = p + step * d;
    st                = out;
  }
  a.px[i]  = p.x;
  a.py[i]  = p.y;
  a.pz[i]  = p.z;
  a.nav[i] = st.nav;
  a.last[i] = st.last;
  if (a.taken) a.taken[i] = s;
}

// Serving-loop kernel: locate each track from the world, then run its own number of
// ComputeStepAndPropagatedState steps (per-track step counts), all in registers.
template <unsigned TM>
__global__ __launch_bounds__(kNavBlock) __attribute__((amdgpu_waves_per_eu(kLoopWaves<TM>))) void k_locate_propagate(GeomView g, ServeArgs a)
{
  VGMI_NAV_STACK;
  const size_t i = size_t(blockIdx.x) * kNavBlock + threadIdx.x;
  if (i >= a.n) return;
  V3 p(a.px[i], a.py[i], a.pz[i]);
  const V3 d(a.dx[i], a.dy[i], a.dz[i]);
  State st{locate_global<TM>(g, p, stk), 0u, false};
  const int nsteps = a.nsteps[i];
  int s            = 0;
  for (; s < nsteps && st.nav != 0; ++s) {
    State out;
    const double step = compute_step<TM>(g, p, d, a.limit, st, out, true, stk);
    p                 = p + step * d;
    st                = out;
  }
  a.opx[i]    = p.x;
  a.opy[i]    = p.y;
  a.opz[i]    = p.z;
  a.onav[i]   = st.nav;
  a.otaken[i] = s;
}

} // namespace nav

vgmi_status launch_locate_propagate(const vgmi_geom *geom, ServeArgs const &a, hipStream_t st)
{
  if (a.n == 0) return VGMI_OK;
  const dim3 grid(unsigned((a.n + nav::kNavBlock - 1) / nav::kNavBlock));
  if ((geom->typeMask & ~nav::kSimpleTypes) == 0u)
    hipLaunchKernelGGL(nav::k_locate_propagate<nav::kSimpleTypes>, grid, dim3(nav::kNavBlock), 0, st, geom->view, a);
  else
    hipLaunchKernelGGL(nav::k_locate_propagate<nav::kAllTypes>, grid, dim3(nav::kNavBlock), 0, st, geom->view, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(VGMI_ERR_DEVICE, "serving launch: %s", hipGetErrorString(e));
  return VGMI_OK;
}

} // namespace vgmi

using namespace vgmi;
using namespace vgmi::nav;

static unsigned nav_grid(size_t n) { return unsigned((n + kNavBlock - 1) / kNavBlock); }

// Launches KERNEL<kSimpleTypes> when the geometry has only simple solids, else KERNEL<kAllTypes>.
#define VGMI_NAV_LAUNCH(KERNEL, GEOM, N, STREAM, ...)                                                                \
  do {                                                                                                               \
    if (((GEOM)->typeMask & ~kSimpleTypes) == 0u)                                                                    \
      hipLaunchKernelGGL(KERNEL<kSimpleTypes>, dim3(nav_grid(N)), dim3(kNavBlock), 0, (hipStream_t)(STREAM),         \
                         __VA_ARGS__);                                                                               \
    else                                                                                                             \
      hipLaunchKernelGGL(KERNEL<kAllTypes>, dim3(nav_grid(N)), dim3(kNavBlock), 0, (hipStream_t)(STREAM),            \
                         __VA_ARGS__);                                                                               \
  } while (0)

static vgmi_status check_geom(const vgmi_geom *g, int precision, const char *who)
{
  if (!g) return set_error(VGMI_ERR_INVALID, "%s: null geometry", who);
  if (!g->device_block) return set_error(VGMI_ERR_INVALID, "%s: geometry was built host-only", who);
  if (precision != VGMI_FP64 && precision != VGMI_MIXED)
    return set_error(VGMI_ERR_INVALID, "%s: unknown precision %d", who, precision);
  return VGMI_OK;
}

static vgmi_status launched(const char *who)
{
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(VGMI_ERR_DEVICE, "%s launch: %s", who, hipGetErrorString(e));
  return VGMI_OK;
}

extern "C" {

vgmi_status vgmi_nav_locate(const vgmi_geom *geom, int precision, const double *px, const double *py,
                            const double *pz, uint32_t *nav_out, size_t n, void *stream)
{
  if (vgmi_status s = check_geom(geom, precision, "vgmi_nav_locate")) return s;
  if (n == 0) return VGMI_OK;
  if (!px || !py || !pz || !nav_out) return set_error(VGMI_ERR_INVALID, "vgmi_nav_locate: null pointer");
  VGMI_NAV_LAUNCH(k_locate, geom, n, stream, geom->view, px, py, pz, nav_out, n);
  return launched("vgmi_nav_locate");
}

vgmi_status vgmi_nav_step(const vgmi_geom *geom, int precision, const double *px, const double *py, const double *pz,
                          const double *dx, const double *dy, const double *dz, const double *step_limit,
                          const uint32_t *nav_in, const uint32_t *last_in, uint32_t *nav_out, uint32_t *last_out,
                          uint8_t *boundary_out, double *step_out, size_t n, int relocate, void *stream)
{
  if (vgmi_status s = check_geom(geom, precision, "vgmi_nav_step")) return s;
  if (n == 0) return VGMI_OK;
  if (!px || !py || !pz || !dx || !dy || !dz || !step_limit || !nav_in || !nav_out || !step_out)
    return set_error(VGMI_ERR_INVALID, "vgmi_nav_step: null pointer");
  StepArgs a{px, py, pz, dx, dy, dz, step_limit, nav_in, last_in, nav_out, last_out, boundary_out, step_out, n,
             relocate};
  VGMI_NAV_LAUNCH(k_step, geom, n, stream, geom->view, a);
  return launched("vgmi_nav_step");
}

vgmi_status vgmi_nav_safety(const vgmi_geom *geom, int precision, const double *px, const double *py,
                            const double *pz, const uint32_t *nav_in, double *safety_out, size_t n, void *stream)
{
  if (vgmi_status s = check_geom(geom, precision, "vgmi_nav_safety")) return s;
  if (n == 0) return VGMI_OK;
  if (!px || !py || !pz || !nav_in || !safety_out) return set_error(VGMI_ERR_INVALID, "vgmi_nav_safety: null pointer");
  VGMI_NAV_LAUNCH(k_safety, geom, n, stream, geom->view, px, py, pz, nav_in, safety_out, n);
  return launched("vgmi_nav_safety");
}

vgmi_status vgmi_nav_propagate(const vgmi_geom *geom, int precision, double *px, double *py, double *pz,
                               const double *dx, const double *dy, const double *dz, double step_limit, uint32_t *nav,
                               uint32_t *last, int32_t *steps_taken, size_t n, int nsteps, void *stream)
{
  if (vgmi_status s = check_geom(geom, precision, "vgmi_nav_propagate")) return s;
  if (n == 0 || nsteps <= 0) return VGMI_OK;
  if (!px || !py || !pz || !dx || !dy || !dz || !nav || !last)
    return set_error(VGMI_ERR_INVALID, "vgmi_nav_propagate: null pointer");
  PropArgs a{px, py, pz, dx, dy, dz, step_limit, nav, last, steps_taken, n, nsteps};
  VGMI_NAV_LAUNCH(k_propagate, geom, n, stream, geom->view, a);
  return launched("vgmi_nav_propagate");
}

} // extern "C"
