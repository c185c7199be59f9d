// Navigation kernels and their C ABI entry points (one lane per track).
#include "dev_nav.h"
#include <algorithm>
#include <mutex>
#include "geom_internal.h"
#include "vgmi_internal.h"

namespace vgmi {
namespace nav {

// Shape-type sets the kernels are instantiated for: a geometry built only from the simple
// solids gets kernels without the polycone/polyhedron/sphere/torus solvers, whose register
// demand would otherwise set the occupancy of every navigation kernel.
constexpr unsigned kSimpleTypes = (1u << kBox) | (1u << kTube) | (1u << kCone) | (1u << kTrapezoid) | (1u << kOrb);
constexpr unsigned kAllTypes    = ~kMixedFlag;
// Waves per SIMD the transport-loop kernels are compiled for: the simple-solid variants are
// held to 128 VGPRs (4 waves, a few bytes of spill); the all-solid variants keep every register.
#ifndef VGMI_LOOP_WAVES
#define VGMI_LOOP_WAVES 4
#endif
template <unsigned TM>
constexpr unsigned kLoopWaves = (TM & ~kMixedFlag) == kAllTypes ? 1u : unsigned(VGMI_LOOP_WAVES);

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

// BVHLevelLocator::LevelLocate / LevelLocateExclVol (VecGeom/navigation/BVHLevelLocator.h:44-77 over
// BVH::LevelLocate, VecGeom/base/BVH.h:219-344): the first daughter of logical volume lv[i] (BVH order)
// whose shape contains the point, skipping the placed volume excl[i]; -1 when none does (or lv[i] is not
// a logical volume).
struct LevelArgs {
  const int32_t *lv, *excl;
  const double *px, *py, *pz;
  int32_t *out;
  double *dx, *dy, *dz; // the point in the daughter's frame (optional)
  size_t n;
};

template <unsigned TM>
__global__ __launch_bounds__(kNavBlock) void k_level_locate(GeomView g, LevelArgs a)
{
  VGMI_NAV_STACK;
  const size_t i = size_t(blockIdx.x) * kNavBlock + threadIdx.x;
  if (i >= a.n) return;
  const int lv = a.lv[i];
  int pv       = -1;
  V3 dlp(0., 0., 0.);
  const V3 lp(a.px[i], a.py[i], a.pz[i]);
  if (lv >= 0 && lv < g.nlogical && g.lvs[lv].nNodes > 0) {
    NavVol M;
    M.nodeBase = g.lvs[lv].nodeBase;
    M.slotBase = g.lvs[lv].dOffset;
    if (!level_locate<TM>(g, M, a.excl ? a.excl[i] : -1, lp, pv, dlp, stk)) pv = -1;
  }
  a.out[i] = pv;
  if (a.dx) {
    if (pv < 0) dlp = lp;
    a.dx[i] = dlp.x;
    a.dy[i] = dlp.y;
    a.dz[i] = dlp.z;
  }
}

struct StepArgs {
  const double *px, *py, *pz, *dx, *dy, *dz, *limit;
  const uint32_t *nav_in, *last_in;
  uint32_t *nav_out, *last_out;
  uint8_t *bnd_out;
  double *step_out;
  size_t n;
  int relocate;
  const uint8_t *calc_safety; // with safety_out: per-track flag, null = every track
  double *safety_out;         // null: no safety
};

template <unsigned TM>
__global__ __launch_bounds__(kNavBlock) void k_step(GeomView g, StepArgs a)
{
  VGMI_NAV_STACK;
  const size_t i = size_t(blockIdx.x) * kNavBlock + threadIdx.x;
  if (i >= a.n) return;
  State in{a.nav_in[i], a.last_in ? a.last_in[i] : 0u, false}, out = in;
  double step = 0., budget = -1.;
  if (in.nav != 0) {
    double safety = 0.;
    step = compute_step<TM>(g, V3(a.px[i], a.py[i], a.pz[i]), V3(a.dx[i], a.dy[i], a.dz[i]), a.limit[i], in, out,
                            a.relocate != 0, stk, budget, a.safety_out ? &safety : nullptr,
                            a.calc_safety ? a.calc_safety[i] != 0 : true);
    if (a.safety_out) a.safety_out[i] = safety;
  } else {
    out.nav = 0;
    if (a.safety_out) a.safety_out[i] = 0.;
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
  const int32_t *tsteps; // per-track step counts (null: nsteps for every track)
  const uint32_t *perm;  // lane j runs track perm[*first + j] (null: track j)
  const uint32_t *first; // number of leading perm entries to skip (tracks outside the world)
  int accumulate;        // add to taken instead of storing
  int lanes;             // tracks per wavefront (1..64); lanes >= this one idle
  const uint32_t *dn;    // device-side track count (<= n, the capacity), or null: n tracks
};

// The transport loop of one track stays in registers for all its steps; only the
// final state goes back to HBM.
template <unsigned TM>
__global__ __launch_bounds__(kNavBlock) __attribute__((amdgpu_waves_per_eu(kLoopWaves<TM>))) void k_propagate(GeomView g, PropArgs a)
{
  VGMI_NAV_STACK;
  const unsigned lane = threadIdx.x & 63u;
  const size_t wave   = (size_t(blockIdx.x) * kNavBlock + threadIdx.x) >> 6;
  const size_t n      = a.dn ? size_t(*a.dn) : a.n;
  unsigned lanes      = unsigned(a.lanes);
  size_t first        = 0;
  if (a.perm) {
    // Regrouped chunk: only the tracks still running are in perm[first, n). They are spread over
    // every launched wavefront, fewest per wavefront (a power of two) that covers them, so the
    // tail of a loop whose tracks finish early runs with few tracks per wavefront. The grid holds
    // ceil(n / a.lanes) wavefronts, so a.lanes always covers them.
    first               = *a.first;
    const size_t active = n > first ? n - first : 0;
    const size_t waves  = (size_t(gridDim.x) * kNavBlock) >> 6;
    unsigned l          = 1;
    while (l < lanes && size_t(l) * waves < active) l <<= 1;
    lanes = l;
  }
  if (lane >= lanes) return;
  size_t i = wave * lanes + lane;
  if (a.perm) {
    i += first;
    if (i >= n) return;
    i = a.perm[i];
  } else if (i >= n) {
    return;
  }
  V3 p(a.px[i], a.py[i], a.pz[i]);
  const V3 d(a.dx[i], a.dy[i], a.dz[i]);
  State st{a.nav[i], a.last[i], false};
  int todo = a.nsteps;
  if (a.tsteps) {
    const int rem = a.tsteps[i] - (a.accumulate ? a.taken[i] : 0);
    todo          = rem < todo ? rem : todo;
  }
  double budget = -1.;
  int s         = 0;
  for (; s < todo && st.nav != 0; ++s) {
    State out;
    const double step = compute_step<TM>(g, p, d, a.limit, st, out, true, stk, budget);
    p                 = p + step * d;
    st                = out;
  }
  a.px[i]  = p.x;
  a.py[i]  = p.y;
  a.pz[i]  = p.z;
  a.nav[i] = st.nav;
  a.last[i] = st.last;
  if (a.taken) a.taken[i] = a.accumulate ? a.taken[i] + s : s;
}

// ---- regrouping: a counting sort of tracks by the logical volume their state is in
// (key 0: outside the world), so that the lanes of a wavefront run the same volume's code.
constexpr int kSortBlock = 256;
constexpr int kLocalBins = 8192; // bins kept in LDS (32 KiB); larger key ranges use global atomics

// Key 0 also takes tracks that have used up their own step count.
struct SortKeys {
  const uint32_t *nav;
  const int32_t *tsteps, *taken; // optional
  VGMI_DEV uint32_t operator()(GeomView const &g, size_t i) const
  {
    const uint32_t s = nav[i];
    if (s == 0 || (tsteps && taken[i] >= tsteps[i])) return 0u;
    return uint32_t(g.vols[top(g, s)].lv) + 1u;
  }
};

__global__ void k_fill_u32(uint32_t *a, size_t n, uint32_t v)
{
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) a[i] = v;
}

__global__ __launch_bounds__(kSortBlock) void k_sort_hist(GeomView g, SortKeys key, size_t cap, const uint32_t *dn,
                                                          uint32_t *hist, int nbins)
{
  const size_t n = dn ? size_t(*dn) : cap;
  extern __shared__ uint32_t sh[];
  const bool local = nbins <= kLocalBins;
  if (local)
    for (int b = threadIdx.x; b < nbins; b += kSortBlock) sh[b] = 0;
  __syncthreads();
  for (size_t i = size_t(blockIdx.x) * kSortBlock + threadIdx.x; i < n; i += size_t(gridDim.x) * kSortBlock) {
    const uint32_t k = key(g, i);
    if (local)
      atomicAdd(&sh[k], 1u);
    else
      atomicAdd(&hist[k], 1u);
  }
  __syncthreads();
  if (local)
    for (int b = threadIdx.x; b < nbins; b += kSortBlock)
      if (sh[b]) atomicAdd(&hist[b], sh[b]);
}

// Exclusive prefix sum of hist[0..nbins) in place, one block of 1024 threads.
__global__ __launch_bounds__(1024) void k_sort_scan(uint32_t *hist, int nbins)
{
  __shared__ uint32_t part[1024];
  const int t   = threadIdx.x;
  const int per = (nbins + 1023) / 1024;
  const int b0 = t * per, b1 = b0 + per < nbins ? b0 + per : nbins;
  uint32_t sum = 0;
  for (int b = b0; b < b1; ++b) sum += hist[b];
  part[t] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const uint32_t v = t >= off ? part[t - off] : 0u;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  uint32_t run = part[t] - sum;
  for (int b = b0; b < b1; ++b) {
    const uint32_t c = hist[b];
    hist[b]          = run;
    run += c;
  }
}

// perm[cursor[key]++] = i, with block-local ranks so that each block reserves one range per key.
__global__ __launch_bounds__(kSortBlock) void k_sort_scatter(GeomView g, SortKeys key, size_t cap, const uint32_t *dn,
                                                             uint32_t *cursor, uint32_t *perm, int nbins)
{
  const size_t n = dn ? size_t(*dn) : cap;
  extern __shared__ uint32_t sh[];
  const size_t i   = size_t(blockIdx.x) * kSortBlock + threadIdx.x;
  const uint32_t k = i < n ? key(g, i) : 0u;
  if (nbins > kLocalBins) {
    if (i < n) perm[atomicAdd(&cursor[k], 1u)] = uint32_t(i);
    return;
  }
  for (int b = threadIdx.x; b < nbins; b += kSortBlock) sh[b] = 0;
  __syncthreads();
  const uint32_t rank = i < n ? atomicAdd(&sh[k], 1u) : 0u;
  __syncthreads();
  for (int b = threadIdx.x; b < nbins; b += kSortBlock) {
    const uint32_t c = sh[b];
    if (c) sh[b] = atomicAdd(&cursor[b], c);
  }
  __syncthreads();
  if (i < n) perm[sh[k] + rank] = uint32_t(i);
}

// Start of a serving launch: locate each track from the world, copy its point to the output
// arrays (which the transport loop then updates in place), clear last and taken.
template <unsigned TM>
__global__ __launch_bounds__(kNavBlock) void k_locate_init(GeomView g, const double *px, const double *py,
                                                           const double *pz, double *opx, double *opy, double *opz,
                                                           uint32_t *nav, uint32_t *last, int32_t *taken, size_t cap,
                                                           const uint32_t *dn)
{
  VGMI_NAV_STACK;
  const size_t i = size_t(blockIdx.x) * kNavBlock + threadIdx.x;
  if (i >= (dn ? size_t(*dn) : cap)) return;
  const V3 p(px[i], py[i], pz[i]);
  opx[i]  = p.x;
  opy[i]  = p.y;
  opz[i]  = p.z;
  nav[i]  = locate_global<TM>(g, p, stk);
  last[i] = 0u;
  taken[i] = 0;
}

// ---- NavStateIndex <-> path interop (one lane per state; level-major arrays, entry k of state i
// at [k * n + i], so every level's entries are written coalesced).

// NavStateIndex::GetCurrentLevel / ValueAt / GetPathAsListOfIndices
// (VecGeom/navigation/NavStateIndex.h:194-211, 341-349; source/NavStateIndex.cpp:40-53).
__global__ __launch_bounds__(256) void k_nav_to_path(GeomView g, const uint32_t *__restrict__ nav, size_t n,
                                                     int maxl, int32_t *levels, int32_t *pvols, int32_t *indices)
{
  const size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  uint32_t s  = nav[i];
  const int L = s ? int(g.navind[s + 2] & 0xffu) : -1;
  levels[i]   = L + 1;
  for (int k = L + 1; k < maxl; ++k) {
    if (pvols) pvols[size_t(k) * n + i] = -1;
    if (indices) indices[size_t(k) * n + i] = -1;
  }
  for (int k = L; k >= 0; --k) {
    const int pv = int(g.navind[s + 1]);
    if (pvols) pvols[size_t(k) * n + i] = pv;
    if (indices) indices[size_t(k) * n + i] = k > 0 ? g.vols[pv].childId : 0;
    s = g.navind[s];
  }
}

// kind 0: ResetPathFromListOfIndices (source/NavStateIndex.cpp:77-88), daughter indices below the
// world (the first entry is not read); kind 1: placed volumes world first (NavStatePath levels pushed
// one by one, NavStateIndex.h:282-287). A path the geometry does not hold gives state 0.
__global__ __launch_bounds__(256) void k_nav_from_path(GeomView g, int kind, const int32_t *__restrict__ levels,
                                                       const int32_t *__restrict__ path, size_t n, int maxl,
                                                       int npvol, uint32_t *nav)
{
  const size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const int L = levels[i];
  uint32_t s  = 0;
  bool ok     = L >= 1 && L <= maxl;
  for (int k = 0; ok && k < L; ++k) {
    const int v = path[size_t(k) * n + i];
    if (k == 0) {
      ok = kind == 0 || v == g.worldPvol;
      s  = 1;
      continue;
    }
    int cid = v;
    if (kind == 1) {
      if (v < 0 || v >= npvol) {
        ok = false;
        break;
      }
      cid = g.vols[v].childId;
    }
    if (cid < 0 || uint32_t(cid) >= (g.navind[s + 2] >> 16)) {
      ok = false;
      break;
    }
    const uint32_t c = g.navind[s + 3 + uint32_t(cid)];
    if (kind == 1 && int(g.navind[c + 1]) != v) {
      ok = false;
      break;
    }
    s = c;
  }
  nav[i] = ok ? s : 0u;
}

// NavStateIndex::Distance (source/NavStateIndex.cpp:150-166). The reference compares the placed
// volumes of both paths from the top and stops at the first difference; equal placed volumes at
// every level down to k is the same as equal states at level k, so both states are lifted to the
// shallower level and popped together until they meet.
__global__ __launch_bounds__(256) void k_nav_distance(GeomView g, const uint32_t *__restrict__ a,
                                                      const uint32_t *__restrict__ b, size_t n, int32_t *out)
{
  const size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  uint32_t sa = a[i], sb = b[i];
  const int la = sa ? int(g.navind[sa + 2] & 0xffu) : -1;
  const int lb = sb ? int(g.navind[sb + 2] & 0xffu) : -1;
  int m        = la < lb ? la : lb;
  for (int k = la; k > m; --k) sa = g.navind[sa];
  for (int k = lb; k > m; --k) sb = g.navind[sb];
  while (m >= 0 && sa != sb) {
    sa = g.navind[sa];
    sb = g.navind[sb];
    --m;
  }
  out[i] = (la - m) + (lb - m);
}

// NavStateIndex::GlobalToLocal(point[, tolevel]) (NavStateIndex.h:262-269, 368-377).
__global__ __launch_bounds__(256) void k_nav_global_to_local(GeomView g, const uint32_t *__restrict__ nav,
                                                             int tolevel, const double *__restrict__ px,
                                                             const double *__restrict__ py,
                                                             const double *__restrict__ pz, double *lx, double *ly,
                                                             double *lz, size_t n)
{
  const size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  uint32_t s = nav[i];
  if (tolevel >= 0 && s) {
    int up = int(g.navind[s + 2] & 0xffu) - tolevel;
    if (up < 0) s = 0;
    for (; s && up > 0; --up) s = g.navind[s];
  }
  Mat m;
  top_matrix(g, s, m);
  const V3 l = to_local(m.t, m.r, V3(px[i], py[i], pz[i]));
  lx[i]      = l.x;
  ly[i]      = l.y;
  lz[i]      = l.z;
}

} // namespace nav

} // namespace vgmi

using namespace vgmi;
using namespace vgmi::nav;

static unsigned nav_grid(size_t n) { return unsigned((n + kNavBlock - 1) / kNavBlock); }

// Below this many tracks a transport loop is not regrouped: the GPU is not full, and the sort
// plus the per-chunk synchronisation cost more than the coherence gains (4096 tracks x 100 steps:
// 2.6 ms unsorted, 4.5 ms sorted every 10 steps; the sort pays from 65536 tracks on).
constexpr size_t kRegroupMinTracks = 65536;

// Launches KERNEL<types | precision>: kSimpleTypes when the geometry has only simple solids,
// else kAllTypes; kMixedFlag for VGMI_MIXED.
#define VGMI_NAV_LAUNCH_TM(KERNEL, TMV, N, STREAM, ...)                                                              \
  hipLaunchKernelGGL(KERNEL<TMV>, dim3(nav_grid(N)), dim3(kNavBlock), 0, (hipStream_t)(STREAM), __VA_ARGS__)
#define VGMI_NAV_LAUNCH(KERNEL, GEOM, PREC, N, STREAM, ...)                                                          \
  do {                                                                                                               \
    const bool simple_ = ((GEOM)->typeMask & ~kSimpleTypes) == 0u;                                                   \
    if ((PREC) == VGMI_MIXED) {                                                                                      \
      if (simple_)                                                                                                   \
        VGMI_NAV_LAUNCH_TM(KERNEL, kSimpleTypes | kMixedFlag, N, STREAM, __VA_ARGS__);                              \
      else                                                                                                           \
        VGMI_NAV_LAUNCH_TM(KERNEL, kAllTypes | kMixedFlag, N, STREAM, __VA_ARGS__);                                 \
    } else {                                                                                                         \
      if (simple_)                                                                                                   \
        VGMI_NAV_LAUNCH_TM(KERNEL, kSimpleTypes, N, STREAM, __VA_ARGS__);                                            \
      else                                                                                                           \
        VGMI_NAV_LAUNCH_TM(KERNEL, kAllTypes, N, STREAM, __VA_ARGS__);                                               \
    }                                                                                                                \
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

namespace vgmi {

// Tracks per wavefront of a transport loop over n tracks. A wavefront runs the union of its
// lanes' code paths, so a lone track steps several times faster than one of 64 divergent
// tracks; while the GPU has idle wavefront slots, tracks are spread over more wavefronts
// (fewer per wavefront) at no cost in throughput. kWaveSlots = 256 CUs x 4 SIMDs x 4 waves.
constexpr size_t kWaveSlots = 4096;
int lanes_per_wave(const vgmi_geom *geom, size_t n)
{
  if (geom->lanesPerWave > 0) return geom->lanesPerWave;
  int l = 1;
  while (l < 64 && size_t(l) * kWaveSlots < n) l <<= 1;
  return l;
}
// Threads to launch so that ceil(n / lanes) wavefronts of 64 lanes exist.
static size_t lane_threads(size_t n, int lanes) { return (n + size_t(lanes) - 1) / size_t(lanes) * 64; }

size_t transport_work_words(const vgmi_geom *geom, size_t n)
{
  if (geom->regroupSteps <= 0 || n < kRegroupMinTracks) return 0;
  return size_t(geom->nlogical) + 1 + n;
}

vgmi_status launch_transport(const vgmi_geom *geom, TransportArgs const &t, uint32_t *work, hipStream_t st)
{
  const int chunk = geom->regroupSteps;
  const size_t n  = t.n;
  if (n == 0 || t.nsteps <= 0) return VGMI_OK;
  if (!work || chunk <= 0 || n < kRegroupMinTracks) {
    const int lanes = lanes_per_wave(geom, n);
    PropArgs a{t.px, t.py, t.pz, t.dx, t.dy, t.dz, t.limit, t.nav, t.last, t.taken, n, t.nsteps, t.track_steps,
               nullptr, nullptr, 0, lanes, t.dn};
    VGMI_NAV_LAUNCH(k_propagate, geom, t.precision, lane_threads(n, lanes), st, geom->view, a);
    return launched("transport loop");
  }
  // Chunks of `chunk` steps; before each, the tracks still running are regrouped by the
  // volume they are in (a counting sort into perm; key 0, outside or done, is skipped).
  const int nbins  = geom->nlogical + 1;
  uint32_t *cursor = work, *perm = work + nbins;
  const size_t lds = nbins <= kLocalBins ? size_t(nbins) * sizeof(uint32_t) : 0;
  const unsigned hgrid = unsigned(std::min<size_t>((n + kSortBlock - 1) / kSortBlock, 2048));
  const unsigned sgrid = unsigned((n + kSortBlock - 1) / kSortBlock);
  const int lanes      = lanes_per_wave(geom, n);
  if (t.taken) hipLaunchKernelGGL(k_fill_u32, dim3(1024), dim3(256), 0, st, reinterpret_cast<uint32_t *>(t.taken), n, 0u);
  const SortKeys key{t.nav, t.track_steps, t.taken};
  for (int done = 0; done < t.nsteps; done += chunk) {
    hipLaunchKernelGGL(k_fill_u32, dim3(unsigned((nbins + 255) / 256)), dim3(256), 0, st, cursor, size_t(nbins), 0u);
    hipLaunchKernelGGL(k_sort_hist, dim3(hgrid), dim3(kSortBlock), lds, st, geom->view, key, n, t.dn, cursor, nbins);
    hipLaunchKernelGGL(k_sort_scan, dim3(1), dim3(1024), 0, st, cursor, nbins);
    hipLaunchKernelGGL(k_sort_scatter, dim3(sgrid), dim3(kSortBlock), lds, st, geom->view, key, n, t.dn, cursor, perm,
                       nbins);
    PropArgs a{t.px, t.py, t.pz, t.dx, t.dy, t.dz, t.limit, t.nav, t.last, t.taken, n, std::min(chunk, t.nsteps - done),
               t.track_steps, perm, cursor, 1, lanes, t.dn};
    VGMI_NAV_LAUNCH(k_propagate, geom, t.precision, lane_threads(n, lanes), st, geom->view, a);
  }
  return launched("transport loop");
}

vgmi_status launch_locate_init(const vgmi_geom *geom, int precision, const double *px, const double *py,
                               const double *pz, double *opx, double *opy, double *opz, uint32_t *nav,
                               uint32_t *last, int32_t *taken, size_t n, hipStream_t st, const uint32_t *dn)
{
  if (n == 0) return VGMI_OK;
  VGMI_NAV_LAUNCH(k_locate_init, geom, precision, n, st, geom->view, px, py, pz, opx, opy, opz, nav, last, taken, n,
                  dn);
  return launched("serving locate");
}

} // namespace vgmi

extern "C" {

vgmi_status vgmi_nav_locate(const vgmi_geom *geom, int precision, const double *px, const double *py,
                            const double *pz, uint32_t *nav_out, size_t n, void *stream)
{
  if (vgmi_status s = check_geom(geom, precision, "vgmi_nav_locate")) return s;
  if (n == 0) return VGMI_OK;
  if (!px || !py || !pz || !nav_out) return set_error(VGMI_ERR_INVALID, "vgmi_nav_locate: null pointer");
  VGMI_NAV_LAUNCH(k_locate, geom, precision, n, stream, geom->view, px, py, pz, nav_out, n);
  return launched("vgmi_nav_locate");
}

vgmi_status vgmi_nav_level_locate(const vgmi_geom *geom, int precision, const int32_t *logical,
                                  const int32_t *excl_placed, const double *px, const double *py, const double *pz,
                                  int32_t *daughter_out, double *dpx, double *dpy, double *dpz, size_t n, void *stream)
{
  if (vgmi_status s = check_geom(geom, precision, "vgmi_nav_level_locate")) return s;
  if (n == 0) return VGMI_OK;
  if (!logical || !px || !py || !pz || !daughter_out)
    return set_error(VGMI_ERR_INVALID, "vgmi_nav_level_locate: null pointer");
  if ((dpx || dpy || dpz) && !(dpx && dpy && dpz))
    return set_error(VGMI_ERR_INVALID, "vgmi_nav_level_locate: give all three daughter-frame arrays or none");
  LevelArgs a{logical, excl_placed, px, py, pz, daughter_out, dpx, dpy, dpz, n};
  VGMI_NAV_LAUNCH(k_level_locate, geom, precision, n, stream, geom->view, a);
  return launched("vgmi_nav_level_locate");
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
             relocate, nullptr, nullptr};
  VGMI_NAV_LAUNCH(k_step, geom, precision, n, stream, geom->view, a);
  return launched("vgmi_nav_step");
}

vgmi_status vgmi_nav_step_safety(const vgmi_geom *geom, int precision, const double *px, const double *py,
                                 const double *pz, const double *dx, const double *dy, const double *dz,
                                 const double *step_limit, const uint32_t *nav_in, const uint32_t *last_in,
                                 const uint8_t *calc_safety, uint32_t *nav_out, uint32_t *last_out,
                                 uint8_t *boundary_out, double *step_out, double *safety_out, size_t n, void *stream)
{
  if (vgmi_status s = check_geom(geom, precision, "vgmi_nav_step_safety")) return s;
  if (n == 0) return VGMI_OK;
  if (!px || !py || !pz || !dx || !dy || !dz || !step_limit || !nav_in || !nav_out || !step_out || !safety_out)
    return set_error(VGMI_ERR_INVALID, "vgmi_nav_step_safety: null pointer");
  StepArgs a{px, py, pz, dx, dy, dz, step_limit, nav_in, last_in, nav_out, last_out, boundary_out, step_out, n,
             1, calc_safety, safety_out};
  VGMI_NAV_LAUNCH(k_step, geom, precision, n, stream, geom->view, a);
  return launched("vgmi_nav_step_safety");
}

static unsigned grid256(size_t n) { return unsigned((n + 255) / 256); }

vgmi_status vgmi_nav_to_path(const vgmi_geom *geom, const uint32_t *nav, size_t n, int max_levels, int32_t *levels,
                             int32_t *pvols, int32_t *indices, void *stream)
{
  if (vgmi_status s = check_geom(geom, VGMI_FP64, "vgmi_nav_to_path")) return s;
  if (n == 0) return VGMI_OK;
  if (!nav || !levels) return set_error(VGMI_ERR_INVALID, "vgmi_nav_to_path: null pointer");
  if (max_levels < geom->host.maxDepth)
    return set_error(VGMI_ERR_CAPACITY, "vgmi_nav_to_path: max_levels %d below the geometry's depth %d", max_levels,
                     geom->host.maxDepth);
  hipLaunchKernelGGL(k_nav_to_path, dim3(grid256(n)), dim3(256), 0, (hipStream_t)stream, geom->view, nav, n,
                     max_levels, levels, pvols, indices);
  return launched("vgmi_nav_to_path");
}

vgmi_status vgmi_nav_from_path(const vgmi_geom *geom, int kind, const int32_t *levels, const int32_t *path,
                               size_t n, int max_levels, uint32_t *nav_out, void *stream)
{
  if (vgmi_status s = check_geom(geom, VGMI_FP64, "vgmi_nav_from_path")) return s;
  if (kind != VGMI_PATH_INDICES && kind != VGMI_PATH_PLACED)
    return set_error(VGMI_ERR_INVALID, "vgmi_nav_from_path: unknown path kind %d", kind);
  if (n == 0) return VGMI_OK;
  if (!levels || !path || !nav_out || max_levels < 1) return set_error(VGMI_ERR_INVALID, "vgmi_nav_from_path: null pointer or no levels");
  hipLaunchKernelGGL(k_nav_from_path, dim3(grid256(n)), dim3(256), 0, (hipStream_t)stream, geom->view, kind, levels,
                     path, n, max_levels, int(geom->host.pvols.size()), nav_out);
  return launched("vgmi_nav_from_path");
}

vgmi_status vgmi_nav_distance(const vgmi_geom *geom, const uint32_t *nav_a, const uint32_t *nav_b, size_t n,
                              int32_t *out, void *stream)
{
  if (vgmi_status s = check_geom(geom, VGMI_FP64, "vgmi_nav_distance")) return s;
  if (n == 0) return VGMI_OK;
  if (!nav_a || !nav_b || !out) return set_error(VGMI_ERR_INVALID, "vgmi_nav_distance: null pointer");
  hipLaunchKernelGGL(k_nav_distance, dim3(grid256(n)), dim3(256), 0, (hipStream_t)stream, geom->view, nav_a, nav_b, n,
                     out);
  return launched("vgmi_nav_distance");
}

vgmi_status vgmi_nav_global_to_local(const vgmi_geom *geom, const uint32_t *nav, int tolevel, const double *px,
                                     const double *py, const double *pz, double *lx, double *ly, double *lz, size_t n,
                                     void *stream)
{
  if (vgmi_status s = check_geom(geom, VGMI_FP64, "vgmi_nav_global_to_local")) return s;
  if (n == 0) return VGMI_OK;
  if (!nav || !px || !py || !pz || !lx || !ly || !lz)
    return set_error(VGMI_ERR_INVALID, "vgmi_nav_global_to_local: null pointer");
  hipLaunchKernelGGL(k_nav_global_to_local, dim3(grid256(n)), dim3(256), 0, (hipStream_t)stream, geom->view, nav,
                     tolevel, px, py, pz, lx, ly, lz, n);
  return launched("vgmi_nav_global_to_local");
}

vgmi_status vgmi_nav_safety(const vgmi_geom *geom, int precision, const double *px, const double *py,
                            const double *pz, const uint32_t *nav_in, double *safety_out, size_t n, void *stream)
{
  if (vgmi_status s = check_geom(geom, precision, "vgmi_nav_safety")) return s;
  if (n == 0) return VGMI_OK;
  if (!px || !py || !pz || !nav_in || !safety_out) return set_error(VGMI_ERR_INVALID, "vgmi_nav_safety: null pointer");
  VGMI_NAV_LAUNCH(k_safety, geom, precision, n, stream, geom->view, px, py, pz, nav_in, safety_out, n);
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
  if (n > 0xffffffffull) return set_error(VGMI_ERR_CAPACITY, "vgmi_nav_propagate: more than 2^32 tracks");
  hipStream_t st = (hipStream_t)stream;
  TransportArgs a{px, py, pz, dx, dy, dz, step_limit, nav, last, steps_taken, nullptr, n, nsteps, precision, nullptr};
  uint32_t *work = nullptr;
  if (const size_t words = transport_work_words(geom, n)) {
    if (hipMallocAsync(reinterpret_cast<void **>(&work), words * sizeof(uint32_t), st) != hipSuccess)
      return set_error(VGMI_ERR_DEVICE, "vgmi_nav_propagate: workspace of %zu words", words);
  }
  const vgmi_status rc = launch_transport(geom, a, work, st);
  if (work) (void)hipFreeAsync(work, st);
  return rc;
}

} // extern "C"

// ------------------------------------------------------------ transport loop without allocation
// Workspace of vgmi_nav_propagate_noalloc: the regrouping histogram and permutation (nlogical + 1 + n
// words) live in this fixed device array instead of a per-call allocation. Each device holds its own
// copy; host code addresses it through the pointer hipGetSymbolAddress returns for the current device
// (never through the array's name, which on the host is only a stand-in).
namespace {
constexpr size_t kNoAllocMaxTracks  = size_t(1) << 21;
constexpr size_t kNoAllocMaxLogical = size_t(1) << 16;
constexpr size_t kNoAllocWords      = kNoAllocMaxTracks + kNoAllocMaxLogical + 1;
constexpr int kMaxDevices           = 64;
__device__ __align__(256) uint32_t scratch[kNoAllocWords];
std::mutex g_scratch_mutex;
uint32_t *g_scratch_ptr[kMaxDevices] = {};

// This device's copy of `scratch`, resolved once per device with that device current.
uint32_t *scratch_for_current_device(vgmi_status &rc)
{
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    rc = set_error(VGMI_ERR_DEVICE, "vgmi_nav_propagate_noalloc: no current device (or device index >= %d)",
                   kMaxDevices);
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void *ptr           = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      rc = set_error(VGMI_ERR_DEVICE, "vgmi_nav_propagate_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_ptr[dev] = static_cast<uint32_t *>(ptr);
  }
  rc = VGMI_OK;
  return g_scratch_ptr[dev];
}
} // namespace

extern "C" {

vgmi_status vgmi_nav_propagate_noalloc(const vgmi_geom *geom, int precision, double *px, double *py, double *pz,
                                       const double *dx, const double *dy, const double *dz, double step_limit,
                                       uint32_t *nav, uint32_t *last, int32_t *steps_taken, size_t n, int nsteps,
                                       void *stream)
{
  if (vgmi_status s = check_geom(geom, precision, "vgmi_nav_propagate_noalloc")) return s;
  if (n == 0 || nsteps <= 0) return VGMI_OK;
  if (!px || !py || !pz || !dx || !dy || !dz || !nav || !last)
    return set_error(VGMI_ERR_INVALID, "vgmi_nav_propagate_noalloc: null pointer");
  if (n > kNoAllocMaxTracks)
    return set_error(VGMI_ERR_CAPACITY, "vgmi_nav_propagate_noalloc: %zu tracks, at most %zu", n, kNoAllocMaxTracks);
  const size_t words = transport_work_words(geom, n);
  if (words > kNoAllocWords)
    return set_error(VGMI_ERR_CAPACITY, "vgmi_nav_propagate_noalloc: %d logical volumes, at most %zu",
                     geom->nlogical, kNoAllocMaxLogical);
  uint32_t *work = nullptr;
  if (words) {
    vgmi_status rc = VGMI_OK;
    work           = scratch_for_current_device(rc);
    if (!work) return rc;
  }
  TransportArgs a{px, py, pz, dx, dy, dz, step_limit, nav, last, steps_taken, nullptr, n, nsteps, precision, nullptr};
  return launch_transport(geom, a, work, (hipStream_t)stream);
}

vgmi_status vgmi_geom_set_lanes_per_wave(vgmi_geom *geom, int lanes)
{
  if (!geom) return set_error(VGMI_ERR_INVALID, "vgmi_geom_set_lanes_per_wave: null geometry");
  if (lanes < 0 || lanes > 64 || (lanes & (lanes - 1)) != 0)
    return set_error(VGMI_ERR_INVALID, "vgmi_geom_set_lanes_per_wave: 0 (automatic) or a power of two up to 64");
  geom->lanesPerWave = lanes;
  return VGMI_OK;
}

vgmi_status vgmi_geom_set_regroup(vgmi_geom *geom, int chunk_steps)
{
  if (!geom) return set_error(VGMI_ERR_INVALID, "vgmi_geom_set_regroup: null geometry");
  if (chunk_steps < 0) return set_error(VGMI_ERR_INVALID, "vgmi_geom_set_regroup: negative chunk");
  geom->regroupSteps = chunk_steps;
  return VGMI_OK;
}

} // extern "C"
