// Tessellated tiles shared by the batched launch and the single-solid launches: output
// pointers of a tile, the LDS budget of a solid's image, the copy of the image to LDS, and the
// tile loop with its queue of deferred safety searches.
#pragma once
#include "dev_shapes.h"

namespace vgmi {

struct BatchOut {
  const double *px, *py, *pz, *dx, *dy, *dz;
  uint8_t *contains;
  int32_t *inside;
  double *din, *dout, *sin, *sout;
};

// Copies the solid's image (imageDoubles doubles) to `lds`; called by all nthreads threads of the
// workgroup, followed by a barrier.
VGMI_DEV void stage_tess_image(TessellatedData const &s, double *lds, int nthreads)
{
  const int n       = s.imageDoubles;
  const double *src = s.base + s.imageOff;
  const double2 *s2 = reinterpret_cast<const double2 *>(src);
  double2 *d2       = reinterpret_cast<double2 *>(lds);
  for (int i = threadIdx.x; i < n / 2; i += nthreads) d2[i] = s2[i];
  if ((n & 1) && threadIdx.x == 0) lds[n - 1] = src[n - 1];
  __syncthreads();
}

// Points of a tessellated tile whose safety search is left to a wavefront (deep inside or far
// from the solid): offset in the tile, side, seed and bound of the search.
constexpr int kTessQueue = 128;
struct TessQueue {
  int n;
  int slot[kTessQueue]; // tile offset, bit 30 set for SafetyToIn
  double seed[kTessQueue], lim[kTessQueue];
};

// Dynamic LDS a workgroup may use for a tessellated solid's image: at most what keeps two
// workgroups on a CU (160 KiB of LDS per CU, less the kernel's static LDS -- tile index, staged
// solid, queue of deferred safety searches -- and the allocation granule); a solid whose image is
// larger is walked in the pool.
constexpr int kLdsPerCU   = 160 * 1024;
constexpr int kStaticLds  = int(sizeof(TessQueue)) + 16 + 512;
constexpr int kTessLdsMax = kLdsPerCU / 2 - kStaticLds;

// The tessellated solid shares one test-ray walk, one safety walk and one walk along d across its
// six methods; the walks read nodes, planes and candidate boxes through `w` (the pool or the LDS
// image). With Contains, Inside and both safeties, a lane whose point is deep (tess::deep_sq) queues its safety search
// and the workgroup's wavefronts run the queued searches together afterwards, each over all facets:
// a deep point's walk would reach most of the tree and hold its wavefront far longer than the
// others. Called by every thread of the workgroup; q->n is zero on entry.
template <class W>
VGMI_DEV void run_tile_tess(TessellatedData const &s, W const &w, int64_t b, int64_t e, BatchOut const &o,
                            int nthreads, TessQueue *q)
{
  const bool fused = o.contains && o.inside && o.sin && o.sout;
  if (fused) {
    const bool withDist = o.din && o.dout;
    const double deep   = tess::deep_sq(s);
    for (int64_t i = b + threadIdx.x; i < e; i += nthreads) {
      const Vec3<double> p(__builtin_nontemporal_load(o.px + i), __builtin_nontemporal_load(o.py + i),
                           __builtin_nontemporal_load(o.pz + i));
      Vec3<double> d(0., 0., 1.);
      if (withDist)
        d = Vec3<double>(__builtin_nontemporal_load(o.dx + i), __builtin_nontemporal_load(o.dy + i),
                         __builtin_nontemporal_load(o.dz + i));
      bool c, toIn;
      int in;
      double si, so, di, dou, seed, lim;
      if (tess::query_all_or_defer(s, w, p, d, c, in, si, so, di, dou, deep, seed, lim, toIn, withDist)) {
        const int k = atomicAdd(&q->n, 1);
        if (k < kTessQueue) {
          q->slot[k] = int(i - b) | (toIn ? 1 << 30 : 0);
          q->seed[k] = seed;
          q->lim[k]  = lim;
        } else { // queue full: the lane searches itself
          const double sq = vmin(seed, tess::safety_sq_threaded(s, w, p, toIn, lim));
          __builtin_nontemporal_store(toIn ? vsqrt(sq) : -1., o.sin + i);
          __builtin_nontemporal_store(toIn ? -1. : vsqrt(sq), o.sout + i);
        }
      } else {
        __builtin_nontemporal_store(si, o.sin + i);
        __builtin_nontemporal_store(so, o.sout + i);
      }
      __builtin_nontemporal_store(uint8_t(c), o.contains + i);
      __builtin_nontemporal_store(int32_t(in), o.inside + i);
      if (withDist) {
        __builtin_nontemporal_store(di, o.din + i);
        __builtin_nontemporal_store(dou, o.dout + i);
      }
    }
    __syncthreads();
    const int nq = vmin(q->n, kTessQueue);
    __syncthreads(); // every wave holds nq before q is reset for the next tile
    const int wave = threadIdx.x / 64, nwaves = nthreads / 64;
    for (int t = wave; t < nq; t += nwaves) {
      const int sl      = q->slot[t];
      const int64_t i   = b + (sl & ((1 << 30) - 1));
      const bool toIn   = (sl >> 30) & 1;
      const Vec3<double> p(o.px[i], o.py[i], o.pz[i]);
      const double sq   = vmin(q->seed[t], tess::safety_sq_wave(s, w, p, toIn, q->lim[t]));
      if (__lane_id() == 0) {
        o.sin[i]  = toIn ? vsqrt(sq) : -1.;
        o.sout[i] = toIn ? -1. : vsqrt(sq);
      }
    }
    if (withDist || (!o.din && !o.dout)) return;
  }
  // the methods one by one (after the fused walk: the one distance it left out)
  for (int64_t i = b + threadIdx.x; i < e; i += nthreads) {
    const Vec3<double> p(__builtin_nontemporal_load(o.px + i), __builtin_nontemporal_load(o.py + i),
                         __builtin_nontemporal_load(o.pz + i));
    if (!fused) {
      if (o.contains) __builtin_nontemporal_store(uint8_t(tess::contains64(s, w, p)), o.contains + i);
      if (o.inside) __builtin_nontemporal_store(int32_t(tess::inside64(s, w, p)), o.inside + i);
      if (o.sin) __builtin_nontemporal_store(tess::safety_in64(s, w, p), o.sin + i);
      if (o.sout) __builtin_nontemporal_store(tess::safety_out64(s, w, p), o.sout + i);
    }
    if (o.din && o.dout) {
      const Vec3<double> d(o.dx[i], o.dy[i], o.dz[i]);
      double din, dout;
      tess::dist_both64(s, w, p, d, din, dout);
      __builtin_nontemporal_store(din, o.din + i);
      __builtin_nontemporal_store(dout, o.dout + i);
    } else if (o.din || o.dout) {
      const Vec3<double> d(o.dx[i], o.dy[i], o.dz[i]);
      if (o.din) __builtin_nontemporal_store(tess::dist_in64(s, w, p, d), o.din + i);
      if (o.dout) __builtin_nontemporal_store(tess::dist_out64(s, w, p, d), o.dout + i);
    }
  }
}

} // namespace vgmi
