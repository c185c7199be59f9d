// Tessellated solid kernels: VecGeom/volumes/kernel/TessellatedImplementation.h:54-300 with the
// one-facet clusters of the scalar backend (VecGeom/volumes/TessellatedCluster.h:232-283 for the
// ray, :498-540 for the safety, GenericKernels.h:224-249 for the edge distance).
//
// The reference visits the facet clusters in the order of their bounding-box hit distance and
// stops at the first box beyond the current distance (HybridNavigator2.h:111-190). Here the
// facets are reached through a threaded BVH built on the host (depth-first order, escape links,
// leaves of up to four facets) that only prunes what cannot change the result: a subtree is
// skipped when the ray misses its enlarged box, when the box lies behind the point, or when it
// starts beyond the current distance. Every facet that passes the reference's single-precision
// box test and can change the result is visited; among equal distances the lowest facet index
// is kept. The FP32 instantiation answers NaN, so the mixed path evaluates this solid in FP64.
#pragma once
#include "dev_common.h"
#include "shape_data.h"
#include <float.h>

namespace vgmi {
namespace tess {

using V3d = Vec3<double>;

VGMI_DEV const double *facet(TessellatedData const &s, int k) { return s.base + s.facetOff + k * kFacetDoubles; }
VGMI_DEV const double *node(TessellatedData const &s, int j) { return s.base + s.nodeOff + j * kNodeDoubles; }
VGMI_DEV V3d at3(const double *a) { return V3d(a[0], a[1], a[2]); }

// BoxImplementation::IntersectCachedKernel2<float, float> on the facet's single-precision box as
// HybridNavigator2::GetHitCandidates_v calls it (:111-150): t0 = 0, t1 = FLT_MAX.
VGMI_DEV bool box_candidate(const double *F, V3d const &p, V3d const &inv)
{
  const float *b = reinterpret_cast<const float *>(F); // lo x y z, hi x y z
  const int sx = inv.x < 0, sy = inv.y < 0, sz = inv.z < 0;
  const float px = float(p.x), py = float(p.y);
  float tmin  = float(double(b[3 * sx + 0] - px) * inv.x);
  float tymax = float(double(b[3 * (1 - sy) + 1] - py) * inv.y);
  if (tmin > tymax) return false;
  float tmax  = float(double(b[3 * (1 - sx) + 0] - px) * inv.x);
  float tymin = float(double(b[3 * sy + 1] - py) * inv.y);
  if (tymin > tmax) return false;
  tmin        = tmin > tymin ? tmin : tymin;
  tmax        = tmax < tymax ? tmax : tymax;
  float tzmin = float((double(b[3 * sz + 2]) - p.z) * inv.z);
  float tzmax = float((double(b[3 * (1 - sz) + 2]) - p.z) * inv.z);
  if (tmin > tzmax || tzmin > tmax) return false;
  tmin = tmin > tzmin ? tmin : tzmin;
  tmax = tmax < tzmax ? tmax : tzmax;
  if (!(tmin <= FLT_MAX && tmax > float(0.0 - kTolerance))) return false;
  return tmin < FLT_MAX;
}

// Ray against an enlarged node box: may the subtree hold a facet hit at a distance <= bound?
VGMI_DEV bool node_reached(const double *N, V3d const &p, V3d const &inv, double bound)
{
  double t0 = (N[0] - p.x) * inv.x, t1 = (N[3] - p.x) * inv.x;
  double u0 = (N[1] - p.y) * inv.y, u1 = (N[4] - p.y) * inv.y;
  double w0 = (N[2] - p.z) * inv.z, w1 = (N[5] - p.z) * inv.z;
  const double tmin = vmax(vmin(t0, t1), vmax(vmin(u0, u1), vmin(w0, w1)));
  const double tmax = vmin(vmax(t0, t1), vmin(vmax(u0, u1), vmax(w0, w1)));
  return tmin <= tmax * (1. + 1e-12) + 1e-12 && tmax > -1e-6 && !(tmin > bound * (1. + 1e-12) + 1e-12);
}

// Squared distance from p to a node box (0 inside).
VGMI_DEV double node_dist2(const double *N, V3d const &p)
{
  const double dx = vmax(vmax(N[0] - p.x, p.x - N[3]), 0.);
  const double dy = vmax(vmax(N[1] - p.y, p.y - N[4]), 0.);
  const double dz = vmax(vmax(N[2] - p.z, p.z - N[5]), 0.);
  return dx * dx + dy * dy + dz * dz;
}

VGMI_DEV int node_first(const double *N) { return reinterpret_cast<const int32_t *>(N + 6)[0]; }
VGMI_DEV int node_count(const double *N) { return reinterpret_cast<const int32_t *>(N + 6)[1]; }
VGMI_DEV int node_escape(const double *N) { return reinterpret_cast<const int32_t *>(N + 7)[0]; }

struct RayResult {
  double dIn, dOut;
  int sIn, sOut;        // facet slots (positions in the facet array), -1: none
  double idxIn, idxOut; // facet indices of those slots
};

// Facet loop of DistanceToSolid: the smallest entering and exiting hits (ties: lowest index).
// `primaryIn` selects which of the two bounds the search (the other only matters below it).
VGMI_DEV RayResult facet_hits(TessellatedData const &s, V3d const &p, V3d const &d, bool primaryIn)
{
  const V3d inv(1.0 / NonZero(d.x), 1.0 / NonZero(d.y), 1.0 / NonZero(d.z));
  RayResult r{kInfLength, kInfLength, -1, -1, 0., 0.};
  int j = 0;
  while (j < s.nnodes) {
    const double *N    = node(s, j);
    const double bound = primaryIn ? r.dIn : r.dOut;
    if (!node_reached(N, p, inv, bound)) {
      j = node_escape(N);
      continue;
    }
    const int cnt = node_count(N);
    if (cnt == 0) {
      ++j;
      continue;
    }
    const int first = node_first(N);
    for (int k = first; k < first + cnt; ++k) {
      const double *F = facet(s, k);
      if (!box_candidate(F, p, inv)) continue;
      const V3d n      = at3(F + 3);
      const double ndd = NonZero(d.Dot(n));
      const double saf = p.Dot(n) + F[6];
      bool vin         = ndd < 0. && saf > -kTolerance;
      bool vout        = ndd > 0. && saf < kTolerance;
      if (!vin && !vout) continue;
      const double dist = -saf / ndd;
      const V3d q       = p + dist * d;
      const bool hit    = (q - at3(F + 7)).Dot(at3(F + 10)) > -kTolerance &&
                       (q - at3(F + 13)).Dot(at3(F + 16)) > -kTolerance &&
                       (q - at3(F + 19)).Dot(at3(F + 22)) > -kTolerance;
      if (!hit) continue;
      const double idx = F[25];
      if (vin) {
        if (dist < r.dIn || (dist == r.dIn && idx < r.idxIn)) r.dIn = dist, r.sIn = k, r.idxIn = idx;
      } else {
        if (dist < r.dOut || (dist == r.dOut && idx < r.idxOut)) r.dOut = dist, r.sOut = k, r.idxOut = idx;
      }
    }
    j = node_escape(N);
  }
  return r;
}

// BoxImplementation::IntersectCachedKernel2<double, double> on the solid's extent,
// t0 = -kTolerance, t1 = kInfLength (TessellatedImplementation.h:181-192).
VGMI_DEV double extent_hit(TessellatedData const &s, V3d const &p, V3d const &d)
{
  const V3d inv(1.0 / NonZero(d.x), 1.0 / NonZero(d.y), 1.0 / NonZero(d.z));
  const int sx = inv.x < 0, sy = inv.y < 0, sz = inv.z < 0;
  const double *lo = s.minExt, *hi = s.maxExt;
  double tmin  = ((sx ? hi : lo)[0] - p.x) * inv.x;
  double tymax = ((sy ? lo : hi)[1] - p.y) * inv.y;
  if (tmin > tymax) return kInfLength;
  double tmax  = ((sx ? lo : hi)[0] - p.x) * inv.x;
  double tymin = ((sy ? hi : lo)[1] - p.y) * inv.y;
  if (tymin > tmax) return kInfLength;
  tmin         = tmin > tymin ? tmin : tymin;
  tmax         = tmax < tymax ? tmax : tymax;
  double tzmin = ((sz ? hi : lo)[2] - p.z) * inv.z;
  double tzmax = ((sz ? lo : hi)[2] - p.z) * inv.z;
  if (tmin > tzmax || tzmin > tmax) return kInfLength;
  tmin = tmin > tzmin ? tmin : tzmin;
  tmax = tmax < tzmax ? tmax : tzmax;
  if (!(tmin <= kInfLength + kTolerance && tmax > -kTolerance - kTolerance)) return kInfLength;
  return tmin;
}

VGMI_DEV double dot_normal(TessellatedData const &s, int slot, V3d const &d) { return d.Dot(at3(facet(s, slot) + 3)); }

// DistanceToSolid<ToIn = false> along the test direction (Contains / Inside)
VGMI_DEV RayResult test_ray(TessellatedData const &s, V3d const &p)
{
  return facet_hits(s, p, at3(s.testDir), false);
}

VGMI_DEV bool contains_of(TessellatedData const &s, RayResult const &r)
{
  if (r.sOut < 0) return false;
  const V3d t = at3(s.testDir);
  return !(r.sIn >= 0 && r.dIn < r.dOut && r.dIn * dot_normal(s, r.sIn, t) < -kTolerance);
}

VGMI_DEV bool contains64(TessellatedData const &s, V3d const &p) { return contains_of(s, test_ray(s, p)); }

VGMI_DEV int inside64(TessellatedData const &s, V3d const &p)
{
  const RayResult r = test_ray(s, p);
  const V3d t       = at3(s.testDir);
  // DistanceToSolid's own "outside" verdict resets the exit surface
  int sOut          = r.sOut;
  double dOut       = r.dOut;
  if (sOut >= 0 && r.sIn >= 0 && r.dIn < dOut && r.dIn * dot_normal(s, r.sIn, t) < -kTolerance) sOut = -1, dOut = -1.;
  if (sOut < 0) return kOutside;
  if (dOut < 0 || dOut * dot_normal(s, sOut, t) < kTolerance) return kSurface;
  if (r.sIn < 0 || dOut < r.dIn) return kInside;
  if (r.dIn < 0 || r.dIn * dot_normal(s, r.sIn, t) > -kTolerance) return kSurface;
  return kOutside;
}

VGMI_DEV double dist_in64(TessellatedData const &s, V3d const &p, V3d const &d)
{
  if (extent_hit(s, p, d) >= kInfLength) return kInfLength;
  const RayResult r = facet_hits(s, p, d, true);
  if (r.sIn < 0) {
    if (r.sOut >= 0 && r.dOut * dot_normal(s, r.sOut, d) > kTolerance) return -1.;
    return kInfLength;
  }
  if (r.sOut >= 0 && r.dOut > kTolerance && r.dOut < r.dIn) return -1.;
  return r.dIn;
}

VGMI_DEV double dist_out64(TessellatedData const &s, V3d const &p, V3d const &d)
{
  const RayResult r = facet_hits(s, p, d, false);
  if (r.sOut < 0) return -1.;
  if (r.sIn >= 0 && r.dIn < r.dOut && r.dIn * dot_normal(s, r.sIn, d) < -kTolerance) return -1.;
  return r.dOut;
}

VGMI_DEV double seg_dist2(V3d const &c0, V3d const &c1, V3d const &p)
{
  const V3d line    = c1 - c0;
  const double dot0 = (p - c0).Dot(line);
  if (dot0 <= 0) return (p - c0).Mag2();
  const double dot1 = line.Mag2();
  if (dot1 <= dot0) return (p - c1).Mag2();
  return ((c0 + (dot0 / dot1) * line) - p).Mag2();
}

// Squared safety of one facet (TessellatedCluster::SafetySq, TessellatedCluster.h:498-540 with one
// facet per cluster): the squared plane distance when the projection lies within the triangle on
// the right side, else the smallest squared distance to its edges. Facets whose squared plane
// distance exceeds `lim` are skipped (their value is at least that, since the projection and every
// edge point lie in the plane; the 1e-9 margin covers the rounding of both quantities).
VGMI_DEV void facet_safety_sq(const double *F, V3d const &p, bool toIn, double lim, double &best)
{
  const V3d n     = at3(F + 3);
  const double sp = p.Dot(n) + F[6];
  if (sp * sp > lim * (1. + 1e-9)) return;
  const V3d q  = p - sp * n;
  const V3d v0 = at3(F + 7), v1 = at3(F + 13), v2 = at3(F + 19);
  bool within = (q - v0).Dot(at3(F + 10)) > -kTolerance && (q - v1).Dot(at3(F + 16)) > -kTolerance &&
                (q - v2).Dot(at3(F + 22)) > -kTolerance;
  within &= toIn ? sp > -kTolerance : sp < kTolerance;
  double sq = sp * sp;
  if (!within) {
    const double e0 = seg_dist2(v0, v1, p), e1 = seg_dist2(v1, v2, p), e2 = seg_dist2(v2, v0, p);
    sq              = vmin(e0, vmin(e1, e2));
  }
  if (sq < best) best = sq;
}

// SafetySq<ToIn>: the smallest per-facet squared safety over all facets (the minimum does not
// depend on the visiting order). `bound` is an upper bound of that minimum; subtrees whose box is
// farther than the current minimum (or the bound) are skipped. A greedy descent into the nearer
// child seeds the minimum, then the threaded depth-first walk (each box read once) prunes against it.
VGMI_DEV double safety_sq_threaded(TessellatedData const &s, V3d const &p, bool toIn, double bound)
{
  double best = kInfLength;
  int j       = 0;
  while (j < s.nnodes) {
    const double *N = node(s, j);
    if (node_dist2(N, p) > vmin(best, bound)) {
      j = node_escape(N);
      continue;
    }
    const int cnt = node_count(N);
    if (cnt == 0) {
      ++j;
      continue;
    }
    const int first = node_first(N);
    for (int k = first; k < first + cnt; ++k) facet_safety_sq(facet(s, k), p, toIn, vmin(best, bound), best);
    j = node_escape(N);
  }
  return best;
}

// The facets of the leaf reached by always descending into the nearer child give a first minimum
// (an upper bound of the result), then the threaded walk prunes against it.
VGMI_DEV double safety_sq_seeded(TessellatedData const &s, V3d const &p, bool toIn, double bound)
{
  double seed = kInfLength;
  int j       = 0;
  for (;;) {
    const double *N = node(s, j);
    const int cnt   = node_count(N);
    if (cnt > 0) {
      const int first = node_first(N);
      for (int k = first; k < first + cnt; ++k) facet_safety_sq(facet(s, k), p, toIn, vmin(seed, bound), seed);
      break;
    }
    const int l = j + 1, r = node_escape(node(s, l));
    j = node_dist2(node(s, l), p) <= node_dist2(node(s, r), p) ? l : r;
  }
  return vmin(seed, safety_sq_threaded(s, p, toIn, vmin(seed, bound)));
}

VGMI_DEV double safety_sq(TessellatedData const &s, V3d const &p, bool toIn, double bound = kInfLength)
{
  return safety_sq_seeded(s, p, toIn, bound);
}

// An upper bound of the smallest per-facet squared safety. That minimum is at most the squared
// distance to any point of the surface (the facet holding the closest surface point yields it,
// by its plane or by its edges), so a vertex of facet 0, or the test-ray hit at distance t,
// bounds it; 1e-6 mm covers rounding and the tolerance of the hit test.
VGMI_DEV double safety_bound(TessellatedData const &s, V3d const &p, double tHit)
{
  double b = vsqrt((p - at3(facet(s, 0) + 7)).Mag2());
  if (tHit >= 0 && tHit < b) b = tHit;
  b += 1e-6;
  return b * b * (1. + 1e-9);
}

// All six methods of one point with one test-ray walk (Contains, Inside and both safeties),
// one safety walk (the other safety is -1) and the two directed walks.
VGMI_DEV void query_all(TessellatedData const &s, V3d const &p, V3d const &d, bool &contains, int &inside,
                        double &sin, double &sout, double &din, double &dout)
{
  const RayResult r = test_ray(s, p);
  const V3d t       = at3(s.testDir);
  int sOut          = r.sOut;
  double dOut       = r.dOut;
  if (sOut >= 0 && r.sIn >= 0 && r.dIn < dOut && r.dIn * dot_normal(s, r.sIn, t) < -kTolerance) sOut = -1, dOut = -1.;
  contains = sOut >= 0;
  if (sOut < 0)
    inside = kOutside;
  else if (dOut < 0 || dOut * dot_normal(s, sOut, t) < kTolerance)
    inside = kSurface;
  else if (r.sIn < 0 || dOut < r.dIn)
    inside = kInside;
  else if (r.dIn < 0 || r.dIn * dot_normal(s, r.sIn, t) > -kTolerance)
    inside = kSurface;
  else
    inside = kOutside;
  if (contains) {
    sin  = -1.;
    sout = vsqrt(safety_sq(s, p, false, safety_bound(s, p, r.dOut)));
  } else {
    sout = -1.;
    sin  = vsqrt(safety_sq(s, p, true, safety_bound(s, p, r.sIn >= 0 ? r.dIn : kInfLength)));
  }
  din  = dist_in64(s, p, d);
  dout = dist_out64(s, p, d);
}

template <class R>
VGMI_DEV bool Contains(TessellatedData const &s, Vec3<R> const &p)
{
  return contains64(s, V3d(p.x, p.y, p.z));
}

template <class R>
VGMI_DEV int Inside(TessellatedData const &s, Vec3<R> const &p)
{
  return inside64(s, V3d(p.x, p.y, p.z));
}

template <class R>
VGMI_DEV R DistanceToIn(TessellatedData const &s, Vec3<R> const &p, Vec3<R> const &d)
{
  if (sizeof(R) == sizeof(float)) return R(__builtin_nanf(""));
  return R(dist_in64(s, V3d(p.x, p.y, p.z), V3d(d.x, d.y, d.z)));
}

template <class R>
VGMI_DEV R DistanceToOut(TessellatedData const &s, Vec3<R> const &p, Vec3<R> const &d)
{
  if (sizeof(R) == sizeof(float)) return R(__builtin_nanf(""));
  return R(dist_out64(s, V3d(p.x, p.y, p.z), V3d(d.x, d.y, d.z)));
}

template <class R>
VGMI_DEV R SafetyToIn(TessellatedData const &s, Vec3<R> const &p)
{
  if (sizeof(R) == sizeof(float)) return R(__builtin_nanf(""));
  const V3d q(p.x, p.y, p.z);
  const RayResult r = test_ray(s, q);
  if (contains_of(s, r)) return R(-1.);
  return R(vsqrt(safety_sq(s, q, true, safety_bound(s, q, r.sIn >= 0 ? r.dIn : kInfLength))));
}

template <class R>
VGMI_DEV R SafetyToOut(TessellatedData const &s, Vec3<R> const &p)
{
  if (sizeof(R) == sizeof(float)) return R(__builtin_nanf(""));
  const V3d q(p.x, p.y, p.z);
  const RayResult r = test_ray(s, q);
  if (!contains_of(s, r)) return R(-1.);
  return R(vsqrt(safety_sq(s, q, false, safety_bound(s, q, r.dOut))));
}

} // namespace tess
} // namespace vgmi
