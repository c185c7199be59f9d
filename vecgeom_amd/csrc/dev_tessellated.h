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
VGMI_DEV V3d at3(const double *a) { return V3d(a[0], a[1], a[2]); }

// Where the walks read the node records, the facet planes (normal[3], offset) and the facets'
// single-precision candidate boxes: the solid's pool (PoolWalk), or the solid's image of them
// that a workgroup copied to LDS (LdsWalk). The image's node boxes are the pool's rounded
// outwards to single precision: a larger box passes every node test the pool's box passes, so
// the walk reaches every facet the pool walk reaches, and facets are decided by the same FP64
// values either way. Facet vertices and sides are always read from the pool, and only for facets
// that pass the plane and box tests.
// The walks read each record whole, before any test on it, so that its loads are in flight
// together instead of one after another behind the tests' early exits.
struct NodeRec {
  V3d lo, hi;
  int first, count, escape;
};
struct FBox {
  float lo[3], hi[3];
};
struct Plane {
  V3d n;
  double off;
};

VGMI_DEV Plane plane_at(const double *P) { return Plane{at3(P), P[3]}; }
VGMI_DEV FBox fbox_at(const float *b) { return FBox{{b[0], b[1], b[2]}, {b[3], b[4], b[5]}}; }

struct PoolWalk {
  const double *nodes, *facets;
  VGMI_DEV NodeRec node(int j) const
  {
    const double *N  = nodes + j * kNodeDoubles;
    const int32_t *I = reinterpret_cast<const int32_t *>(N + 6);
    return NodeRec{at3(N), at3(N + 3), I[0], I[1], I[2]};
  }
  VGMI_DEV Plane plane(int k) const { return plane_at(facets + k * kFacetDoubles + 3); }
  VGMI_DEV FBox fbox(int k) const { return fbox_at(reinterpret_cast<const float *>(facets + k * kFacetDoubles)); }
};

struct LdsWalk {
  const TessNode32 *nodes;
  const double *planes; // kImagePlaneDoubles per facet
  const float *fboxes;  // kImageBoxFloats per facet
  VGMI_DEV NodeRec node(int j) const
  {
    const TessNode32 n = nodes[j];
    return NodeRec{V3d(n.lo[0], n.lo[1], n.lo[2]), V3d(n.hi[0], n.hi[1], n.hi[2]), n.first, n.count, n.escape};
  }
  VGMI_DEV Plane plane(int k) const { return plane_at(planes + k * kImagePlaneDoubles); }
  VGMI_DEV FBox fbox(int k) const { return fbox_at(fboxes + k * kImageBoxFloats); }
};

VGMI_DEV PoolWalk pool_walk(TessellatedData const &s) { return PoolWalk{s.base + s.nodeOff, s.base + s.facetOff}; }

// The LDS walk over an image already copied to `img` (TessellatedData::imageDoubles doubles).
VGMI_DEV LdsWalk lds_walk(TessellatedData const &s, const double *img)
{
  return LdsWalk{reinterpret_cast<const TessNode32 *>(img), img + s.nnodes * kImageNodeDoubles,
                 reinterpret_cast<const float *>(img + s.nnodes * kImageNodeDoubles + s.nfacets * kImagePlaneDoubles)};
}

// BoxImplementation::IntersectCachedKernel2<float, float> on the facet's single-precision box as
// HybridNavigator2::GetHitCandidates_v calls it (:111-150): t0 = 0, t1 = maxstep, the step limit
// of the query as a float (FLT_MAX without one; BVHSortedIntersectionsLooper's limit,
// TessellatedImplementation.h:250).
// The reference's early exits are folded into one conjunction (same value: every test is a
// comparison of values computed the same way whether or not an earlier test failed).
VGMI_DEV bool box_candidate(FBox const &b, V3d const &p, V3d const &inv, float maxstep = FLT_MAX,
                            double stepMax = kInfLength)
{
  const bool sx = inv.x < 0, sy = inv.y < 0, sz = inv.z < 0;
  const float px = float(p.x), py = float(p.y);
  float tmin        = float(double((sx ? b.hi[0] : b.lo[0]) - px) * inv.x);
  const float tymax = float(double((sy ? b.lo[1] : b.hi[1]) - py) * inv.y);
  float tmax        = float(double((sx ? b.lo[0] : b.hi[0]) - px) * inv.x);
  const float tymin = float(double((sy ? b.hi[1] : b.lo[1]) - py) * inv.y);
  const bool xy     = !(tmin > tymax) && !(tymin > tmax);
  tmin              = tmin > tymin ? tmin : tymin;
  tmax              = tmax < tymax ? tmax : tymax;
  const float tzmin = float((double(sz ? b.hi[2] : b.lo[2]) - p.z) * inv.z);
  const float tzmax = float((double(sz ? b.lo[2] : b.hi[2]) - p.z) * inv.z);
  const bool z      = !(tmin > tzmax || tzmin > tmax);
  tmin              = tmin > tzmin ? tmin : tzmin;
  tmax              = tmax < tzmax ? tmax : tzmax;
  bool ok = xy;
  ok &= z;
  ok &= tmin <= float(double(maxstep) + kTolerance) && tmax > float(0.0 - kTolerance);
  ok &= tmin < maxstep;
  ok &= !(double(tmin) > stepMax); // the hook stops at boxes entered beyond stepMax (:213-216)
  return ok;
}

// Ray against an enlarged node box: may the subtree hold a facet hit at a distance <= bound?
VGMI_DEV bool node_reached(V3d const &lo, V3d const &hi, V3d const &p, V3d const &inv, double bound)
{
  double t0 = (lo.x - p.x) * inv.x, t1 = (hi.x - p.x) * inv.x;
  double u0 = (lo.y - p.y) * inv.y, u1 = (hi.y - p.y) * inv.y;
  double w0 = (lo.z - p.z) * inv.z, w1 = (hi.z - p.z) * inv.z;
  const double tmin = vmax(vmin(t0, t1), vmax(vmin(u0, u1), vmin(w0, w1)));
  const double tmax = vmin(vmax(t0, t1), vmin(vmax(u0, u1), vmax(w0, w1)));
  return tmin <= tmax * (1. + 1e-12) + 1e-12 && tmax > -1e-6 && !(tmin > bound * (1. + 1e-12) + 1e-12);
}

// Squared distance from p to a node box (0 inside).
VGMI_DEV double node_dist2(V3d const &lo, V3d const &hi, V3d const &p)
{
  const double dx = vmax(vmax(lo.x - p.x, p.x - hi.x), 0.);
  const double dy = vmax(vmax(lo.y - p.y, p.y - hi.y), 0.);
  const double dz = vmax(vmax(lo.z - p.z, p.z - hi.z), 0.);
  return dx * dx + dy * dy + dz * dz;
}


struct RayResult {
  double dIn, dOut;
  int sIn, sOut;        // facet slots (positions in the facet array), -1: none
  double idxIn, idxOut; // facet indices of those slots
};

// Facet loop of DistanceToSolid: the smallest entering and exiting hits (ties: lowest index).
// `mode` selects the bound of the search: the entering hit (kHitsIn; the exiting hit only matters
// below it), the exiting hit (kHitsOut; likewise), or both (kHitsBoth: a subtree is skipped only
// beyond both current hits, so both come out as the smallest over all facets, and each holds every
// value the single searches would give it where they use it).
constexpr int kHitsIn = 0, kHitsOut = 1, kHitsBoth = 2;

template <class W>
VGMI_DEV RayResult facet_hits(TessellatedData const &s, W const &w, V3d const &p, V3d const &d, int mode,
                              float maxstep = FLT_MAX, double stepMax = kInfLength)
{
  const V3d inv(1.0 / NonZero(d.x), 1.0 / NonZero(d.y), 1.0 / NonZero(d.z));
  RayResult r{kInfLength, kInfLength, -1, -1, 0., 0.};
  // a node entered beyond the step limit holds no candidate box (every facet box inside it is
  // entered later; the margin covers the float rounding of the candidate test)
  const double stepBound = maxstep < FLT_MAX ? double(maxstep) * (1. + 1e-5) + 1e-5 : kInfLength;
  // The walk alternates two phases: every lane first moves over skipped subtrees and inner nodes
  // until it stands on a leaf it must open (or has finished), then the lanes holding a leaf test
  // its facets together; a lane reaching a leaf early waits for the others instead of making the
  // whole wavefront run the facet tests once per lane.
  int j = 0;
  while (j < s.nnodes) {
    const double bound =
        vmin(stepBound, mode == kHitsIn ? r.dIn : (mode == kHitsOut ? r.dOut : vmax(r.dIn, r.dOut)));
    NodeRec nd  = w.node(j);
    bool atLeaf = false;
    for (;;) {
      if (!node_reached(nd.lo, nd.hi, p, inv, bound))
        j = nd.escape;
      else if (nd.count == 0)
        ++j;
      else {
        atLeaf = true;
        break;
      }
      if (j >= s.nnodes) break;
      nd = w.node(j);
    }
    if (!atLeaf) break;
    for (int k = nd.first; k < nd.first + nd.count; ++k) {
      // The candidate box, side and triangle tests are pure predicates, taken here plane first.
      // A hit farther than the current one of its kind cannot replace it (equal distances
      // are decided by the facet index below).
      const Plane P    = w.plane(k);
      const FBox B     = w.fbox(k);
      const double ndd = NonZero(d.Dot(P.n));
      const double saf = p.Dot(P.n) + P.off;
      bool vin         = ndd < 0. && saf > -kTolerance;
      bool vout        = ndd > 0. && saf < kTolerance;
      if (!vin && !vout) continue;
      const double dist = -saf / ndd;
      if (vin ? dist > r.dIn : dist > r.dOut) continue;
      if (!box_candidate(B, p, inv, maxstep, stepMax)) continue;
      const double *F = facet(s, k);
      const V3d v0 = at3(F + 7), s0 = at3(F + 10), v1 = at3(F + 13), s1 = at3(F + 16), v2 = at3(F + 19),
                s2     = at3(F + 22);
      const double idx = F[25];
      const V3d q      = p + dist * d;
      bool hit         = (q - v0).Dot(s0) > -kTolerance;
      hit &= (q - v1).Dot(s1) > -kTolerance;
      hit &= (q - v2).Dot(s2) > -kTolerance;
      if (!hit) continue;
      if (vin) {
        if (dist < r.dIn || (dist == r.dIn && idx < r.idxIn)) r.dIn = dist, r.sIn = k, r.idxIn = idx;
      } else {
        if (dist < r.dOut || (dist == r.dOut && idx < r.idxOut)) r.dOut = dist, r.sOut = k, r.idxOut = idx;
      }
    }
    j = nd.escape;
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
template <class W>
VGMI_DEV RayResult test_ray(TessellatedData const &s, W const &w, V3d const &p)
{
  return facet_hits(s, w, p, at3(s.testDir), kHitsOut);
}

VGMI_DEV bool contains_of(TessellatedData const &s, RayResult const &r)
{
  if (r.sOut < 0) return false;
  const V3d t = at3(s.testDir);
  return !(r.sIn >= 0 && r.dIn < r.dOut && r.dIn * dot_normal(s, r.sIn, t) < -kTolerance);
}

template <class W>
VGMI_DEV bool contains64(TessellatedData const &s, W const &w, V3d const &p)
{
  return contains_of(s, test_ray(s, w, p));
}

template <class W>
VGMI_DEV int inside64(TessellatedData const &s, W const &w, V3d const &p)
{
  const RayResult r = test_ray(s, w, p);
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

// DistanceToSolid<ToIn = true / false> results from the hits of a search along d.
VGMI_DEV double dist_in_of(TessellatedData const &s, RayResult const &r, V3d const &d)
{
  if (r.sIn < 0) {
    if (r.sOut >= 0 && r.dOut * dot_normal(s, r.sOut, d) > kTolerance) return -1.;
    return kInfLength;
  }
  if (r.sOut >= 0 && r.dOut > kTolerance && r.dOut < r.dIn) return -1.;
  return r.dIn;
}

VGMI_DEV double dist_out_of(TessellatedData const &s, RayResult const &r, V3d const &d)
{
  if (r.sOut < 0) return -1.;
  if (r.sIn >= 0 && r.dIn < r.dOut && r.dIn * dot_normal(s, r.sIn, d) < -kTolerance) return -1.;
  return r.dOut;
}

template <class W>
VGMI_DEV double dist_in64(TessellatedData const &s, W const &w, V3d const &p, V3d const &d)
{
  if (extent_hit(s, p, d) >= kInfLength) return kInfLength;
  return dist_in_of(s, facet_hits(s, w, p, d, kHitsIn), d);
}

template <class W>
VGMI_DEV double dist_out64(TessellatedData const &s, W const &w, V3d const &p, V3d const &d)
{
  return dist_out_of(s, facet_hits(s, w, p, d, kHitsOut), d);
}

// DistanceToSolid with a finite step limit (TessellatedImplementation.h:191-199, 213-216, 250): the
// extent box distance is returned at once when it is not below the limit (DistanceToIn), and facet
// boxes entered beyond the limit are no candidates.
template <class W>
VGMI_DEV double dist_in_step64(TessellatedData const &s, W const &w, V3d const &p, V3d const &d, double stepMax)
{
  const double e = extent_hit(s, p, d);
  if (e >= stepMax) return e;
  const float maxstep = stepMax < double(FLT_MAX) ? float(stepMax) : FLT_MAX;
  return dist_in_of(s, facet_hits(s, w, p, d, kHitsIn, maxstep, stepMax), d);
}

template <class W>
VGMI_DEV double dist_out_step64(TessellatedData const &s, W const &w, V3d const &p, V3d const &d, double stepMax)
{
  const float maxstep = stepMax < double(FLT_MAX) ? float(stepMax) : FLT_MAX;
  return dist_out_of(s, facet_hits(s, w, p, d, kHitsOut, maxstep, stepMax), d);
}

// Both distances along d from one search.
template <class W>
VGMI_DEV void dist_both64(TessellatedData const &s, W const &w, V3d const &p, V3d const &d, double &din, double &dout)
{
  const RayResult r = facet_hits(s, w, p, d, kHitsBoth);
  din               = extent_hit(s, p, d) >= kInfLength ? kInfLength : dist_in_of(s, r, d);
  dout              = dist_out_of(s, r, d);
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
// edge point lie in the plane; the 1e-9 margin covers the rounding of both quantities). P is the
// facet's plane, F its record.
VGMI_DEV void facet_safety_sq(Plane const &P, const double *F, V3d const &p, bool toIn, double lim, double &best)
{
  const V3d n     = P.n;
  const double sp = p.Dot(n) + P.off;
  if (sp * sp > lim * (1. + 1e-9)) return;
  const V3d q  = p - sp * n;
  const V3d v0 = at3(F + 7), s0 = at3(F + 10), v1 = at3(F + 13), s1 = at3(F + 16), v2 = at3(F + 19),
            s2 = at3(F + 22);
  bool within = (q - v0).Dot(s0) > -kTolerance;
  within &= (q - v1).Dot(s1) > -kTolerance;
  within &= (q - v2).Dot(s2) > -kTolerance;
  within &= toIn ? sp > -kTolerance : sp < kTolerance;
  double sq = sp * sp;
  if (!within) {
    const double e0 = seg_dist2(v0, v1, p), e1 = seg_dist2(v1, v2, p), e2 = seg_dist2(v2, v0, p);
    sq              = vmin(e0, vmin(e1, e2));
  }
  if (sq < best) best = sq;
}

// The facets of one leaf, in order.
template <class W>
VGMI_DEV void leaf_safety_sq(TessellatedData const &s, W const &w, NodeRec const &nd, V3d const &p, bool toIn,
                             double bound, double &best)
{
  for (int k = nd.first; k < nd.first + nd.count; ++k)
    facet_safety_sq(w.plane(k), facet(s, k), p, toIn, vmin(best, bound), best);
}

// SafetySq<ToIn>: the smallest per-facet squared safety over all facets (the minimum does not
// depend on the visiting order). `bound` is an upper bound of that minimum; subtrees whose box is
// farther than the current minimum (or the bound) are skipped. A greedy descent into the nearer
// child seeds the minimum, then the threaded depth-first walk (each box read once) prunes against it.
template <class W>
VGMI_DEV double safety_sq_threaded(TessellatedData const &s, W const &w, V3d const &p, bool toIn, double bound)
{
  double best = kInfLength;
  int j       = 0;
  // two phases per round, as in facet_hits: advance to the next leaf to open, then open it
  while (j < s.nnodes) {
    const double lim = vmin(best, bound);
    NodeRec nd       = w.node(j);
    bool atLeaf      = false;
    for (;;) {
      if (node_dist2(nd.lo, nd.hi, p) > lim)
        j = nd.escape;
      else if (nd.count == 0)
        ++j;
      else {
        atLeaf = true;
        break;
      }
      if (j >= s.nnodes) break;
      nd = w.node(j);
    }
    if (!atLeaf) break;
    leaf_safety_sq(s, w, nd, p, toIn, bound, best);
    j = nd.escape;
  }
  return best;
}

// SafetySq<false> together with the facet that gives it, for TessellatedImplementation::NormalKernel
// (TessellatedImplementation.h:162-172): the walk of safety_sq_threaded over the pool, keeping among
// equal squared safeties the lowest facet index. Returns the facet's slot in the pool (-1 when the
// solid has no facet).
VGMI_DEV int nearest_facet_slot(TessellatedData const &s, V3d const &p, bool toIn)
{
  const PoolWalk w = pool_walk(s);
  double best      = kInfLength;
  int bestSlot = -1, bestIdx = 0x7fffffff;
  int j = 0;
  while (j < s.nnodes) {
    const NodeRec nd = w.node(j);
    if (node_dist2(nd.lo, nd.hi, p) > best) {
      j = nd.escape;
      continue;
    }
    if (nd.count == 0) {
      ++j;
      continue;
    }
    for (int k = nd.first; k < nd.first + nd.count; ++k) {
      double sq = kInfLength;
      facet_safety_sq(w.plane(k), facet(s, k), p, toIn, best, sq);
      const int idx = int(facet(s, k)[25]);
      if (sq < best || (sq == best && sq < kInfLength && idx < bestIdx)) best = sq, bestSlot = k, bestIdx = idx;
    }
    j = nd.escape;
  }
  return bestSlot;
}

// The facets of the leaf reached by always descending into the nearer child give a first minimum
// (an upper bound of the result) that the full search prunes against.
template <class W>
VGMI_DEV double safety_sq_seed(TessellatedData const &s, W const &w, V3d const &p, bool toIn, double bound)
{
  double seed = kInfLength;
  int j       = 0;
  NodeRec nd  = w.node(0);
  for (;;) {
    if (nd.count > 0) {
      leaf_safety_sq(s, w, nd, p, toIn, bound, seed);
      break;
    }
    const NodeRec l = w.node(j + 1), r = w.node(l.escape);
    const bool left = node_dist2(l.lo, l.hi, p) <= node_dist2(r.lo, r.hi, p);
    j               = left ? j + 1 : l.escape;
    nd              = left ? l : r;
  }
  return seed;
}

template <class W>
VGMI_DEV double safety_sq_seeded(TessellatedData const &s, W const &w, V3d const &p, bool toIn, double bound)
{
  const double seed = safety_sq_seed(s, w, p, toIn, bound);
  return vmin(seed, safety_sq_threaded(s, w, p, toIn, vmin(seed, bound)));
}

// The smallest per-facet squared safety of p, computed by the 64 lanes of a wavefront together
// (lane l takes facets l, l + 64, ...; every lane must call it with the same p, toIn and lim);
// every lane returns it. `lim` bounds it from above (facets whose plane lies beyond are skipped).
template <class W>
VGMI_DEV double safety_sq_wave(TessellatedData const &s, W const &w, V3d const &p, bool toIn, double lim)
{
  double best = kInfLength;
  for (int k = int(__lane_id()); k < s.nfacets; k += 64)
    facet_safety_sq(w.plane(k), facet(s, k), p, toIn, vmin(best, lim), best);
  for (int off = 32; off > 0; off >>= 1) best = vmin(best, __shfl_xor(best, off));
  return best;
}

// Squared seeded minimum beyond which a point counts as deep inside or far from the solid: the
// walk's boxes bound facet distances poorly there, so such a point's search is left to a wavefront
// (two fifths of the solid's smallest extent; lower thresholds send points to the wavefront
// search that the per-lane walk finishes sooner).
VGMI_DEV double deep_sq(TessellatedData const &s)
{
  const double e =
      0.4 * vmin(s.maxExt[0] - s.minExt[0], vmin(s.maxExt[1] - s.minExt[1], s.maxExt[2] - s.minExt[2]));
  return e * e;
}

template <class W>
VGMI_DEV double safety_sq(TessellatedData const &s, W const &w, V3d const &p, bool toIn, double bound = kInfLength)
{
  return safety_sq_seeded(s, w, p, toIn, bound);
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

// query_all, except that a point whose seeded safety minimum exceeds `deep` returns true with
// sin and sout unset and its search left to the caller: the result is then vmin(seed, the smallest
// per-facet squared safety below lim) on the side toIn (safety_sq_wave computes it). Without
// withDist the walk along d is left out (din and dout unset).
template <class W>
VGMI_DEV bool query_all_or_defer(TessellatedData const &s, W const &w, V3d const &p, V3d const &d, bool &contains,
                                 int &inside, double &sin, double &sout, double &din, double &dout, double deep,
                                 double &seed, double &lim, bool &toIn, bool withDist = true)
{
  const RayResult r = test_ray(s, w, p);
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
  toIn = !contains;
  const double bound =
      contains ? safety_bound(s, p, r.dOut) : safety_bound(s, p, r.sIn >= 0 ? r.dIn : kInfLength);
  seed = safety_sq_seed(s, w, p, toIn, bound);
  lim  = vmin(seed, bound);
  if (withDist) dist_both64(s, w, p, d, din, dout);
  if (lim > deep) return true;
  const double sq = vmin(seed, safety_sq_threaded(s, w, p, toIn, lim));
  sin             = contains ? -1. : vsqrt(sq);
  sout            = contains ? vsqrt(sq) : -1.;
  return false;
}

// All six methods of one point with one test-ray walk (Contains, Inside and both safeties),
// one safety walk (the other safety is -1) and one walk along d for both distances.
template <class W>
VGMI_DEV void query_all(TessellatedData const &s, W const &w, V3d const &p, V3d const &d, bool &contains,
                        int &inside, double &sin, double &sout, double &din, double &dout)
{
  double seed, lim;
  bool toIn;
  (void)query_all_or_defer(s, w, p, d, contains, inside, sin, sout, din, dout, __builtin_inf(), seed, lim, toIn);
}

template <class W>
VGMI_DEV double safety_in64(TessellatedData const &s, W const &w, V3d const &q)
{
  const RayResult r = test_ray(s, w, q);
  if (contains_of(s, r)) return -1.;
  return vsqrt(safety_sq(s, w, q, true, safety_bound(s, q, r.sIn >= 0 ? r.dIn : kInfLength)));
}

template <class W>
VGMI_DEV double safety_out64(TessellatedData const &s, W const &w, V3d const &q)
{
  const RayResult r = test_ray(s, w, q);
  if (!contains_of(s, r)) return -1.;
  return vsqrt(safety_sq(s, w, q, false, safety_bound(s, q, r.dOut)));
}

template <class R>
VGMI_DEV bool Contains(TessellatedData const &s, Vec3<R> const &p)
{
  return contains64(s, pool_walk(s), V3d(p.x, p.y, p.z));
}

template <class R>
VGMI_DEV int Inside(TessellatedData const &s, Vec3<R> const &p)
{
  return inside64(s, pool_walk(s), V3d(p.x, p.y, p.z));
}

template <class R>
VGMI_DEV R DistanceToIn(TessellatedData const &s, Vec3<R> const &p, Vec3<R> const &d)
{
  if (sizeof(R) == sizeof(float)) return R(__builtin_nanf(""));
  return R(dist_in64(s, pool_walk(s), V3d(p.x, p.y, p.z), V3d(d.x, d.y, d.z)));
}

template <class R>
VGMI_DEV R DistanceToOut(TessellatedData const &s, Vec3<R> const &p, Vec3<R> const &d)
{
  if (sizeof(R) == sizeof(float)) return R(__builtin_nanf(""));
  return R(dist_out64(s, pool_walk(s), V3d(p.x, p.y, p.z), V3d(d.x, d.y, d.z)));
}

template <class R>
VGMI_DEV R SafetyToIn(TessellatedData const &s, Vec3<R> const &p)
{
  if (sizeof(R) == sizeof(float)) return R(__builtin_nanf(""));
  return R(safety_in64(s, pool_walk(s), V3d(p.x, p.y, p.z)));
}

template <class R>
VGMI_DEV R SafetyToOut(TessellatedData const &s, Vec3<R> const &p)
{
  if (sizeof(R) == sizeof(float)) return R(__builtin_nanf(""));
  return R(safety_out64(s, pool_walk(s), V3d(p.x, p.y, p.z)));
}

} // namespace tess
} // namespace vgmi
