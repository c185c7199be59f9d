// Tessellated solid restated from VecGeom/volumes/Tile.h, TessellatedCluster.h, TessellatedStruct.h
// and kernel/TessellatedImplementation.h (scalar backend: kVecSize = 1, one facet per cluster).
// TEST INFRASTRUCTURE ONLY (see common.h).
#include "shapes.h"
#include <cfloat>
#include <cmath>

namespace vgo {

static V3 Normalized(V3 const &v) { return v * (1. / v.Mag()); } // Vector3D.h:202

// Tile<3>::AddVertex (Tile.h:73-140): normal from the first two non-degenerate edges, plane offset
// from the first non-degenerate edge; TessellatedCluster::AddFacet (:170-200) recomputes the side
// vectors as normal x (v[i+1] - v[i]), normalised, for every edge.
static bool MakeFacet(V3 const &a, V3 const &b, V3 const &c, TriFacet &t)
{
  t.v[0] = a, t.v[1] = b, t.v[2] = c;
  int nvert = 3;
  for (int i = 0; i < 3; ++i)
    if ((t.v[(i + 1) % 3] - t.v[i]).Mag2() < kTolerance) nvert--;
  if (nvert < 3) return false;
  bool degenerated = true;
  for (int i = 0; i < 2 && degenerated; ++i) {
    V3 e1 = t.v[i + 1] - t.v[i];
    if (e1.Mag2() < kTolerance) continue;
    for (int j = i + 1; j < 3; ++j) {
      V3 e2 = t.v[(j + 1) % 3] - t.v[j];
      if (e2.Mag2() < kTolerance) continue;
      t.n = e1.Cross(e2);
      if (t.n.Mag2() < kTolerance) continue;
      t.n         = t.n * (1. / t.n.Mag()); // Normalize(), Vector3D.h:198
      degenerated = false;
      break;
    }
  }
  if (degenerated) return false;
  for (int i = 0; i < 3; ++i) {
    V3 e1 = t.v[(i + 1) % 3] - t.v[i];
    if (e1.Mag2() < kTolerance) continue;
    t.dist = -t.n.Dot(t.v[i]);
    break;
  }
  for (int i = 0; i < 3; ++i) t.side[i] = Normalized(t.n.Cross(t.v[(i + 1) % 3] - t.v[i]));
  for (int k = 0; k < 3; ++k) {
    double lo = t.v[0][k], hi = t.v[0][k];
    for (int i = 1; i < 3; ++i) {
      if (t.v[i][k] < lo) lo = t.v[i][k];
      if (t.v[i][k] > hi) hi = t.v[i][k];
    }
    t.blo[k] = float(lo), t.bhi[k] = float(hi);
  }
  return true;
}

V3 TessellatedTestDir(std::vector<TriFacet> const &f)
{
  static const double cand[][3] = {{0.3917183, 0.5276331, 0.7537210},
                                   {-0.6417143, 0.2311529, 0.7312417},
                                   {0.1873307, -0.8215501, 0.5384472},
                                   {0.7741139, 0.4109378, -0.4813220}};
  for (auto const &c : cand) {
    const V3 d = Normalized(V3(c[0], c[1], c[2]));
    bool good  = true;
    for (auto const &t : f)
      if (std::fabs(t.n.Dot(d)) < 1.e-8) good = false;
    if (good) return d;
  }
  return Normalized(V3(cand[0][0], cand[0][1], cand[0][2]));
}

Tessellated::Tessellated(const double *tri, int ntri)
{
  minExt = V3(kInfLength, kInfLength, kInfLength);
  maxExt = V3(-kInfLength, -kInfLength, -kInfLength);
  for (int i = 0; i < ntri; ++i) {
    const double *q = tri + 9 * i;
    TriFacet t;
    if (!MakeFacet(V3(q[0], q[1], q[2]), V3(q[3], q[4], q[5]), V3(q[6], q[7], q[8]), t)) continue;
    f.push_back(t);
    for (int k = 0; k < 3; ++k) // TessellatedStruct::AddFacet, :215-230
      for (int j = 0; j < 3; ++j) {
        minExt[j] = std::min(minExt[j], t.v[k][j]);
        maxExt[j] = std::max(maxExt[j], t.v[k][j]);
      }
  }
  testDir = TessellatedTestDir(f);
  ok      = !f.empty();
}

// BoxImplementation::IntersectCachedKernel2<double,double> on (fMinExtent, fMaxExtent),
// BoxImplementation.h:326-368, with t0 = -kTolerance, t1 = kInfLength.
static double BoxHit(V3 const &lo, V3 const &hi, V3 const &p, V3 const &inv, double t0, double t1)
{
  const V3 c[2] = {lo, hi};
  const int sx = inv.x < 0, sy = inv.y < 0, sz = inv.z < 0;
  double tmin  = (c[sx].x - p.x) * inv.x;
  double tymax = (c[1 - sy].y - p.y) * inv.y;
  if (tmin > tymax) return kInfLength;
  double tmax  = (c[1 - sx].x - p.x) * inv.x;
  double tymin = (c[sy].y - p.y) * inv.y;
  if (tymin > tmax) return kInfLength;
  tmin         = tmin > tymin ? tmin : tymin;
  tmax         = tmax < tymax ? tmax : tymax;
  double tzmin = (c[sz].z - p.z) * inv.z;
  double tzmax = (c[1 - sz].z - p.z) * inv.z;
  if (tmin > tzmax || tzmin > tmax) return kInfLength;
  tmin = tmin > tzmin ? tmin : tzmin;
  tmax = tmax < tzmax ? tmax : tzmax;
  if (!(tmin <= t1 + kTolerance && tmax > t0 - kTolerance)) return kInfLength;
  return tmin;
}

// The same kernel instantiated <float, float> on a facet's single-precision box, as
// HybridNavigator2::GetHitCandidates_v calls it (:111-150): x and y differences in float,
// the z difference in double, each product by the double inverse direction rounded to float;
// t0 = 0, t1 = maxstep (float). Returns whether the box is a candidate.
// The user hook then stops at the first box (in distance order) entered beyond
// min(stepMax, distance) (TessellatedImplementation.h:213-216): a box entered beyond stepMax is
// no candidate (boxes beyond the current distance cannot change the result).
static bool FacetBoxCandidate(TriFacet const &t, V3 const &p, V3 const &inv, float maxstep, double stepMax)
{
  const int sx = inv.x < 0, sy = inv.y < 0, sz = inv.z < 0;
  const float cx[2] = {t.blo[0], t.bhi[0]}, cy[2] = {t.blo[1], t.bhi[1]}, cz[2] = {t.blo[2], t.bhi[2]};
  float tmin  = float(double(cx[sx] - float(p.x)) * inv.x);
  float tymax = float(double(cy[1 - sy] - float(p.y)) * inv.y);
  if (tmin > tymax) return false;
  float tmax  = float(double(cx[1 - sx] - float(p.x)) * inv.x);
  float tymin = float(double(cy[sy] - float(p.y)) * inv.y);
  if (tymin > tmax) return false;
  tmin        = tmin > tymin ? tmin : tymin;
  tmax        = tmax < tymax ? tmax : tymax;
  float tzmin = float((double(cz[sz]) - p.z) * inv.z);
  float tzmax = float((double(cz[1 - sz]) - p.z) * inv.z);
  if (tmin > tzmax || tzmin > tmax) return false;
  tmin = tmin > tzmin ? tmin : tzmin;
  tmax = tmax < tzmax ? tmax : tzmax;
  if (!(tmin <= float(double(maxstep) + kTolerance) && tmax > float(0.0 - kTolerance))) return false;
  return tmin < maxstep && !(double(tmin) > stepMax); // "hit = distance < maxstep", then the hook
}

// TessellatedCluster::DistanceToCluster for one facet (:232-283)
static void DistanceToFacet(TriFacet const &t, V3 const &p, V3 const &d, double &toIn, double &toOut, bool &vin,
                            bool &vout)
{
  const double ndd = NonZero(d.Dot(t.n));
  const double saf = p.Dot(t.n) + t.dist;
  vin              = ndd < 0. && saf > -kTolerance;
  vout             = ndd > 0. && saf < kTolerance;
  const double dist = -saf / ndd;
  const V3 q        = p + dist * d;
  bool hit          = true;
  for (int i = 0; i < 3; ++i) hit &= (q - t.v[i]).Dot(t.side[i]) > -kTolerance;
  vin &= hit;
  vout &= hit;
  toIn = toOut = dist;
}

void Tessellated::DistanceToSolid(bool toIn, V3 const &p, V3 const &d, double stepMax, double &distance, int &isurf,
                                  double &distother, int &isurfother) const
{
  isurf = isurfother = -1;
  const V3 inv(1.0 / NonZero(d.x), 1.0 / NonZero(d.y), 1.0 / NonZero(d.z));
  if (toIn) {
    distance = BoxHit(minExt, maxExt, p, inv, -kTolerance, kInfLength);
    if (distance >= stepMax) return;
  }
  distance = distother = kInfLength;
  double dIn = kInfLength, dOut = kInfLength;
  int sIn = -1, sOut = -1;
  const float maxstep = std::min(float(stepMax), FLT_MAX);
  for (int k = 0; k < int(f.size()); ++k) {
    if (!FacetBoxCandidate(f[k], p, inv, maxstep, stepMax)) continue;
    double ci, co;
    bool vin, vout;
    DistanceToFacet(f[k], p, d, ci, co, vin, vout);
    if (vin) {
      if (ci < dIn) dIn = ci, sIn = k;
    } else if (vout) {
      if (co < dOut) dOut = co, sOut = k;
    }
  }
  if (toIn) {
    distance = dIn, isurf = sIn, distother = dOut, isurfother = sOut;
    if (sIn < 0) {
      if (sOut >= 0 && dOut * d.Dot(f[sOut].n) > kTolerance) distance = -1.;
    } else {
      if (sOut >= 0 && dOut > kTolerance && dOut < dIn) distance = -1.;
    }
  } else {
    distance = dOut, isurf = sOut, distother = dIn, isurfother = sIn;
    if (sOut < 0)
      distance = -1.;
    else if (sIn >= 0 && dIn < dOut && dIn * d.Dot(f[sIn].n) < -kTolerance) {
      distance = -1.;
      isurf    = -1;
    }
  }
}

// GenericKernels.h:224-249 (one lane)
static double SegmentDistSq(V3 const &c0, V3 const &c1, V3 const &p)
{
  const V3 line = c1 - c0;
  const double dot0 = (p - c0).Dot(line);
  if (dot0 <= 0) return (p - c0).Mag2();
  const double dot1 = line.Mag2();
  if (dot1 <= dot0) return (p - c1).Mag2();
  return ((c0 + (dot0 / dot1) * line) - p).Mag2();
}

// TessellatedCluster::SafetySq<ToIn> for one facet (:498-540), minimised over all facets
double Tessellated::SafetySq(bool toIn, V3 const &p, int &isurf) const
{
  double best = kInfLength;
  isurf       = -1;
  for (int k = 0; k < int(f.size()); ++k) {
    TriFacet const &t = f[k];
    double s          = p.Dot(t.n) + t.dist;
    const V3 q        = p - s * t.n;
    bool within       = true;
    for (int i = 0; i < 3; ++i) within &= (q - t.v[i]).Dot(t.side[i]) > -kTolerance;
    within &= toIn ? s > -kTolerance : s < kTolerance;
    double sq = s * s;
    if (!within) {
      const double e0 = SegmentDistSq(t.v[0], t.v[1], p), e1 = SegmentDistSq(t.v[1], t.v[2], p),
                   e2 = SegmentDistSq(t.v[2], t.v[0], p);
      sq = std::min(e0, std::min(e1, e2)); // Vector3D::Min()
    }
    if (sq < best) best = sq, isurf = k;
  }
  return best;
}

bool Tessellated::Contains(V3 const &p) const // TessellatedImplementation.h:54-67
{
  double dOut, dIn;
  int sOut, sIn;
  DistanceToSolid(false, p, testDir, kInfLength, dOut, sOut, dIn, sIn);
  return sOut >= 0;
}

int Tessellated::Inside(V3 const &p) const // :72-97
{
  double dOut, dIn;
  int sOut, sIn;
  DistanceToSolid(false, p, testDir, kInfLength, dOut, sOut, dIn, sIn);
  if (sOut < 0) return kOutside;
  if (dOut < 0 || dOut * testDir.Dot(f[sOut].n) < kTolerance) return kSurface;
  if (sIn < 0 || dOut < dIn) return kInside;
  if (dIn < 0 || dIn * testDir.Dot(f[sIn].n) > -kTolerance) return kSurface;
  return kOutside;
}

double Tessellated::DistanceToIn(V3 const &p, V3 const &d, double step_max) const
{
  double dist, other;
  int s, so;
  DistanceToSolid(true, p, d, step_max, dist, s, other, so);
  return dist;
}

double Tessellated::DistanceToOut(V3 const &p, V3 const &d, double step_max) const
{
  double dist, other;
  int s, so;
  DistanceToSolid(false, p, d, step_max, dist, s, other, so);
  return dist;
}

double Tessellated::SafetyToIn(V3 const &p) const // :128-141
{
  if (Contains(p)) return -1.;
  int s;
  return std::sqrt(SafetySq(true, p, s));
}

double Tessellated::SafetyToOut(V3 const &p) const // :143-156
{
  if (!Contains(p)) return -1.;
  int s;
  return std::sqrt(SafetySq(false, p, s));
}

} // namespace vgo
