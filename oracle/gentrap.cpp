// Oracle: GenTrap, planar case (see shapes.h for the reference files followed).
// TEST INFRASTRUCTURE ONLY (see common.h).
#include "shapes.h"

namespace vgo {
namespace {
inline double Mx(double a, double b) { return a > b ? a : b; }
inline double Mn(double a, double b) { return a < b ? a : b; }
inline V3 Normalized(V3 const &v) { return v * (1. / v.Mag()); } // Vector3D::Normalize (Vector3D.h:198)

// SegmentsCrossing (GenTrapStruct.h:296-317)
bool SegmentsCrossing(V3 p, V3 p1, V3 q, V3 q1)
{
  const V3 r = p1 - p, s = q1 - q;
  const V3 rxs = r.Cross(s);
  if (rxs.Mag2() < kTolerance) return false;
  const double t = (q - p).Cross(s).z / rxs.z;
  if (t < kTolerance || t > 1. - kTolerance) return false;
  const double u = (q - p).Cross(r).z / rxs.z;
  if (u < kTolerance || u > 1. - kTolerance) return false;
  return true;
}

// Tile<4>::AddVertex checks and plane (VecGeom/volumes/Tile.h:70-130): false for a degenerate quadrilateral
bool QuadPlane(V3 const v[4], V3 &normal, double &distance)
{
  int nvert = 4;
  for (int i = 0; i < 4; ++i)
    if ((v[(i + 1) % 4] - v[i]).Mag2() < kTolerance) nvert--;
  if (nvert < 3) return false;
  bool degenerated = true;
  for (int i = 0; i < 3 && degenerated; ++i) {
    const V3 e1 = v[i + 1] - v[i];
    if (e1.Mag2() < kTolerance) continue;
    for (int j = i + 1; j < 4; ++j) {
      const V3 e2 = v[(j + 1) % 4] - v[j];
      if (e2.Mag2() < kTolerance) continue;
      normal = e1.Cross(e2);
      if (normal.Mag2() < kTolerance) continue;
      normal      = Normalized(normal);
      degenerated = false;
      break;
    }
  }
  if (degenerated) return false;
  for (int i = 0; i < 4; ++i) {
    if ((v[(i + 1) % 4] - v[i]).Mag2() < kTolerance) continue;
    distance = -normal.Dot(v[i]);
    break;
  }
  return true;
}
} // namespace

// GenTrapStruct::Initialize (GenTrapStruct.h:57-188)
GenTrap::GenTrap(const double *x, const double *y, double h)
{
  dz            = h;
  halfInverseDz = 0.5 / h;
  // SecondOrderSurfaceShell::Initialize on the vertices as given (before any reversal)
  for (int i = 0; i < 4; ++i) {
    const int j = (i + 1) % 4;
    const V3 va(x[i], y[i], -h), vb(x[i + 4], y[i + 4], h), vc(x[j], y[j], -h), vd(x[j + 4], y[j + 4], h);
    V3 n = (vb - va).Cross(vc - va);
    if (n.Mag2() < kTolerance) {
      n = (vb - va).Cross(vd - vb);
      if (n.Mag2() < kTolerance) n = V3(0., 0., 1.);
    }
    shellNormal[i] = Normalized(n);
  }
  for (int i = 0; i < 8; ++i) vertices[i] = V3(x[i], y[i], i < 4 ? -h : h);
  double sum1 = 0., sum2 = 0.;
  for (int i = 0; i < 4; ++i) {
    const int j = (i + 1) % 4;
    sum1 += vertices[i].x * vertices[j].y - vertices[j].x * vertices[i].y;
    sum2 += vertices[i + 4].x * vertices[j + 4].y - vertices[j + 4].x * vertices[i + 4].y;
  }
  if (sum1 * sum2 < -kTolerance) return; // opposite orientation of the two polygons
  if (sum1 > kTolerance) {               // anticlockwise: reverted to clockwise
    std::swap(vertices[1], vertices[3]);
    std::swap(vertices[5], vertices[7]);
  }
  for (int i = 0; i < 4; ++i) {
    connX[i] = (vertices[i] - vertices[i + 4]).x;
    connY[i] = (vertices[i] - vertices[i + 4]).y;
  }
  for (int i = 0; i < 8; ++i) vx[i] = vertices[i].x, vy[i] = vertices[i].y;
  double deltaX[8], deltaY[8];
  for (int i = 0; i < 4; ++i) {
    const int j    = (i + 1) % 4;
    deltaX[i]     = vx[j] - vx[i];
    deltaX[i + 4] = vx[j + 4] - vx[i + 4];
    deltaY[i]     = vy[j] - vy[i];
    deltaY[i + 4] = vy[j + 4] - vy[i + 4];
  }
  if (SegmentsCrossing(vertices[0], vertices[1], vertices[3], vertices[2]) ||
      SegmentsCrossing(vertices[1], vertices[2], vertices[0], vertices[3]) ||
      SegmentsCrossing(vertices[4], vertices[5], vertices[7], vertices[6]) ||
      SegmentsCrossing(vertices[5], vertices[6], vertices[4], vertices[7]))
    return;
  for (int i = 0; i < 4; ++i) { // ComputeIsConvexQuadrilaterals (:218-232)
    const int j = (i + 1) % 4;
    if (deltaX[i] * deltaY[j] - deltaY[i] * deltaX[j] > kTolerance) return;
    if (deltaX[i + 4] * deltaY[j + 4] - deltaY[i + 4] * deltaX[j + 4] > kTolerance) return;
  }
  for (int i = 0; i < 4; ++i) { // ComputeIsTwisted (:246-276)
    const double dx1 = vertices[(i + 1) % 4].x - vertices[i].x, dy1 = vertices[(i + 1) % 4].y - vertices[i].y;
    if (dx1 == 0 && dy1 == 0) continue;
    const double dx2 = vertices[4 + (i + 1) % 4].x - vertices[4 + i].x;
    const double dy2 = vertices[4 + (i + 1) % 4].y - vertices[4 + i].y;
    if (dx2 == 0 && dy2 == 0) continue;
    if (std::fabs(dy1 * dx2 - dx1 * dy2) < kTolerance) continue;
    twisted = true;
  }
  if (twisted) return;
  // the TessellatedSection's lateral quadrilaterals (:178-184); degenerate ones are not added
  const int quad[4][4] = {{0, 4, 5, 1}, {1, 5, 6, 2}, {2, 6, 7, 3}, {3, 7, 4, 0}};
  for (int f = 0; f < 4; ++f) {
    const V3 v[4] = {vertices[quad[f][0]], vertices[quad[f][1]], vertices[quad[f][2]], vertices[quad[f][3]]};
    V3 n;
    double d = 0.;
    if (QuadPlane(v, n, d)) fnormal[nfacets] = n, fdist[nfacets] = d, ++nfacets;
  }
  ok = true;
}

// Contains<Precision, bool> (GenTrapImplementation.h:330-348) + TessellatedSection::Contains
bool GenTrap::Contains(V3 const &p) const
{
  if (std::fabs(p.z) > dz) return false;
  for (int i = 0; i < nfacets; ++i)
    if (DistPlane(i, p) > 0) return false;
  return true;
}

// Inside<Precision, Inside_t> (:376-396) + TessellatedSection::Inside
int GenTrap::Inside(V3 const &p) const
{
  const double safZ = std::fabs(p.z) - dz;
  if (safZ > kTolerance) return kOutside;
  const bool insideZ = safZ < -kTolerance;
  bool in = true, out = false;
  for (int i = 0; i < nfacets; ++i) {
    const double d = DistPlane(i, p);
    out            = out || d > kTolerance;
    in             = in && d < -kTolerance;
  }
  const int inside = out ? kOutside : (in ? kInside : kSurface);
  if (insideZ || inside == kOutside) return inside;
  return kSurface;
}

// DistanceToIn (:461-476) + TessellatedSection::DistanceToIn<false> + TessellatedCluster::DistanceToInConvex
double GenTrap::DistanceToIn(V3 const &p, V3 const &d, double stepMax) const
{
  const double invdirz = 1. / NonZero(d.z);
  const double pz      = p.z;
  if ((std::fabs(pz) - dz) > -kTolerance && pz * d.z >= 0) return kInfLength;
  const double ddz   = std::copysign(dz, invdirz);
  const double distz = -(pz + ddz) * invdirz;
  double distance    = distz;
  double limit       = Mn(-(pz - ddz) * invdirz, stepMax);
  for (int i = 0; i < nfacets; ++i) {
    const double proj        = NonZero(d.Dot(fnormal[i]));
    const bool moving_away   = proj > -kTolerance;
    const double pdist       = DistPlane(i, p);
    const bool side_correct  = pdist > -kTolerance;
    if (side_correct && moving_away) return kInfLength;
    const bool from_outside = side_correct && !moving_away;
    const bool from_inside  = !side_correct && moving_away;
    double dmin = -kInfLength, dmax = kInfLength;
    const double dist = -pdist / NonZero(proj);
    if (from_outside) dmin = dist;
    if (from_inside) dmax = dist;
    distance = Mx(distance, dmin);
    limit    = Mn(limit, dmax);
    if (!(distance < limit - kTolerance)) return kInfLength;
  }
  return distance;
}

// DistanceToOut (:519-536) + TessellatedSection::DistanceToOut + TessellatedCluster::DistanceToOutConvex
double GenTrap::DistanceToOut(V3 const &p, V3 const &d, double) const
{
  const double pz   = p.z;
  const double safz = std::fabs(pz) - dz;
  if (safz > kTolerance) return -kTolerance;
  const double vz = d.z;
  double distance = (std::copysign(dz, vz) - pz) / NonZero(vz);
  for (int i = 0; i < nfacets; ++i) {
    const double saf = DistPlane(i, p);
    if (!(saf < kTolerance)) return -kTolerance;
    const double ndd = NonZero(d.Dot(fnormal[i]));
    const double dist = ndd > 0. ? -saf / ndd : kInfLength;
    if (dist < distance) distance = dist;
  }
  return distance;
}

// SafetyToIn / SafetyToOut (:571-608) + TessellatedSection::SafetyToIn / SafetyToOut (fZ = 0)
double GenTrap::SafetyToIn(V3 const &p) const
{
  double safety = Mx(0. - dz - p.z, p.z - 0. - dz);
  for (int i = 0; i < nfacets; ++i) {
    const double s = DistPlane(i, p);
    if (s > safety) safety = s;
  }
  return safety;
}

double GenTrap::SafetyToOut(V3 const &p) const
{
  double safety = Mx(0. - dz - p.z, p.z - 0. - dz);
  for (int i = 0; i < nfacets; ++i) {
    const double s = DistPlane(i, p);
    if (s > safety) safety = s;
  }
  return -safety;
}

// NormalKernel (:610-680) and GetClosestEdge (:683-735): the z planes within 10 kTolerance, else the shell
// normal of the closest edge of the section at the point's z (the shell keeps the vertex order as given,
// the section uses the reverted one)
bool GenTrap::Normal(V3 const &p, V3 &n) const
{
  V3 normal(0., 0., 0.);
  const double safz = std::fabs(dz - std::fabs(p.z));
  const bool onZ    = safz < 10. * kTolerance;
  if (onZ && p.z > 0.) normal.z = 1.;
  if (onZ && p.z < 0.) normal.z = -1.;
  if (onZ) {
    n = normal;
    return true;
  }
  const double cf = halfInverseDz * (dz - p.z);
  double vertexX[4], vertexY[4];
  for (int i = 0; i < 4; i++) {
    vertexX[i] = vx[i + 4] + cf * connX[i];
    vertexY[i] = vy[i + 4] + cf * connY[i];
  }
  double iseg = 0., fraction = -1., safe = kInfLength, ssq = kInfLength;
  for (int i = 0; i < 4; ++i) {
    const int j = (i + 1) % 4;
    const double dx = vertexX[j] - vertexX[i], dy = vertexY[j] - vertexY[i];
    double dpx = p.x - vertexX[i], dpy = p.y - vertexY[i];
    const double lsq     = dx * dx + dy * dy;
    const bool collapsed = lsq < kTolerance;
    if (collapsed) {
      ssq = dpx * dpx + dpy * dpy;
      if (ssq < safe) iseg = double(i), fraction = -1., safe = ssq;
      continue;
    }
    double u = (dpx * dx + dpy * dy) / NonZero(lsq);
    if (u > 1) dpx = p.x - vertexX[j], dpy = p.y - vertexY[j];
    if (u >= 0 && u <= 1) dpx = dpx - u * dx, dpy = dpy - u * dy;
    if (u > 1 || u < 0) u = -1.;
    ssq = dpx * dpx + dpy * dpy;
    if (ssq < safe) iseg = double(i), fraction = u, safe = ssq;
  }
  (void)fraction;
  n = shellNormal[int(iseg)];
  return true;
}

// GenTrapStruct::Extent (:199-214)
void GenTrap::Extent(V3 &lo, V3 &hi) const
{
  lo = hi = vertices[0];
  lo.z = -dz, hi.z = dz;
  for (int i = 0; i < 4; ++i) {
    if (lo.x > vertices[i].x) lo.x = vertices[i].x;
    if (hi.x < vertices[i].x) hi.x = vertices[i].x;
    if (lo.y > vertices[i].y) lo.y = vertices[i].y;
    if (hi.y < vertices[i].y) hi.y = vertices[i].y;
    if (lo.x > vertices[i + 4].x) lo.x = vertices[i + 4].x;
    if (hi.x < vertices[i + 4].x) hi.x = vertices[i + 4].x;
    if (lo.y > vertices[i + 4].y) lo.y = vertices[i + 4].y;
    if (hi.y < vertices[i + 4].y) hi.y = vertices[i + 4].y;
  }
}

} // namespace vgo
