// Oracle: Trapezoid (general G4Trap parametrisation). TEST INFRASTRUCTURE ONLY (see common.h).
#include "shapes.h"

namespace vgo {

// volumes/UnplacedTrapezoid.h:56-63 (tan of the alpha angles) and TrapezoidStruct::CalculateCached
Trapezoid::Trapezoid(double dz_, double theta_, double phi_, double dy1_, double dx1_, double dx2_, double alpha1,
                     double dy2_, double dx3_, double dx4_, double alpha2)
    : dz(dz_), theta(theta_), phi(phi_), dy1(dy1_), dx1(dx1_), dx2(dx2_), tanAlpha1(std::tan(alpha1)), dy2(dy2_),
      dx3(dx3_), dx4(dx4_), tanAlpha2(std::tan(alpha2))
{
  tthetaCphi = std::tan(theta) * std::cos(phi);
  tthetaSphi = std::tan(theta) * std::sin(phi);
  // source/UnplacedTrapezoid.cpp:447-465 FromParametersToCorners
  double dxdyDy1 = tanAlpha1 * dy1, dxdyDy2 = tanAlpha2 * dy2;
  double dxdzDz = tthetaCphi * dz, dydzDz = tthetaSphi * dz;
  V3 pt[8];
  pt[0] = V3(-dxdzDz - dxdyDy1 - dx1, -dydzDz - dy1, -dz);
  pt[1] = V3(-dxdzDz - dxdyDy1 + dx1, -dydzDz - dy1, -dz);
  pt[2] = V3(-dxdzDz + dxdyDy1 - dx2, -dydzDz + dy1, -dz);
  pt[3] = V3(-dxdzDz + dxdyDy1 + dx2, -dydzDz + dy1, -dz);
  pt[4] = V3(+dxdzDz - dxdyDy2 - dx3, +dydzDz - dy2, +dz);
  pt[5] = V3(+dxdzDz - dxdyDy2 + dx3, +dydzDz - dy2, +dz);
  pt[6] = V3(+dxdzDz + dxdyDy2 - dx4, +dydzDz + dy2, +dz);
  pt[7] = V3(+dxdzDz + dxdyDy2 + dx4, +dydzDz + dy2, +dz);
  SetPlanes(pt);
}

// UnplacedTrapezoid(TrapCorners): source/UnplacedTrapezoid.cpp:129-134 → fromCornersToParameters :496-523
// (parameters recovered from the corners), then the side planes of the given corners (MakePlanes(pt)).
Trapezoid::Trapezoid(V3 const pt[8])
{
  dz             = pt[7].z;
  double DzRecip = 1.0 / dz;
  dy1            = 0.50 * (pt[2].y - pt[0].y);
  dx1            = 0.50 * (pt[1].x - pt[0].x);
  dx2            = 0.50 * (pt[3].x - pt[2].x);
  tanAlpha1      = 0.25 * (pt[2].x + pt[3].x - pt[1].x - pt[0].x) / dy1;
  dy2            = 0.50 * (pt[6].y - pt[4].y);
  dx3            = 0.50 * (pt[5].x - pt[4].x);
  dx4            = 0.50 * (pt[7].x - pt[6].x);
  tanAlpha2      = 0.25 * (pt[6].x + pt[7].x - pt[5].x - pt[4].x) / dy2;
  tthetaCphi     = (pt[4].x + dy2 * tanAlpha2 + dx3) * DzRecip;
  tthetaSphi     = (pt[4].y + dy2) * DzRecip;
  theta          = std::atan(std::sqrt(tthetaSphi * tthetaSphi + tthetaCphi * tthetaCphi));
  phi            = std::atan2(tthetaSphi, tthetaCphi);
  SetPlanes(pt);
}

// source/UnplacedTrapezoid.cpp:630-666 and MakeAPlane :575-619
void Trapezoid::SetPlanes(V3 const pt[8])
{
  auto plane = [&](int i, V3 const &p1, V3 const &p2, V3 const &p3, V3 const &p4) {
    V3 n       = (p3 - p1).Cross(p4 - p2);
    double inv = 1. / n.Mag();
    n          = n * inv;
    V3 c       = 0.25 * (((p1 + p2) + p3) + p4);
    A[i]       = n.x;
    B[i]       = n.y;
    C[i]       = n.z;
    D[i]       = -n.Dot(c);
  };
  plane(0, pt[0], pt[1], pt[5], pt[4]);
  plane(1, pt[2], pt[6], pt[7], pt[3]);
  plane(2, pt[0], pt[4], pt[6], pt[2]);
  plane(3, pt[1], pt[3], pt[7], pt[5]);
}

// source/UnplacedTrapezoid.cpp:298-337 Extent via fromPlanesToCorners :467-494
void Trapezoid::Extent(V3 &lo, V3 &hi) const
{
  V3 pt[8];
  double mz = -dz, pz = dz;
  pt[0].y = pt[1].y = -(D[0] - dz * C[0]) / B[0];
  pt[2].y = pt[3].y = -(D[1] - dz * C[1]) / B[1];
  pt[4].y = pt[5].y = -(D[0] + dz * C[0]) / B[0];
  pt[6].y = pt[7].y = -(D[1] + dz * C[1]) / B[1];
  pt[0].x = ((B[2] / B[0]) * (D[0] - dz * C[0]) - (D[2] - dz * C[2])) / A[2];
  pt[1].x = ((B[3] / B[0]) * (D[0] - dz * C[0]) - (D[3] - dz * C[3])) / A[3];
  pt[2].x = ((B[2] / B[1]) * (D[1] - dz * C[1]) - (D[2] - dz * C[2])) / A[2];
  pt[3].x = ((B[3] / B[1]) * (D[1] - dz * C[1]) - (D[3] - dz * C[3])) / A[3];
  pt[4].x = ((B[2] / B[0]) * (D[0] + dz * C[0]) - (D[2] + dz * C[2])) / A[2];
  pt[5].x = ((B[3] / B[0]) * (D[0] + dz * C[0]) - (D[3] + dz * C[3])) / A[3];
  pt[6].x = ((B[2] / B[1]) * (D[1] + dz * C[1]) - (D[2] + dz * C[2])) / A[2];
  pt[7].x = ((B[3] / B[1]) * (D[1] + dz * C[1]) - (D[3] + dz * C[3])) / A[3];
  auto mx = [](double a, double b) { return std::max(a, b); };
  auto mn = [](double a, double b) { return std::min(a, b); };
  hi.x = mx(mx(mx(pt[0].x, pt[1].x), mx(pt[2].x, pt[3].x)), mx(mx(pt[4].x, pt[5].x), mx(pt[6].x, pt[7].x)));
  lo.x = mn(mn(mn(pt[0].x, pt[1].x), mn(pt[2].x, pt[3].x)), mn(mn(pt[4].x, pt[5].x), mn(pt[6].x, pt[7].x)));
  hi.y = mx(mx(mx(pt[0].y, pt[1].y), mx(pt[2].y, pt[3].y)), mx(mx(pt[4].y, pt[5].y), mx(pt[6].y, pt[7].y)));
  lo.y = mn(mn(mn(pt[0].y, pt[1].y), mn(pt[2].y, pt[3].y)), mn(mn(pt[4].y, pt[5].y), mn(pt[6].y, pt[7].y)));
  lo.z = mz;
  hi.z = pz;
}

// TrapezoidImplementation.h:126-155 + PlaneShell::GenericKernelForContainsAndInside
void Trapezoid::ContainsInside(V3 const &p, bool forInside, bool &cin, bool &cout) const
{
  cout = std::fabs(p.z) > (forInside ? dz + kHalfTolerance : dz);
  cin  = forInside && std::fabs(p.z) < dz - kHalfTolerance;
  double dist[4];
  for (int i = 0; i < 4; ++i) dist[i] = A[i] * p.x + B[i] * p.y + C[i] * p.z + D[i];
  for (int i = 0; i < 4; ++i) {
    cout = cout || (dist[i] > (forInside ? kHalfTolerance : 0.));
    if (forInside) cin = cin && (dist[i] < -kHalfTolerance);
  }
}

bool Trapezoid::Contains(V3 const &p) const
{
  bool ci = false, co = false;
  ContainsInside(p, false, ci, co);
  return !co;
}

int Trapezoid::Inside(V3 const &p) const
{
  bool ci = false, co = false;
  ContainsInside(p, true, ci, co);
  int r = kSurface;
  if (co) r = kOutside;
  if (ci) r = kInside;
  return r;
}

// TrapezoidImplementation.h:157-205 + PlaneShell::DistanceToIn (PlaneShell.h:150-198)
double Trapezoid::DistanceToIn(V3 const &p, V3 const &d, double) const
{
  double signZ = Sign(d.z);
  double mx    = signZ * dz - p.z;
  if (signZ * mx < kHalfTolerance) return kInfLength;
  double invdir = 1.0 / NonZero(d.z);
  double smax   = mx * invdir;
  double smin   = -(signZ * dz + p.z) * invdir;
  double pdist[4], proj[4], vdist[4];
  for (int i = 0; i < 4; ++i) {
    pdist[i] = A[i] * p.x + B[i] * p.y + C[i] * p.z + D[i];
    proj[i]  = A[i] * d.x + B[i] * d.y + C[i] * d.z;
    vdist[i] = -pdist[i] / NonZero(proj[i]);
  }
  for (int i = 0; i < 4; ++i) {
    if (pdist[i] > kHalfTolerance && proj[i] >= 0.) return kInfLength;
    if (pdist[i] > -kHalfTolerance && proj[i] > 0.) return kInfLength;
  }
  for (int i = 0; i < 4; ++i) {
    bool posPoint = pdist[i] > -kHalfTolerance;
    bool posDir   = proj[i] > 0;
    bool fromIn   = (!posPoint && posDir);
    bool fromOut  = (posPoint && !posDir);
    if (fromIn && vdist[i] < smin) return kInfLength;
    if (fromOut && vdist[i] > smax) return kInfLength;
    if (fromOut && vdist[i] > smin) smin = vdist[i];
    if (fromIn && vdist[i] < smax) smax = vdist[i];
  }
  return smin <= smax ? smin : kInfLength;
}

// TrapezoidImplementation.h:207-245 + PlaneShell::DistanceToOut (PlaneShell.h:200-225)
double Trapezoid::DistanceToOut(V3 const &p, V3 const &d, double) const
{
  if (std::fabs(p.z) > dz + kHalfTolerance) return -1.0;
  double dist  = kInfLength;
  double distz = (Sign(d.z) * dz - p.z) / NonZero(d.z);
  if (d.z != 0.) dist = distz;
  double out = kInfLength;
  for (int i = 0; i < 4; ++i) {
    double pd = A[i] * p.x + B[i] * p.y + C[i] * p.z + D[i];
    double pr = A[i] * d.x + B[i] * d.y + C[i] * d.z;
    double vd = -pd / NonZero(pr);
    if (pd > kHalfTolerance) out = -1.;
    if (pr > 0. && vd < out) out = vd;
  }
  if (out < dist) dist = out;
  return dist;
}

// TrapezoidImplementation.h:247-260 + PlaneShell::SafetyToIn (N == 4 branch)
double Trapezoid::SafetyToIn(V3 const &p) const
{
  double safety = std::fabs(p.z) - dz;
  for (int i = 0; i < 4; ++i) {
    double dd = A[i] * p.x + B[i] * p.y + C[i] * p.z + D[i];
    if (dd > safety) safety = dd;
  }
  return safety;
}

// TrapezoidImplementation.h:262-276 + PlaneShell::SafetyToOut (N == 4 branch)
double Trapezoid::SafetyToOut(V3 const &p) const
{
  double safety = dz - std::fabs(p.z);
  for (int i = 0; i < 4; ++i) {
    double dd = -(A[i] * p.x + B[i] * p.y + C[i] * p.z + D[i]);
    if (dd < safety) safety = dd;
  }
  return safety;
}

} // namespace vgo
