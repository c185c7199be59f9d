// Oracle: CutTube (see shapes.h for the reference files followed).
// TEST INFRASTRUCTURE ONLY (see common.h).
#include "shapes.h"

namespace vgo {
namespace {
inline double Mx(double a, double b) { return a > b ? a : b; }
inline double Mn(double a, double b) { return a < b ? a : b; }

// Vector3D::Unit (Vector3D.h:284-290): divided by sqrt(mag2 + the smallest normal double)
V3 Unit(V3 const &v)
{
  const double m = std::sqrt(v.Mag2() + DBL_MIN);
  return V3(v.x / m, v.y / m, v.z / m);
}
} // namespace

// CutTubeStruct (CutTubeStruct.h:36-47) and Plane::Set (Plane.h:44-51: zeroes below kTolerance, then
// normalised; distance = -n.origin)
CutTube::CutTube(double rmin, double rmax, double z, double sphi, double dphi, V3 const &bottom, V3 const &top)
    : tube(rmin, rmax, kInfLength, sphi, dphi), dz(z)
{
  const V3 u[2] = {Unit(bottom), Unit(top)};
  const double oz[2] = {-z, z};
  for (int i = 0; i < 2; ++i) {
    V3 f = u[i];
    for (int k = 0; k < 3; ++k)
      if (std::fabs(f[k]) < kTolerance) f[k] = 0.0;
    const double inv = 1. / f.Mag();
    n[i]             = f * inv;
    dist[i]          = inv * -f.Dot(V3(0., 0., oz[i]));
  }
  cosPhi1 = std::cos(sphi);
  sinPhi1 = std::sin(sphi);
  cosPhi2 = std::cos(sphi + dphi);
  sinPhi2 = std::sin(sphi + dphi);
  maxVal  = Mx(rmax, z);
  ok      = !(bottom.z >= 0 || top.z <= 0) && rmax > 0 && z > 0 && dphi > 0; // UnplacedCutTube.h:41-45
}

// CutPlanes::Inside (CutPlanes.h:133-144)
int CutTube::CutInside(V3 const &p) const
{
  const double d0 = DistPlane(0, p), d1 = DistPlane(1, p);
  int in = (d0 < 0. && d1 < 0.) ? kInside : kOutside;
  if (std::fabs(d0) < kTolerance || std::fabs(d1) < kTolerance) in = kSurface;
  return in;
}

// CutTubeImplementation.h:84-99
bool CutTube::Contains(V3 const &p) const
{
  if (!(DistPlane(0, p) < 0. && DistPlane(1, p) < 0.)) return false;
  return tube.Contains(p);
}

// CutTubeImplementation.h:102-120
int CutTube::Inside(V3 const &p) const
{
  const int ic = CutInside(p);
  if (ic == kOutside) return kOutside; // the tube is not asked: the blend below gives the cut result
  int in = tube.Inside(p);
  if (ic == kSurface && in != kOutside) in = ic;
  return in;
}

// CutTubeImplementation.h:123-145: a point beyond 100 times the larger of rmax and dz is first moved
// along the direction to that radius
double CutTube::DistanceToIn(V3 const &pt, V3 const &d, double step_max) const
{
  V3 p                = pt;
  const double ptDist = p.Mag();
  double move         = 0.;
  if (ptDist > 100. * maxVal) {
    move = ptDist - 100. * maxVal;
    p    = p + d * move;
  }
  return DistanceToInKernel(p, d, step_max) + move;
}

// CutTubeImplementation.h:148-244 on the scalar backend (early returns taken)
double CutTube::DistanceToInKernel(V3 const &p, V3 const &d, double stepMax) const
{
  double distance = kInfLength;
  int ic          = CutInside(p);
  int instart     = tube.Inside(p);
  if (ic == kOutside || (ic == kSurface && instart != kOutside)) instart = ic;
  if (instart == kInside) distance = -1.;
  // CutPlanes::DistanceToIn (CutPlanes.h:146-160) with Plane::DistanceToIn (Plane.h:128-147)
  double dpl[2];
  for (int i = 0; i < 2; ++i) {
    const double ndd = NonZero(d.Dot(n[i]));
    const double saf = DistPlane(i, p);
    dpl[i]           = (ndd < 0. && saf > -kTolerance) ? -saf / ndd : -kInfLength;
    if (d.Dot(n[i]) > 0.) dpl[i] = -kInfLength;
  }
  double dplanes         = Mx(dpl[0], dpl[1]);
  const double safplanes = Mx(DistPlane(0, p), DistPlane(1, p)); // CutPlanes::SafetyToIn
  const bool hitplanes   = std::fabs(dplanes) < stepMax && safplanes > -kTolerance;
  if (ic != kInside && !hitplanes) return distance;
  V3 propagated = p;
  if (!hitplanes) dplanes = 0.;
  if (hitplanes) {
    propagated = propagated + d * dplanes;
    ic         = CutInside(propagated);
  }
  bool done = ic == kOutside;
  done      = done || instart == kInside;
  if (done) return distance;
  instart           = tube.Inside(propagated);
  const bool inside = hitplanes && instart != kOutside;
  if (inside && !done) distance = dplanes;
  done = done || inside;
  if (done) {
    if (std::fabs(distance) < kTolerance) distance = 0.;
    return distance;
  }
  // CutPlanes::DistanceToOut (CutPlanes.h:162-171) with Plane::DistanceToOut (Plane.h:149-171)
  double dexit = kInfLength;
  for (int i = 0; i < 2; ++i) {
    const double ndd = NonZero(d.Dot(n[i]));
    const double saf = DistPlane(i, propagated);
    double di        = kInfLength;
    if (saf > kTolerance) di = -kInfLength;
    if (ndd > 0. && saf < kTolerance) di = -saf / ndd;
    dexit = i == 0 ? di : Mn(dexit, di);
  }
  double dtube = tube.DistanceToInKernel(propagated, d);
  if (dtube < 0.) dtube = 0.;
  if (dexit < dtube) dtube = kInfLength;
  if (!done && (dtube + dplanes) < stepMax) distance = dtube + dplanes;
  return distance;
}

// CutTubeImplementation.h:248-266
double CutTube::DistanceToOut(V3 const &p, V3 const &d, double step_max) const
{
  double distance = kInfLength;
  for (int i = 0; i < 2; ++i) {
    const double ndd = NonZero(d.Dot(n[i]));
    const double saf = DistPlane(i, p);
    double di        = kInfLength;
    if (saf > kTolerance) di = -kInfLength;
    if (ndd > 0. && saf < kTolerance) di = -saf / ndd;
    distance = i == 0 ? di : Mn(distance, di);
  }
  const double dtube = tube.DistanceToOut(p, d, step_max);
  if (dtube < distance) distance = dtube;
  return distance;
}

// CutTubeImplementation.h:269-282: the larger of the signed plane and tube safeties
double CutTube::SafetyToIn(V3 const &p) const
{
  double safety       = Mx(DistPlane(0, p), DistPlane(1, p));
  const double stube  = tube.SafetyToIn(p);
  if (stube > safety) safety = stube;
  return safety;
}

// CutTubeImplementation.h:285-298
double CutTube::SafetyToOut(V3 const &p) const
{
  double safety      = Mn(-DistPlane(0, p), -DistPlane(1, p));
  const double stube = tube.SafetyToOut(p);
  if (stube < safety) safety = stube;
  return safety;
}

// CutTubeImplementation.h:301-322: the tube's normal where the tube surface is nearer than the planes,
// else the cut plane of the point's z half
bool CutTube::Normal(V3 const &p, V3 &norm) const
{
  const double safcut  = Mn(-DistPlane(0, p), -DistPlane(1, p));
  const double saftube = tube.SafetyToOut(p);
  if (std::fabs(saftube) < std::fabs(safcut)) return tube.Normal(p, norm);
  norm = p.z < 0 ? n[0] : n[1];
  return true;
}

// UnplacedCutTube.cpp:128-172 with ZlimitTop/ZlimitBottom (UnplacedCutTube.h:133-147)
void CutTube::Extent(V3 &lo, V3 &hi) const
{
  const double rmax = tube.rmax, rmin = tube.rmin, sphi = tube.sphi, dphi = tube.dphi;
  const V3 &T = n[1], &B = n[0];
  const double dztop    = rmax * std::sqrt(1. - T.z * T.z) / T.z;
  const double dzbottom = -rmax * std::sqrt(1. - B.z * B.z) / B.z;
  lo                    = V3(-rmax, -rmax, -dz - dzbottom);
  hi                    = V3(rmax, rmax, dz + dztop);
  if (dphi == kTwoPi) return;
  auto zTop = [&](double r, double phi) { return dz - (r / T.z) * (T.x * std::cos(phi) + T.y * std::sin(phi)); };
  auto zBot = [&](double r, double phi) { return -dz - (r / B.z) * (B.x * std::cos(phi) + B.y * std::sin(phi)); };
  if (!tube.wedge.Contains(V3(-T.x, -T.y, 0.))) {
    hi.z = Mx(zTop(rmax, sphi), zTop(rmax, sphi + dphi));
    hi.z = Mx(hi.z, Mx(zTop(rmin, sphi), zTop(rmin, sphi + dphi)));
  }
  if (!tube.wedge.Contains(V3(-B.x, -B.y, 0.))) {
    lo.z = Mn(zBot(rmax, sphi), zBot(rmax, sphi + dphi));
    lo.z = Mn(lo.z, Mn(zBot(rmin, sphi), zBot(rmin, sphi + dphi)));
  }
  if (!tube.wedge.Contains(V3(1., 0., 0.))) {
    hi.x = Mx(rmax * cosPhi1, rmax * cosPhi2);
    hi.x = Mx(hi.x, Mx(rmin * cosPhi1, rmin * cosPhi2));
  }
  if (!tube.wedge.Contains(V3(-1., 0., 0.))) {
    lo.x = Mn(rmax * cosPhi1, rmax * cosPhi2);
    lo.x = Mn(lo.x, Mn(rmin * cosPhi1, rmin * cosPhi2));
  }
  if (!tube.wedge.Contains(V3(0., 1., 0.))) {
    hi.y = Mx(rmax * sinPhi1, rmax * sinPhi2);
    hi.y = Mx(hi.y, Mx(rmin * sinPhi1, rmin * sinPhi2));
  }
  if (!tube.wedge.Contains(V3(0., -1., 0.))) {
    lo.y = Mn(rmax * sinPhi1, rmax * sinPhi2);
    lo.y = Mn(lo.y, Mn(rmin * sinPhi1, rmin * sinPhi2));
  }
}

} // namespace vgo
