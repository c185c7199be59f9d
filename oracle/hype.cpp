// Oracle: Hype (UniversalHype). TEST INFRASTRUCTURE ONLY (see common.h).
//
// Scalar backend of VecGeom/volumes/kernel/HypeImplementation.h with the helpers of
// VecGeom/volumes/HypeUtilities.h: masks are plain bools and every vecCore::MaskFull
// early return of the reference becomes a return, as the scalar backend evaluates it.
#include "shapes.h"

namespace vgo {

// HypeStruct.h:54-83 CalculateCached
Hype::Hype(double rmin_, double rmax_, double stIn_, double stOut_, double dz_)
    : rmin(rmin_), rmax(rmax_), stIn(stIn_), stOut(stOut_), dz(dz_)
{
  tIn             = std::tan(stIn);
  tOut            = std::tan(stOut);
  tIn2            = tIn * tIn;
  tOut2           = tOut * tOut;
  rmin2           = rmin * rmin;
  rmax2           = rmax * rmax;
  dz2             = dz * dz;
  endInnerRadius2 = tIn2 * dz2 + rmin2;
  endOuterRadius2 = tOut2 * dz2 + rmax2;
  endInnerRadius  = std::sqrt(endInnerRadius2);
  endOuterRadius  = std::sqrt(endOuterRadius2);
  zTol            = kTolerance * dz;
  innerRadTol     = kTolerance * endInnerRadius;
  outerRadTol     = kTolerance * endOuterRadius;
}

// HypeImplementation.h:98-124 GenericKernelForContainsAndInside
void Hype::ContainsInside(V3 const &p, bool forInside, bool &cin, bool &cout) const
{
  const double r2    = p.Perp2();
  const double oRad2 = rmax2 + tOut2 * p.z * p.z;
  double iRad2       = 0.;
  cout               = std::fabs(p.z) > dz + zTol;
  if (cout) return;
  cout = r2 > oRad2 + outerRadTol;
  if (cout) return;
  if (inner()) {
    iRad2 = rmin2 + tIn2 * p.z * p.z;
    cout  = r2 < iRad2 - innerRadTol;
  }
  if (cout) return;
  if (forInside) {
    cin = std::fabs(p.z) < dz - zTol && r2 < oRad2 - outerRadTol;
    if (inner()) cin = cin && r2 > iRad2 + innerRadTol;
  }
}

bool Hype::Contains(V3 const &p) const
{
  bool ci = false, co = false;
  ContainsInside(p, false, ci, co);
  return !co;
}

int Hype::Inside(V3 const &p) const
{
  bool ci = false, co = false;
  ContainsInside(p, true, ci, co);
  return co ? kOutside : (ci ? kInside : kSurface);
}

// HypeUtilities.h:28-49 IsCompletelyOutside
bool Hype::CompletelyOutside(V3 const &p) const
{
  const double r2    = p.Perp2();
  const double oRad2 = rmax2 + tOut2 * p.z * p.z;
  if (std::fabs(p.z) > dz + zTol) return true;
  if (r2 > oRad2 + outerRadTol) return true;
  if (inner()) {
    const double iRad2 = rmin2 + tIn2 * p.z * p.z;
    return r2 < iRad2 - innerRadTol;
  }
  return false;
}

// HypeUtilities.h:51-67 IsCompletelyInside
bool Hype::CompletelyInside(V3 const &p) const
{
  const double r2    = p.Perp2();
  const double oRad2 = rmax2 + tOut2 * p.z * p.z;
  bool in            = std::fabs(p.z) < dz - zTol && r2 < oRad2 - outerRadTol;
  if (inner()) {
    const double iRad2 = rmin2 + tIn2 * p.z * p.z;
    in                 = in && r2 > iRad2 + innerRadTol;
  }
  return in;
}

// HypeUtilities.h:80-106 IsPointMovingInside{Outer,Inner}Surface
static bool moving_in_outer(Hype const &h, V3 const &p, V3 const &d)
{
  double pz = p.z, vz = d.z;
  if (pz < 0.) vz = -vz, pz = -pz;
  return (p.x * d.x + p.y * d.y - pz * h.tOut2 * vz) < 0.;
}
static bool moving_in_inner(Hype const &h, V3 const &p, V3 const &d)
{
  double pz = p.z, vz = d.z;
  if (pz < 0.) vz = -vz, pz = -pz;
  return (p.x * d.x + p.y * d.y - pz * h.tIn2 * vz) > 0.;
}

// HypeUtilities.h:167-194 IsPointMovingOutside{Outer,Inner}Surface: the normal at the point
// (the sign flips of pz/vz there do not enter the result)
static bool moving_out_outer(Hype const &h, V3 const &p, V3 const &d)
{
  return V3(p.x, p.y, -p.z * h.tOut2).Dot(d) > 0.;
}
static bool moving_out_inner(Hype const &h, V3 const &p, V3 const &d)
{
  return V3(-p.x, -p.y, p.z * h.tIn2).Dot(d) > 0.;
}

// HypeUtilities.h:108-147 IsPointOnSurfaceAndMovingInside
bool Hype::OnSurfaceMovingIn(V3 const &p, V3 const &d) const
{
  const double rho2  = p.Perp2();
  const double radI2 = rmin2 + tIn2 * p.z * p.z;
  const double radO2 = rmax2 + tOut2 * p.z * p.z;
  const bool zSurf   = (rho2 - endOuterRadius2) < kTolerance && (endInnerRadius2 - rho2) < kTolerance &&
                     std::fabs(std::fabs(p.z) - dz) < kTolerance;
  if (zSurf) return p.z * d.z < 0.;
  const bool outerSurf = std::fabs(radO2 - rho2) < outerRadTol;
  if (outerSurf) return moving_in_outer(*this, p, d);
  if (inner()) {
    const bool innerSurf = std::fabs(radI2 - rho2) < innerRadTol;
    if (innerSurf) return moving_in_inner(*this, p, d);
  }
  return false;
}

// HypeUtilities.h:230-274 IsPointOnSurfaceAndMovingOutside
bool Hype::OnSurfaceMovingOut(V3 const &p, V3 const &d) const
{
  const double rho2  = p.Perp2();
  const double radI2 = rmin2 + tIn2 * p.z * p.z;
  const double radO2 = rmax2 + tOut2 * p.z * p.z;
  const bool zSurf   = (rho2 - endOuterRadius2) < kTolerance && (endInnerRadius2 - rho2) < kTolerance &&
                     std::fabs(std::fabs(p.z) - dz) < kTolerance;
  if (zSurf && p.z * d.z > 0.) return true;
  if (std::fabs(radO2 - rho2) < outerRadTol && moving_out_outer(*this, p, d)) return true;
  if (inner() && std::fabs(radI2 - rho2) < innerRadTol && moving_out_inner(*this, p, d)) return true;
  return false;
}

// HypeUtilities.h:196-228 IsPointOn{Outer,Inner}SurfaceAndMovingOutside
bool Hype::OnOuterMovingOut(V3 const &p, V3 const &d) const
{
  const double rho2  = p.x * p.x + p.y * p.y;
  const double absZ  = std::fabs(p.z);
  const double radO2 = rmax2 + tOut2 * p.z * p.z;
  return std::fabs(radO2 - rho2) < outerRadTol && absZ >= 0. && absZ < dz && moving_out_outer(*this, p, d);
}
bool Hype::OnInnerMovingOut(V3 const &p, V3 const &d) const
{
  if (!inner()) return false;
  const double rho2  = p.x * p.x + p.y * p.y;
  const double absZ  = std::fabs(p.z);
  const double radI2 = rmin2 + tIn2 * p.z * p.z;
  return std::fabs(radI2 - rho2) < innerRadTol && absZ >= 0. && absZ < dz && moving_out_inner(*this, p, d);
}

// HypeUtilities.h:149-165 GetPointOfIntersectionWithZPlane
bool Hype::ZPlane(V3 const &p, V3 const &d, bool toIn, double &zDist) const
{
  zDist           = (Sign(toIn ? p.z : d.z) * dz - p.z) / NonZero(d.z);
  const double r2 = (p + zDist * d).Perp2();
  if (!inner()) return r2 < endOuterRadius2;
  return r2 < endOuterRadius2 && r2 > endInnerRadius2;
}

// HypeUtilities.h:374-436 HypeHelpers<Precision, ForDistToIn, ForInnerSurface>: leaves dist
// untouched when the quadratic has no real root
bool Hype::Hyperbolic(V3 const &p, V3 const &d, bool toIn, bool innerSurf, double &dist) const
{
  const double t2 = innerSurf ? tIn2 : tOut2;
  const double r2 = innerSurf ? rmin2 : rmax2;
  const double a  = d.Perp2() - t2 * d.z * d.z;
  const double b  = (d.x * p.x + d.y * p.y - t2 * d.z * p.z);
  const double c  = p.Perp2() - t2 * p.z * p.z - r2;
  if (!(b * b - a * c > 0.)) return false;
  const double sq = std::sqrt(b * b - a * c);
  if (innerSurf) {
    if (toIn)
      dist = b < 0. ? (-b + sq) / a : c / (-b - sq);
    else
      dist = b > 0. ? (-b - sq) / a : c / (-b + sq);
  } else {
    if (toIn)
      dist = b >= 0. ? (-b - sq) / a : c / (-b + sq);
    else
      dist = b < 0. ? (-b + sq) / a : c / (-b - sq);
  }
  if (dist < 0.) dist = kInfLength;
  const double newPtZ = p.z + dist * d.z;
  return std::fabs(newPtZ) <= dz;
}

// HypeImplementation.h:126-199
double Hype::DistanceToIn(V3 const &p, V3 const &d, double) const
{
  const double absZ = std::fabs(p.z);
  double zDist = kInfLength, dist = kInfLength;
  const double r = p.Perp2();
  if (OnSurfaceMovingIn(p, d)) return 0.;
  if (CompletelyInside(p)) return -1.;
  const bool hitZ    = ZPlane(p, d, true, zDist);
  const bool aboveIn = absZ > dz && p.z * d.z < 0.;
  if (aboveIn && hitZ) return zDist;
  const V3 newPt   = p + zDist * d;
  const double rp2 = newPt.Perp2();
  const bool outerFromOutsideZ = aboveIn && !hitZ && rp2 >= endOuterRadius2;
  const bool outerFromWithinZ  = (r > ((rmax2 + tOut2 * absZ * absZ) + kHalfTolerance)) && absZ >= 0. && absZ <= dz;
  bool cond = (outerFromOutsideZ || outerFromWithinZ || OnOuterMovingOut(p, d)) && Hyperbolic(p, d, true, false, dist);
  if (cond) return dist;
  if (inner()) {
    const bool innerFromOutsideZ = aboveIn && !hitZ && rp2 <= endInnerRadius2;
    const bool innerFromWithinZ  = (r < ((rmin2 + tIn2 * absZ * absZ) - kHalfTolerance)) && absZ >= 0. && absZ <= dz;
    cond = (innerFromOutsideZ || innerFromWithinZ || OnInnerMovingOut(p, d)) && Hyperbolic(p, d, true, true, dist);
    if (cond) return dist;
  }
  return kInfLength;
}

// HypeImplementation.h:201-236
double Hype::DistanceToOut(V3 const &p, V3 const &d, double) const
{
  double zDist = kInfLength, dist = kInfLength;
  if (OnSurfaceMovingOut(p, d)) return 0.;
  if (CompletelyOutside(p)) return -1.;
  ZPlane(p, d, false, zDist);
  if (zDist < 0.) zDist = kInfLength;
  Hyperbolic(p, d, false, false, dist);
  if (dist < 0.) dist = kInfLength;
  double distance = std::min(zDist, dist);
  if (inner()) {
    Hyperbolic(p, d, false, true, dist);
    if (dist < 0.) dist = kInfLength;
    distance = std::min(distance, dist);
  }
  return distance;
}

// HypeUtilities.h:276-313 ApproxDistOutside / ApproxDistInside
static double approx_dist_outside(double pr, double pz, double r0, double tanPhi)
{
  const double r1  = std::sqrt(r0 * r0 + tanPhi * tanPhi * pz * pz);
  const double z1  = pz;
  const double r2  = pr;
  const double z2  = std::sqrt((pr * pr - r0 * r0) / (tanPhi * tanPhi));
  const double ddz = z2 - z1;
  const double dr  = r2 - r1;
  const double r3  = std::sqrt(dr * dr + ddz * ddz);
  return r3 < DBL_MIN ? (r2 - r1) : (r2 - r1) * ddz / r3;
}
static double approx_dist_inside(double pr, double pz, double r0, double tan2Phi)
{
  if (tan2Phi < DBL_MIN) return r0 - pr;
  const double rh  = std::sqrt(r0 * r0 + pz * pz * tan2Phi);
  const double dr  = -rh;
  const double ddz = pz * tan2Phi;
  const double len = std::sqrt(dr * dr + ddz * ddz);
  return std::fabs((pr - rh) * dr) / len;
}

// HypeImplementation.h:238-295
double Hype::SafetyToIn(V3 const &p) const
{
  const double absZ = std::fabs(p.z);
  const double r2   = p.Perp2();
  const double r    = std::sqrt(r2);
  bool cin = false, cout = false;
  ContainsInside(p, true, cin, cout);
  if (!cin && !cout) return 0.;
  if (cin) return -1.;
  const double sigz = absZ - dz;
  if (sigz > kHalfTolerance && r < endOuterRadius && r > endInnerRadius) return sigz;
  if (sigz > kHalfTolerance && r > endOuterRadius)
    return std::sqrt((r - endOuterRadius) * (r - endOuterRadius) + (sigz) * (sigz));
  if (sigz > kHalfTolerance && r < endInnerRadius)
    return std::sqrt((r - endInnerRadius) * (r - endInnerRadius) + (sigz) * (sigz));
  if ((r2 > ((rmax2 + tOut2 * absZ * absZ) + kHalfTolerance)) && absZ > 0. && absZ < dz)
    return approx_dist_outside(r, absZ, rmax, tOut);
  if ((r2 < ((rmin2 + tIn2 * absZ * absZ) - kHalfTolerance)) && absZ > 0. && absZ < dz)
    return approx_dist_inside(r, absZ, rmin, tIn2);
  return 0.;
}

// HypeImplementation.h:297-330
double Hype::SafetyToOut(V3 const &p) const
{
  const double r    = std::sqrt(p.x * p.x + p.y * p.y);
  const double absZ = std::fabs(p.z);
  bool cin = false, cout = false;
  ContainsInside(p, true, cin, cout);
  if (!cin && !cout) return 0.;
  if (cout) return -1.;
  const double distZ = std::fabs(std::fabs(p.z) - dz);
  double distInner   = 0.;
  if (inner() && stIn != 0.) distInner = approx_dist_outside(r, absZ, rmin, tIn);
  if (inner() && stIn == 0.) distInner = r - rmin;
  if (!inner() && stIn == 0.) distInner = kInfLength;
  const double distOuter = approx_dist_inside(r, absZ, rmax, tOut2);
  double safety          = std::min(distInner, distOuter);
  return std::min(safety, distZ);
}

} // namespace vgo
