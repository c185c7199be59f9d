// Oracle: Orb and Cone. TEST INFRASTRUCTURE ONLY (see common.h).
#include "shapes.h"

namespace vgo {

// ============================================================== Orb
// OrbImplementation.h:86-99 GenericKernelForContainsAndInside
bool Orb::Contains(V3 const &p) const
{
  double rad2 = p.Mag2();
  double tolR = r + kTolerance;
  return !(rad2 >= tolR * tolR);
}

int Orb::Inside(V3 const &p) const
{
  double rad2  = p.Mag2();
  double tolRi = r - kTolerance;
  bool cin     = rad2 <= tolRi * tolRi;
  double tolRo = r + kTolerance;
  bool cout    = rad2 >= tolRo * tolRo;
  int res      = kSurface;
  if (cout) res = kOutside;
  if (cin) res = kInside;
  return res;
}

// OrbImplementation.h:101-124 and DetectIntersectionAndCalculateDistance :198-220
double Orb::DistanceToIn(V3 const &p, V3 const &d, double) const
{
  double rad = p.Mag();
  if (rad < r - kTolerance) return -1.;
  double pDotV = p.Dot(d);
  bool onSurf  = (rad >= r - kTolerance) && (rad <= r + kTolerance);
  if (onSurf && pDotV < 0.) return 0.;
  double c  = p.Mag2() - r * r;
  double d2 = pDotV * pDotV - c;
  if (d2 >= 0. && pDotV <= 0.) return -pDotV - std::sqrt(std::fabs(d2));
  return kInfLength;
}

// OrbImplementation.h:126-150
double Orb::DistanceToOut(V3 const &p, V3 const &d, double) const
{
  double rad = p.Mag();
  if (rad > r + kTolerance) return -1.;
  double pDotV = p.Dot(d);
  bool onSurf  = (rad >= r - kTolerance) && (rad <= r + kTolerance);
  if (onSurf && pDotV > 0.) return 0.;
  double c  = p.Mag2() - r * r;
  double d2 = pDotV * pDotV - c;
  if (d2 >= 0.) return -pDotV + std::sqrt(std::fabs(d2));
  return kInfLength;
}

// OrbImplementation.h:152-170
double Orb::SafetyToIn(V3 const &p) const
{
  double rad = p.Mag();
  if (rad < r - kTolerance) return -1.;
  if (rad > r - kTolerance && rad < r + kTolerance) return 0.;
  return rad - r;
}

// OrbImplementation.h:172-190
double Orb::SafetyToOut(V3 const &p) const
{
  double rad = p.Mag();
  if (rad > r + kTolerance) return -1.;
  if (rad > r - kTolerance && rad < r + kTolerance) return 0.;
  return r - rad;
}

// ============================================================== Cone
// ConeStruct.h:415-426 constructor, SetAndCheckDPhiAngle/SPhiAngle :309-346, CalculateCached :76-140
Cone::Cone(double rmin1_, double rmax1_, double rmin2_, double rmax2_, double dz_, double sphi_, double dphi_)
    : rmin1(rmin1_ < 0.0 ? 0.0 : rmin1_), rmax1(rmax1_), rmin2(rmin2_ < 0.0 ? 0.0 : rmin2_), rmax2(rmax2_), dz(dz_),
      sphi(sphi_), dphi(dphi_)
{
  wedge.Set(dphi_, sphi_); // the wedge keeps the vectors of the raw angles
  if (dphi_ >= kTwoPi - 0.5 * kAngTolerance) {
    dphi = kTwoPi;
    sphi = 0;
  } else if (dphi_ > 0) {
    dphi = dphi_;
  }
  wedge.dphi = dphi;
  // SetAndCheckSPhiAngle(_sphi)
  if (sphi_ < 0)
    sphi = kTwoPi - std::fmod(std::fabs(sphi_), kTwoPi);
  else
    sphi = std::fmod(sphi_, kTwoPi);
  if (sphi + dphi > kTwoPi) sphi -= kTwoPi;
  a1x = std::cos(sphi);
  a1y = std::sin(sphi);
  a2x = std::cos(sphi + dphi);
  a2y = std::sin(sphi + dphi);
  // CalculateCached
  origRmax1 = rmax1;
  origRmax2 = rmax2;
  origRmin1 = rmin1_;
  origRmin2 = rmin2_;
  if (rmin1 == rmax1) rmax1 += kConeTolerance;
  if (rmin2 == rmax2) rmax2 += kConeTolerance;
  sqRmin1     = rmin1 * rmin1;
  sqRmax1     = rmax1 * rmax1;
  sqRmin2     = rmin2 * rmin2;
  sqRmax2     = rmax2 * rmax2;
  tanRMin     = (rmin2 - rmin1) * 0.5 / dz;
  secRMin     = std::sqrt(1.0 + tanRMin * tanRMin);
  invSecRMin  = 1. / NonZero(secRMin);
  tanRMax     = (rmax2 - rmax1) * 0.5 / dz;
  secRMax     = std::sqrt(1.0 + tanRMax * tanRMax);
  invSecRMax  = 1. / NonZero(secRMax);
  innerSlope  = -(rmin1 - rmin2) / (2. * dz);
  outerSlope  = -(rmax1 - rmax2) / (2. * dz);
  innerOffset = rmin2 - innerSlope * dz;
  outerOffset = rmax2 - outerSlope * dz;
  innerTol    = kConeTolerance * secRMin;
  outerTol    = kConeTolerance * secRMax;
  if (rmin2 > rmin1) {
    innerApex    = 2 * dz * rmin1 / (rmin2 - rmin1);
    tanInnerApex = rmin2 / (2 * dz + innerApex);
  } else {
    innerApex    = 2 * dz * rmin2 / NonZero(rmin1 - rmin2);
    tanInnerApex = rmin1 / (2 * dz + innerApex);
  }
  if (rmin1 == 0. || rmin2 == 0.) innerApex = 0.;
  if (rmin1 == 0.) tanInnerApex = rmin2 / (2 * dz);
  if (rmin2 == 0.) tanInnerApex = rmin1 / (2 * dz);
  if (rmax2 > rmax1) {
    outerApex    = 2 * dz * rmax1 / (rmax2 - rmax1);
    tanOuterApex = rmax2 / (2 * dz + outerApex);
  } else {
    outerApex    = 2 * dz * rmax2 / NonZero(rmax1 - rmax2);
    tanOuterApex = rmax1 / (2 * dz + outerApex);
  }
  if (rmax1 == 0. || rmax2 == 0.) outerApex = 0.;
  if (rmax1 == 0.) tanOuterApex = rmax2 / (2 * dz);
  if (rmax2 == 0.) tanOuterApex = rmax1 / (2 * dz);
  zNormInner = tanRMin / NonZero(secRMin);
  zNormOuter = -tanRMax / NonZero(secRMax);
  tolIz      = dz - kHalfTolerance;
  tolOz      = dz + kHalfTolerance;
}

// source/UnplacedCone.cpp:149-213 (constructor radii, phi-reduced box)
void Cone::Extent(V3 &lo, V3 &hi) const
{
  const double mx = origRmax1 > origRmax2 ? origRmax1 : origRmax2;
  const double mn = origRmin1 > origRmin2 ? origRmin2 : origRmin1;
  lo              = V3(-mx, -mx, -dz);
  hi              = V3(mx, mx, dz);
  if (dphi == kTwoPi) return;
  PhiExtent(wedge, mx, mn, sphi, dphi, lo, hi);
}

// ConeUtilities.h:118-135 GetRadiusOfConeAtPoint
double Cone::RadiusAt(bool inner, double z) const
{
  if (inner) return rmin1 == rmin2 ? rmin1 : innerSlope * z + innerOffset;
  return origRmax1 == origRmax2 ? origRmax1 : outerSlope * z + outerOffset;
}

// ConeUtilities.h:62-100 PointInCyclicalSector (cone tolerance)
bool Cone::InSector(double x, double y, bool includeSurface) const
{
  double startCheck = (-x * a1y) + (y * a1x);
  double endCheck   = (-a2x * y) + (a2y * x);
  double t          = includeSurface ? -kConeTolerance : kConeTolerance;
  if (dphi <= M_PI) return (startCheck >= t) && (endCheck >= t);
  return (startCheck >= t) || (endCheck >= t);
}

// ConeUtilities.h:186-210 GetNormal
V3 Cone::Normal(bool inner, V3 const &p) const
{
  double rho = p.Perp();
  if (inner) {
    if (rmin1 == rmin2 && rmin1 != 0.) return V3(-p.x, -p.y, 0.);
    return V3(-p.x, -p.y, zNormInner * (rho * secRMin));
  }
  if (rmax1 == rmax2 && rmax1 != 0.) return V3(p.x, p.y, 0.);
  return V3(p.x, p.y, zNormOuter * (rho * secRMax));
}

// ConeUtilities.h:156-184 PhiPlaneTrajectoryIntersection (PositiveDirectionOfPhiVector = true)
bool Cone::PhiPlaneHit(double ax, double ay, double nx, double ny, V3 const &p, V3 const &d, bool toIn,
                       double &dist) const
{
  dist    = kInfLength;
  bool ok = toIn ? ((d.x * nx) + (d.y * ny) > 0.) : ((d.x * nx) + (d.y * ny) < 0.);
  double dirDotXY = (d.y * ax) - (d.x * ay);
  if (dirDotXY != 0) dist = ((ay * p.x) - (ax * p.y)) / NonZero(dirDotXY);
  ok = ok && dist > -kConeTolerance;
  double hitx = p.x + dist * d.x;
  double hity = p.y + dist * d.y;
  if (toIn) {
    double hitz = p.z + dist * d.z;
    double r2   = (hitx * hitx) + (hity * hity);
    double ri   = RadiusAt(true, hitz) + kTolerance;
    double ro   = RadiusAt(false, hitz) - kTolerance;
    ok          = ok && std::fabs(hitz) <= tolIz && (r2 >= ri * ri) && (r2 <= ro * ro);
    ok          = ok && ((hitx * ax) + (hity * ay)) > 0.;
  } else {
    ok = ok && ((hitx * ax) + (hity * ay)) >= 0.;
  }
  return ok;
}

// scalar ConeHelpers<Precision>::DetectIntersectionAndCalculateDistanceToConicalSurface (ConeUtilities.h:604-760)
bool Cone::ConicalHit(bool toIn, bool inner, V3 const &p, V3 const &d, double &distance) const
{
  distance       = kInfLength;
  double rho2    = p.Perp2();
  double coneRad = RadiusAt(inner, p.z);
  double tol     = inner ? innerTol : outerTol;
  bool onSurf    = (rho2 >= (coneRad * coneRad - tol * coneRad)) && (rho2 <= (coneRad * coneRad + tol * coneRad)) &&
                (std::fabs(p.z) < (dz + kConeTolerance));
  V3 n         = Normal(inner, p);
  bool tangent = std::fabs(d.Dot(n)) == 0.;
  if (onSurf && tangent) return false;
  if (onSurf) {
    bool moving = toIn ? (d.Dot(n) <= 0.) : (d.Dot(n) >= 0.);
    if (!hasPhi()) {
      if (moving) return distance = 0., true;
    } else if (InSector(p.x, p.y, true) && moving) {
      return distance = 0., true;
    }
  }
  bool ok        = false;
  double pDotV2D = p.x * d.x + p.y * d.y;
  double a, b, c;
  if (inner) {
    if (rmin1 == rmin2) {
      b = pDotV2D;
      a = d.Perp2();
      c = p.Perp2() - rmin2 * rmin2;
    } else {
      double newPz = rmin2 > rmin1 ? (p.z + dz + innerApex) * tanInnerApex : (p.z - dz - innerApex) * tanInnerApex;
      double dirT  = d.z * tanInnerApex;
      a            = (d.x * d.x) + (d.y * d.y) - dirT * dirT;
      b            = pDotV2D - (newPz * dirT);
      c            = p.Perp2() - (newPz * newPz);
    }
    double b2 = b * b, ac = a * c;
    if (b2 < ac) return false;
    double d2 = b2 - ac, delta = std::sqrt(d2);
    if (toIn) {
      distance = b >= 0. ? (c / NonZero(-b - delta)) : (-b + delta) / NonZero(a);
    } else {
      if (b == 0. && delta == 0.) return false;
      distance = b >= 0. ? (-b - delta) / NonZero(a) : (c / NonZero(-b + delta));
    }
    if (distance < 0.) return false;
    double newZ = p.z + (d.z * distance);
    ok          = std::fabs(newZ) < dz;
  } else {
    if (origRmax1 == origRmax2) {
      a = d.Perp2();
      b = pDotV2D;
      c = (p.Perp2() - origRmax2 * origRmax2);
    } else {
      double newPz = origRmax2 > origRmax1 ? (p.z + dz + outerApex) * tanOuterApex
                                           : (p.z - dz - outerApex) * tanOuterApex;
      double dirT = d.z * tanOuterApex;
      a           = d.x * d.x + d.y * d.y - dirT * dirT;
      b           = (pDotV2D - (newPz * dirT));
      c           = (p.Perp2() - (newPz * newPz));
    }
    double b2 = b * b, ac = a * c, d2 = b2 - ac;
    if (d2 < 0) return false;
    double delta = std::sqrt(d2);
    if (toIn) {
      if (b == 0. && delta == 0.) return false;
      distance    = b > 0. ? (-b - delta) / NonZero(a) : (c / NonZero(-b + delta));
      double newZ = p.z + (d.z * distance);
      ok          = std::fabs(newZ) < dz + kHalfTolerance;
    } else {
      distance = b < 0. ? (-b + delta) / NonZero(a) : (c / NonZero(-b - delta));
      ok       = distance > 0.;
    }
    if (distance < 0.) return false;
  }
  if (hasPhi()) {
    double hitx = 0, hity = 0;
    if (distance < kInfLength) {
      hitx = p.x + distance * d.x;
      hity = p.y + distance * d.y;
    }
    ok = ok && InSector(hitx, hity, true) && (distance < kInfLength);
  }
  return ok;
}

// scalar ConeHelpers<Precision>::GenericKernelForContainsAndInside (ConeUtilities.h:762-800)
void Cone::ContainsInside(V3 const &p, bool forInside, bool &cin, bool &cout) const
{
  double absz = std::fabs(p.z);
  cout        = absz > (forInside ? dz + kConeTolerance : dz);
  cin         = forInside && absz < dz - kConeTolerance;
  if (cout) return;
  double r2   = p.x * p.x + p.y * p.y;
  double rmax = origRmax1 == origRmax2 ? origRmax1 : outerSlope * p.z + outerOffset;
  cout        = cout || (r2 > (forInside ? rmax * (rmax + 2.0 * outerTol) : rmax * rmax));
  if (forInside) cin = cin && (r2 < std::fabs(rmax * (rmax - 2.0 * outerTol)));
  if (cout) return;
  if (hasRmin()) {
    double rmin = innerSlope * p.z + innerOffset;
    cout        = cout || (r2 <= (forInside ? std::fabs(rmin * (rmin - 2.0 * innerTol)) : rmin * rmin));
    if (forInside) cin = cin && (r2 > rmin * (rmin + 2.0 * innerTol));
    if (cout) return;
  }
  if (hasPhi()) {
    bool wi = false, wo = false;
    wedge.ContainsInside(p, forInside, wi, wo);
    cout = cout || wo;
    if (forInside) cin = cin && wi;
  }
}

bool Cone::Contains(V3 const &p) const
{
  bool ci = false, co = false;
  ContainsInside(p, false, ci, co);
  return !co;
}

int Cone::Inside(V3 const &p) const
{
  bool ci = false, co = false;
  ContainsInside(p, true, ci, co);
  int r = kSurface;
  if (co) r = kOutside;
  if (ci) r = kInside;
  return r;
}

// ConeImplementation.h:118-255
double Cone::DistanceToIn(V3 const &p, V3 const &d, double) const
{
  double distz = std::fabs(p.z) - dz;
  if ((distz > kConeTolerance && (p.z * d.z) >= 0.) || (std::fabs(distz) < kConeTolerance && (p.z * d.z) > 0.))
    return kInfLength;
  double outerRad = RadiusAt(false, p.z);
  double rsq      = p.Perp2();
  if ((rsq > std::fabs(outerRad * (outerRad - 2.0 * outerTol))) && (d.Dot(Normal(false, p)) >= 0.))
    return kInfLength;
  bool inside = distz < -kConeTolerance;
  inside      = inside && rsq < std::fabs(outerRad * (outerRad - 2.0 * outerTol));
  if (hasRmin()) {
    double innerRad = RadiusAt(true, p.z);
    inside          = inside && rsq > innerRad * (innerRad + 2.0 * innerTol);
  }
  if (hasPhi()) inside = inside && InSector(p.x, p.y, false);
  if (inside) return -1.0;
  double distance = kInfLength;
  distz /= NonZero(std::fabs(d.z));
  double hitx = p.x + distz * d.x;
  double hity = p.y + distz * d.y;
  double r2   = (hitx * hitx) + (hity * hity);
  bool top    = (p.z >= tolIz) && (r2 <= sqRmax2 + kTolerance);
  bool bottom = (p.z <= -tolIz) && (r2 <= sqRmax1 + kTolerance);
  bool okz    = top || bottom;
  if (hasRmin()) {
    top    = top && (r2 >= sqRmin2 - kTolerance);
    bottom = bottom && (r2 >= sqRmin1 - kTolerance);
    okz    = okz && (top || bottom);
  }
  if (hasPhi()) okz = okz && InSector(hitx, hity, true);
  if (okz) return distz;
  double dOuter = kInfLength;
  bool okOuter  = ConicalHit(true, false, p, d, dOuter);
  if (okOuter && dOuter < distance) return dOuter;
  if (hasRmin()) {
    double dInner = kInfLength;
    bool okInner  = ConicalHit(true, true, p, d, dInner);
    if (okInner && dInner < distance) distance = dInner;
  }
  if (hasPhi()) {
    double dp = kInfLength;
    bool okp  = PhiPlaneHit(a1x, a1y, wedge.normal1.x, wedge.normal1.y, p, d, true, dp);
    if (okp && dp < distance) distance = dp;
    okp = PhiPlaneHit(a2x, a2y, wedge.normal2.x, wedge.normal2.y, p, d, true, dp);
    if (okp && dp < distance) distance = dp;
  }
  return distance;
}

// ConeImplementation.h:257-370
double Cone::DistanceToOut(V3 const &p, V3 const &d, double) const
{
  double distz    = std::fabs(p.z) - dz;
  bool outside    = distz > kConeTolerance;
  double rsq      = p.Perp2();
  double outerRad = RadiusAt(false, p.z);
  outside         = outside || rsq > outerRad * (outerRad + 2.0 * outerTol);
  if (outside) return -1.0;
  if (hasRmin()) {
    double innerRad = RadiusAt(true, p.z);
    outside         = outside || rsq < std::fabs(innerRad * (innerRad - 2.0 * innerTol));
    if (outside) return -1.0;
  }
  // the phi "outside" test only runs when some lane is already outside: never for a scalar
  bool up = d.z > 0., down = d.z < 0.;
  if ((up && p.z > 0. && std::fabs(distz) < kConeTolerance) || (down && p.z < 0. && std::fabs(distz) < kConeTolerance))
    return distz;
  double distance = kInfLength;
  double dirZInv  = 1. / NonZero(d.z);
  if (up) distance = (dz - p.z) * dirZInv;
  if (down) distance = (-dz - p.z) * dirZInv;
  double dOuter = kInfLength;
  bool okOuter  = ConicalHit(false, false, p, d, dOuter);
  if (okOuter && dOuter < distance) distance = dOuter;
  if (hasRmin()) {
    double dInner = kInfLength;
    bool okInner  = ConicalHit(false, true, p, d, dInner);
    if (okInner && dInner < distance) distance = dInner;
  }
  if (hasPhi()) {
    bool onStart = wedge.IsOnSurfaceGeneric(wedge.along1, wedge.normal1, p);
    bool onEnd   = wedge.IsOnSurfaceGeneric(wedge.along2, wedge.normal2, p);
    double c1    = d.x * -wedge.normal1.x + d.y * -wedge.normal1.y + d.z * -wedge.normal1.z;
    double c2    = d.x * -wedge.normal2.x + d.y * -wedge.normal2.y + d.z * -wedge.normal2.z;
    if ((onStart && c1 > 0.) || (onEnd && c2 > 0.)) return 0.;
    double dp = kInfLength;
    bool okp  = PhiPlaneHit(a1x, a1y, wedge.normal1.x, wedge.normal1.y, p, d, false, dp);
    if (okp && dp < distance) distance = dp;
    okp = PhiPlaneHit(a2x, a2y, wedge.normal2.x, wedge.normal2.y, p, d, false, dp);
    if (okp && dp < distance) distance = dp;
  }
  if (distance < 0. && std::fabs(distance) < kTolerance) distance = 0.;
  return distance;
}

// ConeImplementation.h:498-515 SafeDistanceToConicalSurface
double Cone::SafeDistConical(bool inner, V3 const &p) const
{
  double rho = p.Perp();
  if (inner) {
    double pRMin = tanRMin * p.z + (rmin1 + rmin2) * 0.5;
    return (rho - pRMin) * invSecRMin;
  }
  double pRMax = tanRMax * p.z + (rmax1 + rmax2) * 0.5;
  return (pRMax - rho) * invSecRMax;
}

// ConeImplementation.h:372-425
double Cone::SafetyToIn(V3 const &p) const
{
  double distz    = std::fabs(p.z) - dz;
  bool inside     = distz < -kConeTolerance;
  double outerRad = RadiusAt(false, p.z);
  double rsq      = p.Perp2();
  inside          = inside && rsq < std::fabs(outerRad * (outerRad - 2.0 * outerTol));
  if (hasRmin()) {
    double innerRad = RadiusAt(true, p.z);
    inside          = inside && rsq > innerRad * (innerRad + 2.0 * innerTol);
  }
  if (hasPhi() && inside) inside = inside && InSector(p.x, p.y, false);
  if (inside) return -1.0;
  double safeZ  = std::fabs(p.z) - dz;
  double sOuter = -SafeDistConical(false, p);
  double sInner = -kInfLength;
  if (hasRmin()) sInner = -SafeDistConical(true, p);
  double safety = std::max(safeZ, std::max(sOuter, sInner));
  if (hasPhi()) safety = std::max(wedge.SafetyToIn(p), safety);
  if (std::fabs(safety) < kTolerance) safety = 0.;
  return safety;
}

// ConeImplementation.h:427-478
double Cone::SafetyToOut(V3 const &p) const
{
  double distz    = std::fabs(p.z) - dz;
  double rsq      = p.Perp2();
  bool outside    = distz > kConeTolerance;
  double outerRad = RadiusAt(false, p.z);
  outside         = outside || rsq > outerRad * (outerRad + 2.0 * outerTol);
  if (hasRmin()) {
    double innerRad = RadiusAt(true, p.z);
    outside         = outside || rsq < std::fabs(innerRad * (innerRad - 2.0 * innerTol));
  }
  if (outside) return -1.0;
  double safeZ  = dz - std::fabs(p.z);
  double sOuter = SafeDistConical(false, p);
  double sInner = kInfLength;
  if (hasRmin()) sInner = SafeDistConical(true, p);
  double safety = std::min(safeZ, std::min(sOuter, sInner));
  if (hasPhi()) safety = std::min(wedge.SafetyToOut(p), safety);
  if (std::fabs(safety) < kTolerance) safety = 0.;
  return safety;
}

} // namespace vgo
