// Oracle: Sphere (rmin/rmax shells, phi wedge, theta cones) restated from VecGeom's
// scalar backend. TEST INFRASTRUCTURE ONLY (see common.h).
//
// Sources: VecGeom/volumes/SphereStruct.h (constructor checks), VecGeom/volumes/ThetaCone_Evolution.h
// (theta section), VecGeom/volumes/Wedge_Evolution.h (phi section), VecGeom/volumes/SphereUtilities.h
// (surface/inside predicates), VecGeom/volumes/kernel/SphereImplementation.h (the kernels),
// source/UnplacedSphere.cpp:76-165 (extent). Early returns follow the scalar backend.
#include "shapes.h"

namespace vgo {

namespace {
inline double Min(double a, double b) { return std::min(a, b); }
inline double Max(double a, double b) { return std::max(a, b); }
} // namespace

// ================================================================ Wedge additions
// Wedge_Evolution.h:235-247
bool Wedge::OnSurfaceMoving(bool start, bool movingOut, V3 const &p, V3 const &d) const
{
  V3 const &n = start ? normal1 : normal2;
  if (!IsOnSurfaceGeneric(start ? along1 : along2, n, p)) return false;
  const double c = d.Dot(-n);
  return movingOut ? (c > 0.005 * kHalfTolerance) : (c < 0.005 * kHalfTolerance);
}

// Wedge_Evolution.h:394-414
void Wedge::DistanceToIn(V3 const &p, V3 const &d, double &d1, double &d2) const
{
  d1 = kInfLength;
  d2 = kInfLength;
  const double comp1 = d.x * normal1.x + d.y * normal1.y;
  const double comp2 = d.x * normal2.x + d.y * normal2.y;
  if (comp1 > 0.) {
    const double t = -(p.x * normal1.x + p.y * normal1.y) / comp1;
    if (t > 0.) d1 = t;
  }
  if (comp2 > 0.) {
    const double t = -(p.x * normal2.x + p.y * normal2.y) / comp2;
    if (t > 0.) d2 = t;
  }
}

// Wedge_Evolution.h:416-436
void Wedge::DistanceToOut(V3 const &p, V3 const &d, double &d1, double &d2) const
{
  const double comp1 = d.x * normal1.x + d.y * normal1.y;
  const double comp2 = d.x * normal2.x + d.y * normal2.y;
  d1                 = kInfLength;
  d2                 = kInfLength;
  if (comp1 < 0.) {
    const double t = -(p.x * normal1.x + p.y * normal1.y) / comp1;
    if (t > 0.) d1 = t;
  }
  if (comp2 < 0.) {
    const double t = -(p.x * normal2.x + p.y * normal2.y) / comp2;
    if (t > 0.) d2 = t;
  }
}

// ================================================================ ThetaCone
// ThetaCone_Evolution.h:62-83 (the kHalfPi macro there is 0.5 * kPi)
void ThetaCone::Set(double s, double d)
{
  sTheta     = s;
  dTheta     = d;
  angTol     = kTolerance;
  eTheta     = sTheta + dTheta;
  halfAngTol = 0.5 * angTol;
  double ts = sTheta, te = eTheta;
  if (sTheta > kPi / 2) ts = kPi - sTheta;
  if (eTheta > kPi / 2) te = kPi - eTheta;
  tanS   = std::tan(ts);
  tanS2  = tanS * tanS;
  tanE   = std::tan(te);
  tanE2  = tanE * tanE;
  tanBis = std::tan(ts + (dTheta / 2));
  if (sTheta > kPi / 2 && eTheta > kPi / 2) tanBis = std::tan(ts - (dTheta / 2));
  slope1 = std::tan(kPi / 2 - sTheta);
  slope2 = std::tan(kPi / 2 - eTheta);
}

// ThetaCone_Evolution.h:150-172
V3 ThetaCone::Normal(bool start, V3 const &p) const
{
  if (start) {
    V3 n(p.x, p.y, -tanS2 * p.z);
    return sTheta <= 0.5 * kPi ? -n : n;
  }
  V3 n(p.x, p.y, -tanE2 * p.z);
  return eTheta <= 0.5 * kPi ? n : -n;
}

// ThetaCone_Evolution.h:183-196
bool ThetaCone::IsOnSurface(bool start, V3 const &p) const
{
  const double rhs  = start ? std::fabs(tanS * p.z) : std::fabs(tanE * p.z);
  const double rho2 = p.Perp2();
  return rho2 >= MinusTolSq(rhs) && rho2 <= PlusTolSq(rhs);
}

// ThetaCone_Evolution.h:222-233
bool ThetaCone::OnSurfaceMoving(bool start, bool movingOut, V3 const &p, V3 const &d) const
{
  if (!IsOnSurface(start, p)) return false;
  const double c = d.Dot(Normal(start, p));
  return movingOut ? c > 0. : c < 0.;
}

// ThetaCone_Evolution.h:640-707
bool ThetaCone::IsCompletelyInside(V3 const &p) const
{
  const double hp    = 0.5 * kPi;
  const double rho   = std::sqrt(p.Mag2() - (p.z * p.z));
  const double c1r   = std::fabs(p.z * tanS);
  const double c2r   = std::fabs(p.z * tanE);
  const bool onZ     = p.z != 0. && p.x == 0. && p.y == 0.;
  const bool onXY    = p.z == 0. && (p.x != 0. || p.y != 0.);
  const double tol20 = 2 * angTol * 10.;
  bool ci = onZ && ((sTheta == 0. && eTheta == kPi) || (p.z > 0. && sTheta == 0.) || (p.z < 0. && eTheta == kPi));
  ci = ci || (onXY && (sTheta < hp && eTheta > hp && (hp - sTheta) > angTol && (eTheta - hp) > kTolerance));
  if (sTheta < hp + halfAngTol) {
    if (eTheta < hp + halfAngTol && sTheta < eTheta) {
      const double mn = c1r + tol20, mx = c2r - tol20;
      ci = ci || (((rho <= mx) && (rho >= mn) && (p.z > 0.) && (sTheta != 0.)) ||
                  ((rho <= mx) && (sTheta == 0.) && (p.z > 0.)));
    }
    if (eTheta > hp + halfAngTol && sTheta < eTheta) {
      const double mn = c1r + tol20, mx = c2r + tol20;
      ci = ci || (((rho >= mn) && (p.z > 0.)) || ((rho >= mx) && (p.z < 0.)));
    }
    if (eTheta >= hp - halfAngTol && eTheta <= hp + halfAngTol) ci = ci && !(std::fabs(p.z) < halfAngTol);
  }
  if (eTheta > hp + halfAngTol) {
    if (sTheta >= hp - halfAngTol && sTheta <= hp + halfAngTol) ci = ci && !(std::fabs(p.z) < halfAngTol);
    if (sTheta > hp + halfAngTol && sTheta < eTheta) {
      const double mn = c1r - tol20, mx = c2r + tol20;
      ci = ci || (((rho <= mn) && (rho >= mx) && (p.z < 0.) && (eTheta != kPi)) ||
                  ((rho <= mn) && (p.z < 0.) && (eTheta == kPi)));
    }
  }
  return ci;
}

// ThetaCone_Evolution.h:711-788
bool ThetaCone::IsCompletelyOutside(V3 const &p) const
{
  const double hp    = 0.5 * kPi;
  const double rho   = std::sqrt(p.Perp2());
  const double c1r   = std::fabs(p.z * tanS);
  const double c2r   = std::fabs(p.z * tanE);
  const bool onZ     = p.z != 0. && p.x == 0. && p.y == 0.;
  const bool onXY    = p.z == 0. && (p.x != 0. || p.y != 0.);
  const double tol20 = 2 * angTol * 10.;
  bool co = onZ && ((sTheta != 0. && eTheta != kPi) || (p.z > 0. && sTheta != 0.) || (p.z < 0. && eTheta != kPi));
  co = co || (onXY && (((sTheta < hp) && (eTheta < hp) && ((hp - sTheta) > angTol) && ((hp - eTheta) > kTolerance)) ||
                       ((sTheta > hp && eTheta > hp) && ((sTheta - hp) > angTol) && ((eTheta - hp) > kTolerance))));
  if (sTheta < hp + halfAngTol) {
    if (eTheta < hp + halfAngTol && sTheta < eTheta) {
      const double mn = c1r - tol20, mx = c2r + tol20;
      co = co || ((rho < mn) || (rho > mx) || (p.z < 0.));
    }
    if (eTheta > hp + halfAngTol && sTheta < eTheta) {
      const double mn = c1r - tol20, mx = c2r - tol20;
      co = co || (((rho < mn) && (p.z > 0.)) || ((rho < mx) && (p.z < 0.)));
    }
    if (eTheta >= hp - halfAngTol && eTheta <= hp + halfAngTol) co = co && !(std::fabs(p.z) < halfAngTol);
  }
  if (eTheta > hp + halfAngTol) {
    if (sTheta >= hp - halfAngTol && sTheta <= hp + halfAngTol) co = co && !(std::fabs(p.z) < halfAngTol);
    if (sTheta > hp + halfAngTol && sTheta < eTheta) {
      const double mn = c1r + tol20, mx = c2r - tol20;
      co = co || ((rho < mx) || (rho > mn) || (p.z > 0.));
    }
  }
  return co;
}

// ThetaCone_Evolution.h:372-377
double ThetaCone::DistanceToLine(double slope, double x, double y) const
{
  const double dist = (y - slope * x) / std::sqrt(1. + slope * slope);
  return std::fabs(dist);
}

// ThetaCone_Evolution.h:283-325
double ThetaCone::SafetyToIn(V3 const &p) const
{
  const double rad = std::sqrt(p.x * p.x + p.y * p.y);
  const double s1  = DistanceToLine(slope1, rad, p.z);
  const double s2  = DistanceToLine(slope2, rad, p.z);
  double safe      = Min(s1, s2);
  if (Contains(p)) return 0.;
  if (sTheta < kPi / 2 + halfAngTol) {
    if (eTheta < kPi / 2 + halfAngTol && sTheta < eTheta && p.z < 0.) safe = s2;
    if (eTheta > kPi / 2 + halfAngTol && sTheta < eTheta) {
      if (p.z > 0.) safe = s1;
      if (p.z < 0.) safe = s2;
    }
  }
  if (eTheta > kPi / 2 + halfAngTol && sTheta > kPi / 2 + halfAngTol && sTheta < eTheta && p.z > 0.) safe = s1;
  return safe;
}

// ThetaCone_Evolution.h:332-369
double ThetaCone::SafetyToOut(V3 const &p) const
{
  const double rad = std::sqrt(p.x * p.x + p.y * p.y);
  const double bis = std::fabs(p.z * tanBis);
  bool cond        = false;
  const double s1  = DistanceToLine(slope1, rad, p.z);
  const double s2  = DistanceToLine(slope2, rad, p.z);
  if (sTheta < kPi / 2 + halfAngTol) {
    if (eTheta < kPi / 2 + halfAngTol && sTheta < eTheta) cond = (rad < bis) && (sTheta != 0.);
    if (eTheta > kPi / 2 + halfAngTol && sTheta < eTheta) cond = s1 < s2;
  }
  if (eTheta > kPi / 2 + halfAngTol && sTheta > kPi / 2 + halfAngTol && sTheta < eTheta)
    cond = !((rad < bis) && (eTheta != kPi));
  return cond ? s1 : s2;
}

// ThetaCone_Evolution.h:383-497. d1/d2/i1/i2 keep the caller's values where no branch assigns them.
void ThetaCone::DistanceToIn(V3 const &p, V3 const &d, double &d1, double &d2, bool &i1, bool &i2) const
{
  const double hp  = 0.5 * kPi;
  const double t3  = 3. * kTolerance;
  bool done        = false;
  double r1        = kInfLength, r2 = kInfLength;
  const double pv  = p.x * d.x + p.y * d.y;
  const double rho2 = p.x * p.x + p.y * p.y;
  const double dr2 = d.Perp2();
  const double b   = pv - p.z * d.z * tanS2;
  const double a   = dr2 - d.z * d.z * tanS2;
  const double c   = rho2 - p.z * p.z * tanS2;
  const double dd  = b * b - a * c;
  const double ai  = 1. / NonZero(a);
  if (dd > 0.) r1 = (-b + std::sqrt(std::fabs(dd))) * ai;
  done = done || (std::fabs(r1) < t3);
  if (std::fabs(r1) < t3) r1 = 0.;
  if (!done && r1 < 0.) r1 = kInfLength;
  const double b2  = pv - p.z * d.z * tanE2;
  const double a2  = dr2 - d.z * d.z * tanE2;
  const double c2  = rho2 - p.z * p.z * tanE2;
  const double dd2 = b2 * b2 - a2 * c2;
  const double a2i = 1. / NonZero(a2);
  if (dd2 > 0.) r2 = (-b2 - std::sqrt(std::fabs(dd2))) * a2i;
  if (!done && std::fabs(r2) < t3) r2 = 0.;
  done = done || (std::fabs(r2) < t3);
  if (!done && r2 < 0.) r2 = kInfLength;

  if (sTheta < hp + halfAngTol) {
    if (eTheta < hp + halfAngTol && sTheta < eTheta) {
      d1              = r1;
      d2              = r2;
      const double z1 = p.z + d1 * d.z;
      const double z2 = p.z + d2 * d.z;
      i1              = (dd > 0.) && (z1 > 0.);
      i2              = (dd2 > 0.) && (z2 > 0.);
    }
    if (eTheta >= hp - halfAngTol && eTheta <= hp + halfAngTol) {
      if (d.z > 0.) d2 = -p.z / d.z;
      const double z2 = p.z + d2 * d.z;
      i2              = (d2 != kInfLength) && (std::fabs(z2) < halfAngTol);
    }
    if (eTheta > hp + halfAngTol && sTheta < eTheta) {
      d1 = r1;
      if (dd2 > 0.) r2 = (-b2 + std::sqrt(std::fabs(dd2))) * a2i;
      done = std::fabs(r2) < t3;
      if (std::fabs(r2) < t3) r2 = 0.;
      if (!done && r2 < 0.) r2 = kInfLength;
      d2              = r2;
      const double z1 = p.z + d1 * d.z;
      const double z2 = p.z + d2 * d.z;
      i1              = (dd > 0) && (d1 != kInfLength) && (z1 > 0.);
      i2              = (dd2 > 0) && (d2 != kInfLength) && (z2 < 0.);
    }
  }
  if (sTheta >= hp - halfAngTol && eTheta > hp + halfAngTol && sTheta < eTheta) {
    if (dd > 0.) r1 = (-b - std::sqrt(std::fabs(dd))) * ai;
    done = std::fabs(r1) < t3;
    if (std::fabs(r1) < t3) r1 = 0.;
    if (!done && r1 < 0.) r1 = kInfLength;
    d1 = r1;
    if (dd2 > 0.) r2 = (-b2 + std::sqrt(std::fabs(dd2))) * a2i;
    done = std::fabs(r2) < t3;
    if (std::fabs(r2) < t3) r2 = 0.;
    if (!done && r2 < 0.) r2 = kInfLength;
    d2              = r2;
    const double z1 = p.z + d1 * d.z;
    const double z2 = p.z + d2 * d.z;
    i1              = (dd > 0) && (d1 != kInfLength) && (z1 < 0.);
    i2              = (dd2 > 0) && (d2 != kInfLength) && (z2 < 0.);
  }
  if (sTheta >= hp - halfAngTol && sTheta <= hp + halfAngTol) {
    if (d.z < 0.) d1 = -p.z / d.z;
    const double z1 = p.z + d1 * d.z;
    i1              = (d1 != kInfLength) && (std::fabs(z1) < halfAngTol);
  }
}

// ThetaCone_Evolution.h:503-636
void ThetaCone::DistanceToOut(V3 const &p, V3 const &d, double &d1, double &d2, bool &i1, bool &i2) const
{
  const double inf  = kInfLength;
  double r1         = kInfLength, r2 = kInfLength;
  const double pv   = p.x * d.x + p.y * d.y;
  const double rho2 = p.x * p.x + p.y * p.y;
  const double b    = pv - p.z * d.z * tanS2;
  const double a    = d.x * d.x + d.y * d.y - d.z * d.z * tanS2;
  const double c    = rho2 - p.z * p.z * tanS2;
  double dd         = b * b - a * c;
  if (dd < 0. && std::fabs(dd) < kHalfTolerance) dd = 0.;
  if ((dd >= 0.) && b >= 0. && a != 0.) r1 = (-b - std::sqrt(std::fabs(dd))) / NonZero(a);
  if ((dd >= 0.) && b < 0.) r1 = c / NonZero(-b + std::sqrt(std::fabs(dd)));
  if (r1 < 0.) r1 = inf;
  const double b2 = p.x * d.x + p.y * d.y - p.z * d.z * tanE2;
  const double a2 = d.x * d.x + d.y * d.y - d.z * d.z * tanE2;
  const double c2 = p.x * p.x + p.y * p.y - p.z * p.z * tanE2;
  double dd2      = (b2 * b2) - (a2 * c2);
  if (dd2 < 0. && std::fabs(dd2) < kHalfTolerance) dd2 = 0.;
  if ((dd2 >= 0.) && b2 >= 0.) r2 = c2 / NonZero(-b2 - std::sqrt(std::fabs(dd2)));
  if ((dd2 >= 0.) && b2 < 0. && a2 != 0.) r2 = (-b2 + std::sqrt(std::fabs(dd2))) / NonZero(a2);
  if (r2 < 0. && std::fabs(r2) > kTolerance) r2 = inf;
  if (std::fabs(r2) < kTolerance) r2 = 0.;

  if (sTheta < kPi / 2 + halfAngTol) {
    if (eTheta < kPi / 2 + halfAngTol && sTheta < eTheta) {
      d1              = r1;
      d2              = r2;
      const double z1 = p.z + d1 * d.z;
      const double z2 = p.z + d2 * d.z;
      i1              = (dd > 0) && (d1 != kInfLength) && (z1 > -kHalfTolerance);
      i2              = (dd2 > 0) && (d2 != kInfLength) && (z2 > -kHalfTolerance);
      const double dr2 = d.x * d.x + d.y * d.y;
      double zs = kInfLength, ze = kInfLength;
      if (sTheta != 0.) zs = dr2 / tanS;
      if (eTheta != 0.) ze = dr2 / tanE;
      const bool cond = (p.x == 0. && p.y == 0. && p.z == 0. && d.z < zs && d.z < ze);
      if (cond) d1 = 0., d2 = 0.;
      i1 = i1 || cond;
      i2 = i2 || cond;
    }
    if (eTheta >= kPi / 2 - halfAngTol && eTheta <= kPi / 2 + halfAngTol) {
      d1 = r1;
      d2 = inf;
      if (d.z < 0.) d2 = -1. * p.z / d.z;
      const double z2 = p.z + d2 * d.z;
      const double z1 = p.z + d1 * d.z;
      i2              = (d2 != kInfLength) && (std::fabs(z2) < kHalfTolerance) && !(d.z == 0.);
      i1              = (dd >= 0) && (d1 != kInfLength) && (std::fabs(z1) < kHalfTolerance) && !(d.z == 0.);
    }
    if (eTheta > kPi / 2 + halfAngTol && sTheta < eTheta) {
      d1 = r1;
      if ((dd2 >= 0.) && b2 > 0. && a2 != 0.) r2 = (-b2 - std::sqrt(std::fabs(dd2))) / NonZero(a2);
      if ((dd2 >= 0.) && b2 <= 0.) r2 = c2 / NonZero(-b2 + std::sqrt(std::fabs(dd2)));
      if (r2 < 0.) r2 = inf;
      d2              = r2;
      const double z1 = p.z + d1 * d.z;
      const double z2 = p.z + d2 * d.z;
      i1              = (dd >= 0) && (d1 != kInfLength) && (z1 > -kHalfTolerance);
      i2              = (dd2 >= 0) && (d2 != kInfLength) && (z2 < kHalfTolerance);
    }
  }
  if (eTheta > kPi / 2 + halfAngTol && sTheta < eTheta) {
    r2 = kInfLength;
    if ((dd2 >= 0.) && b2 > 0. && a2 != 0.) r2 = (-b2 - std::sqrt(std::fabs(dd2))) / NonZero(a2);
    if ((dd2 >= 0.) && b2 <= 0.) r2 = c2 / NonZero(-b2 + std::sqrt(std::fabs(dd2)));
    d2 = r2;
    if (sTheta > kPi / 2 + halfAngTol) {
      if ((dd >= 0.) && b > 0.) r1 = c / NonZero(-b - std::sqrt(std::fabs(dd)));
      if ((dd >= 0.) && b <= 0. && a != 0.) r1 = (-b + std::sqrt(std::fabs(dd))) / NonZero(a);
      if (r1 < 0.) r1 = inf;
      d1               = r1;
      const double z1  = p.z + d1 * d.z;
      i1               = (dd > 0) && (d1 != kInfLength) && (z1 < kHalfTolerance);
      const double z2  = p.z + d2 * d.z;
      i2               = (dd2 > 0) && (d2 != kInfLength) && (z2 < kHalfTolerance);
      const double dr2 = d.x * d.x + d.y * d.y;
      double zs = -kInfLength, ze = -kInfLength;
      if (tanS != 0.) zs = -dr2 / tanS;
      if (tanE != 0.) ze = -dr2 / tanE;
      const bool cond = (p.x == 0. && p.y == 0. && p.z == 0. && d.z > zs && d.z > ze);
      if (cond) d1 = 0., d2 = 0.;
      i1 = i1 || cond;
      i2 = i2 || cond;
    }
  }
  if (sTheta >= kPi / 2 - halfAngTol && sTheta <= kPi / 2 + halfAngTol) {
    d2 = r2;
    d1 = kInfLength;
    if (d.z > 0.) d1 = -1. * p.z / NonZero(d.z);
    const double z1 = p.z + d1 * d.z;
    const double z2 = p.z + d2 * d.z;
    i1              = (d1 != kInfLength) && (std::fabs(z1) < kHalfTolerance) && (d.z != 0.);
    i2              = (dd2 >= 0) && (d2 != kInfLength) && (std::fabs(z2) < kHalfTolerance) && (d.z != 0.);
  }
}

// ================================================================ Sphere
// SphereStruct.h:245-266 with CheckPhiAngles (:128-150, :180-199) and CheckThetaAngles (:100-126)
Sphere::Sphere(double rmin_, double rmax_, double sphi_, double dphi_, double stheta_, double dtheta_)
    : rmin(rmin_), rmax(rmax_), sphi(sphi_), dphi(dphi_), stheta(stheta_), dtheta(dtheta_)
{
  wedge.Set(dphi_, sphi_);
  theta.Set(stheta_, dtheta_);
  // CheckDPhiAngle
  if (dphi_ >= 2 * kPi - kAngTolerance * 0.5) {
    dphi    = 2 * kPi;
    sphi    = 0;
    fullPhi = true;
  } else {
    fullPhi = false;
    if (dphi_ > 0) dphi = dphi_;
  }
  // CheckSPhiAngle
  if (!fullPhi) {
    if (sphi_ < 0)
      sphi = 2 * kPi - std::fmod(std::fabs(sphi_), 2 * kPi);
    else
      sphi = std::fmod(sphi_, 2 * kPi);
    if (sphi + dphi > 2 * kPi) sphi -= 2 * kPi;
  }
  full = fullPhi && fullTheta;
  // CheckThetaAngles
  if (!((stheta_ < 0) || (stheta_ > kPi))) stheta = stheta_;
  if (dtheta_ + stheta_ >= kPi)
    dtheta = kPi - stheta_;
  else if (dtheta_ > 0)
    dtheta = dtheta_;
  fullTheta = !(dtheta - stheta < kPi);
  full      = fullPhi && fullTheta;
  etheta    = stheta + dtheta;
}

// SphereImplementation.h:87-126
void Sphere::ContainsInside(V3 const &p, bool forInside, bool &ci, bool &co) const
{
  const double r2 = p.Mag2();
  if (rmin != 0.)
    ci = r2 <= MinusTolSq(rmax) && r2 >= PlusTolSq(rmin);
  else
    ci = r2 <= MinusTolSq(rmax);
  if (rmin != 0.)
    co = r2 >= PlusTolSq(rmax) || r2 <= MinusTolSq(rmin);
  else
    co = r2 >= PlusTolSq(rmax);
  if (!fullPhi) {
    bool cip = false, cop = false;
    wedge.ContainsInside(p, forInside, cip, cop);
    co = co || cop;
    if (forInside) ci = ci && cip;
  }
  if (!fullTheta) {
    const bool cit = forInside ? theta.IsCompletelyInside(p) : false;
    const bool cot = theta.IsCompletelyOutside(p);
    co             = co || cot;
    if (forInside) ci = ci && cit;
  }
}

// SphereUtilities.h:127-157 (the wedge is evaluated also for a full-phi sphere)
bool Sphere::CompletelyOutside(V3 const &p) const
{
  const double rad = p.Mag();
  const bool outR  = (rad > (rmax + kTolerance)) || (rad < (rmin - kTolerance));
  bool inPhi = false, outPhi = false;
  wedge.ContainsInside(p, true, inPhi, outPhi);
  const bool outTheta = theta.IsCompletelyOutside(p);
  return outR || outPhi || outTheta;
}

bool Sphere::CompletelyInside(V3 const &p) const
{
  const double rad = p.Mag();
  const bool inR   = (rad < (rmax - kTolerance)) && (rad > (rmin + kTolerance));
  bool inPhi = false, outPhi = false;
  wedge.ContainsInside(p, true, inPhi, outPhi);
  const bool inTheta = theta.IsCompletelyInside(p);
  return inR && inPhi && inTheta;
}

// SphereUtilities.h:159-195 with IsPointOnRadialSurfaceAndMovingOut (:99-118)
bool Sphere::OnSurfaceMoving(bool out, V3 const &p, V3 const &d) const
{
  auto onRadius = [&](double r) {
    const double m2 = p.Mag2();
    return m2 <= PlusTolSq(r) && m2 >= MinusTolSq(r);
  };
  const V3 mp(-p.x, -p.y, -p.z);
  bool outerR = onRadius(rmax) && (out ? (d.Dot(p) > kSqrtTolerance) : (d.Dot(p) < -kSqrtTolerance));
  bool innerR = false, sPhi = false, ePhi = false, sTh = false, eTh = false;
  if (rmin != 0.) innerR = onRadius(rmin) && (out ? (d.Dot(mp) > kSqrtTolerance) : (d.Dot(mp) < -kSqrtTolerance));
  if (dphi < (kTwoPi - kHalfTolerance)) {
    sPhi = wedge.OnSurfaceMoving(true, out, p, d);
    ePhi = wedge.OnSurfaceMoving(false, out, p, d);
  }
  if (dtheta < (kPi - kHalfTolerance)) {
    sTh = theta.OnSurfaceMoving(true, out, p, d);
    eTh = theta.OnSurfaceMoving(false, out, p, d);
  }
  const double m2 = p.Mag2();
  return ((outerR || innerR) && wedge.Contains(p) && theta.Contains(p)) ||
         ((sPhi || ePhi) && (m2 >= rmin * rmin) && (m2 <= rmax * rmax) && theta.Contains(p)) ||
         ((sTh || eTh) && (m2 >= rmin * rmin) && (m2 <= rmax * rmax) && wedge.Contains(p));
}

// SphereImplementation.h:343-383 (done is always false where the scalar kernels call it)
void Sphere::MinDistFromPhi(bool toIn, V3 const &p, V3 const &d, double &distance) const
{
  double dp1 = kInfLength, dp2 = kInfLength;
  if (toIn)
    wedge.DistanceToIn(p, d, dp1, dp2);
  else
    wedge.DistanceToOut(p, d, dp1, dp2);
  double dist = Min(dp1, dp2);
  V3 t        = p + dist * d;
  double r2   = t.Mag2();
  bool tc     = ((dist == dp1) && wedge.IsOnSurfaceGeneric(wedge.along1, wedge.normal1, t)) ||
            ((dist == dp2) && wedge.IsOnSurfaceGeneric(wedge.along2, wedge.normal2, t));
  const bool c1 = tc && (r2 > rmin * rmin) && (r2 < rmax * rmax) && theta.Contains(t);
  if (c1) distance = Min(dist, distance);
  dist = Max(dp1, dp2);
  t    = p + dist * d;
  r2   = t.Mag2();
  tc   = ((dist == dp1) && wedge.IsOnSurfaceGeneric(wedge.along1, wedge.normal1, t)) ||
       ((dist == dp2) && wedge.IsOnSurfaceGeneric(wedge.along2, wedge.normal2, t));
  const bool c2 = tc && (r2 > rmin * rmin) && (r2 < rmax * rmax) && theta.Contains(t);
  if (!c1 && c2) distance = Min(dist, distance);
}

// SphereImplementation.h:58-76
bool Sphere::Contains(V3 const &p) const
{
  bool ci = false, co = false;
  ContainsInside(p, false, ci, co);
  return !co;
}

int Sphere::Inside(V3 const &p) const
{
  bool ci = false, co = false;
  ContainsInside(p, true, ci, co);
  int r = kSurface;
  if (co) r = kOutside;
  if (ci) r = kInside;
  return r;
}

// SphereImplementation.h:131-239
double Sphere::DistanceToIn(V3 const &p, V3 const &d, double) const
{
  const double rad2 = p.Mag2();
  const double pv   = p.Dot(d);
  double c          = rad2 - rmax * rmax;
  if (CompletelyInside(p)) return -1.0;
  if (OnSurfaceMoving(false, p, d)) return 0.0;
  double sd1 = kInfLength, sd2 = kInfLength;
  double d2  = (pv * pv - c);
  if (d2 < 0. || ((c > 0.) && (pv > 0.))) return kInfLength;
  if (d2 >= 0.) sd1 = (-pv - std::sqrt(std::fabs(d2)));
  double outerDist = kInfLength, innerDist = kInfLength;
  if (full) {
    if (sd1 >= 0.) outerDist = sd1;
  } else {
    const V3 t = p + sd1 * d;
    if (wedge.Contains(t) && theta.Contains(t) && (sd1 >= 0.)) outerDist = sd1;
  }
  if (rmin != 0.) {
    c  = rad2 - rmin * rmin;
    d2 = pv * pv - c;
    if (d2 >= 0.) sd2 = (-pv + std::sqrt(std::fabs(d2)));
    if (full) {
      if (sd2 >= 0.) innerDist = sd2;
    } else {
      const V3 t = p + sd2 * d;
      if ((sd2 >= 0.) && wedge.Contains(t) && theta.Contains(t)) innerDist = sd2;
    }
  }
  double distance = Min(outerDist, innerDist);
  if (!fullPhi) MinDistFromPhi(true, p, d, distance);
  double dThetaMin = kInfLength;
  if (!fullTheta) {
    bool i1 = false, i2 = false;
    double dt1 = kInfLength, dt2 = kInfLength;
    theta.DistanceToIn(p, d, dt1, dt2, i1, i2);
    const V3 q1      = p + dt1 * d;
    const double m1  = q1.Mag2();
    const V3 q2      = p + dt2 * d;
    const double m2  = q2.Mag2();
    bool v1          = (m1 >= rmin * rmin && m1 <= rmax * rmax) && i1;
    bool v2          = (m2 >= rmin * rmin && m2 <= rmax * rmax) && i2;
    if (!fullPhi) {
      v1 = v1 && wedge.Contains(q1);
      v2 = v2 && wedge.Contains(q2);
    }
    if (v2 && !v1) dThetaMin = dt2;
    if (v1 && !v2) dThetaMin = dt1;
    if (v1 && v2) dThetaMin = Min(dt1, dt2);
  }
  distance          = Min(dThetaMin, distance);
  const V3 dd       = V3(0., 0., 0.) - p;
  const double nd   = dd.Mag();
  const double m2   = dd.Mag2();
  const double inv  = std::sqrt(m2 + kMinimum); // Vector3D::Unit, base/Vector3D.h:284-290
  const V3 u(dd.x / inv, dd.y / inv, dd.z / inv);
  if ((stheta > kHalfTolerance || etheta < (kPi - kHalfTolerance)) && (std::fabs(u.x - d.x) < kHalfTolerance) &&
      (std::fabs(u.y - d.y) < kHalfTolerance) && (std::fabs(u.z - d.z) < kHalfTolerance))
    distance = Min(distance, nd);
  return distance;
}

// SphereImplementation.h:244-340
double Sphere::DistanceToOut(V3 const &p, V3 const &d, double) const
{
  const double pv   = p.Dot(d);
  const double rad2 = p.Mag2();
  double c          = rad2 - rmax * rmax;
  double sd1 = kInfLength, sd2 = kInfLength;
  if (CompletelyOutside(p)) return -1.0;
  if (OnSurfaceMoving(true, p, d)) return 0.0;
  double d2 = (pv * pv - c);
  if (d2 >= 0.) sd1 = (-pv + std::sqrt(std::fabs(d2)));
  if (rmin != 0.) {
    c  = rad2 - rmin * rmin;
    d2 = (pv * pv - c);
    if ((d2 >= 0.) && (pv < 0.)) sd2 = (-pv - std::sqrt(std::fabs(d2)));
  }
  const double snxt = Min(sd1, sd2);
  const bool semi   = ((stheta == 0. && etheta == kPi / 2.) && d.z >= 0.) ||
                    ((stheta == kPi / 2. && etheta == kPi) && d.z <= 0.);
  if (semi) return snxt;
  double dThetaMin = kInfLength;
  if (!fullTheta) {
    bool i1 = false, i2 = false;
    double dt1 = kInfLength, dt2 = kInfLength;
    theta.DistanceToOut(p, d, dt1, dt2, i1, i2);
    if (i2 && !i1) dThetaMin = dt2;
    if (!i2 && i1) dThetaMin = dt1;
    if (i2 && i1) dThetaMin = Min(dt1, dt2);
  }
  double distance = Min(dThetaMin, snxt);
  if (!fullPhi) {
    if (dphi <= kPi) {
      double dp1, dp2;
      wedge.DistanceToOut(p, d, dp1, dp2);
      distance = Min(Min(dp1, dp2), distance);
    } else {
      MinDistFromPhi(false, p, d, distance);
    }
  }
  return distance;
}

// SphereImplementation.h:388-431
double Sphere::SafetyToIn(V3 const &p) const
{
  const double rad = p.Mag();
  bool ci = false, co = false;
  ContainsInside(p, true, ci, co);
  if (ci) return -1.0;
  if (!ci && !co) return 0.0;
  double safety;
  if (rmin != 0.) {
    const double sMin = rmin - rad, sMax = rad - rmax;
    safety = (sMin > sMax) ? sMin : sMax;
  } else {
    safety = rad - rmax;
  }
  if (!fullPhi) safety = Max(wedge.SafetyToIn(p), safety);
  if (!fullTheta) safety = Max(theta.SafetyToIn(p), safety);
  return safety;
}

// SphereImplementation.h:436-474
double Sphere::SafetyToOut(V3 const &p) const
{
  const double rad = p.Mag();
  bool ci = false, co = false;
  ContainsInside(p, true, ci, co);
  if (co) return -1.0;
  if (!ci && !co) return 0.0;
  double safety;
  if (rmin != 0.) {
    const double sMin = rad - rmin, sMax = rmax - rad;
    safety = (sMin < sMax) ? sMin : sMax;
  } else {
    safety = rmax - rad;
  }
  if (!fullPhi) safety = Min(wedge.SafetyToOut(p), safety);
  if (!fullTheta) safety = Min(theta.SafetyToOut(p), safety);
  return safety;
}

// source/UnplacedSphere.cpp:76-165
void Sphere::Extent(V3 &lo, V3 &hi) const
{
  lo = V3(-rmax, -rmax, -rmax);
  hi = V3(rmax, rmax, rmax);
  if (full) return;
  double st1 = 0., st2 = 0.;
  if (!fullTheta) {
    hi.z = ((stheta <= kPi / 2.) ? rmax : rmin) * std::cos(stheta);
    lo.z = ((etheta <= kPi / 2.) ? rmin : rmax) * std::cos(etheta);
    st1  = std::sin(stheta);
    st2  = std::sin(etheta);
    if (stheta <= kPi / 2.) {
      if (etheta < kPi / 2.) {
        hi.x = rmax * st2;
        lo.x = -rmax * st2;
        hi.y = rmax * st2;
        lo.y = -rmax * st2;
      }
    } else {
      hi.x = rmax * st1;
      lo.x = -rmax * st1;
      hi.y = rmax * st1;
      lo.y = -rmax * st1;
    }
  }
  if (!fullPhi) {
    double Rmax = rmax;
    if (stheta > kPi / 2.) Rmax *= st1;
    if (etheta < kPi / 2.) Rmax *= st2;
    const double Rmin = rmin * std::min(st1, st2);
    const double Rin  = 0.5 * (Rmax + Rmin);
    const bool o0     = !wedge.Contains(V3(Rin, 0, 0));
    const bool o90    = !wedge.Contains(V3(0, Rin, 0));
    const bool o180   = !wedge.Contains(V3(-Rin, 0, 0));
    const bool o270   = !wedge.Contains(V3(0, -Rin, 0));
    if (!(o0 || o90 || o180 || o270)) return;
    const double Cx = Rmax * std::cos(sphi), Dx = Rmax * std::cos(sphi + dphi);
    const double Cy = Rmax * std::sin(sphi), Dy = Rmax * std::sin(sphi + dphi);
    if (o0) hi.x = std::max(Cx, Dx);
    if (o90) hi.y = std::max(Cy, Dy);
    if (o180) lo.x = std::min(Cx, Dx);
    if (o270) lo.y = std::min(Cy, Dy);
    if (dphi >= kPi) return;
    const double Ax = Rmin * std::cos(sphi), Bx = Rmin * std::cos(sphi + dphi);
    const double Ay = Rmin * std::sin(sphi), By = Rmin * std::sin(sphi + dphi);
    double t = std::max(Ax, Bx);
    hi.x     = t > hi.x ? t : hi.x;
    t        = std::max(Ay, By);
    hi.y     = t > hi.y ? t : hi.y;
    t        = std::min(Ax, Bx);
    lo.x     = t < lo.x ? t : lo.x;
    t        = std::min(Ay, By);
    lo.y     = t < lo.y ? t : lo.y;
  }
}

} // namespace vgo
