// Oracle: Torus2 restated from VecGeom's scalar backend. TEST INFRASTRUCTURE ONLY (see common.h).
//
// Sources: VecGeom/volumes/TorusStruct2.h:22-45 and VecGeom/volumes/UnplacedTorus2.h:64-75 (bounding
// tube), VecGeom/volumes/kernel/TorusImplementation2.h:14-190 (SolveCubic, Sort4, SolveQuartic),
// :192-660 (kernels). Cbrt/Pow/ATan2/cos are the C library's, as in the scalar backend.
#include "shapes.h"

namespace vgo {

// TorusImplementation2.h:14-56
unsigned SolveCubic(double a, double b, double c, double *x)
{
  const double ott     = 1. / 3.;
  const double sq3     = std::sqrt(3.);
  const double inv6sq3 = 1. / (6. * sq3);
  unsigned ireal       = 1;
  double p             = b - a * a * ott;
  double q             = c - a * b * ott + 2. * a * a * a * ott * ott * ott;
  double delta         = 4 * p * p * p + 27. * q * q;
  double t, u;
  if (delta >= 0) {
    delta = std::sqrt(delta);
    t     = (-3 * q * sq3 + delta) * inv6sq3;
    u     = (3 * q * sq3 + delta) * inv6sq3;
    x[0]  = std::copysign(1., t) * std::cbrt(std::fabs(t)) - std::copysign(1., u) * std::cbrt(std::fabs(u)) - a * ott;
  } else {
    delta = std::sqrt(-delta);
    t     = -0.5 * q;
    u     = delta * inv6sq3;
    x[0]  = 2. * std::pow(t * t + u * u, 0.5 * ott) * std::cos(ott * std::atan2(u, t));
    x[0] -= a * ott;
  }
  t     = x[0] * x[0] + a * x[0] + b;
  u     = a + x[0];
  delta = u * u - 4. * t;
  if (delta >= 0) {
    ireal = 3;
    delta = std::sqrt(delta);
    x[1]  = 0.5 * (-u - delta);
    x[2]  = 0.5 * (-u + delta);
  }
  return ireal;
}

namespace {
// TorusImplementation2.h:58-80
inline void CmpAndSwap(double *a, int i, int j)
{
  if (a[i] > a[j]) {
    const double c = a[j];
    a[j]           = a[i];
    a[i]           = c;
  }
}
inline void Sort4(double *a)
{
  CmpAndSwap(a, 0, 2);
  CmpAndSwap(a, 1, 3);
  CmpAndSwap(a, 0, 1);
  CmpAndSwap(a, 2, 3);
  CmpAndSwap(a, 1, 2);
}
} // namespace

// TorusImplementation2.h:82-171
int SolveQuartic(double a, double b, double c, double d, double *x)
{
  double e     = b - 3. * a * a / 8.;
  double f     = c + a * a * a / 8. - 0.5 * a * b;
  double g     = d - 3. * a * a * a * a / 256. + a * a * b / 16. - a * c / 4.;
  double xx[4] = {kInfLength, kInfLength, kInfLength, kInfLength};
  double delta;
  double h       = 0.;
  unsigned ireal = 0;
  if (std::fabs(f) < 1E-6) {
    delta = e * e - 4. * g;
    if (delta < 0.) return 0;
    delta = std::sqrt(delta);
    h     = 0.5 * (-e - delta);
    if (h >= 0) {
      h          = std::sqrt(h);
      x[ireal++] = -h - 0.25 * a;
      x[ireal++] = h - 0.25 * a;
    }
    h = 0.5 * (-e + delta);
    if (h >= 0) {
      h          = std::sqrt(h);
      x[ireal++] = -h - 0.25 * a;
      x[ireal++] = h - 0.25 * a;
    }
    Sort4(x);
    return ireal;
  }
  if (std::fabs(g) < 1E-6) {
    x[ireal++]     = -0.25 * a;
    unsigned ncube = SolveCubic(0, e, f, xx);
    for (unsigned i = 0; i < ncube; i++) x[ireal++] = xx[i] - 0.25 * a;
    Sort4(x);
    return ireal;
  }
  ireal = SolveCubic(2. * e, e * e - 4. * g, -f * f, xx);
  if (ireal == 1) {
    if (xx[0] <= 0) return 0;
    h = std::sqrt(xx[0]);
  } else {
    for (unsigned i = 0; i < 3; i++) {
      h = xx[i];
      if (h >= 0) break;
    }
    if (h <= 0) return 0;
    h = std::sqrt(h);
  }
  double j = 0.5 * (e + h * h - f / h);
  ireal    = 0;
  delta    = h * h - 4. * j;
  if (delta >= 0) {
    delta      = std::sqrt(delta);
    x[ireal++] = 0.5 * (-h - delta) - 0.25 * a;
    x[ireal++] = 0.5 * (-h + delta) - 0.25 * a;
  }
  delta = h * h - 4. * g / j;
  if (delta >= 0) {
    delta      = std::sqrt(delta);
    x[ireal++] = 0.5 * (h - delta) - 0.25 * a;
    x[ireal++] = 0.5 * (h + delta) - 0.25 * a;
  }
  Sort4(x);
  return ireal;
}

// TorusStruct2.h:36-43, UnplacedTorus2.h:64-75
Torus2::Torus2(double rmin_, double rmax_, double rtor_, double sphi_, double dphi_)
    : rmin(rmin_), rmax(rmax_), rtor(rtor_), sphi(sphi_), dphi(dphi_), rmin2(rmin_ * rmin_), rmax2(rmax_ * rmax_),
      rtor2(rtor_ * rtor_)
{
  wedge.Set(dphi_, sphi_);
  bounding = Tube(rtor - rmax - kTolerance, rtor + rmax + kTolerance, rmax, sphi_, dphi_);
}

// TorusImplementation2.h:192-198
double Torus2::DistSqrToTorusR(V3 const &point, V3 const &dir, double dist) const
{
  const V3 p       = point + dir * dist;
  const double rxy = p.Perp();
  return (rxy - rtor) * (rxy - rtor) + p.z * p.z;
}

// TorusImplementation2.h:317-365 (early returns of the scalar backend after the rmax and rmin tests)
void Torus2::ContainsInside(V3 const &p, bool forInside, bool notForDisk, bool &ci, bool &co) const
{
  const double tol   = 100. * kTolerance;
  const double rxy   = std::sqrt(p[0] * p[0] + p[1] * p[1]);
  const double radsq = (rxy - rtor) * (rxy - rtor) + p[2] * p[2];
  if (forInside) {
    co = radsq > (tol * rmax + rmax2);
    ci = radsq < (-tol * rmax + rmax2);
  } else {
    co = radsq > rmax2;
  }
  if (co) return;
  if (forInside) {
    co = co || radsq < (-tol * rmin + rmin2);
    ci = ci && radsq > (tol * rmin + rmin2);
  } else {
    co = co || radsq < rmin2;
  }
  if (co) return;
  if ((dphi < kTwoPi) && notForDisk) {
    bool cip = false, cop = false;
    wedge.ContainsInside(p, forInside, cip, cop);
    co = co || cop;
    if (forInside) ci = ci && cip;
  }
}

// TorusImplementation2.h:520-528 (GenericKernel<ForInside = true, notForDisk = false>)
bool Torus2::Contains(V3 const &p) const
{
  bool ci = false, co = false;
  ContainsInside(p, true, false, ci, co);
  return !co;
}

// TorusImplementation2.h:302-314
int Torus2::Inside(V3 const &p) const
{
  bool ci = false, co = false;
  ContainsInside(p, true, true, ci, co);
  int r = kSurface;
  if (co) r = kOutside;
  if (ci) r = kInside;
  return r;
}

// TorusImplementation2.h:367-493
double Torus2::ToBoundary(V3 const &pt, V3 const &dir, double radius, bool out) const
{
  double s          = kInfLength;
  const double tol  = 100. * kTolerance;
  const double r0sq = pt[0] * pt[0] + pt[1] * pt[1] + pt[2] * pt[2];
  const double rdn  = pt[0] * dir[0] + pt[1] * dir[1] + pt[2] * dir[2];
  const double rsum = rtor2 + radius * radius;
  const double a    = 4. * rdn;
  const double b    = 2. * (r0sq + 2. * rdn * rdn - rsum + 2. * rtor2 * dir[2] * dir[2]);
  const double c    = 4. * (r0sq * rdn - rsum * rdn + 2. * rtor2 * pt[2] * dir[2]);
  const double d    = r0sq * r0sq - 2. * r0sq * rsum + 4. * rtor2 * pt[2] * pt[2] +
                   (rtor2 - radius * radius) * (rtor2 - radius * radius);
  double x[4] = {kInfLength, kInfLength, kInfLength, kInfLength};
  int nsol    = 0;
  if (std::fabs(dir[2]) < 1E-3 && std::fabs(pt[2]) < 0.1 * radius) {
    double r0              = rtor - std::sqrt((radius - pt[2]) * (radius + pt[2]));
    const double invdirxy2 = 1. / (1 - dir.z * dir.z);
    const double b0        = (pt[0] * dir[0] + pt[1] * dir[1]) * invdirxy2;
    double c0              = (pt[0] * pt[0] + (pt[1] - r0) * (pt[1] + r0)) * invdirxy2;
    double delta           = b0 * b0 - c0;
    if (delta > 0) {
      x[nsol] = -b0 - std::sqrt(delta);
      if (x[nsol] > -tol) nsol++;
      x[nsol] = -b0 + std::sqrt(delta);
      if (x[nsol] > -tol) nsol++;
    }
    r0    = rtor + std::sqrt((radius - pt[2]) * (radius + pt[2]));
    c0    = (pt[0] * pt[0] + (pt[1] - r0) * (pt[1] + r0)) * invdirxy2;
    delta = b0 * b0 - c0;
    if (delta > 0) {
      x[nsol] = -b0 - std::sqrt(delta);
      if (x[nsol] > -tol) nsol++;
      x[nsol] = -b0 + std::sqrt(delta);
      if (x[nsol] > -tol) nsol++;
    }
    if (nsol) Sort4(x);
  } else {
    nsol = SolveQuartic(a, b, c, d, x);
  }
  if (!nsol) return kInfLength;
  const bool inner = std::fabs(radius - rmin) < kTolerance;
  for (int i = 0; i < nsol; i++) {
    if (x[i] < -10) continue;
    V3 r0   = pt + x[i] * dir;
    V3 norm = r0;
    r0.z    = 0.;
    r0      = r0 * (1. / r0.Mag()); // Vector3D::Normalize
    r0      = r0 * rtor;
    norm    = norm - r0;
    const double ndotd = norm.Dot(dir);
    if (inner ^ out) {
      if (ndotd < 0) continue;
    } else {
      if (ndotd > 0) continue;
    }
    if (dphi < kTwoPi) {
      bool ci = false, co = false;
      wedge.ContainsInside(r0, true, ci, co); // ContainsWithBoundary = !completelyoutside
      if (co) continue;
    }
    s             = x[i];
    double eps    = kInfLength;
    double delta  = s * s * s * s + a * s * s * s + b * s * s + c * s + d;
    double eps0   = -delta / (4. * s * s * s + 3. * a * s * s + 2. * b * s + c);
    int ntry      = 0;
    while (std::fabs(eps) > kTolerance) {
      if (std::fabs(eps0) > 100) break;
      s += eps0;
      if (std::fabs(s + eps0) < kTolerance) break;
      delta = s * s * s * s + a * s * s * s + b * s * s + c * s + d;
      eps   = -delta / (4. * s * s * s + 3. * a * s * s + 2. * b * s + c);
      if (std::fabs(eps) >= std::fabs(eps0)) break;
      ntry++;
      if (ntry > 100) break;
      eps0 = eps;
    }
    if (s < -tol) continue;
    return std::max(0., s);
  }
  return kInfLength;
}

// TorusImplementation2.h:199-299
double Torus2::DistanceToOut(V3 const &p, V3 const &d, double) const
{
  const bool hasphi  = dphi < kTwoPi;
  const bool hasrmin = rmin > 0;
  bool done          = false;
  const double distz = std::fabs(p.z) - rmax;
  done               = done || distz > kHalfTolerance;
  const double rsq   = p.x * p.x + p.y * p.y;
  const double outerExcl = rtor + rmax + kHalfTolerance;
  done                   = done || rsq > outerExcl * outerExcl;
  const double innerExcl = rtor - rmax - kHalfTolerance;
  done                   = done || rsq < innerExcl * innerExcl;
  if (done) return -1.;
  if (Inside(p) == kOutside) return -1.;
  const double dout = ToBoundary(p, d, rmax, true);
  double din        = kInfLength;
  if (hasrmin) din = ToBoundary(p, d, rmin, true);
  double distance = std::min(dout, din);
  if (hasphi) {
    double dp1, dp2;
    wedge.DistanceToOut(p, d, dp1, dp2);
    bool smaller = dp1 < distance;
    if (smaller) {
      const V3 ip = p + d * dp1;
      bool ci = false, co = false;
      ContainsInside(ip, false, false, ci, co); // UnplacedContainsDisk
      if (!co) {
        const double diri = ip.x * wedge.along1.x + ip.y * wedge.along1.y;
        if (diri >= 0 && smaller) distance = dp1;
      }
    }
    smaller = dp2 < distance;
    if (smaller) {
      const V3 ip = p + d * dp2;
      bool ci = false, co = false;
      ContainsInside(ip, false, false, ci, co);
      if (!co) {
        const double diri2 = ip.x * wedge.along2.x + ip.y * wedge.along2.y;
        if (diri2 >= 0 && (dp2 < distance) && smaller) distance = dp2;
      }
    }
  }
  if (distance >= kInfLength) distance = -1.;
  return distance;
}

// TorusImplementation2.h:538-628 (NO_SPECIALIZATION: UniversalTube bounding-tube kernels)
double Torus2::DistanceToIn(V3 const &point, V3 const &direction, double stepMax) const
{
  V3 lp           = point;
  const V3 ld     = direction;
  double distance = kInfLength;
  double tubeDist = kInfLength;
  const bool inBounds = bounding.Contains(lp);
  if (!inBounds)
    tubeDist = bounding.DistanceToIn(lp, ld, stepMax);
  else
    tubeDist = 0.;
  if (inBounds) {
    if (Inside(point) == kInside) return -1.;
  } else {
    if (tubeDist == kInfLength) return kInfLength;
  }
  lp = lp + tubeDist * ld;
  if (dphi < kTwoPi) {
    double d1, d2;
    wedge.DistanceToIn(lp, ld, d1, d2);
    if (d1 != kInfLength) {
      const double daxis = DistSqrToTorusR(lp, ld, d1);
      if (daxis >= rmin2 && daxis < rmax2) {
        distance = d1;
        if (std::fabs(distance) < kTolerance) return distance + tubeDist;
      }
    }
    if (d2 != kInfLength) {
      const double daxis = DistSqrToTorusR(lp, ld, d2);
      if (daxis >= rmin2 && daxis < rmax2) {
        distance = std::min(d2, distance);
        if (std::fabs(distance) < kTolerance) return distance + tubeDist;
      }
    }
    distance = kInfLength;
  }
  double dd = ToBoundary(lp, ld, rmax, false);
  if (rmin > 0.) dd = std::min(dd, ToBoundary(lp, ld, rmin, false));
  distance = std::min(distance, dd);
  distance += tubeDist;
  if (std::fabs(distance) > kInfLength) distance = kInfLength;
  return distance;
}

// TorusImplementation2.h:630-650
double Torus2::SafetyToIn(V3 const &p) const
{
  const double rxy = std::sqrt(p[0] * p[0] + p[1] * p[1]);
  const double rad = std::sqrt((rxy - rtor) * (rxy - rtor) + p[2] * p[2]);
  double safety    = rad - rmax;
  if (rmin != 0.) safety = std::max(rmin - rad, rad - rmax);
  if (dphi < kTwoPi && rxy != 0.) safety = std::max(wedge.SafetyToIn(p), safety);
  return safety;
}

// TorusImplementation2.h:495-513
double Torus2::SafetyToOut(V3 const &p) const
{
  const double rxy = std::sqrt(p[0] * p[0] + p[1] * p[1]);
  const double rad = std::sqrt((rxy - rtor) * (rxy - rtor) + p[2] * p[2]);
  double safety    = rmax - rad;
  if (rmin != 0.) safety = std::min(rad - rmin, rmax - rad);
  if (dphi < kTwoPi) safety = std::min(wedge.SafetyToOut(p), safety);
  return safety;
}

} // namespace vgo
