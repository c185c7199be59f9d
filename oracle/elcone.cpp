// Oracle: EllipticalCone (see shapes.h for the reference files followed).
// TEST INFRASTRUCTURE ONLY (see common.h).
#include "shapes.h"

namespace vgo {
namespace {
inline double Mx(double a, double b) { return a > b ? a : b; }
inline double Mn(double a, double b) { return a < b ? a : b; }
V3 Unit(V3 const &v) // Vector3D::Unit (Vector3D.h:284-290)
{
  const double m = std::sqrt(v.Mag2() + DBL_MIN);
  return V3(v.x / m, v.y / m, v.z / m);
}
} // namespace

// UnplacedEllipticalCone::CheckParameters (source/UnplacedEllipticalCone.cpp:28-63): values below
// 2 kTolerance raised to it, zcut at most h
EllipticalCone::EllipticalCone(double a, double b, double h, double zcut) : dx(a), dy(b), dz(h), zCut(zcut)
{
  const double tol = 2. * kTolerance;
  if (dx < tol || dy < tol || dz < tol || zCut < tol) {
    if (dx < tol) dx = tol;
    if (dy < tol) dy = tol;
    if (dz < tol) dz = tol;
    if (zCut < tol) zCut = tol;
  }
  zCut              = Mn(dz, zCut);
  const double h2   = dz + zCut;
  const double xmax = dx * h2, ymax = dy * h2, m = Mn(dx, dy);
  rsph       = std::sqrt(xmax * xmax + ymax * ymax + zCut * zCut);
  invDx      = 1. / dx;
  invDy      = 1. / dy;
  cosAxisMin = m / std::sqrt(1. + m * m);
  dApex      = kHalfTolerance / cosAxisMin;
}

// GenericKernelForContainsAndInside (:90-115)
double EllipticalCone::Safety(V3 const &p) const
{
  const double px = p.x * invDx, py = p.y * invDy, pz = p.z;
  const double hp = std::sqrt(px * px + py * py) + pz;
  return Mx((hp - dz) * cosAxisMin, std::fabs(pz) - zCut);
}

bool EllipticalCone::Contains(V3 const &p) const { return !(Safety(p) > kHalfTolerance); }

int EllipticalCone::Inside(V3 const &p) const
{
  const double s = Safety(p);
  return s <= -kHalfTolerance ? kInside : (s > kHalfTolerance ? kOutside : kSurface);
}

// :118-205 (far points first moved to 2 rsph from the origin)
double EllipticalCone::DistanceToIn(V3 const &point, V3 const &d, double) const
{
  const double kTwoEpsilon = 2. * kEpsilon;
  double distance          = kInfLength, offset = 0.;
  V3 p                     = point;
  if (p.Mag2() > 1024. * rsph * rsph && d.Dot(p) < 0.) offset = p.Mag() - 2. * rsph;
  p               = p + d * offset;
  const double px = p.x * invDx, py = p.y * invDy, pz = p.z, pz0 = p.z - dz;
  const double vx = d.x * invDx, vy = d.y * invDy;
  double vz       = d.z;
  const double Ar = vx * vx + vy * vy, Br = px * vx + py * vy, Cr = px * px + py * py;
  const double vzvz = vz * vz;
  const bool tinyA  = std::fabs(Ar - vzvz) < kTwoEpsilon * vzvz;
  if (tinyA) vz = vz + std::fabs(vz) * kTwoEpsilon;
  const double Az = vz * vz, Bz = pz0 * vz, Cz = pz0 * pz0;
  const double A = Ar - Az, B = Br - Bz, B0 = Br - pz0 * d.z, C = Cr - Cz;
  double D         = B * B - A * C;
  const double sfz = std::fabs(pz) - zCut;
  double nz        = std::sqrt(Cr);
  const double sfr = (nz + pz0) * cosAxisMin;
  if (std::fabs(p.x) + std::fabs(p.y) < 0.1 * kHalfTolerance) nz = 1.;
  const double pzA = pz0 + dApex;
  bool done = (sfz >= -kHalfTolerance && pz * vz >= 0.) || (sfr >= -kHalfTolerance && Br + nz * vz >= 0.) ||
              (pz0 * cosAxisMin > -kHalfTolerance && (Cr - pzA * pzA) <= 0. && A >= 0.);
  if (sfr <= 0. && D < 0.) D = 0.;
  done = done || (D < 0.) || ((D < kTwoEpsilon * B * B) && (A >= 0.));
  const double invz = -1. / NonZero(vz);
  const double ddz  = std::copysign(zCut, invz);
  const double tzin = (pz - ddz) * invz, tzout = (pz + ddz) * invz;
  double tmp = 0., t1 = 0., t2 = 0.;
  if (!done) tmp = -B - std::copysign(std::sqrt(D), B);
  if (!done) t1 = tmp / A;
  if (!done && tmp != 0) t2 = C / tmp;
  if (!done && tinyA && B != 0.) t2 = -C / (2. * B0);
  const double tmin = Mn(t1, t2), tmax = Mx(t1, t2);
  double trin = tmin, trout = tmax;
  done = done || (A >= 0. && pz0 + vz * tmin >= 0.);
  if (!done && A < 0.) trin = -kInfLength, trout = kInfLength;
  if (!done && A < 0. && vz < 0.) trin = tmax;
  if (!done && A < 0. && vz > 0.) trout = tmin;
  const double tin = Mx(tzin, trin), tout = Mn(tzout, trout);
  if (!done && (tout - tin) >= kHalfTolerance) distance = tin + offset;
  return distance;
}

// :208-270
double EllipticalCone::DistanceToOut(V3 const &p, V3 const &d, double) const
{
  const double kTwoEpsilon = 2. * kEpsilon;
  double distance          = 0.;
  const double px = p.x * invDx, py = p.y * invDy, pz = p.z, pz0 = pz - dz;
  const double hp      = std::sqrt(px * px + py * py) + pz;
  const double sfr     = (hp - dz) * cosAxisMin;
  const double sfz     = std::fabs(pz) - zCut;
  const bool outside   = Mx(sfr, sfz) > kHalfTolerance;
  if (outside) distance = -1.;
  bool done = outside;
  const double vx = d.x * invDx, vy = d.y * invDy;
  double vz       = d.z;
  const double Ar = vx * vx + vy * vy, Br = px * vx + py * vy, Cr = px * px + py * py;
  const bool tinyA = std::fabs(Ar - vz * vz) < kTwoEpsilon * vz * vz;
  if (tinyA) vz = vz + std::fabs(vz) * kTwoEpsilon;
  const double Az = vz * vz, Bz = pz0 * vz, Cz = pz0 * pz0;
  const double A = Ar - Az, B = Br - Bz, B0 = Br - pz0 * d.z, C = Cr - Cz;
  double D = B * B - A * C;
  if (sfr <= 0. && D < 0.) D = 0.;
  done         = done || (D < 0.) || (D < kTwoEpsilon * B * B && A >= 0.);
  double tzout = DBL_MAX; // kMaximum
  if (vz != 0.) tzout = (std::copysign(zCut, vz) - pz) / d.z;
  double tmp = 0., t1 = 0., t2 = 0.;
  if (!done) tmp = -B - std::copysign(std::sqrt(D), B);
  if (!done) t1 = tmp / A;
  if (!done && tmp != 0.) t2 = C / tmp;
  if (!done && tinyA && B0 != 0.) t2 = -C / (2. * B0);
  const double tmin = Mn(t1, t2), tmax = Mx(t1, t2);
  double trout = tmax;
  done = done || ((A >= 0. && pz0 + vz * tmax >= 0.) || (pz0 >= 0. && vz >= 0.));
  if (!done && A < 0.) trout = kInfLength;
  if (!done && A < 0. && vz > 0.) trout = tmin;
  if (!done) distance = Mn(tzout, trout);
  return distance;
}

// :273-287
double EllipticalCone::SafetyToIn(V3 const &p) const
{
  double s = Safety(p);
  if (std::fabs(s) <= kHalfTolerance) s = 0.;
  return s;
}

// :290-304
double EllipticalCone::SafetyToOut(V3 const &p) const
{
  const double px = p.x * invDx, py = p.y * invDy, pz = p.z;
  const double hp = std::sqrt(px * px + py * py) + pz;
  double s        = Mn((dz - hp) * cosAxisMin, zCut - std::fabs(pz));
  if (std::fabs(s) <= kHalfTolerance) s = 0.;
  return s;
}

// NormalKernel (:307-333)
bool EllipticalCone::Normal(V3 const &p, V3 &n) const
{
  V3 normal(0., 0., 0.);
  bool valid     = true;
  const double z = std::fabs(p.z) - zCut;
  if (std::fabs(z) <= kHalfTolerance) normal.z = Sign(p.z);
  const double nx = p.x * invDx * invDx, ny = p.y * invDy * invDy;
  double nz       = std::sqrt(p.x * nx + p.y * ny);
  if ((nx * nx + ny * ny) == 0.) nz = 1.;
  const V3 nside(nx, ny, nz);
  const double ds = (nz + p.z - dz) * cosAxisMin;
  if (std::fabs(ds) <= kHalfTolerance) normal = Unit(normal + Unit(nside));
  if (!(normal.Mag2() > 0.)) {
    valid    = false;
    normal.z = Sign(p.z);
    if (ds > z) normal = Unit(nside);
  }
  n = normal;
  return valid;
}

// source/UnplacedEllipticalCone.cpp:67-75
void EllipticalCone::Extent(V3 &lo, V3 &hi) const
{
  const double xmax = dx * (dz + zCut), ymax = dy * (dz + zCut);
  lo = V3(-xmax, -ymax, -zCut);
  hi = V3(xmax, ymax, zCut);
}

} // namespace vgo
