// Oracle: EllipticalTube and Ellipsoid (see shapes.h for the reference files followed).
// TEST INFRASTRUCTURE ONLY (see common.h).
#include "shapes.h"

namespace vgo {
namespace {
inline double Mx(double a, double b) { return a > b ? a : b; }
inline double Mn(double a, double b) { return a < b ? a : b; }
inline V3 Unit(V3 const &v)
{
  const double m = std::sqrt(v.Mag2() + kMinimum); // Vector3D::Unit (Vector3D.h:284-290)
  return V3(v.x / m, v.y / m, v.z / m);
}
constexpr double kMaximum = DBL_MAX; // base/Math.h:42
} // namespace

// ---------------------------------------------------------------- EllipticalTube
// UnplacedEllipticalTube(dx, dy, dz) -> SetParameters -> CheckParameters (source/UnplacedEllipticalTube.cpp:
// 20-62): semi-axes below 2 kTolerance (the reference tests dy twice and not dz) reset all three
EllipticalTube::EllipticalTube(double X, double Y, double Z) : dx(X), dy(Y), dz(Z)
{
  const double tol = 2. * kTolerance;
  if (dx < tol || dy < tol || dy < tol) dx = dy = dz = tol;
  const double R = Mn(dx, dy);
  rsph           = std::sqrt(dx * dx + dy * dy + dz * dz);
  ddx            = dx * dx;
  ddy            = dy * dy;
  sx             = R / dx;
  sy             = R / dy;
  r              = R;
  q1             = 0.5 / R;
  q2             = 0.5 * (R + kHalfTolerance * kHalfTolerance / R);
  scratch        = 2. * R * R * kEpsilon;
}

// GenericKernelForContainsAndInside (EllipticalTubeImplementation.h:85-107)
double EllipticalTube::Safety(V3 const &p) const
{
  const double x = p.x * sx, y = p.y * sy;
  const double distR = q1 * (x * x + y * y) - q2;
  const double distZ = std::fabs(p.z) - dz;
  return Mx(distR, distZ);
}

bool EllipticalTube::Contains(V3 const &p) const { return !(Safety(p) > kHalfTolerance); }

int EllipticalTube::Inside(V3 const &p) const
{
  const double s = Safety(p);
  int in         = kSurface;
  if (s > kHalfTolerance) in = kOutside;
  if (s <= -kHalfTolerance) in = kInside;
  return in;
}

// :109-160
double EllipticalTube::DistanceToIn(V3 const &point, V3 const &direction, double) const
{
  double distance = kInfLength, offset = 0.;
  V3 pcur = point;
  const double Rfar2 = 1024. * rsph * rsph;
  if (pcur.Mag2() > Rfar2 && direction.Dot(point) < 0.) pcur = pcur + (offset = pcur.Mag() - 2. * rsph) * direction;
  const double px = pcur.x * sx, py = pcur.y * sy, pz = pcur.z;
  const double vx = direction.x * sx, vy = direction.y * sy, vz = direction.z;
  const double invz  = -1. / NonZero(vz);
  const double ddz   = std::copysign(dz, invz);
  const double tzmin = (pz - ddz) * invz;
  const double tzmax = (pz + ddz) * invz;
  const double rr = px * px + py * py;
  const double A  = vx * vx + vy * vy;
  const double B  = px * vx + py * vy;
  const double C  = rr - r * r;
  const double D  = B * B - A * C;
  const double distZ     = std::fabs(pz) - dz;
  const double distR     = q1 * rr - q2;
  const bool parallelToZ = (A < kEpsilon || std::fabs(vz) >= 1.);
  const bool leaving     = (distZ >= -kHalfTolerance && pz * vz >= 0.) || (distR >= -kHalfTolerance && (B >= 0. || parallelToZ));
  if (!leaving && parallelToZ) distance = tzmin + offset;
  const bool done = leaving || parallelToZ || D <= A * A * scratch;
  double tmp = 0., t1 = 0., t2 = 0.;
  if (!done) tmp = -B - std::copysign(std::sqrt(D), B);
  if (!done) t1 = tmp / A;
  if (!done) t2 = C / tmp;
  const double trmin = Mn(t1, t2), trmax = Mx(t1, t2);
  const double tin = Mx(tzmin, trmin), tout = Mn(tzmax, trmax);
  if (!done && (tout - tin) >= kHalfTolerance) distance = tin + offset;
  return distance;
}

// :162-203
double EllipticalTube::DistanceToOut(V3 const &point, V3 const &direction, double) const
{
  const double px = point.x * sx, py = point.y * sy, pz = point.z;
  const double vx = direction.x * sx, vy = direction.y * sy, vz = direction.z;
  const double rr      = px * px + py * py;
  const double distR   = q1 * rr - q2;
  const double distZ   = std::fabs(pz) - dz;
  const bool outside   = Mx(distR, distZ) > kHalfTolerance;
  double distance      = 0.;
  if (outside) distance = -1.;
  double tzmax = kMaximum;
  if (vz != 0.) tzmax = (std::copysign(dz, vz) - pz) / vz;
  const double A = vx * vx + vy * vy;
  const double B = px * vx + py * vy;
  const double C = rr - r * r;
  const double D = B * B - A * C;
  const bool parallelToZ = (A < kEpsilon || std::fabs(vz) >= 1.);
  if (!outside && parallelToZ) distance = tzmax;
  const bool done = outside || parallelToZ || D <= 0.;
  if (!done && B >= 0.) distance = Mn(tzmax, -C / (std::sqrt(D) + B));
  if (!done && B < 0.) distance = Mn(tzmax, (std::sqrt(D) - B) / A);
  return distance;
}

// :205-231
double EllipticalTube::SafetyToIn(V3 const &p) const
{
  const double x = p.x * sx, y = p.y * sy;
  const double distR = std::sqrt(x * x + y * y) - r;
  const double distZ = std::fabs(p.z) - dz;
  double safety      = Mx(distR, distZ);
  if (std::fabs(safety) <= kHalfTolerance) safety = 0.;
  return safety;
}

double EllipticalTube::SafetyToOut(V3 const &p) const
{
  const double x = p.x * sx, y = p.y * sy;
  const double distR = r - std::sqrt(x * x + y * y);
  const double distZ = dz - std::fabs(p.z);
  double safety      = Mn(distR, distZ);
  if (std::fabs(safety) <= kHalfTolerance) safety = 0.;
  return safety;
}

// NormalKernel (:233-268)
bool EllipticalTube::Normal(V3 const &point, V3 &n) const
{
  V3 normal(0., 0., 0.);
  const double x = point.x * sx, y = point.y * sy;
  double distR = q1 * (x * x + y * y) - q2;
  if (std::fabs(distR) <= kHalfTolerance) normal = Unit(V3(point.x * ddy, point.y * ddx, 0.));
  const double distZ = std::fabs(point.z) - dz;
  if (std::fabs(distZ) <= kHalfTolerance) normal.z = Sign(point.z);
  if (normal.Mag2() > 1.) normal = Unit(normal);
  if (normal.Mag2() > 0.) {
    n = normal;
    return true;
  }
  normal.z = Sign(point.z);
  distR    = std::sqrt(x * x + y * y) - r;
  if (distR > distZ && (x * x + y * y) > 0.) normal = Unit(V3(point.x * ddy, point.y * ddx, 0.));
  n = normal;
  return false;
}

// ---------------------------------------------------------------- Ellipsoid
// UnplacedEllipsoid(dx, dy, dz, zBottomCut, zTopCut) -> CheckParameters (source/UnplacedEllipsoid.cpp:27-116)
Ellipsoid::Ellipsoid(double A_, double B_, double C_, double zb, double zt)
    : dx(A_), dy(B_), dz(C_), zBottomCut(zb), zTopCut(zt)
{
  const double tol = 2. * kTolerance;
  if (dx < tol || dy < tol || dy < tol) dx = dy = dz = tol;
  const double A = dx, B = dy, C = dz;
  if (zBottomCut == 0 && zTopCut == 0) zBottomCut = -C, zTopCut = C;
  if (zBottomCut >= C || zTopCut <= -C || zBottomCut >= zTopCut) zBottomCut = -C, zTopCut = C;
  zBottomCut = Mx(zBottomCut, -dz);
  zTopCut    = Mn(zTopCut, dz);
  xmax = A, ymax = B;
  if (zBottomCut > 0.) {
    const double ratio = zBottomCut / C, scale = std::sqrt((1. - ratio) * (1 + ratio));
    xmax = A * scale, ymax = B * scale;
  }
  if (zTopCut < 0.) {
    const double ratio = zTopCut / C, scale = std::sqrt((1. - ratio) * (1 + ratio));
    xmax = A * scale, ymax = B * scale;
  }
  rsph = Mx(Mx(A, B), C);
  r    = Mn(Mn(A, B), C);
  sx   = r / A;
  sy   = r / B;
  sz   = r / C;
  const double scZBottomCut = zBottomCut * sz, scZTopCut = zTopCut * sz;
  scZMidCut = 0.5 * (scZTopCut + scZBottomCut);
  scZDimCut = 0.5 * (scZTopCut - scZBottomCut);
  q1        = 0.5 / r;
  q2        = 0.5 * (r + kHalfTolerance * kHalfTolerance / r);
}

// GenericKernelForContainsAndInside (EllipsoidImplementation.h:85-108)
double Ellipsoid::Safety(V3 const &p) const
{
  const double x = p.x * sx, y = p.y * sy, z = p.z * sz;
  const double distZ = std::fabs(z - scZMidCut) - scZDimCut;
  const double distR = q1 * (x * x + y * y + z * z) - q2;
  return Mx(distZ, distR);
}

bool Ellipsoid::Contains(V3 const &p) const { return !(Safety(p) > kHalfTolerance); }

int Ellipsoid::Inside(V3 const &p) const
{
  const double s = Safety(p);
  int in         = kSurface;
  if (s > kHalfTolerance) in = kOutside;
  if (s <= -kHalfTolerance) in = kInside;
  return in;
}

// :110-161
double Ellipsoid::DistanceToIn(V3 const &point, V3 const &direction, double) const
{
  double distance = kInfLength, offset = 0.;
  V3 pcur = point;
  const double Rfar2 = 1024. * rsph * rsph;
  if (pcur.Mag2() > Rfar2 && direction.Dot(point) < 0.) pcur = pcur + (offset = pcur.Mag() - 2. * rsph) * direction;
  const double px = pcur.x * sx, py = pcur.y * sy, pz = pcur.z * sz;
  const double vx = direction.x * sx, vy = direction.y * sy, vz = direction.z * sz;
  const double pzcut = pz - scZMidCut, dzcut = scZDimCut;
  const double distZ = std::fabs(pzcut) - dzcut;
  const double rr = px * px + py * py + pz * pz;
  const double vv = vx * vx + vy * vy + vz * vz;
  const double pv = px * vx + py * vy + pz * vz;
  const double distR = q1 * rr - q2;
  const bool leaving = (distZ >= -kHalfTolerance && pzcut * vz >= 0.) || (distR >= -kHalfTolerance && pv >= 0.);
  const double invz  = -1. / NonZero(vz);
  const double ddz   = std::copysign(dzcut, invz);
  const double tzmin = (pzcut - ddz) * invz;
  const double tzmax = (pzcut + ddz) * invz;
  const double A = vv, B = pv, C = (rr - r * r);
  const double D     = B * B - A * C;
  const double sqrtD = std::sqrt(std::fabs(D));
  const double trmin = (-B - sqrtD) / A;
  const double trmax = (-B + sqrtD) / A;
  const double tmin = Mx(tzmin, trmin), tmax = Mn(tzmax, trmax);
  const double EPS = 2. * rr * vv * kEpsilon;
  const bool done  = leaving || (D <= EPS) || ((tmax - tmin) <= kHalfTolerance);
  if (!done) distance = tmin + offset;
  return distance;
}

// :163-201
double Ellipsoid::DistanceToOut(V3 const &point, V3 const &direction, double) const
{
  const double px = point.x * sx, py = point.y * sy, pz = point.z * sz;
  const double vx = direction.x * sx, vy = direction.y * sy, vz = direction.z * sz;
  const double pzcut = pz - scZMidCut, dzcut = scZDimCut;
  const double distZ = std::fabs(pzcut) - dzcut;
  const double rr = px * px + py * py + pz * pz;
  const double vv = vx * vx + vy * vy + vz * vz;
  const double pv = px * vx + py * vy + pz * vz;
  const double distR = q1 * rr - q2;
  const bool outside = Mx(distR, distZ) > kHalfTolerance;
  double distance    = 0.;
  if (outside) distance = -1.;
  double tzmax = kMaximum;
  if (vz != 0.) tzmax = (std::copysign(dzcut, vz) - pzcut) / vz;
  const double B = pv / vv, C = (rr - r * r) / vv;
  const double D     = B * B - C;
  const double sqrtD = std::sqrt(std::fabs(D));
  const double trmax = -B + sqrtD;
  const double EPS   = 2. * rr * vv * kEpsilon;
  const bool done    = outside || (D <= EPS);
  if (!done) distance = Mn(tzmax, trmax);
  return distance;
}

// :203-243 (a positive safety is raised to the distance from the cut planes and the x-y extent)
double Ellipsoid::SafetyToIn(V3 const &p) const
{
  const double x = p.x * sx, y = p.y * sy, z = p.z * sz;
  const double rad   = std::sqrt(x * x + y * y + z * z);
  const double safeZ = std::fabs(z - scZMidCut) - scZDimCut;
  const double safeR = rad - r;
  double safety      = Mx(safeZ, safeR);
  if (std::fabs(safety) <= kHalfTolerance) safety = 0.;
  const double distZ  = Mx(p.z - zTopCut, zBottomCut - p.z);
  const double distXY = Mx(std::fabs(p.x) - xmax, std::fabs(p.y) - ymax);
  if (safety > 0.) safety = Mx(Mx(safety, distZ), distXY);
  return safety;
}

double Ellipsoid::SafetyToOut(V3 const &p) const
{
  const double x = p.x * sx, y = p.y * sy, z = p.z * sz;
  const double safeR = r - std::sqrt(x * x + y * y + z * z);
  const double safeZ = scZDimCut - std::fabs(z - scZMidCut);
  double safety      = Mn(safeZ, safeR);
  if (std::fabs(safety) <= kHalfTolerance) safety = 0.;
  const double distZ = Mn(zTopCut - p.z, p.z - zBottomCut);
  if (safety > 0.) safety = Mn(safeR, distZ);
  return safety;
}

// NormalKernel (:245-287)
bool Ellipsoid::Normal(V3 const &point, V3 &n) const
{
  V3 normal(0., 0., 0.);
  const double px = point.x, py = point.y, pz = point.z;
  const double A = dx, B = dy, C = dz;
  const double x = px * sx, y = py * sy, z = pz * sz;
  const double mag2 = x * x + y * y + z * z;
  double distR      = q1 * mag2 - q2;
  if (std::fabs(distR) <= kHalfTolerance) normal = Unit(V3(px / (A * A), py / (B * B), pz / (C * C)));
  const double distZ = std::fabs(z - scZMidCut) - scZDimCut;
  if (std::fabs(distZ) <= kHalfTolerance) normal.z = normal.z + Sign(z - scZMidCut);
  if (normal.Mag2() > 1.) normal = Unit(normal);
  if (normal.Mag2() > 0.) {
    n = normal;
    return true;
  }
  normal.z = Sign(z - scZMidCut);
  distR    = std::sqrt(mag2) - r;
  if (distR > distZ && mag2 > 0.) normal = Unit(V3(px / (A * A), py / (B * B), pz / (C * C)));
  n = normal;
  return false;
}

} // namespace vgo
