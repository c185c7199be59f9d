// Oracle: Trd and Paraboloid. TEST INFRASTRUCTURE ONLY (see common.h).
#include "shapes.h"

#include <algorithm>
#include <cmath>

namespace vgo {

// ============================================================== Trd (UniversalTrd)
// TrdStruct.h:102-121 CalculateCached. The y tolerance is computed from the x slope, as there.
Trd::Trd(double x1, double x2, double y1, double y2, double z, bool trd1)
    : dx1(x1), dx2(x2), dy1(y1), dy2(y2), dz(z), plainY(trd1)
{
  x2minusx1    = dx2 - dx1;
  y2minusy1    = dy2 - dy1;
  halfx1plusx2 = 0.5 * (dx1 + dx2);
  halfy1plusy2 = 0.5 * (dy1 + dy2);
  fx           = 0.5 * (dx1 - dx2) / dz;
  fy           = 0.5 * (dy1 - dy2) / dz;
  secxz        = std::sqrt(1 + fx * fx);
  secyz        = std::sqrt(1 + fy * fy);
  calfx        = 1. / std::sqrt(1.0 + fx * fx);
  calfy        = 1. / std::sqrt(1.0 + fy * fy);
  tolx         = kTolerance * std::sqrt(x2minusx1 * x2minusx1 + 4 * dz * dz);
  toly         = kTolerance * std::sqrt(x2minusx1 * x2minusx1 + 4 * dz * dz);
}

// TrdUtilities::UnplacedInside (TrdImplementation.h:171-210): the lateral tests compare the cross
// product of (|x| - dx1, z + dz) with the side's direction (dx2 - dx1, 2 dz).
void Trd::InsideFlags(V3 const &p, bool surface, bool &cin, bool &cout) const
{
  const double pz = p.z + dz;
  cout            = std::fabs(p.z) > (surface ? dz + kHalfTolerance : dz);
  if (surface) cin = std::fabs(p.z) < dz - kHalfTolerance;
  double cross = x2minusx1 * pz - 2.0 * dz * (std::fabs(p.x) - dx1);
  if (surface) {
    cout = cout || cross < -tolx;
    cin  = cin && cross > tolx;
  } else {
    cout = cout || cross < 0;
  }
  if (plainY) { // TrdImplementation.h:201-205 (HasVaryingY kNo)
    cout = cout || std::fabs(p.y) > (surface ? dy1 + kHalfTolerance : dy1);
    if (surface) cin = cin && std::fabs(p.y) < dy1 - kHalfTolerance;
    return;
  }
  cross = y2minusy1 * pz - 2.0 * dz * (std::fabs(p.y) - dy1);
  if (surface) {
    cout = cout || cross < -toly;
    cin  = cin && cross > toly;
  } else {
    cout = cout || cross < 0;
  }
}

bool Trd::Contains(V3 const &p) const
{
  bool cin = false, cout = false;
  InsideFlags(p, false, cin, cout);
  return !cout;
}

int Trd::Inside(V3 const &p) const
{
  bool cin = false, cout = false;
  InsideFlags(p, true, cin, cout);
  int r = kSurface;
  if (cin) r = kInside;
  if (cout) r = kOutside;
  return r;
}

// TrdUtilities::FaceTrajectoryIntersection (TrdImplementation.h:75-130): the ray against the lateral face
// +x (forY false) or +y, or its mirror image (mirrored), entering (toIn) or leaving. `dist` is written
// only when the face's orientation allows a crossing.
bool Trd::Face(bool forY, bool mirrored, bool toIn, V3 const &p, V3 const &d, double &dist) const
{
  double alongV, v1, posV, posK, dirV, dirK, fK, fV, halfKplus;
  if (forY) {
    alongV = y2minusy1, v1 = dy1, posV = p.y, posK = p.x, dirV = d.y, dirK = d.x, fK = fx, fV = fy;
    halfKplus = halfx1plusx2;
  } else {
    alongV = x2minusx1, v1 = dx1, posV = p.x, posK = p.y, dirV = d.x, dirK = d.y, fK = fy, fV = fx;
    halfKplus = halfy1plusy2;
  }
  if (mirrored) posV *= -1., dirV *= -1.;
  const double ndotv = dirV + fV * d.z;
  bool ok            = toIn ? ndotv < 0. : ndotv > 0.;
  if (!ok) return false;
  const double alongZ = 2.0 * dz;
  dist = (alongZ * (posV - v1) - alongV * (p.z + dz)) / (d.z * alongV - dirV * alongZ + kTiny);
  ok   = dist > 0. - kHalfTolerance;
  if (ok) {
    const double hitz = p.z + dist * d.z;
    ok                = std::fabs(hitz) <= dz;
    const double hitk = posK + dist * dirK;
    const double dK   = halfKplus - fK * hitz;
    ok                = ok && std::fabs(hitk) <= dK;
    if (ok && std::fabs(dist) < kHalfTolerance) dist = 0.0;
  }
  return ok;
}

// TrdImplementation.h:299-385
double Trd::DistanceToIn(V3 const &p, V3 const &d, double) const
{
  double distance   = kInfLength;
  const bool inz    = std::fabs(p.z) < dz - kHalfTolerance;
  double distx      = halfx1plusx2 - fx * p.z;
  const bool inx    = (distx - std::fabs(p.x)) * calfx > 0. + kHalfTolerance;
  double disty;
  bool iny;
  if (varyingY()) {
    disty = halfy1plusy2 - fy * p.z;
    iny   = (disty - std::fabs(p.y)) * calfy > 0. + kHalfTolerance;
  } else {
    disty = std::fabs(p.y) - dy1;
    iny   = disty < 0. - kHalfTolerance;
  }
  if (inx && iny && inz) return -1.; // the final |distance| < kHalfTolerance rule leaves -1
  bool okz = p.z * d.z < 0. && !inz;
  if (okz) {
    const double distz = (std::fabs(p.z) - dz) / std::fabs(d.z);
    const double hitx  = std::fabs(p.x + distz * d.x);
    const double hity  = std::fabs(p.y + distz * d.y);
    const bool okzt    = p.z > (dz - kHalfTolerance) && hitx <= dx2 && hity <= dy2;
    const bool okzb    = p.z < (-dz + kHalfTolerance) && hitx <= dx1 && hity <= dy1;
    okz                = okzt || okzb;
    if (okz) distance = distz;
  }
  auto finish = [](double dist) { return std::fabs(dist) < kHalfTolerance ? 0.0 : dist; };
  if (okz) return finish(distance);
  bool okx = false;
  if (!inx) {
    if (Face(false, false, true, p, d, distx)) distance = distx, okx = true;
    okx = false;
    if (Face(false, true, true, p, d, distx)) distance = distx, okx = true;
  }
  if (okx) return finish(distance);
  if (varyingY()) {
    if (!iny) {
      if (Face(true, false, true, p, d, disty)) distance = disty;
      if (Face(true, true, true, p, d, disty)) distance = disty;
    }
  } else if (!iny) {
    disty /= std::fabs(d.y);
    const double zhit = p.z + disty * d.z;
    const double xhit = p.x + disty * d.x;
    const double dx   = halfx1plusx2 - fx * zhit;
    const bool oky    = p.y * d.y < 0 && disty > -kHalfTolerance && std::fabs(xhit) < dx && std::fabs(zhit) < dz;
    if (oky) distance = disty;
  }
  return finish(distance);
}

// TrdImplementation.h:387-495
double Trd::DistanceToOut(V3 const &p, V3 const &d, double) const
{
  double distance     = 0.0;
  const double invdir = 1. / std::fabs(d.z + kTiny);
  const double safz   = dz - std::fabs(p.z);
  bool out            = safz < 0. - kHalfTolerance;
  double distx        = halfx1plusx2 - fx * p.z;
  out                 = out || (distx - std::fabs(p.x)) * calfx < 0. - kHalfTolerance;
  double disty;
  if (varyingY()) {
    disty = halfy1plusy2 - fy * p.z;
    out   = out || (disty - std::fabs(p.y)) * calfy < 0. - kHalfTolerance;
  } else {
    disty = dy1 - std::fabs(p.y);
    out   = out || disty < 0. - kHalfTolerance;
  }
  if (out) return -1.;
  auto finish = [](double dist) { return std::fabs(dist) < kHalfTolerance ? 0.0 : dist; };
  if (d.z > 0.) {
    const double distz = (dz - p.z) * invdir;
    const double hitx  = std::fabs(p.x + distz * d.x);
    const double hity  = std::fabs(p.y + distz * d.y);
    if (hitx <= dx2 && hity <= dy2) return finish(distz);
  }
  if (d.z < 0.) {
    const double distz = (p.z + dz) * invdir;
    const double hitx  = std::fabs(p.x + distz * d.x);
    const double hity  = std::fabs(p.y + distz * d.y);
    if (hitx <= dx1 && hity <= dy1) return finish(distz);
  }
  if (Face(false, false, false, p, d, distx)) return finish(distx);
  if (Face(false, true, false, p, d, distx)) return finish(distx);
  if (varyingY()) {
    if (Face(true, false, false, p, d, disty)) return finish(disty);
    if (Face(true, true, false, p, d, disty)) distance = disty;
  } else {
    const double plane = d.y < 0. ? -dy1 : dy1;
    disty              = (plane - p.y) / d.y;
    const double zhit  = p.z + disty * d.z;
    const double xhit  = p.x + disty * d.x;
    const double dx    = halfx1plusx2 - fx * zhit;
    if (std::fabs(xhit) < dx && std::fabs(zhit) < dz) distance = disty;
  }
  return finish(distance);
}

// TrdUtilities::Safety (TrdImplementation.h:132-165)
double Trd::Safety(V3 const &p, bool inside) const
{
  double dist       = dz - std::fabs(p.z);
  const double dxz  = halfx1plusx2 - fx * p.z;
  const double safx = (dxz - std::fabs(p.x)) * calfx;
  if (dxz >= 0 && safx < dist) dist = safx;
  if (varyingY()) {
    const double dyz  = halfy1plusy2 - fy * p.z;
    const double safy = (dyz - std::fabs(p.y)) * calfy;
    if (dyz >= 0 && safy < dist) dist = safy;
  } else {
    const double safy = dy1 - std::fabs(p.y);
    if (safy < dist) dist = safy;
  }
  return inside ? dist : -dist;
}

double Trd::SafetyToIn(V3 const &p) const { return Safety(p, false); }
double Trd::SafetyToOut(V3 const &p) const { return Safety(p, true); }

// ============================================================== Paraboloid
// ParaboloidStruct.h:48-70 CalculateCached and ComputeBoundingBox
Paraboloid::Paraboloid(double rlo_, double rhi_, double dz_) : rlo(rlo_), rhi(rhi_), dz(dz_)
{
  rlo2            = rlo * rlo;
  rhi2            = rhi * rhi;
  const double dd = 1. / (rhi2 - rlo2);
  a               = 2. * dz * dd;
  b               = -dz * (rlo2 + rhi2) * dd;
  k1              = (rhi2 - rlo2) * 0.5 / dz;
  k2              = (rhi2 + rlo2) * 0.5;
  invA            = 1 / a;
  invB            = 1 / b;
  a2              = a * a;
  b2              = b * b;
  dx              = rhi > rlo ? rhi : rlo;
  dy              = dx;
}

// ParaboloidImplementation.h:101-116 GenericKernelForContainsAndInside
void Paraboloid::Flags(V3 const &p, bool forInside, bool &cin, bool &cout) const
{
  cin               = false;
  const double diff = p.Perp2() - (k1 * p.z + k2);
  const double absZ = std::fabs(p.z);
  cout              = absZ > dz + kTolerance || diff > kTolerance;
  if (cout) return;
  if (forInside) cin = absZ < dz - kTolerance && diff < -kTolerance;
}

bool Paraboloid::Contains(V3 const &p) const
{
  bool cin, cout;
  Flags(p, false, cin, cout);
  return !cout;
}

int Paraboloid::Inside(V3 const &p) const
{
  bool cin, cout;
  Flags(p, true, cin, cout);
  int r = kSurface;
  if (cout) r = kOutside;
  if (cin) r = kInside;
  return r;
}

// ParaboloidImplementation.h:118-131 IsOnZPlane
bool Paraboloid::OnZPlane(bool top, V3 const &p) const
{
  const double rho2 = p.Perp2();
  if (top) return std::fabs(p.z - dz) < kTolerance && rho2 < rhi2 + kHalfTolerance;
  return std::fabs(p.z + dz) < kTolerance && rho2 < rlo2 + kHalfTolerance;
}

// ParaboloidImplementation.h:147-262. Far points are first moved along d to 2 * Rsph from the origin
// (the offset is added back to every distance).
double Paraboloid::DistanceToIn(V3 const &point, V3 const &d, double) const
{
  double offset = 0.;
  V3 p          = point;
  const double rsph  = 1.5 * (dx > dz ? dx : dz);
  const double rfar2 = 1024. * rsph * rsph;
  if (p.Mag2() > rfar2 && d.Dot(p) < 0.) offset = p.Mag() - 2. * rsph;
  p = p + d * offset;
  const double absZ = std::fabs(p.z);
  const double rho2 = p.Perp2();
  if (absZ > dz && p.z * d.z >= 0.) return kInfLength;
  const double diff = rho2 - (k1 * p.z + k2);
  if (absZ < dz - kTolerance && diff < -kTolerance) return -1.;
  if ((OnZPlane(true, point) && d.z < 0.) || (OnZPlane(false, point) && d.z > 0.)) return 0.;
  const V3 normal(p.x, p.y, -k1 * 0.5);
  if (diff > -kTolerance && diff < kTolerance && d.Dot(normal) < 0.) return 0.;
  double distance = kInfLength;
  if (!(absZ < dz)) {
    double distZ         = kInfLength;
    const bool bottom    = p.z < -dz && d.z > 0;
    const bool top       = p.z > dz && d.z < 0;
    if (top) distZ = (dz - p.z) / NonZero(d.z);
    if (bottom) distZ = (-dz - p.z) / NonZero(d.z);
    const double xHit    = p.x + distZ * d.x;
    const double yHit    = p.y + distZ * d.y;
    const double rhoHit2 = xHit * xHit + yHit * yHit;
    if (top && rhoHit2 <= rhi2) distance = distZ + offset;
    if (top && rhoHit2 < rhi2) return distance;
    if (bottom && rhoHit2 <= rlo2) return distZ + offset;
  }
  if (!(p.z > dz && rho2 < rhi2)) {
    const double dirRho2 = d.Perp2();
    const double pDotV2D = p.x * d.x + p.y * d.y;
    const double qa      = a * dirRho2;
    const double qb      = 0.5 * d.z - a * pDotV2D;
    const double qc      = b + a * p.Perp2() - p.z;
    const double d2      = qb * qb - qa * qc;
    if (d2 < 0.) return distance;
    const double distParab = qb <= 0. ? (qb - std::sqrt(d2)) / NonZero(qa) : qc / NonZero(qb + std::sqrt(d2));
    const double zHit      = p.z + distParab * d.z;
    if (std::fabs(zHit) <= dz && distParab > 0.) distance = distParab + offset;
  }
  return distance;
}

// ParaboloidImplementation.h:264-330
double Paraboloid::DistanceToOut(V3 const &p, V3 const &d, double) const
{
  if (std::fabs(p.z) > dz + kTolerance) return -1.;
  const double rho2  = p.Perp2();
  const double value = rho2 - (k1 * p.z + k2);
  if (value > kHalfTolerance) return -1.;
  if ((OnZPlane(true, p) && d.z > 0.) || (OnZPlane(false, p) && d.z < 0.)) return 0.;
  const V3 normal(p.x, p.y, -k1 * 0.5);
  if (value > -kTolerance && value < kTolerance && d.Dot(normal) > 0.) return 0.;
  const double dirZinv = 1. / NonZero(d.z);
  const double distZ   = d.z < 0 ? -(dz + p.z) * dirZinv : (dz - p.z) * dirZinv;
  const double dirRho2 = d.Perp2();
  const double pDotV2D = p.x * d.x + p.y * d.y;
  const double qa      = a * dirRho2;
  const double qb      = 0.5 * d.z - a * pDotV2D;
  const double qc      = b + a * p.Perp2() - p.z;
  const double d2      = qb * qb - qa * qc;
  double distParab     = kInfLength;
  if (d2 >= 0. && qb > 0.) distParab = (qb + std::sqrt(d2)) * (1. / NonZero(qa));
  if (d2 >= 0. && qb <= 0.) distParab = qc / NonZero(qb - std::sqrt(d2));
  return std::min(distParab, distZ);
}

// ParaboloidImplementation.h:332-376
double Paraboloid::SafetyToIn(V3 const &p) const
{
  const double absZ  = std::fabs(p.z);
  const double safeZ = absZ - dz;
  const double rho2  = p.Perp2();
  const double value = a * rho2 + b - p.z;
  if (absZ < dz - kTolerance && value < -kHalfTolerance) return -1.;
  if (std::fabs(std::fabs(p.z) - dz) < kTolerance && (rho2 < rhi2 + kHalfTolerance || rho2 < rlo2 + kHalfTolerance))
    return 0.;
  if (value > -kTolerance && value < kTolerance) return 0.;
  const double r0sq = (p.z - b) * invA;
  if (r0sq < 0) return safeZ;
  const double ro2 = p.x * p.x + p.y * p.y;
  const double dr  = std::sqrt(ro2) - std::sqrt(r0sq);
  if (dr < 1.E-8) return safeZ;
  const double talf  = -2. * a * std::sqrt(r0sq);
  const double salf  = talf / std::sqrt(1. + talf * talf);
  const double safeR = std::fabs(dr * salf);
  return std::max(safeR, safeZ);
}

// ParaboloidImplementation.h:378-428
double Paraboloid::SafetyToOut(V3 const &p) const
{
  const double absZ = std::fabs(p.z);
  const double safZ = dz - absZ;
  if (absZ > dz + kTolerance) return -1.;
  const double rho2  = p.Perp2();
  const double value = a * rho2 + b - p.z;
  if (value > kHalfTolerance) return -1.;
  if (std::fabs(std::fabs(p.z) - dz) < kTolerance && (rho2 < rhi2 + kHalfTolerance || rho2 < rlo2 + kHalfTolerance))
    return 0.;
  if (value > -kTolerance && value < kTolerance) return 0.;
  const double r0sq = (p.z - b) * invA;
  if (r0sq < 0) return 0.;
  const double ro2 = p.x * p.x + p.y * p.y;
  const double z0  = a * ro2 + b;
  const double dr  = std::sqrt(ro2) - std::sqrt(r0sq);
  if (dr > -1.E-8) return 0.;
  const double dzp  = std::fabs(p.z - z0);
  const double safR = -dr * dzp / std::sqrt(dr * dr + dzp * dzp);
  return std::min(safR, safZ);
}

} // namespace vgo
