// Oracle: Box, phi Wedge and Tube. TEST INFRASTRUCTURE ONLY (see common.h).
#include "shapes.h"

namespace vgo {

// ============================================================== Box
// VecGeom/volumes/kernel/BoxImplementation.h:76-82
bool Box::Contains(V3 const &p) const { return (p.Abs() - h).Max() < 0.0; }

// BoxImplementation.h:84-93
int Box::Inside(V3 const &p) const
{
  double dist = (p.Abs() - h).Max();
  int in      = dist < 0.0 ? kInside : kOutside;
  if (std::fabs(dist) < kHalfTolerance) in = kSurface;
  return in;
}

// BoxImplementation.h:106-131
double Box::DistanceToIn(V3 const &p, V3 const &d, double) const
{
  V3 inv(1.0 / NonZero(d.x), 1.0 / NonZero(d.y), 1.0 / NonZero(d.z));
  V3 sgn(Sign(d.x), Sign(d.y), Sign(d.z));
  V3 tin        = -sgn * h - p;
  V3 tout       = sgn * h - p;
  double orthog = std::fabs((sgn * tout).Min());
  double dout   = (tout * inv).Min();
  double dist   = (tin * inv).Max();
  if (dist >= dout || dout <= kHalfTolerance || orthog <= kHalfTolerance) dist = kInfLength;
  return dist;
}

// BoxImplementation.h:133-151
double Box::DistanceToOut(V3 const &p, V3 const &d, double) const
{
  V3 inv(1.0 / NonZero(d.x), 1.0 / NonZero(d.y), 1.0 / NonZero(d.z));
  V3 sgn(Sign(d.x), Sign(d.y), Sign(d.z));
  double safeIn = (p.Abs() - h).Max();
  V3 tout       = sgn * h - p;
  double dist   = (tout * inv).Min();
  if (safeIn > kHalfTolerance) dist = -1.0;
  return dist;
}

// BoxImplementation.h:153-167
double Box::SafetyToIn(V3 const &p) const { return (p.Abs() - h).Max(); }
double Box::SafetyToOut(V3 const &p) const { return (h - p.Abs()).Min(); }

// ============================================================== Wedge
// source/Wedge_Evolution.cpp:17-128 (only the along/normal vectors and fDPhi are used)
void Wedge::Set(double angle, double zeroangle)
{
  dphi    = angle;
  along1  = V3(std::cos(zeroangle), std::sin(zeroangle), 0);
  along2  = V3(std::cos(zeroangle + angle), std::sin(zeroangle + angle), 0);
  normal1 = V3(-std::sin(zeroangle), std::cos(zeroangle), 0);
  normal2 = V3(std::sin(zeroangle + angle), -std::cos(zeroangle + angle), 0);
}

// Wedge_Evolution.h:340-352
bool Wedge::IsOnSurfaceGeneric(V3 const &along, V3 const &normal, V3 const &p) const
{
  bool c1 = along.x * p.x + along.y * p.y >= 0.;
  if (!c1) return false;
  return std::fabs(normal.x * p.x + normal.y * p.y) < kTolerance;
}

// Wedge_Evolution.h:305-336
void Wedge::ContainsInside(V3 const &p, bool forInside, bool &cin, bool &cout) const
{
  double startCheck = (-p.x * along1.y + p.y * along1.x);
  double endCheck   = (-along2.x * p.y + along2.y * p.x);
  cout              = startCheck < 0.;
  if (dphi < kPi)
    cout = cout || (endCheck < 0.);
  else
    cout = cout && (endCheck < 0.);
  if (forInside) {
    bool onSurf = IsOnSurfaceGeneric(along1, normal1, p) || IsOnSurfaceGeneric(along2, normal2, p);
    cout        = cout && !onSurf;
    cin         = !onSurf && !cout;
  }
}

// Wedge_Evolution.h:354-372
double Wedge::SafetyToOut(V3 const &p) const
{
  double d1 = p.x * normal1.x + p.y * normal1.y;
  double d2 = p.x * normal2.x + p.y * normal2.y;
  return dphi < kPi ? std::min(d1, d2) : std::max(d1, d2);
}

// Wedge_Evolution.h:374-392
double Wedge::SafetyToIn(V3 const &p) const
{
  double d1 = p.x * normal1.x + p.y * normal1.y;
  double d2 = p.x * normal2.x + p.y * normal2.y;
  return dphi < kPi ? std::max(-1 * d1, -1 * d2) : std::min(-1 * d1, -1 * d2);
}

// ============================================================== Tube
// VecGeom/volumes/TubeStruct.h:44-175 (ctor: raw sphi/dphi, CalculateCached)
Tube::Tube(double rmin_, double rmax_, double dz, double sphi_, double dphi_)
    : rmin(rmin_ < 0.0 ? 0.0 : rmin_), rmax(rmax_), z(dz), sphi(sphi_), dphi(dphi_)
{
  wedge.Set(dphi, sphi);
  tolIz     = z - kHalfTolerance;
  tolOz     = z + kHalfTolerance;
  rmin2     = rmin * rmin;
  tolOrmin  = rmin - kHalfTolerance;
  tolIrmin  = rmin + kHalfTolerance;
  tolOrmin2 = tolOrmin * tolOrmin;
  tolIrmin2 = tolIrmin * tolIrmin;
  rmax2     = rmax * rmax;
  tolOrmax  = rmax + kHalfTolerance;
  tolIrmax  = rmax - kHalfTolerance;
  tolOrmax2 = tolOrmax * tolOrmax;
  tolIrmax2 = tolIrmax * tolIrmax;
  a1x       = std::cos(sphi);
  a1y       = std::sin(sphi);
  a2x       = std::cos(sphi + dphi);
  a2y       = std::sin(sphi + dphi);
  maxVal    = std::max(rmax, z);
}

// source/UnplacedTube.cpp:214-262
void Tube::Extent(V3 &lo, V3 &hi) const
{
  lo = V3(-rmax, -rmax, -z);
  hi = V3(rmax, rmax, z);
  if (dphi == kTwoPi) return;
  PhiExtent(wedge, rmax, rmin, sphi, dphi, lo, hi);
}

void PhiExtent(Wedge const &w, double rOut, double rIn, double sphi, double dphi, V3 &lo, V3 &hi)
{
  const double Rin = 0.5 * (rOut + rIn);
  const bool o0    = !w.Contains(V3(Rin, 0, 0));
  const bool o90   = !w.Contains(V3(0, Rin, 0));
  const bool o180  = !w.Contains(V3(-Rin, 0, 0));
  const bool o270  = !w.Contains(V3(0, -Rin, 0));
  if (!(o0 || o90 || o180 || o270)) return;
  const double Cx = rOut * std::cos(sphi), Dx = rOut * std::cos(sphi + dphi);
  const double Cy = rOut * std::sin(sphi), Dy = rOut * std::sin(sphi + dphi);
  if (o0) hi.x = std::max(Cx, Dx);
  if (o90) hi.y = std::max(Cy, Dy);
  if (o180) lo.x = std::min(Cx, Dx);
  if (o270) lo.y = std::min(Cy, Dy);
  if (dphi >= kPi) return;
  const double Ax = rIn * std::cos(sphi), Bx = rIn * std::cos(sphi + dphi);
  const double Ay = rIn * std::sin(sphi), By = rIn * std::sin(sphi + dphi);
  double t = std::max(Ax, Bx);
  hi.x     = t > hi.x ? t : hi.x;
  t        = std::max(Ay, By);
  hi.y     = t > hi.y ? t : hi.y;
  t        = std::min(Ax, Bx);
  lo.x     = t < lo.x ? t : lo.x;
  t        = std::min(Ay, By);
  lo.y     = t < lo.y ? t : lo.y;
}

// TubeImplementation.h:58-99 PointInCyclicalSector (onSurface=false)
bool Tube::InSector(double x, double y, bool includeSurface) const
{
  double startCheck = (-x * a1y + y * a1x);
  double endCheck   = (-a2x * y + a2y * x);
  double t          = includeSurface ? -kHalfTolerance : kHalfTolerance;
  if (dphi <= M_PI) return (startCheck >= t) && (endCheck >= t);
  return (startCheck >= t) || (endCheck >= t);
}

// TubeImplementation.h:101-151 CircleTrajectoryIntersection
bool Tube::CircleHit(double b, double c, V3 const &p, V3 const &d, bool largest, bool sectorCheck,
                     double &dist) const
{
  double delta = b * b - c;
  bool ok      = delta > 0.;
  if (largest) ok = ok || (delta == 0.);
  if (!ok) delta = 0.;
  delta = std::sqrt(delta);
  if (!largest) delta = -delta;
  dist = -b + delta;
  ok   = ok && (dist >= -2 * kTolerance);
  if (!ok) return false;
  if (sectorCheck) {
    double hitz = p.z + dist * d.z;
    ok          = ok && (std::fabs(hitz) <= z);
    if (!ok) return false;
    if (hasPhi()) {
      double hitx = p.x + dist * d.x;
      double hity = p.y + dist * d.y;
      ok          = ok && InSector(hitx, hity, true);
    }
  }
  return ok;
}

// TubeImplementation.h:255-305 PhiPlaneTrajectoryIntersection
bool Tube::PhiPlaneHit(double ax, double ay, double nx, double ny, V3 const &p, V3 const &d, bool posDir, bool toIn,
                       double &dist) const
{
  double dirDotNorm = d.x * nx + d.y * ny;
  bool ok           = toIn ? (dirDotNorm > 0.) : (dirDotNorm < 0.);
  double dirDotXY   = (d.y * ax - d.x * ay);
  dist              = (ay * p.x - ax * p.y) / NonZero(dirDotXY);
  ok                = ok && ((dist * std::fabs(dirDotNorm)) > -kHalfTolerance);
  if (toIn) {
    double hitx = p.x + dist * d.x;
    double hity = p.y + dist * d.y;
    double hitz = p.z + dist * d.z;
    double r2   = hitx * hitx + hity * hity;
    ok          = ok && std::fabs(hitz) <= tolOz && (r2 >= tolOrmin2) && (r2 <= tolOrmax2);
    if (posDir) ok = ok && (hitx * ax + hity * ay) > 0.;
  } else if (posDir) {
    double hitx = p.x + dist * d.x;
    double hity = p.y + dist * d.y;
    ok          = ok && (hitx * ax + hity * ay) >= 0.;
  }
  return ok;
}

// TubeImplementation.h:206-237 PhiPlaneSafety<inside=false>
double Tube::PhiPlaneSafetyOutside(V3 const &p) const
{
  double safety = dphi > M_PI ? std::sqrt(p.x * p.x + p.y * p.y) : kInfLength;
  double phi1   = p.x * a1y - p.y * a1x;
  if (phi1 > -kHalfTolerance && phi1 < safety) safety = phi1;
  double phi2 = p.x * a2y - p.y * a2x;
  phi2 *= -1;
  if (phi2 > -kHalfTolerance && phi2 < safety) safety = phi2;
  return safety;
}

// TubeImplementation.h:389-447 GenericKernelForContainsAndInside, with
// MakePlus/MinusTolerant(Square) from volumes/kernel/GenericKernels.h:26-82
void Tube::ContainsInside(V3 const &p, bool forInside, bool &cin, bool &cout) const
{
  double absz = std::fabs(p.z);
  cout        = absz > (forInside ? z + kHalfTolerance : z);
  if (forInside) cin = absz < z - kHalfTolerance;
  double r2 = p.x * p.x + p.y * p.y;
  cout      = cout || (r2 > (forInside ? rmax * (rmax + 2.0 * kTolerance) : rmax * rmax));
  if (forInside) cin = cin && (r2 < std::fabs(rmax * (rmax - 2.0 * kTolerance)));
  if (hasRmin()) {
    cout = cout || (r2 <= (forInside ? std::fabs(rmin * (rmin - 2.0 * kTolerance)) : rmin * rmin));
    if (forInside) cin = cin && (r2 > rmin * (rmin + 2.0 * kTolerance));
  }
  if (hasPhi()) {
    bool wi = false, wo = false;
    wedge.ContainsInside(p, forInside, wi, wo);
    cout = cout || wo;
    if (forInside) cin = cin && wi;
  }
}

bool Tube::Contains(V3 const &p) const
{
  bool ci = false, co = false;
  ContainsInside(p, false, ci, co);
  return !co;
}

int Tube::Inside(V3 const &p) const
{
  bool ci = false, co = false;
  ContainsInside(p, true, ci, co);
  int in = kSurface;
  if (co) in = kOutside;
  if (ci) in = kInside;
  return in;
}

// TubeImplementation.h:475-502: far points are first moved to 100*maxVal from the origin
double Tube::DistanceToIn(V3 const &pt, V3 const &d, double) const
{
  V3 p          = pt;
  double ptDist = p.Mag();
  double move   = 0.;
  if (ptDist > 100. * maxVal) {
    move = ptDist - 100. * maxVal;
    p    = p + move * d;
  }
  double dist = DistanceToInKernel(p, d);
  return dist + move;
}

// TubeImplementation.h:504-681 DistanceToInKernel
double Tube::DistanceToInKernel(V3 const &p, V3 const &d) const
{
  double distz = std::fabs(p.z) - z;
  if (distz > kHalfTolerance && p.z * d.z >= 0) return kInfLength;
  double rsq   = p.x * p.x + p.y * p.y;
  double rdotn = p.x * d.x + p.y * d.y;
  if (rsq > tolIrmax2 && rdotn >= 0) return kInfLength;

  // inside -> -1
  bool inside = distz < -kHalfTolerance;
  inside      = inside && rsq < tolIrmax2;
  if (hasRmin()) inside = inside && rsq > tolIrmin2;
  if (hasPhi() && inside) inside = inside && InSector(p.x, p.y, false);
  if (inside) return -1.0;

  double distance = kInfLength;
  // z-plane entry
  distz /= NonZeroAbs(d.z);
  double hitx = p.x + distz * d.x;
  double hity = p.y + distz * d.y;
  double r2   = hitx * hitx + hity * hity;
  bool okz    = distz > -kHalfTolerance && (p.z * d.z < 0);
  okz         = okz && (r2 <= rmax2);
  if (hasRmin()) okz = okz && (rmin2 <= r2);
  if (hasPhi() && okz) okz = okz && InSector(hitx, hity, true);
  if (okz) return distz;

  // on the rmax (or rmin) surface and moving in: TubeImplementation.h:310-345
  auto onSurfMovingIn = [&](bool inner) {
    double rho = p.Perp2();
    bool on;
    if (inner)
      on = (rho >= tolOrmin2) && (rho <= tolIrmin2) && (std::fabs(p.z) < (z + kTolerance));
    else
      on = (rho >= tolIrmax2) && (rho <= tolOrmax2) && (std::fabs(p.z) < (z + kTolerance));
    V3 n    = inner ? V3(-p.x, -p.y, 0.) : V3(p.x, p.y, 0.);
    return on && (d.Dot(n) <= 0.);
  };
  bool surfIn = onSurfMovingIn(false);
  if (hasRmin()) surfIn = surfIn || onSurfMovingIn(true);
  if (!hasPhi()) {
    if (surfIn) return 0.;
  } else {
    if (InSector(p.x, p.y, true) && surfIn) return 0.;
  }

  double invnsq = 1. / NonZero(1. - d.z * d.z);
  double b      = invnsq * rdotn;
  double crmax  = invnsq * (rsq - rmax2);
  double drmax  = 0.;
  bool okrmax   = CircleHit(b, crmax, p, d, false, true, drmax);
  okrmax        = okrmax && drmax < distance;
  if (okrmax) return drmax;

  if (hasRmin()) {
    double crmin = invnsq * (rsq - rmin2);
    double drmin = -kInfLength;
    bool okrmin  = CircleHit(b, crmin, p, d, true, true, drmin);
    okrmin       = okrmin && drmin < distance;
    if (okrmin) distance = drmin;
  }

  if (hasPhi()) {
    double dphi1 = 0.;
    // SectorType is kUnknownAngle for UniversalTube, so the positive-direction check is on
    bool okphi = PhiPlaneHit(a1x, a1y, wedge.normal1.x, wedge.normal1.y, p, d, true, true, dphi1);
    okphi      = okphi && dphi1 < distance;
    if (okphi) distance = dphi1;
    // the second plane is not masked by the first plane's hit (TubeImplementation.h:676-679)
    double dphi2 = 0.;
    okphi        = PhiPlaneHit(a2x, a2y, wedge.normal2.x, wedge.normal2.y, p, d, true, true, dphi2);
    if (okphi && dphi2 < distance) distance = dphi2;
  }
  return distance;
}

// TubeImplementation.h:686-821 DistanceToOut (UniversalTube takes the "bigger than pi or unknown" phi branch)
double Tube::DistanceToOut(V3 const &p, V3 const &d, double) const
{
  double distz = z - std::fabs(p.z);
  if (distz < -kHalfTolerance) return -1.;
  double rsq   = p.x * p.x + p.y * p.y;
  double rdotn = d.x * p.x + d.y * p.y;
  double crmax = rsq - rmax2;
  double crmin = rsq;
  if (crmax > 2.0 * kTolerance * rmax) return -1.;
  if (hasRmin()) {
    crmin -= rmin2;
    if (crmin < -2.0 * kTolerance * rmin) return -1.;
  }
  if (hasPhi()) {
    bool wi = false, wo = false;
    wedge.ContainsInside(p, true, wi, wo);
    if (wo) return -1.;
  }
  double distance = kInfLength;
  double invdirz  = 1. / NonZero(d.z);
  distz           = (z - p.z) * invdirz;
  if (d.z < 0) distz = (-z - p.z) * invdirz;
  if (d.z != 0. && distz < distance) distance = distz;

  double invnsq = 1. / NonZero(1. - d.z * d.z);
  double b      = invnsq * rdotn;
  if (hasRmin()) {
    double drmin = kInfLength;
    crmin *= invnsq;
    bool ok = CircleHit(b, crmin, p, d, false, false, drmin);
    if (ok && drmin < distance) distance = drmin;
  }
  double drmax = kInfLength;
  crmax *= invnsq;
  bool ok = CircleHit(b, crmax, p, d, true, false, drmax);
  if (ok && drmax < distance) distance = drmax;

  if (hasPhi()) {
    double dp = kInfLength;
    bool okp  = PhiPlaneHit(a1x, a1y, wedge.normal1.x, wedge.normal1.y, p, d, true, false, dp);
    if (okp && dp < distance) distance = dp;
    okp = PhiPlaneHit(a2x, a2y, wedge.normal2.x, wedge.normal2.y, p, d, true, false, dp);
    if (okp && dp < distance) distance = dp;
  }
  return distance;
}

// TubeImplementation.h:909-941 SafetyToInOld
double Tube::SafetyToIn(V3 const &p) const
{
  double safety = std::fabs(p.z) - z;
  double r      = std::sqrt(p.x * p.x + p.y * p.y);
  double sr     = r - rmax;
  if (sr > safety) safety = sr;
  if (hasRmin()) {
    double s = rmin - r;
    if (s > safety) safety = s;
  }
  if (hasPhi()) {
    bool insector = InSector(p.x, p.y, false);
    if (insector) return safety;
    double sphi_ = PhiPlaneSafetyOutside(p);
    if (sphi_ < kInfLength && sphi_ > safety) safety = sphi_;
  }
  return safety;
}

// TubeImplementation.h:943-964 SafetyToOutOld
double Tube::SafetyToOut(V3 const &p) const
{
  double safety = z - std::fabs(p.z);
  double r      = std::sqrt(p.x * p.x + p.y * p.y);
  double sr     = rmax - r;
  if (sr < safety) safety = sr;
  if (hasRmin()) {
    double s = r - rmin;
    if (s < safety) safety = s;
  }
  if (hasPhi()) {
    double s = wedge.SafetyToOut(p);
    if (s < safety) safety = s;
  }
  return safety;
}

} // namespace vgo
