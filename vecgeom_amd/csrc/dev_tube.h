// Tube kernels: the run-time-specialised "UniversalTube" of
// VecGeom/volumes/kernel/TubeImplementation.h:389-964 (rmin and phi treatment
// are record flags, VecGeom/volumes/kernel/shapetypes/TubeTypes.h:92-144),
// phi helpers TubeImplementation.h:58-308 and the evolution wedge
// VecGeom/volumes/Wedge_Evolution.h:305-392.
#pragma once
#include "dev_common.h"
#include "shape_data.h"

namespace vgmi {
namespace tube {

// PointInCyclicalSector (onSurface = false)
template <class R>
VGMI_DEV bool InSector(TubeData const &t, R x, R y, bool includeSurface)
{
  R startCheck = (-x * R(t.a1y) + y * R(t.a1x));
  R endCheck   = (-R(t.a2x) * y + R(t.a2y) * x);
  R tol        = includeSurface ? R(-kHalfTolerance) : R(kHalfTolerance);
  if (t.phiLePi) return (startCheck >= tol) && (endCheck >= tol);
  return (startCheck >= tol) || (endCheck >= tol);
}

// Wedge::IsOnSurfaceGeneric
template <class R>
VGMI_DEV bool OnWedgePlane(R ax, R ay, R nx, R ny, Vec3<R> const &p)
{
  if (!(ax * p.x + ay * p.y >= R(0))) return false;
  return vabs(nx * p.x + ny * p.y) < R(kTolerance);
}

// Wedge::GenericKernelForContainsAndInside
template <class R>
VGMI_DEV void WedgeContainsInside(TubeData const &t, Vec3<R> const &p, bool forInside, bool &cin, bool &cout)
{
  R startCheck = (-p.x * R(t.a1y) + p.y * R(t.a1x));
  R endCheck   = (-R(t.a2x) * p.y + R(t.a2y) * p.x);
  cout         = startCheck < R(0);
  if (t.wedgeLtPi)
    cout = cout || (endCheck < R(0));
  else
    cout = cout && (endCheck < R(0));
  if (forInside) {
    bool on = OnWedgePlane(R(t.a1x), R(t.a1y), R(t.n1x), R(t.n1y), p) ||
              OnWedgePlane(R(t.a2x), R(t.a2y), R(t.n2x), R(t.n2y), p);
    cout = cout && !on;
    cin  = !on && !cout;
  }
}

// CircleTrajectoryIntersection
template <class R>
VGMI_DEV bool CircleHit(TubeData const &t, R b, R c, Vec3<R> const &p, Vec3<R> const &d, bool largest,
                        bool sectorCheck, R &dist)
{
  R delta = b * b - c;
  bool ok = delta > R(0);
  if (largest) ok = ok || (delta == R(0));
  if (!ok) delta = R(0);
  delta = vsqrt(delta);
  if (!largest) delta = -delta;
  dist = -b + delta;
  ok   = ok && (dist >= R(-2 * kTolerance));
  if (!ok) return false;
  if (sectorCheck) {
    R hitz = p.z + dist * d.z;
    if (!(vabs(hitz) <= R(t.z))) return false;
    if (t.hasPhi) {
      R hitx = p.x + dist * d.x;
      R hity = p.y + dist * d.y;
      ok     = InSector(t, hitx, hity, true);
    }
  }
  return ok;
}

// PhiPlaneTrajectoryIntersection with PositiveDirectionOfPhiVector = true
template <class R>
VGMI_DEV bool PhiPlaneHit(TubeData const &t, R ax, R ay, R nx, R ny, Vec3<R> const &p, Vec3<R> const &d, bool toIn,
                          R &dist)
{
  R dirDotNorm = d.x * nx + d.y * ny;
  bool ok      = toIn ? (dirDotNorm > R(0)) : (dirDotNorm < R(0));
  R dirDotXY   = (d.y * ax - d.x * ay);
  dist         = (ay * p.x - ax * p.y) / NonZero(dirDotXY);
  ok           = ok && ((dist * vabs(dirDotNorm)) > R(-kHalfTolerance));
  R hitx       = p.x + dist * d.x;
  R hity       = p.y + dist * d.y;
  if (toIn) {
    R hitz = p.z + dist * d.z;
    R r2   = hitx * hitx + hity * hity;
    ok     = ok && vabs(hitz) <= R(t.z + kHalfTolerance) && (r2 >= R(t.tolOrmin2)) && (r2 <= R(t.tolOrmax2));
    ok     = ok && (hitx * ax + hity * ay) > R(0);
  } else {
    ok = ok && (hitx * ax + hity * ay) >= R(0);
  }
  return ok;
}

// GenericKernelForContainsAndInside with MakePlus/MinusTolerant(Square)
// (VecGeom/volumes/kernel/GenericKernels.h:26-82)
template <class R>
VGMI_DEV void ContainsInside(TubeData const &t, Vec3<R> const &p, bool forInside, bool &cin, bool &cout)
{
  const R z = R(t.z), rmax = R(t.rmax), rmin = R(t.rmin);
  R absz    = vabs(p.z);
  cout      = absz > (forInside ? z + R(kHalfTolerance) : z);
  cin       = forInside && absz < z - R(kHalfTolerance);
  R r2      = p.x * p.x + p.y * p.y;
  cout      = cout || (r2 > (forInside ? rmax * (rmax + R(2.0 * kTolerance)) : rmax * rmax));
  if (forInside) cin = cin && (r2 < vabs(rmax * (rmax - R(2.0 * kTolerance))));
  if (t.hasRmin) {
    cout = cout || (r2 <= (forInside ? vabs(rmin * (rmin - R(2.0 * kTolerance))) : rmin * rmin));
    if (forInside) cin = cin && (r2 > rmin * (rmin + R(2.0 * kTolerance)));
  }
  if (t.hasPhi) {
    bool wi = false, wo = false;
    WedgeContainsInside(t, p, forInside, wi, wo);
    cout = cout || wo;
    if (forInside) cin = cin && wi;
  }
}

template <class R>
VGMI_DEV bool Contains(TubeData const &t, Vec3<R> const &p)
{
  bool ci, co;
  ContainsInside(t, p, false, ci, co);
  return !co;
}

template <class R>
VGMI_DEV int Inside(TubeData const &t, Vec3<R> const &p)
{
  bool ci, co;
  ContainsInside(t, p, true, ci, co);
  return ci ? kInside : (co ? kOutside : kSurface);
}

// DistanceToInKernel (TubeImplementation.h:504-681)
template <class R>
VGMI_DEV R DistanceToInKernel(TubeData const &t, Vec3<R> const &p, Vec3<R> const &d)
{
  const R z = R(t.z);
  R distz   = vabs(p.z) - z;
  if (distz > R(kHalfTolerance) && p.z * d.z >= R(0)) return InfLen<R>();
  R rsq   = p.x * p.x + p.y * p.y;
  R rdotn = p.x * d.x + p.y * d.y;
  const R tolIrmax2 = R(t.tolIrmax2);
  if (rsq > tolIrmax2 && rdotn >= R(0)) return InfLen<R>();

  bool inside = distz < R(-kHalfTolerance) && rsq < tolIrmax2;
  if (t.hasRmin) inside = inside && rsq > R(t.tolIrmin2);
  if (t.hasPhi && inside) inside = InSector(t, p.x, p.y, false);
  if (inside) return R(-1);

  // entry through a z plane
  distz /= NonZeroAbs(d.z);
  {
    R hitx   = p.x + distz * d.x;
    R hity   = p.y + distz * d.y;
    R r2     = hitx * hitx + hity * hity;
    bool okz = distz > R(-kHalfTolerance) && (p.z * d.z < R(0)) && (r2 <= R(t.rmax2));
    if (t.hasRmin) okz = okz && (R(t.rmin2) <= r2);
    if (t.hasPhi && okz) okz = InSector(t, hitx, hity, true);
    if (okz) return distz;
  }

  // on a cylindrical surface and moving inwards (IsMovingInsideTubeSurface)
  {
    R rho     = rsq;
    bool zok  = vabs(p.z) < z + R(kTolerance);
    bool onO  = (rho >= tolIrmax2) && (rho <= R(t.tolOrmax2)) && zok && (d.x * p.x + d.y * p.y + d.z * R(0) <= R(0));
    bool surf = onO;
    if (t.hasRmin) {
      bool onI = (rho >= R(t.tolOrmin2)) && (rho <= R(t.tolIrmin2)) && zok &&
                 (d.x * -p.x + d.y * -p.y + d.z * R(0) <= R(0));
      surf = surf || onI;
    }
    if (t.hasPhi) surf = surf && InSector(t, p.x, p.y, true);
    if (surf) return R(0);
  }

  R invnsq = R(1) / NonZero(R(1) - d.z * d.z);
  R b      = invnsq * rdotn;
  R dist   = InfLen<R>();
  {
    R crmax = invnsq * (rsq - R(t.rmax2));
    R dr;
    if (CircleHit(t, b, crmax, p, d, false, true, dr) && dr < dist) return dr;
  }
  if (t.hasRmin) {
    R crmin = invnsq * (rsq - R(t.rmin2));
    R dr;
    if (CircleHit(t, b, crmin, p, d, true, true, dr) && dr < dist) dist = dr;
  }
  if (t.hasPhi) {
    R dp;
    if (PhiPlaneHit(t, R(t.a1x), R(t.a1y), R(t.n1x), R(t.n1y), p, d, true, dp) && dp < dist) dist = dp;
    if (PhiPlaneHit(t, R(t.a2x), R(t.a2y), R(t.n2x), R(t.n2y), p, d, true, dp) && dp < dist) dist = dp;
  }
  return dist;
}

// DistanceToIn with the far-point shift (TubeImplementation.h:475-502)
template <class R>
VGMI_DEV R DistanceToIn(TubeData const &t, Vec3<R> const &pt, Vec3<R> const &d)
{
  Vec3<R> p = pt;
  R ptDist  = vsqrt(p.Mag2());
  R move    = R(0);
  if (ptDist > R(100.) * R(t.maxVal)) {
    move = ptDist - R(100.) * R(t.maxVal);
    p    = p + move * d;
  }
  return DistanceToInKernel(t, p, d) + move;
}

// DistanceToOut (TubeImplementation.h:686-821, UniversalTube phi branch)
template <class R>
VGMI_DEV R DistanceToOut(TubeData const &t, Vec3<R> const &p, Vec3<R> const &d)
{
  const R z = R(t.z);
  R distz   = z - vabs(p.z);
  if (distz < R(-kHalfTolerance)) return R(-1);
  R rsq   = p.x * p.x + p.y * p.y;
  R rdotn = d.x * p.x + d.y * p.y;
  R crmax = rsq - R(t.rmax2);
  R crmin = rsq;
  if (crmax > R(2.0 * kTolerance * t.rmax)) return R(-1);
  if (t.hasRmin) {
    crmin -= R(t.rmin2);
    if (crmin < R(-2.0 * kTolerance * t.rmin)) return R(-1);
  }
  if (t.hasPhi) {
    bool wi, wo;
    WedgeContainsInside(t, p, true, wi, wo);
    if (wo) return R(-1);
  }
  R dist    = InfLen<R>();
  R invdirz = R(1) / NonZero(d.z);
  distz     = d.z < R(0) ? (-z - p.z) * invdirz : (z - p.z) * invdirz;
  if (d.z != R(0) && distz < dist) dist = distz;

  R invnsq = R(1) / NonZero(R(1) - d.z * d.z);
  R b      = invnsq * rdotn;
  if (t.hasRmin) {
    R dr;
    if (CircleHit(t, b, crmin * invnsq, p, d, false, false, dr) && dr < dist) dist = dr;
  }
  {
    R dr;
    if (CircleHit(t, b, crmax * invnsq, p, d, true, false, dr) && dr < dist) dist = dr;
  }
  if (t.hasPhi) {
    R dp;
    if (PhiPlaneHit(t, R(t.a1x), R(t.a1y), R(t.n1x), R(t.n1y), p, d, false, dp) && dp < dist) dist = dp;
    if (PhiPlaneHit(t, R(t.a2x), R(t.a2y), R(t.n2x), R(t.n2y), p, d, false, dp) && dp < dist) dist = dp;
  }
  return dist;
}

// SafetyToInOld (TubeImplementation.h:909-941) with PhiPlaneSafety<inside=false>
template <class R>
VGMI_DEV R SafetyToIn(TubeData const &t, Vec3<R> const &p)
{
  R safety = vabs(p.z) - R(t.z);
  R r      = vsqrt(p.x * p.x + p.y * p.y);
  R sr     = r - R(t.rmax);
  if (sr > safety) safety = sr;
  if (t.hasRmin) {
    R s = R(t.rmin) - r;
    if (s > safety) safety = s;
  }
  if (t.hasPhi) {
    if (InSector(t, p.x, p.y, false)) return safety;
    R sp   = t.phiGtPi ? r : InfLen<R>();
    R phi1 = p.x * R(t.a1y) - p.y * R(t.a1x);
    if (phi1 > R(-kHalfTolerance) && phi1 < sp) sp = phi1;
    R phi2 = -(p.x * R(t.a2y) - p.y * R(t.a2x));
    if (phi2 > R(-kHalfTolerance) && phi2 < sp) sp = phi2;
    if (sp < InfLen<R>() && sp > safety) safety = sp;
  }
  return safety;
}

// SafetyToOutOld (TubeImplementation.h:943-964) with Wedge::SafetyToOut
template <class R>
VGMI_DEV R SafetyToOut(TubeData const &t, Vec3<R> const &p)
{
  R safety = R(t.z) - vabs(p.z);
  R r      = vsqrt(p.x * p.x + p.y * p.y);
  R sr     = R(t.rmax) - r;
  if (sr < safety) safety = sr;
  if (t.hasRmin) {
    R s = r - R(t.rmin);
    if (s < safety) safety = s;
  }
  if (t.hasPhi) {
    R d1 = p.x * R(t.n1x) + p.y * R(t.n1y);
    R d2 = p.x * R(t.n2x) + p.y * R(t.n2y);
    R s  = t.wedgeLtPi ? vmin(d1, d2) : vmax(d1, d2);
    if (s < safety) safety = s;
  }
  return safety;
}

} // namespace tube
} // namespace vgmi
