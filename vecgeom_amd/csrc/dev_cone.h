// Cone kernels: the run-time specialised "UniversalCone" of
// VecGeom/volumes/kernel/ConeImplementation.h:95-521 with the scalar
// ConeHelpers<Precision> conical-surface solver and inside kernel
// (VecGeom/volumes/ConeUtilities.h:580-820), which is what the reference's
// scalar backend evaluates. Phi helpers ConeUtilities.h:62-240.
#pragma once
#include "dev_common.h"
#include "shape_data.h"

namespace vgmi {
namespace cone {

template <class R>
VGMI_DEV R RadiusAt(ConeData const &c, bool inner, R z)
{
  if (inner) return c.rmin1 == c.rmin2 ? R(c.rmin1) : R(c.innerSlope) * z + R(c.innerOffset);
  return c.origRmax1 == c.origRmax2 ? R(c.origRmax1) : R(c.outerSlope) * z + R(c.outerOffset);
}

template <class R>
VGMI_DEV bool InSector(ConeData const &c, R x, R y, bool includeSurface)
{
  R startCheck = (-x * R(c.a1y)) + (y * R(c.a1x));
  R endCheck   = (-R(c.a2x) * y) + (R(c.a2y) * x);
  R t          = includeSurface ? R(-kConeTolerance) : R(kConeTolerance);
  if (c.phiLePi) return (startCheck >= t) && (endCheck >= t);
  return (startCheck >= t) || (endCheck >= t);
}

template <class R>
VGMI_DEV bool OnWedgePlane(R ax, R ay, R nx, R ny, Vec3<R> const &p)
{
  if (!(ax * p.x + ay * p.y >= R(0))) return false;
  return vabs(nx * p.x + ny * p.y) < R(kTolerance);
}

template <class R>
VGMI_DEV void WedgeContainsInside(ConeData const &c, Vec3<R> const &p, bool forInside, bool &cin, bool &cout)
{
  R startCheck = (-p.x * R(c.w1y) + p.y * R(c.w1x));
  R endCheck   = (-R(c.w2x) * p.y + R(c.w2y) * p.x);
  cout         = startCheck < R(0);
  if (c.wedgeLtPi)
    cout = cout || (endCheck < R(0));
  else
    cout = cout && (endCheck < R(0));
  if (forInside) {
    bool on = OnWedgePlane(R(c.w1x), R(c.w1y), R(c.n1x), R(c.n1y), p) ||
              OnWedgePlane(R(c.w2x), R(c.w2y), R(c.n2x), R(c.n2y), p);
    cout = cout && !on;
    cin  = !on && !cout;
  }
}

template <class R>
VGMI_DEV Vec3<R> Normal(ConeData const &c, bool inner, Vec3<R> const &p)
{
  R rho = vsqrt(p.Perp2());
  if (inner) {
    if (c.rmin1 == c.rmin2 && c.rmin1 != 0.) return Vec3<R>(-p.x, -p.y, R(0));
    return Vec3<R>(-p.x, -p.y, R(c.zNormInner) * (rho * R(c.secRMin)));
  }
  if (c.rmax1 == c.rmax2 && c.rmax1 != 0.) return Vec3<R>(p.x, p.y, R(0));
  return Vec3<R>(p.x, p.y, R(c.zNormOuter) * (rho * R(c.secRMax)));
}

template <class R>
VGMI_DEV bool PhiPlaneHit(ConeData const &c, R ax, R ay, R nx, R ny, Vec3<R> const &p, Vec3<R> const &d, bool toIn,
                          R &dist)
{
  dist       = InfLen<R>();
  bool ok    = toIn ? ((d.x * nx) + (d.y * ny) > R(0)) : ((d.x * nx) + (d.y * ny) < R(0));
  R dirDotXY = (d.y * ax) - (d.x * ay);
  if (dirDotXY != R(0)) dist = ((ay * p.x) - (ax * p.y)) / NonZero(dirDotXY);
  ok     = ok && dist > R(-kConeTolerance);
  R hitx = p.x + dist * d.x;
  R hity = p.y + dist * d.y;
  if (toIn) {
    R hitz = p.z + dist * d.z;
    R r2   = (hitx * hitx) + (hity * hity);
    R ri   = RadiusAt(c, true, hitz) + R(kTolerance);
    R ro   = RadiusAt(c, false, hitz) - R(kTolerance);
    ok     = ok && vabs(hitz) <= R(c.tolIz) && (r2 >= ri * ri) && (r2 <= ro * ro);
    ok     = ok && ((hitx * ax) + (hity * ay)) > R(0);
  } else {
    ok = ok && ((hitx * ax) + (hity * ay)) >= R(0);
  }
  return ok;
}

template <class R>
VGMI_DEV bool ConicalHit(ConeData const &c, bool toIn, bool inner, Vec3<R> const &p, Vec3<R> const &d, R &distance)
{
  distance  = InfLen<R>();
  R rho2    = p.Perp2();
  R coneRad = RadiusAt(c, inner, p.z);
  R tol     = inner ? R(c.innerTol) : R(c.outerTol);
  bool on   = (rho2 >= (coneRad * coneRad - tol * coneRad)) && (rho2 <= (coneRad * coneRad + tol * coneRad)) &&
            (vabs(p.z) < (R(c.dz) + R(kConeTolerance)));
  Vec3<R> n  = Normal(c, inner, p);
  R dn       = d.Dot(n);
  if (on && vabs(dn) == R(0)) return false;
  if (on) {
    bool moving = toIn ? (dn <= R(0)) : (dn >= R(0));
    if (!c.hasPhi) {
      if (moving) return distance = R(0), true;
    } else if (InSector(c, p.x, p.y, true) && moving) {
      return distance = R(0), true;
    }
  }
  bool ok   = false;
  R pDotV2D = p.x * d.x + p.y * d.y;
  R a, b, cc;
  const R dz = R(c.dz);
  if (inner) {
    if (c.rmin1 == c.rmin2) {
      b  = pDotV2D;
      a  = d.Perp2();
      cc = p.Perp2() - R(c.rmin2) * R(c.rmin2);
    } else {
      R newPz = c.rmin2 > c.rmin1 ? (p.z + dz + R(c.innerApex)) * R(c.tanInnerApex)
                                  : (p.z - dz - R(c.innerApex)) * R(c.tanInnerApex);
      R dirT = d.z * R(c.tanInnerApex);
      a      = (d.x * d.x) + (d.y * d.y) - dirT * dirT;
      b      = pDotV2D - (newPz * dirT);
      cc     = p.Perp2() - (newPz * newPz);
    }
    R b2 = b * b, ac = a * cc;
    if (b2 < ac) return false;
    R d2 = b2 - ac, delta = vsqrt(d2);
    if (toIn) {
      distance = b >= R(0) ? (cc / NonZero(-b - delta)) : (-b + delta) / NonZero(a);
    } else {
      if (b == R(0) && delta == R(0)) return false;
      distance = b >= R(0) ? (-b - delta) / NonZero(a) : (cc / NonZero(-b + delta));
    }
    if (distance < R(0)) return false;
    R newZ = p.z + (d.z * distance);
    ok     = vabs(newZ) < dz;
  } else {
    if (c.origRmax1 == c.origRmax2) {
      a  = d.Perp2();
      b  = pDotV2D;
      cc = (p.Perp2() - R(c.origRmax2) * R(c.origRmax2));
    } else {
      R newPz = c.origRmax2 > c.origRmax1 ? (p.z + dz + R(c.outerApex)) * R(c.tanOuterApex)
                                          : (p.z - dz - R(c.outerApex)) * R(c.tanOuterApex);
      R dirT = d.z * R(c.tanOuterApex);
      a      = d.x * d.x + d.y * d.y - dirT * dirT;
      b      = (pDotV2D - (newPz * dirT));
      cc     = (p.Perp2() - (newPz * newPz));
    }
    R b2 = b * b, ac = a * cc, d2 = b2 - ac;
    if (d2 < R(0)) return false;
    R delta = vsqrt(d2);
    if (toIn) {
      if (b == R(0) && delta == R(0)) return false;
      distance = b > R(0) ? (-b - delta) / NonZero(a) : (cc / NonZero(-b + delta));
      R newZ   = p.z + (d.z * distance);
      ok       = vabs(newZ) < dz + R(kHalfTolerance);
    } else {
      distance = b < R(0) ? (-b + delta) / NonZero(a) : (cc / NonZero(-b - delta));
      ok       = distance > R(0);
    }
    if (distance < R(0)) return false;
  }
  if (c.hasPhi) {
    R hitx = R(0), hity = R(0);
    if (distance < InfLen<R>()) {
      hitx = p.x + distance * d.x;
      hity = p.y + distance * d.y;
    }
    ok = ok && InSector(c, hitx, hity, true) && (distance < InfLen<R>());
  }
  return ok;
}

template <class R>
VGMI_DEV void ContainsInside(ConeData const &c, Vec3<R> const &p, bool forInside, bool &cin, bool &cout)
{
  const R dz = R(c.dz);
  R absz     = vabs(p.z);
  cout       = absz > (forInside ? dz + R(kConeTolerance) : dz);
  cin        = forInside && absz < dz - R(kConeTolerance);
  if (cout) return;
  R r2       = p.x * p.x + p.y * p.y;
  R rmax     = c.origRmax1 == c.origRmax2 ? R(c.origRmax1) : R(c.outerSlope) * p.z + R(c.outerOffset);
  const R ot = R(c.outerTol);
  cout       = cout || (r2 > (forInside ? rmax * (rmax + R(2.0) * ot) : rmax * rmax));
  if (forInside) cin = cin && (r2 < vabs(rmax * (rmax - R(2.0) * ot)));
  if (cout) return;
  if (c.hasRmin) {
    R rmin     = R(c.innerSlope) * p.z + R(c.innerOffset);
    const R it = R(c.innerTol);
    cout       = cout || (r2 <= (forInside ? vabs(rmin * (rmin - R(2.0) * it)) : rmin * rmin));
    if (forInside) cin = cin && (r2 > rmin * (rmin + R(2.0) * it));
    if (cout) return;
  }
  if (c.hasPhi) {
    bool wi = false, wo = false;
    WedgeContainsInside(c, p, forInside, wi, wo);
    cout = cout || wo;
    if (forInside) cin = cin && wi;
  }
}

template <class R>
VGMI_DEV bool Contains(ConeData const &c, Vec3<R> const &p)
{
  bool ci, co;
  ContainsInside(c, p, false, ci, co);
  return !co;
}

template <class R>
VGMI_DEV int Inside(ConeData const &c, Vec3<R> const &p)
{
  bool ci, co;
  ContainsInside(c, p, true, ci, co);
  return ci ? kInside : (co ? kOutside : kSurface);
}

template <class R>
VGMI_DEV R DistanceToIn(ConeData const &c, Vec3<R> const &p, Vec3<R> const &d)
{
  const R dz = R(c.dz);
  R distz    = vabs(p.z) - dz;
  if ((distz > R(kConeTolerance) && (p.z * d.z) >= R(0)) || (vabs(distz) < R(kConeTolerance) && (p.z * d.z) > R(0)))
    return InfLen<R>();
  R outerRad = RadiusAt(c, false, p.z);
  R rsq      = p.Perp2();
  R ot       = R(c.outerTol);
  R outerIn  = vabs(outerRad * (outerRad - R(2.0) * ot));
  if (rsq > outerIn && d.Dot(Normal(c, false, p)) >= R(0)) return InfLen<R>();
  bool inside = distz < R(-kConeTolerance) && rsq < outerIn;
  if (c.hasRmin) {
    R innerRad = RadiusAt(c, true, p.z);
    inside     = inside && rsq > innerRad * (innerRad + R(2.0) * R(c.innerTol));
  }
  if (c.hasPhi) inside = inside && InSector(c, p.x, p.y, false);
  if (inside) return R(-1);
  R distance = InfLen<R>();
  distz /= NonZero(vabs(d.z));
  {
    R hitx      = p.x + distz * d.x;
    R hity      = p.y + distz * d.y;
    R r2        = (hitx * hitx) + (hity * hity);
    bool top    = (p.z >= R(c.tolIz)) && (r2 <= R(c.sqRmax2) + R(kTolerance));
    bool bottom = (p.z <= -R(c.tolIz)) && (r2 <= R(c.sqRmax1) + R(kTolerance));
    bool okz    = top || bottom;
    if (c.hasRmin) {
      top    = top && (r2 >= R(c.sqRmin2) - R(kTolerance));
      bottom = bottom && (r2 >= R(c.sqRmin1) - R(kTolerance));
      okz    = okz && (top || bottom);
    }
    if (c.hasPhi) okz = okz && InSector(c, hitx, hity, true);
    if (okz) return distz;
  }
  {
    R dOuter;
    if (ConicalHit(c, true, false, p, d, dOuter) && dOuter < distance) return dOuter;
  }
  if (c.hasRmin) {
    R dInner;
    if (ConicalHit(c, true, true, p, d, dInner) && dInner < distance) distance = dInner;
  }
  if (c.hasPhi) {
    R dp;
    if (PhiPlaneHit(c, R(c.a1x), R(c.a1y), R(c.n1x), R(c.n1y), p, d, true, dp) && dp < distance) distance = dp;
    if (PhiPlaneHit(c, R(c.a2x), R(c.a2y), R(c.n2x), R(c.n2y), p, d, true, dp) && dp < distance) distance = dp;
  }
  return distance;
}

template <class R>
VGMI_DEV R DistanceToOut(ConeData const &c, Vec3<R> const &p, Vec3<R> const &d)
{
  const R dz = R(c.dz);
  R distz    = vabs(p.z) - dz;
  R rsq      = p.Perp2();
  R outerRad = RadiusAt(c, false, p.z);
  if (distz > R(kConeTolerance) || rsq > outerRad * (outerRad + R(2.0) * R(c.outerTol))) return R(-1);
  if (c.hasRmin) {
    R innerRad = RadiusAt(c, true, p.z);
    if (rsq < vabs(innerRad * (innerRad - R(2.0) * R(c.innerTol)))) return R(-1);
  }
  const bool up = d.z > R(0), down = d.z < R(0);
  if ((up && p.z > R(0) && vabs(distz) < R(kConeTolerance)) || (down && p.z < R(0) && vabs(distz) < R(kConeTolerance)))
    return distz;
  R distance    = InfLen<R>();
  const R dzInv = R(1) / NonZero(d.z);
  if (up) distance = (dz - p.z) * dzInv;
  if (down) distance = (-dz - p.z) * dzInv;
  {
    R dOuter;
    if (ConicalHit(c, false, false, p, d, dOuter) && dOuter < distance) distance = dOuter;
  }
  if (c.hasRmin) {
    R dInner;
    if (ConicalHit(c, false, true, p, d, dInner) && dInner < distance) distance = dInner;
  }
  if (c.hasPhi) {
    bool onStart = OnWedgePlane(R(c.w1x), R(c.w1y), R(c.n1x), R(c.n1y), p);
    bool onEnd   = OnWedgePlane(R(c.w2x), R(c.w2y), R(c.n2x), R(c.n2y), p);
    R c1         = d.x * -R(c.n1x) + d.y * -R(c.n1y) + d.z * -R(0);
    R c2         = d.x * -R(c.n2x) + d.y * -R(c.n2y) + d.z * -R(0);
    if ((onStart && c1 > R(0)) || (onEnd && c2 > R(0))) return R(0);
    R dp;
    if (PhiPlaneHit(c, R(c.a1x), R(c.a1y), R(c.n1x), R(c.n1y), p, d, false, dp) && dp < distance) distance = dp;
    if (PhiPlaneHit(c, R(c.a2x), R(c.a2y), R(c.n2x), R(c.n2y), p, d, false, dp) && dp < distance) distance = dp;
  }
  if (distance < R(0) && vabs(distance) < R(kTolerance)) distance = R(0);
  return distance;
}

template <class R>
VGMI_DEV R SafeDistConical(ConeData const &c, bool inner, Vec3<R> const &p)
{
  R rho = vsqrt(p.Perp2());
  if (inner) {
    R pRMin = R(c.tanRMin) * p.z + (R(c.rmin1) + R(c.rmin2)) * R(0.5);
    return (rho - pRMin) * R(c.invSecRMin);
  }
  R pRMax = R(c.tanRMax) * p.z + (R(c.rmax1) + R(c.rmax2)) * R(0.5);
  return (pRMax - rho) * R(c.invSecRMax);
}

template <class R>
VGMI_DEV R SafetyToIn(ConeData const &c, Vec3<R> const &p)
{
  R distz     = vabs(p.z) - R(c.dz);
  R outerRad  = RadiusAt(c, false, p.z);
  R rsq       = p.Perp2();
  bool inside = distz < R(-kConeTolerance) && rsq < vabs(outerRad * (outerRad - R(2.0) * R(c.outerTol)));
  if (c.hasRmin) {
    R innerRad = RadiusAt(c, true, p.z);
    inside     = inside && rsq > innerRad * (innerRad + R(2.0) * R(c.innerTol));
  }
  if (c.hasPhi && inside) inside = InSector(c, p.x, p.y, false);
  if (inside) return R(-1);
  R safeZ  = vabs(p.z) - R(c.dz);
  R sOuter = -SafeDistConical(c, false, p);
  R sInner = -InfLen<R>();
  if (c.hasRmin) sInner = -SafeDistConical(c, true, p);
  R safety = vmax(safeZ, vmax(sOuter, sInner));
  if (c.hasPhi) {
    R d1  = p.x * R(c.n1x) + p.y * R(c.n1y);
    R d2  = p.x * R(c.n2x) + p.y * R(c.n2y);
    R sph = c.wedgeLtPi ? vmax(-1 * d1, -1 * d2) : vmin(-1 * d1, -1 * d2);
    safety = vmax(sph, safety);
  }
  if (vabs(safety) < R(kTolerance)) safety = R(0);
  return safety;
}

template <class R>
VGMI_DEV R SafetyToOut(ConeData const &c, Vec3<R> const &p)
{
  R distz      = vabs(p.z) - R(c.dz);
  R rsq        = p.Perp2();
  R outerRad   = RadiusAt(c, false, p.z);
  bool outside = distz > R(kConeTolerance) || rsq > outerRad * (outerRad + R(2.0) * R(c.outerTol));
  if (c.hasRmin) {
    R innerRad = RadiusAt(c, true, p.z);
    outside    = outside || rsq < vabs(innerRad * (innerRad - R(2.0) * R(c.innerTol)));
  }
  if (outside) return R(-1);
  R safeZ  = R(c.dz) - vabs(p.z);
  R sOuter = SafeDistConical(c, false, p);
  R sInner = InfLen<R>();
  if (c.hasRmin) sInner = SafeDistConical(c, true, p);
  R safety = vmin(safeZ, vmin(sOuter, sInner));
  if (c.hasPhi) {
    R d1  = p.x * R(c.n1x) + p.y * R(c.n1y);
    R d2  = p.x * R(c.n2x) + p.y * R(c.n2y);
    R sph = c.wedgeLtPi ? vmin(d1, d2) : vmax(d1, d2);
    safety = vmin(sph, safety);
  }
  if (vabs(safety) < R(kTolerance)) safety = R(0);
  return safety;
}

} // namespace cone
} // namespace vgmi
