// Orb kernels (VecGeom/volumes/kernel/OrbImplementation.h:65-220).
#pragma once
#include "dev_common.h"
#include "shape_data.h"

namespace vgmi {
namespace orb {

template <class R>
VGMI_DEV bool Contains(OrbData const &o, Vec3<R> const &p)
{
  R tolR = R(o.r) + R(kTolerance);
  return !(p.Mag2() >= tolR * tolR);
}

template <class R>
VGMI_DEV int Inside(OrbData const &o, Vec3<R> const &p)
{
  R rad2 = p.Mag2();
  R ti = R(o.r) - R(kTolerance), to = R(o.r) + R(kTolerance);
  if (rad2 <= ti * ti) return kInside;
  return rad2 >= to * to ? kOutside : kSurface;
}

template <class R>
VGMI_DEV R DistanceToIn(OrbData const &o, Vec3<R> const &p, Vec3<R> const &d)
{
  const R r = R(o.r);
  R rad     = vsqrt(p.Mag2());
  if (rad < r - R(kTolerance)) return R(-1);
  R pDotV = p.Dot(d);
  if (rad >= r - R(kTolerance) && rad <= r + R(kTolerance) && pDotV < R(0)) return R(0);
  R c  = p.Mag2() - r * r;
  R d2 = pDotV * pDotV - c;
  if (d2 >= R(0) && pDotV <= R(0)) return -pDotV - vsqrt(vabs(d2));
  return InfLen<R>();
}

template <class R>
VGMI_DEV R DistanceToOut(OrbData const &o, Vec3<R> const &p, Vec3<R> const &d)
{
  const R r = R(o.r);
  R rad     = vsqrt(p.Mag2());
  if (rad > r + R(kTolerance)) return R(-1);
  R pDotV = p.Dot(d);
  if (rad >= r - R(kTolerance) && rad <= r + R(kTolerance) && pDotV > R(0)) return R(0);
  R c  = p.Mag2() - r * r;
  R d2 = pDotV * pDotV - c;
  if (d2 >= R(0)) return -pDotV + vsqrt(vabs(d2));
  return InfLen<R>();
}

template <class R>
VGMI_DEV R SafetyToIn(OrbData const &o, Vec3<R> const &p)
{
  const R r = R(o.r);
  R rad     = vsqrt(p.Mag2());
  if (rad < r - R(kTolerance)) return R(-1);
  if (rad > r - R(kTolerance) && rad < r + R(kTolerance)) return R(0);
  return rad - r;
}

template <class R>
VGMI_DEV R SafetyToOut(OrbData const &o, Vec3<R> const &p)
{
  const R r = R(o.r);
  R rad     = vsqrt(p.Mag2());
  if (rad > r + R(kTolerance)) return R(-1);
  if (rad > r - R(kTolerance) && rad < r + R(kTolerance)) return R(0);
  return r - rad;
}

} // namespace orb
} // namespace vgmi
