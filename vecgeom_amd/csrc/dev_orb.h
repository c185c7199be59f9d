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

// Certified FP32 paths (see dev_common.h): the FP32 point is within ep = 2u |p|_1 of the FP64 one,
// p.d within 4u sum|p_i d_i| + ep, |p|^2 - r^2 within 4u (|p|^2 + r^2) + 2 |p| ep; a root is decided
// once the discriminant's sign and the sign of p.d are.
struct OrbF {
  float rad, pd, epd, c, ec;
};
VGMI_DEV OrbF orb_f32(OrbData const &o, Vec3<float> const &p, Vec3<float> const &d)
{
  const float r  = float(o.r);
  const float ep = 2.f * kU32 * (vabs(p.x) + vabs(p.y) + vabs(p.z));
  const float m2 = p.Mag2();
  OrbF f;
  f.rad = vsqrt(m2);
  f.pd  = p.Dot(d);
  f.epd = 4.f * kU32 * (vabs(p.x * d.x) + vabs(p.y * d.y) + vabs(p.z * d.z)) + ep;
  f.c   = m2 - r * r;
  f.ec  = 4.f * kU32 * (m2 + r * r) + 2.f * f.rad * ep + ep * ep;
  return f;
}

VGMI_DEV float DistanceToInF32(OrbData const &o, Vec3<float> const &p, Vec3<float> const &d)
{
  const OrbF f = orb_f32(o, p, d);
  const int in = cmpGT(0.f, f.c, f.ec); // strictly inside: the reference answers -1
  if (in == kYes) return -1.f;
  if (in == kUnsure) return Undecided();
  const float d2 = f.pd * f.pd - f.c;
  const float e2 = 2.f * vabs(f.pd) * f.epd + f.ec + 3.f * kU32 * (f.pd * f.pd + vabs(f.c));
  const int hit  = and3(not3(cmpGT(0.f, d2, e2)), not3(cmpGT(f.pd, 0.f, f.epd)));
  if (hit == kNo) return InfLen<float>();
  if (hit == kUnsure) return Undecided();
  const float sq = vsqrt(vabs(d2));
  const float t  = -f.pd - sq;
  const float et = f.epd + e2 / (sq + vsqrt(vmax(d2 - e2, 0.f))) + 2.f * kU32 * vabs(t);
  return et <= kMixDistErr ? t : Undecided();
}

VGMI_DEV float DistanceToOutF32(OrbData const &o, Vec3<float> const &p, Vec3<float> const &d)
{
  const OrbF f  = orb_f32(o, p, d);
  const int out = cmpGT(f.c, 0.f, f.ec);
  if (out == kYes) return -1.f;
  if (out == kUnsure) return Undecided();
  const float d2 = f.pd * f.pd - f.c; // >= |c| > 0 inside
  const float e2 = 2.f * vabs(f.pd) * f.epd + f.ec + 3.f * kU32 * (f.pd * f.pd + vabs(f.c));
  if (d2 <= e2) return Undecided();
  const float sq = vsqrt(d2);
  const float t  = -f.pd + sq;
  const float et = f.epd + e2 / (sq + vsqrt(d2 - e2)) + 2.f * kU32 * (vabs(f.pd) + sq);
  return et <= kMixDistErr ? t : Undecided();
}

template <class R>
VGMI_DEV R DistanceToIn(OrbData const &o, Vec3<R> const &p, Vec3<R> const &d)
{
  if constexpr (sizeof(R) == sizeof(float)) return DistanceToInF32(o, p, d);
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
  if constexpr (sizeof(R) == sizeof(float)) return DistanceToOutF32(o, p, d);
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
