// EllipticalTube (VecGeom/volumes/kernel/EllipticalTubeImplementation.h:61-231) and Ellipsoid
// (VecGeom/volumes/kernel/EllipsoidImplementation.h:61-243) kernels: each shape is scaled to a circle or
// sphere of radius r; templated on the arithmetic type.
#pragma once
#include "dev_common.h"
#include "shape_data.h"

namespace vgmi {
namespace eltu {

template <class R>
VGMI_DEV R Mx(R a, R b)
{
  return a > b ? a : b;
}
template <class R>
VGMI_DEV R Mn(R a, R b)
{
  return a < b ? a : b;
}

// GenericKernelForContainsAndInside (:85-107)
template <class R>
VGMI_DEV R safety(EllipticalTubeData const &s, Vec3<R> const &p)
{
  const R x = p.x * R(s.sx), y = p.y * R(s.sy);
  return Mx(R(s.q1) * (x * x + y * y) - R(s.q2), vabs(p.z) - R(s.dz));
}

template <class R>
VGMI_DEV bool Contains(EllipticalTubeData const &s, Vec3<R> const &p)
{
  return !(safety(s, p) > R(kHalfTolerance));
}

template <class R>
VGMI_DEV int Inside(EllipticalTubeData const &s, Vec3<R> const &p)
{
  const R v = safety(s, p);
  return v <= -R(kHalfTolerance) ? kInside : (v > R(kHalfTolerance) ? kOutside : kSurface);
}

// :109-160 (far points are first moved to 2 rsph from the origin along the ray)
template <class R>
VGMI_DEV R DistanceToIn(EllipticalTubeData const &s, Vec3<R> const &point, Vec3<R> const &d)
{
  R distance = InfLen<R>(), offset = R(0.);
  Vec3<R> pc = point;
  const R rsph = R(s.rsph);
  if (pc.Mag2() > R(1024.) * rsph * rsph && d.Dot(point) < R(0.)) {
    offset = vsqrt(pc.Mag2()) - R(2.) * rsph;
    pc     = pc + offset * d;
  }
  const R px = pc.x * R(s.sx), py = pc.y * R(s.sy), pz = pc.z;
  const R vx = d.x * R(s.sx), vy = d.y * R(s.sy), vz = d.z;
  const R invz  = R(-1.) / NonZero(vz);
  const R dz    = vcopysign(R(s.dz), invz);
  const R tzmin = (pz - dz) * invz;
  const R tzmax = (pz + dz) * invz;
  const R rr = px * px + py * py;
  const R A  = vx * vx + vy * vy;
  const R B  = px * vx + py * vy;
  const R C  = rr - R(s.r) * R(s.r);
  const R D  = B * B - A * C;
  const R distZ          = vabs(pz) - R(s.dz);
  const R distR          = R(s.q1) * rr - R(s.q2);
  const bool parallelToZ = A < R(DBL_EPSILON) || vabs(vz) >= R(1.);
  const bool leaving     = (distZ >= -R(kHalfTolerance) && pz * vz >= R(0.)) ||
                       (distR >= -R(kHalfTolerance) && (B >= R(0.) || parallelToZ));
  if (!leaving && parallelToZ) distance = tzmin + offset;
  if (leaving || parallelToZ || D <= A * A * R(s.scratch)) return distance;
  const R tmp  = -B - vcopysign(vsqrt(D), B);
  const R t1   = tmp / A, t2 = C / tmp;
  const R tin  = Mx(tzmin, Mn(t1, t2));
  const R tout = Mn(tzmax, Mx(t1, t2));
  return (tout - tin) >= R(kHalfTolerance) ? tin + offset : distance;
}

// :162-203 (kMaximum, the z distance of a ray parallel to the z planes, is DBL_MAX = kInfLength)
template <class R>
VGMI_DEV R DistanceToOut(EllipticalTubeData const &s, Vec3<R> const &p, Vec3<R> const &d)
{
  const R px = p.x * R(s.sx), py = p.y * R(s.sy), pz = p.z;
  const R vx = d.x * R(s.sx), vy = d.y * R(s.sy), vz = d.z;
  const R rr = px * px + py * py;
  if (Mx(R(s.q1) * rr - R(s.q2), vabs(pz) - R(s.dz)) > R(kHalfTolerance)) return R(-1.);
  const R tzmax = vz != R(0.) ? (vcopysign(R(s.dz), vz) - pz) / vz : InfLen<R>();
  const R A = vx * vx + vy * vy;
  const R B = px * vx + py * vy;
  const R C = rr - R(s.r) * R(s.r);
  const R D = B * B - A * C;
  if (A < R(DBL_EPSILON) || vabs(vz) >= R(1.)) return tzmax;
  if (D <= R(0.)) return R(0.);
  return B >= R(0.) ? Mn(tzmax, -C / (vsqrt(D) + B)) : Mn(tzmax, (vsqrt(D) - B) / A);
}

// :205-231
template <class R>
VGMI_DEV R SafetyToIn(EllipticalTubeData const &s, Vec3<R> const &p)
{
  const R x = p.x * R(s.sx), y = p.y * R(s.sy);
  const R v = Mx(vsqrt(x * x + y * y) - R(s.r), vabs(p.z) - R(s.dz));
  return vabs(v) <= R(kHalfTolerance) ? R(0.) : v;
}

template <class R>
VGMI_DEV R SafetyToOut(EllipticalTubeData const &s, Vec3<R> const &p)
{
  const R x = p.x * R(s.sx), y = p.y * R(s.sy);
  const R v = Mn(R(s.r) - vsqrt(x * x + y * y), R(s.dz) - vabs(p.z));
  return vabs(v) <= R(kHalfTolerance) ? R(0.) : v;
}

} // namespace eltu

namespace ells {

using eltu::Mn;
using eltu::Mx;

// GenericKernelForContainsAndInside (:85-108)
template <class R>
VGMI_DEV R safety(EllipsoidData const &s, Vec3<R> const &p)
{
  const R x = p.x * R(s.sx), y = p.y * R(s.sy), z = p.z * R(s.sz);
  const R distZ = vabs(z - R(s.scZMidCut)) - R(s.scZDimCut);
  const R distR = R(s.q1) * (x * x + y * y + z * z) - R(s.q2);
  return Mx(distZ, distR);
}

template <class R>
VGMI_DEV bool Contains(EllipsoidData const &s, Vec3<R> const &p)
{
  return !(safety(s, p) > R(kHalfTolerance));
}

template <class R>
VGMI_DEV int Inside(EllipsoidData const &s, Vec3<R> const &p)
{
  const R v = safety(s, p);
  return v <= -R(kHalfTolerance) ? kInside : (v > R(kHalfTolerance) ? kOutside : kSurface);
}

// :110-161
template <class R>
VGMI_DEV R DistanceToIn(EllipsoidData const &s, Vec3<R> const &point, Vec3<R> const &d)
{
  R offset   = R(0.);
  Vec3<R> pc = point;
  const R rsph = R(s.rsph);
  if (pc.Mag2() > R(1024.) * rsph * rsph && d.Dot(point) < R(0.)) {
    offset = vsqrt(pc.Mag2()) - R(2.) * rsph;
    pc     = pc + offset * d;
  }
  const R px = pc.x * R(s.sx), py = pc.y * R(s.sy), pz = pc.z * R(s.sz);
  const R vx = d.x * R(s.sx), vy = d.y * R(s.sy), vz = d.z * R(s.sz);
  const R pzcut = pz - R(s.scZMidCut), dzcut = R(s.scZDimCut);
  const R distZ = vabs(pzcut) - dzcut;
  const R rr = px * px + py * py + pz * pz;
  const R vv = vx * vx + vy * vy + vz * vz;
  const R pv = px * vx + py * vy + pz * vz;
  const R distR      = R(s.q1) * rr - R(s.q2);
  const bool leaving = (distZ >= -R(kHalfTolerance) && pzcut * vz >= R(0.)) ||
                       (distR >= -R(kHalfTolerance) && pv >= R(0.));
  const R invz  = R(-1.) / NonZero(vz);
  const R dz    = vcopysign(dzcut, invz);
  const R tzmin = (pzcut - dz) * invz;
  const R tzmax = (pzcut + dz) * invz;
  const R C     = rr - R(s.r) * R(s.r);
  const R D     = pv * pv - vv * C;
  const R sqrtD = vsqrt(vabs(D));
  const R tmin  = Mx(tzmin, (-pv - sqrtD) / vv);
  const R tmax  = Mn(tzmax, (-pv + sqrtD) / vv);
  const R EPS   = R(2.) * rr * vv * R(DBL_EPSILON);
  return leaving || D <= EPS || (tmax - tmin) <= R(kHalfTolerance) ? InfLen<R>() : tmin + offset;
}

// :163-201
template <class R>
VGMI_DEV R DistanceToOut(EllipsoidData const &s, Vec3<R> const &p, Vec3<R> const &d)
{
  const R px = p.x * R(s.sx), py = p.y * R(s.sy), pz = p.z * R(s.sz);
  const R vx = d.x * R(s.sx), vy = d.y * R(s.sy), vz = d.z * R(s.sz);
  const R pzcut = pz - R(s.scZMidCut), dzcut = R(s.scZDimCut);
  const R distZ = vabs(pzcut) - dzcut;
  const R rr = px * px + py * py + pz * pz;
  const R vv = vx * vx + vy * vy + vz * vz;
  const R pv = px * vx + py * vy + pz * vz;
  if (Mx(R(s.q1) * rr - R(s.q2), distZ) > R(kHalfTolerance)) return R(-1.);
  const R tzmax = vz != R(0.) ? (vcopysign(dzcut, vz) - pzcut) / vz : InfLen<R>();
  const R B     = pv / vv;
  const R C     = (rr - R(s.r) * R(s.r)) / vv;
  const R D     = B * B - C;
  const R trmax = -B + vsqrt(vabs(D));
  const R EPS   = R(2.) * rr * vv * R(DBL_EPSILON);
  return D <= EPS ? R(0.) : Mn(tzmax, trmax);
}

// :203-243
template <class R>
VGMI_DEV R SafetyToIn(EllipsoidData const &s, Vec3<R> const &p)
{
  const R x = p.x * R(s.sx), y = p.y * R(s.sy), z = p.z * R(s.sz);
  const R r = vsqrt(x * x + y * y + z * z);
  R v       = Mx(vabs(z - R(s.scZMidCut)) - R(s.scZDimCut), r - R(s.r));
  if (vabs(v) <= R(kHalfTolerance)) v = R(0.);
  const R distZ  = Mx(p.z - R(s.zTopCut), R(s.zBottomCut) - p.z);
  const R distXY = Mx(vabs(p.x) - R(s.xmax), vabs(p.y) - R(s.ymax));
  return v > R(0.) ? Mx(Mx(v, distZ), distXY) : v;
}

template <class R>
VGMI_DEV R SafetyToOut(EllipsoidData const &s, Vec3<R> const &p)
{
  const R x = p.x * R(s.sx), y = p.y * R(s.sy), z = p.z * R(s.sz);
  const R safeR = R(s.r) - vsqrt(x * x + y * y + z * z);
  R v           = Mn(R(s.scZDimCut) - vabs(z - R(s.scZMidCut)), safeR);
  if (vabs(v) <= R(kHalfTolerance)) v = R(0.);
  return v > R(0.) ? Mn(safeR, Mn(R(s.zTopCut) - p.z, p.z - R(s.zBottomCut))) : v;
}

} // namespace ells
} // namespace vgmi
