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

// EllipticalCone (VecGeom/volumes/kernel/EllipticalConeImplementation.h:63-304): the cone
// sqrt((x/dx)^2 + (y/dy)^2) + z = h cut at |z| = zCut; distances solve the scaled quadric.
namespace elco {

template <class R>
VGMI_DEV R safety(EllipticalConeData const &s, Vec3<R> const &p)
{
  const R px = p.x * R(s.invDx), py = p.y * R(s.invDy), pz = p.z;
  const R hp = vsqrt(px * px + py * py) + pz;
  return eltu::Mx((hp - R(s.dz)) * R(s.cosAxisMin), vabs(pz) - R(s.zCut));
}

template <class R>
VGMI_DEV bool Contains(EllipticalConeData const &s, Vec3<R> const &p)
{
  return !(safety(s, p) > R(kHalfTolerance));
}

template <class R>
VGMI_DEV int Inside(EllipticalConeData const &s, Vec3<R> const &p)
{
  const R v = safety(s, p);
  return v <= -R(kHalfTolerance) ? kInside : (v > R(kHalfTolerance) ? kOutside : kSurface);
}

// :118-205 (far points are first moved to 2 rsph from the origin along the ray)
template <class R>
VGMI_DEV R DistanceToIn(EllipticalConeData const &s, Vec3<R> const &point, Vec3<R> const &d)
{
  const R kTwoEps = R(2.) * R(DBL_EPSILON); // 2 kEpsilon
  R distance = InfLen<R>(), offset = R(0.);
  Vec3<R> p  = point;
  const R rsph = R(s.rsph);
  if (p.Mag2() > R(1024.) * rsph * rsph && d.Dot(p) < R(0.)) offset = vsqrt(p.Mag2()) - R(2.) * rsph;
  p          = p + d * offset;
  const R px = p.x * R(s.invDx), py = p.y * R(s.invDy), pz = p.z, pz0 = p.z - R(s.dz);
  const R vx = d.x * R(s.invDx), vy = d.y * R(s.invDy);
  R vz       = d.z;
  const R Ar = vx * vx + vy * vy, Br = px * vx + py * vy, Cr = px * px + py * py;
  const R vzvz     = vz * vz;
  const bool tinyA = vabs(Ar - vzvz) < kTwoEps * vzvz;
  if (tinyA) vz = vz + vabs(vz) * kTwoEps;
  const R Az = vz * vz, Bz = pz0 * vz, Cz = pz0 * pz0;
  const R A = Ar - Az, B = Br - Bz, B0 = Br - pz0 * d.z, C = Cr - Cz;
  R D         = B * B - A * C;
  const R sfz = vabs(pz) - R(s.zCut);
  R nz        = vsqrt(Cr);
  const R sfr = (nz + pz0) * R(s.cosAxisMin);
  if (vabs(p.x) + vabs(p.y) < R(0.1) * R(kHalfTolerance)) nz = R(1.);
  const R pzA = pz0 + R(s.dApex);
  bool done   = (sfz >= -R(kHalfTolerance) && pz * vz >= R(0.)) || (sfr >= -R(kHalfTolerance) && Br + nz * vz >= R(0.)) ||
              (pz0 * R(s.cosAxisMin) > -R(kHalfTolerance) && (Cr - pzA * pzA) <= R(0.) && A >= R(0.));
  if (sfr <= R(0.) && D < R(0.)) D = R(0.);
  done = done || (D < R(0.)) || ((D < kTwoEps * B * B) && (A >= R(0.)));
  const R invz = R(-1.) / NonZero(vz);
  const R ddz  = vcopysign(R(s.zCut), invz);
  const R tzin = (pz - ddz) * invz, tzout = (pz + ddz) * invz;
  R tmp = R(0.), t1 = R(0.), t2 = R(0.);
  if (!done) tmp = -B - vcopysign(vsqrt(D), B);
  if (!done) t1 = tmp / A;
  if (!done && tmp != R(0.)) t2 = C / tmp;
  if (!done && tinyA && B != R(0.)) t2 = -C / (R(2.) * B0);
  const R tmin = eltu::Mn(t1, t2), tmax = eltu::Mx(t1, t2);
  R trin = tmin, trout = tmax;
  done = done || (A >= R(0.) && pz0 + vz * tmin >= R(0.));
  if (!done && A < R(0.)) trin = -InfLen<R>(), trout = InfLen<R>();
  if (!done && A < R(0.) && vz < R(0.)) trin = tmax;
  if (!done && A < R(0.) && vz > R(0.)) trout = tmin;
  const R tin = eltu::Mx(tzin, trin), tout = eltu::Mn(tzout, trout);
  if (!done && (tout - tin) >= R(kHalfTolerance)) distance = tin + offset;
  return distance;
}

// :208-270
template <class R>
VGMI_DEV R DistanceToOut(EllipticalConeData const &s, Vec3<R> const &p, Vec3<R> const &d)
{
  const R kTwoEps = R(2.) * R(DBL_EPSILON); // 2 kEpsilon
  R distance      = R(0.);
  const R px = p.x * R(s.invDx), py = p.y * R(s.invDy), pz = p.z, pz0 = pz - R(s.dz);
  const R hp         = vsqrt(px * px + py * py) + pz;
  const R sfr        = (hp - R(s.dz)) * R(s.cosAxisMin);
  const R sfz        = vabs(pz) - R(s.zCut);
  const bool outside = eltu::Mx(sfr, sfz) > R(kHalfTolerance);
  if (outside) distance = R(-1.);
  bool done  = outside;
  const R vx = d.x * R(s.invDx), vy = d.y * R(s.invDy);
  R vz       = d.z;
  const R Ar = vx * vx + vy * vy, Br = px * vx + py * vy, Cr = px * px + py * py;
  const bool tinyA = vabs(Ar - vz * vz) < kTwoEps * vz * vz;
  if (tinyA) vz = vz + vabs(vz) * kTwoEps;
  const R Az = vz * vz, Bz = pz0 * vz, Cz = pz0 * pz0;
  const R A = Ar - Az, B = Br - Bz, B0 = Br - pz0 * d.z, C = Cr - Cz;
  R D = B * B - A * C;
  if (sfr <= R(0.) && D < R(0.)) D = R(0.);
  done    = done || (D < R(0.)) || (D < kTwoEps * B * B && A >= R(0.));
  R tzout = sizeof(R) == sizeof(float) ? R(FLT_MAX) : R(DBL_MAX); // kMaximum
  if (vz != R(0.)) tzout = (vcopysign(R(s.zCut), vz) - pz) / d.z;
  R tmp = R(0.), t1 = R(0.), t2 = R(0.);
  if (!done) tmp = -B - vcopysign(vsqrt(D), B);
  if (!done) t1 = tmp / A;
  if (!done && tmp != R(0.)) t2 = C / tmp;
  if (!done && tinyA && B0 != R(0.)) t2 = -C / (R(2.) * B0);
  const R tmin = eltu::Mn(t1, t2), tmax = eltu::Mx(t1, t2);
  R trout = tmax;
  done    = done || ((A >= R(0.) && pz0 + vz * tmax >= R(0.)) || (pz0 >= R(0.) && vz >= R(0.)));
  if (!done && A < R(0.)) trout = InfLen<R>();
  if (!done && A < R(0.) && vz > R(0.)) trout = tmin;
  if (!done) distance = eltu::Mn(tzout, trout);
  return distance;
}

// :273-287
template <class R>
VGMI_DEV R SafetyToIn(EllipticalConeData const &s, Vec3<R> const &p)
{
  const R v = safety(s, p);
  return vabs(v) <= R(kHalfTolerance) ? R(0.) : v;
}

// :290-304
template <class R>
VGMI_DEV R SafetyToOut(EllipticalConeData const &s, Vec3<R> const &p)
{
  const R px = p.x * R(s.invDx), py = p.y * R(s.invDy), pz = p.z;
  const R hp = vsqrt(px * px + py * py) + pz;
  const R v  = eltu::Mn((R(s.dz) - hp) * R(s.cosAxisMin), R(s.zCut) - vabs(pz));
  return vabs(v) <= R(kHalfTolerance) ? R(0.) : v;
}

} // namespace elco
} // namespace vgmi
