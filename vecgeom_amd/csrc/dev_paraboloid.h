// Paraboloid kernels (VecGeom/volumes/kernel/ParaboloidImplementation.h:70-428). The solid is
// |z| <= dz and rho^2 <= k1 z + k2, i.e. z >= a rho^2 + b.
#pragma once
#include "dev_common.h"
#include "shape_data.h"

namespace vgmi {
namespace para {

// GenericKernelForContainsAndInside (ParaboloidImplementation.h:101-116)
template <class R>
VGMI_DEV void flags(ParaboloidData const &q, Vec3<R> const &p, bool forInside, bool &cin, bool &cout)
{
  const R diff = p.Perp2() - (R(q.k1) * p.z + R(q.k2));
  const R absZ = vabs(p.z);
  cout         = absZ > R(q.dz + kTolerance) || diff > R(kTolerance);
  cin          = !cout && forInside && absZ < R(q.dz - kTolerance) && diff < -R(kTolerance);
}

template <class R>
VGMI_DEV bool Contains(ParaboloidData const &q, Vec3<R> const &p)
{
  bool ci, co;
  flags(q, p, false, ci, co);
  return !co;
}

template <class R>
VGMI_DEV int Inside(ParaboloidData const &q, Vec3<R> const &p)
{
  bool ci, co;
  flags(q, p, true, ci, co);
  return co ? kOutside : (ci ? kInside : kSurface);
}

// IsOnZPlane (ParaboloidImplementation.h:118-131)
template <class R>
VGMI_DEV bool on_z_plane(ParaboloidData const &q, bool top, Vec3<R> const &p)
{
  const R rho2 = p.Perp2();
  if (top) return vabs(p.z - R(q.dz)) < R(kTolerance) && rho2 < R(q.rhi2) + R(kHalfTolerance);
  return vabs(p.z + R(q.dz)) < R(kTolerance) && rho2 < R(q.rlo2) + R(kHalfTolerance);
}

// ParaboloidImplementation.h:147-262: a far point approaching the solid is first moved along d to
// 2 Rsph from the origin (the offset is added back to the distance).
template <class R>
VGMI_DEV R DistanceToIn(ParaboloidData const &q, Vec3<R> const &point, Vec3<R> const &d)
{
  const R tol   = R(kTolerance);
  const R rsph  = R(1.5 * (q.dx > q.dz ? q.dx : q.dz));
  const R rfar2 = R(1024.) * rsph * rsph;
  R offset      = R(0.);
  if (point.Mag2() > rfar2 && d.Dot(point) < R(0.)) offset = vsqrt(point.Mag2()) - R(2.) * rsph;
  const Vec3<R> p = point + d * offset;
  const R absZ    = vabs(p.z);
  const R rho2    = p.Perp2();
  if (absZ > R(q.dz) && p.z * d.z >= R(0.)) return InfLen<R>();
  const R diff = rho2 - (R(q.k1) * p.z + R(q.k2));
  if (absZ < R(q.dz - kTolerance) && diff < -tol) return R(-1.);
  if ((on_z_plane(q, true, point) && d.z < R(0.)) || (on_z_plane(q, false, point) && d.z > R(0.))) return R(0.);
  const Vec3<R> normal(p.x, p.y, -R(q.k1) * R(0.5));
  if (diff > -tol && diff < tol && d.Dot(normal) < R(0.)) return R(0.);
  R distance = InfLen<R>();
  if (!(absZ < R(q.dz))) {
    const bool bottom = p.z < -R(q.dz) && d.z > R(0);
    const bool top    = p.z > R(q.dz) && d.z < R(0);
    R distZ           = InfLen<R>();
    if (top) distZ = (R(q.dz) - p.z) / NonZero(d.z);
    if (bottom) distZ = (-R(q.dz) - p.z) / NonZero(d.z);
    const R xHit    = p.x + distZ * d.x;
    const R yHit    = p.y + distZ * d.y;
    const R rhoHit2 = xHit * xHit + yHit * yHit;
    if (top && rhoHit2 <= R(q.rhi2)) distance = distZ + offset;
    if (top && rhoHit2 < R(q.rhi2)) return distance;
    if (bottom && rhoHit2 <= R(q.rlo2)) return distZ + offset;
  }
  if (!(p.z > R(q.dz) && rho2 < R(q.rhi2))) {
    const R qa = R(q.a) * d.Perp2();
    const R qb = R(0.5) * d.z - R(q.a) * (p.x * d.x + p.y * d.y);
    const R qc = R(q.b) + R(q.a) * p.Perp2() - p.z;
    const R d2 = qb * qb - qa * qc;
    if (d2 < R(0.)) return distance;
    const R dp   = qb <= R(0.) ? (qb - vsqrt(d2)) / NonZero(qa) : qc / NonZero(qb + vsqrt(d2));
    const R zHit = p.z + dp * d.z;
    if (vabs(zHit) <= R(q.dz) && dp > R(0.)) distance = dp + offset;
  }
  return distance;
}

// ParaboloidImplementation.h:264-330
template <class R>
VGMI_DEV R DistanceToOut(ParaboloidData const &q, Vec3<R> const &p, Vec3<R> const &d)
{
  const R tol = R(kTolerance);
  if (vabs(p.z) > R(q.dz + kTolerance)) return R(-1.);
  const R value = p.Perp2() - (R(q.k1) * p.z + R(q.k2));
  if (value > R(kHalfTolerance)) return R(-1.);
  if ((on_z_plane(q, true, p) && d.z > R(0.)) || (on_z_plane(q, false, p) && d.z < R(0.))) return R(0.);
  const Vec3<R> normal(p.x, p.y, -R(q.k1) * R(0.5));
  if (value > -tol && value < tol && d.Dot(normal) > R(0.)) return R(0.);
  const R dirZinv = R(1.) / NonZero(d.z);
  const R distZ   = d.z < R(0) ? -(R(q.dz) + p.z) * dirZinv : (R(q.dz) - p.z) * dirZinv;
  const R qa      = R(q.a) * d.Perp2();
  const R qb      = R(0.5) * d.z - R(q.a) * (p.x * d.x + p.y * d.y);
  const R qc      = R(q.b) + R(q.a) * p.Perp2() - p.z;
  const R d2      = qb * qb - qa * qc;
  R dp            = InfLen<R>();
  if (d2 >= R(0.) && qb > R(0.)) dp = (qb + vsqrt(d2)) * (R(1.) / NonZero(qa));
  if (d2 >= R(0.) && qb <= R(0.)) dp = qc / NonZero(qb - vsqrt(d2));
  return vmin(dp, distZ);
}

// ParaboloidImplementation.h:332-376
template <class R>
VGMI_DEV R SafetyToIn(ParaboloidData const &q, Vec3<R> const &p)
{
  const R tol   = R(kTolerance);
  const R absZ  = vabs(p.z);
  const R safeZ = absZ - R(q.dz);
  const R rho2  = p.Perp2();
  const R value = R(q.a) * rho2 + R(q.b) - p.z;
  if (absZ < R(q.dz - kTolerance) && value < -R(kHalfTolerance)) return R(-1.);
  if (vabs(vabs(p.z) - R(q.dz)) < tol &&
      (rho2 < R(q.rhi2) + R(kHalfTolerance) || rho2 < R(q.rlo2) + R(kHalfTolerance)))
    return R(0.);
  if (value > -tol && value < tol) return R(0.);
  const R r0sq = (p.z - R(q.b)) * R(q.invA);
  if (r0sq < R(0)) return safeZ;
  const R dr = vsqrt(p.x * p.x + p.y * p.y) - vsqrt(r0sq);
  if (dr < R(1.E-8)) return safeZ;
  const R talf = R(-2.) * R(q.a) * vsqrt(r0sq);
  const R salf = talf / vsqrt(R(1.) + talf * talf);
  return vmax(vabs(dr * salf), safeZ);
}

// ParaboloidImplementation.h:378-428
template <class R>
VGMI_DEV R SafetyToOut(ParaboloidData const &q, Vec3<R> const &p)
{
  const R tol  = R(kTolerance);
  const R absZ = vabs(p.z);
  const R safZ = R(q.dz) - absZ;
  if (absZ > R(q.dz + kTolerance)) return R(-1.);
  const R rho2  = p.Perp2();
  const R value = R(q.a) * rho2 + R(q.b) - p.z;
  if (value > R(kHalfTolerance)) return R(-1.);
  if (vabs(vabs(p.z) - R(q.dz)) < tol &&
      (rho2 < R(q.rhi2) + R(kHalfTolerance) || rho2 < R(q.rlo2) + R(kHalfTolerance)))
    return R(0.);
  if (value > -tol && value < tol) return R(0.);
  const R r0sq = (p.z - R(q.b)) * R(q.invA);
  if (r0sq < R(0)) return R(0.);
  const R ro2 = p.x * p.x + p.y * p.y;
  const R z0  = R(q.a) * ro2 + R(q.b);
  const R dr  = vsqrt(ro2) - vsqrt(r0sq);
  if (dr > R(-1.E-8)) return R(0.);
  const R dzp  = vabs(p.z - z0);
  const R safR = -dr * dzp / vsqrt(dr * dr + dzp * dzp);
  return vmin(safR, safZ);
}

} // namespace para
} // namespace vgmi
