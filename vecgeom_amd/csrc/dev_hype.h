// Hype kernels (UniversalHype): VecGeom/volumes/kernel/HypeImplementation.h:40-330 with the
// helpers of VecGeom/volumes/HypeUtilities.h:25-438 as the scalar backend evaluates them
// (every vecCore::MaskFull early return is a return; the quadratic helper is the scalar
// HypeHelpers<Precision> specialisation, :374-436). The FP32 pass of the mixed path solves
// the hyperbolic-surface quadratics in FP64 (root polishing) and everything else in FP32.
#pragma once
#include "dev_common.h"
#include "shape_data.h"

namespace vgmi {
namespace hype {

template <class R>
VGMI_DEV R rad2o(HypeData const &h, R z) { return R(h.rmax2) + R(h.tOut2) * z * z; }
template <class R>
VGMI_DEV R rad2i(HypeData const &h, R z) { return R(h.rmin2) + R(h.tIn2) * z * z; }

// GenericKernelForContainsAndInside (HypeImplementation.h:98-124)
template <class R>
VGMI_DEV void contains_inside(HypeData const &h, Vec3<R> const &p, bool forInside, bool &cin, bool &cout)
{
  const R r2    = p.Perp2();
  const R oRad2 = rad2o(h, p.z);
  R iRad2       = R(0);
  cin           = false;
  cout          = vabs(p.z) > R(h.dz) + R(h.zTol);
  if (cout) return;
  cout = r2 > oRad2 + R(h.outerRadTol);
  if (cout) return;
  if (h.inner) {
    iRad2 = rad2i(h, p.z);
    cout  = r2 < iRad2 - R(h.innerRadTol);
  }
  if (cout) return;
  if (forInside) {
    cin = vabs(p.z) < R(h.dz) - R(h.zTol) && r2 < oRad2 - R(h.outerRadTol);
    if (h.inner) cin = cin && r2 > iRad2 + R(h.innerRadTol);
  }
}

template <class R>
VGMI_DEV bool Contains(HypeData const &h, Vec3<R> const &p)
{
  bool ci, co;
  contains_inside(h, p, false, ci, co);
  return !co;
}

template <class R>
VGMI_DEV int Inside(HypeData const &h, Vec3<R> const &p)
{
  bool ci, co;
  contains_inside(h, p, true, ci, co);
  return co ? kOutside : (ci ? kInside : kSurface);
}

// HypeUtilities.h:28-67 IsCompletelyOutside / IsCompletelyInside
template <class R>
VGMI_DEV bool completely_outside(HypeData const &h, Vec3<R> const &p)
{
  const R r2 = p.Perp2();
  if (vabs(p.z) > R(h.dz) + R(h.zTol)) return true;
  if (r2 > rad2o(h, p.z) + R(h.outerRadTol)) return true;
  return h.inner && r2 < rad2i(h, p.z) - R(h.innerRadTol);
}
template <class R>
VGMI_DEV bool completely_inside(HypeData const &h, Vec3<R> const &p)
{
  const R r2 = p.Perp2();
  bool in    = vabs(p.z) < R(h.dz) - R(h.zTol) && r2 < rad2o(h, p.z) - R(h.outerRadTol);
  if (h.inner) in = in && r2 > rad2i(h, p.z) + R(h.innerRadTol);
  return in;
}

// HypeUtilities.h:80-106 (moving in) and :167-194 (moving out: the normal at the point)
template <class R>
VGMI_DEV bool moving_in(HypeData const &h, Vec3<R> const &p, Vec3<R> const &d, bool innerSurf)
{
  R pz = p.z, vz = d.z;
  if (pz < R(0)) vz = -vz, pz = -pz;
  const R s = p.x * d.x + p.y * d.y - pz * R(innerSurf ? h.tIn2 : h.tOut2) * vz;
  return innerSurf ? s > R(0) : s < R(0);
}
template <class R>
VGMI_DEV bool moving_out(HypeData const &h, Vec3<R> const &p, Vec3<R> const &d, bool innerSurf)
{
  if (innerSurf) return Vec3<R>(-p.x, -p.y, p.z * R(h.tIn2)).Dot(d) > R(0);
  return Vec3<R>(p.x, p.y, -p.z * R(h.tOut2)).Dot(d) > R(0);
}

template <class R>
VGMI_DEV bool z_surface(HypeData const &h, Vec3<R> const &p, R rho2)
{
  return (rho2 - R(h.endOuterRadius2)) < R(kTolerance) && (R(h.endInnerRadius2) - rho2) < R(kTolerance) &&
         vabs(vabs(p.z) - R(h.dz)) < R(kTolerance);
}

// HypeUtilities.h:108-147 IsPointOnSurfaceAndMovingInside
template <class R>
VGMI_DEV bool on_surface_moving_in(HypeData const &h, Vec3<R> const &p, Vec3<R> const &d)
{
  const R rho2 = p.Perp2();
  if (z_surface(h, p, rho2)) return p.z * d.z < R(0);
  if (vabs(rad2o(h, p.z) - rho2) < R(h.outerRadTol)) return moving_in(h, p, d, false);
  if (h.inner && vabs(rad2i(h, p.z) - rho2) < R(h.innerRadTol)) return moving_in(h, p, d, true);
  return false;
}

// HypeUtilities.h:230-274 IsPointOnSurfaceAndMovingOutside
template <class R>
VGMI_DEV bool on_surface_moving_out(HypeData const &h, Vec3<R> const &p, Vec3<R> const &d)
{
  const R rho2 = p.Perp2();
  if (z_surface(h, p, rho2) && p.z * d.z > R(0)) return true;
  if (vabs(rad2o(h, p.z) - rho2) < R(h.outerRadTol) && moving_out(h, p, d, false)) return true;
  return h.inner && vabs(rad2i(h, p.z) - rho2) < R(h.innerRadTol) && moving_out(h, p, d, true);
}

// HypeUtilities.h:196-228 IsPointOn{Outer,Inner}SurfaceAndMovingOutside
template <class R>
VGMI_DEV bool on_surface_moving_out_of(HypeData const &h, Vec3<R> const &p, Vec3<R> const &d, bool innerSurf)
{
  if (innerSurf && !h.inner) return false;
  const R rho2 = p.x * p.x + p.y * p.y;
  const R absZ = vabs(p.z);
  const R rad2 = innerSurf ? rad2i(h, p.z) : rad2o(h, p.z);
  return vabs(rad2 - rho2) < R(innerSurf ? h.innerRadTol : h.outerRadTol) && absZ >= R(0) && absZ < R(h.dz) &&
         moving_out(h, p, d, innerSurf);
}

// HypeUtilities.h:149-165 GetPointOfIntersectionWithZPlane
template <class R>
VGMI_DEV bool z_plane(HypeData const &h, Vec3<R> const &p, Vec3<R> const &d, bool toIn, R &zDist)
{
  zDist        = (Sign(toIn ? p.z : d.z) * R(h.dz) - p.z) / NonZero(d.z);
  const R r2   = (p + zDist * d).Perp2();
  if (!h.inner) return r2 < R(h.endOuterRadius2);
  return r2 < R(h.endOuterRadius2) && r2 > R(h.endInnerRadius2);
}

// HypeUtilities.h:374-436 HypeHelpers<Precision>: the quadratic of one hyperbolic surface,
// always solved in FP64; dist is left untouched when there is no real root.
template <class R>
VGMI_DEV bool hyperbolic(HypeData const &h, Vec3<R> const &pr, Vec3<R> const &dr, bool toIn, bool innerSurf, R &distR)
{
  const Vec3<double> p(pr.x, pr.y, pr.z), d(dr.x, dr.y, dr.z);
  const double t2 = innerSurf ? h.tIn2 : h.tOut2;
  const double a  = d.Perp2() - t2 * d.z * d.z;
  const double b  = (d.x * p.x + d.y * p.y - t2 * d.z * p.z);
  const double c  = p.Perp2() - t2 * p.z * p.z - (innerSurf ? h.rmin2 : h.rmax2);
  if (!(b * b - a * c > 0.)) return false;
  const double sq = vsqrt(b * b - a * c);
  double dist;
  if (innerSurf) {
    if (toIn)
      dist = b < 0. ? (-b + sq) / a : c / (-b - sq);
    else
      dist = b > 0. ? (-b - sq) / a : c / (-b + sq);
  } else {
    if (toIn)
      dist = b >= 0. ? (-b - sq) / a : c / (-b + sq);
    else
      dist = b < 0. ? (-b + sq) / a : c / (-b - sq);
  }
  if (dist < 0.) dist = kInfLength;
  const double newPtZ = p.z + dist * d.z;
  distR               = dist >= kInfLength ? InfLen<R>() : R(dist);
  return vabs(newPtZ) <= h.dz;
}

// HypeImplementation.h:126-199
template <class R>
VGMI_DEV R DistanceToIn(HypeData const &h, Vec3<R> const &p, Vec3<R> const &d)
{
  const R absZ = vabs(p.z);
  R zDist = InfLen<R>(), dist = InfLen<R>();
  const R r = p.Perp2();
  if (on_surface_moving_in(h, p, d)) return R(0);
  if (completely_inside(h, p)) return R(-1);
  const bool hitZ    = z_plane(h, p, d, true, zDist);
  const bool aboveIn = absZ > R(h.dz) && p.z * d.z < R(0);
  if (aboveIn && hitZ) return zDist;
  const R rp2 = (p + zDist * d).Perp2();
  const bool outerOut = aboveIn && !hitZ && rp2 >= R(h.endOuterRadius2);
  const bool outerIn  = (r > (rad2o(h, absZ) + R(kHalfTolerance))) && absZ >= R(0) && absZ <= R(h.dz);
  if ((outerOut || outerIn || on_surface_moving_out_of(h, p, d, false)) && hyperbolic(h, p, d, true, false, dist))
    return dist;
  if (h.inner) {
    const bool innerOut = aboveIn && !hitZ && rp2 <= R(h.endInnerRadius2);
    const bool innerIn  = (r < (rad2i(h, absZ) - R(kHalfTolerance))) && absZ >= R(0) && absZ <= R(h.dz);
    if ((innerOut || innerIn || on_surface_moving_out_of(h, p, d, true)) && hyperbolic(h, p, d, true, true, dist))
      return dist;
  }
  return InfLen<R>();
}

// HypeImplementation.h:201-236
template <class R>
VGMI_DEV R DistanceToOut(HypeData const &h, Vec3<R> const &p, Vec3<R> const &d)
{
  R zDist = InfLen<R>(), dist = InfLen<R>();
  if (on_surface_moving_out(h, p, d)) return R(0);
  if (completely_outside(h, p)) return R(-1);
  z_plane(h, p, d, false, zDist);
  if (zDist < R(0)) zDist = InfLen<R>();
  hyperbolic(h, p, d, false, false, dist);
  if (dist < R(0)) dist = InfLen<R>();
  R distance = vmin(zDist, dist);
  if (h.inner) {
    hyperbolic(h, p, d, false, true, dist);
    if (dist < R(0)) dist = InfLen<R>();
    distance = vmin(distance, dist);
  }
  return distance;
}

// HypeUtilities.h:276-313 ApproxDistOutside / ApproxDistInside
template <class R>
VGMI_DEV R approx_dist_outside(R pr, R pz, R r0, R tanPhi)
{
  const R r1  = vsqrt(r0 * r0 + tanPhi * tanPhi * pz * pz);
  const R z1  = pz;
  const R r2  = pr;
  const R z2  = vsqrt((pr * pr - r0 * r0) / (tanPhi * tanPhi));
  const R ddz = z2 - z1;
  const R dr  = r2 - r1;
  const R r3  = vsqrt(dr * dr + ddz * ddz);
  const R tiny = sizeof(R) == sizeof(float) ? R(FLT_MIN) : R(DBL_MIN);
  return r3 < tiny ? (r2 - r1) : (r2 - r1) * ddz / r3;
}
template <class R>
VGMI_DEV R approx_dist_inside(R pr, R pz, R r0, R tan2Phi)
{
  const R tiny = sizeof(R) == sizeof(float) ? R(FLT_MIN) : R(DBL_MIN);
  if (tan2Phi < tiny) return r0 - pr;
  const R rh  = vsqrt(r0 * r0 + pz * pz * tan2Phi);
  const R dr  = -rh;
  const R ddz = pz * tan2Phi;
  const R len = vsqrt(dr * dr + ddz * ddz);
  return vabs((pr - rh) * dr) / len;
}

// HypeImplementation.h:238-295
template <class R>
VGMI_DEV R SafetyToIn(HypeData const &h, Vec3<R> const &p)
{
  const R absZ = vabs(p.z);
  const R r2   = p.Perp2();
  const R r    = vsqrt(r2);
  bool cin, cout;
  contains_inside(h, p, true, cin, cout);
  if (!cin && !cout) return R(0);
  if (cin) return R(-1);
  const R sigz = absZ - R(h.dz);
  const R eo = R(h.endOuterRadius), ei = R(h.endInnerRadius);
  if (sigz > R(kHalfTolerance) && r < eo && r > ei) return sigz;
  if (sigz > R(kHalfTolerance) && r > eo) return vsqrt((r - eo) * (r - eo) + (sigz) * (sigz));
  if (sigz > R(kHalfTolerance) && r < ei) return vsqrt((r - ei) * (r - ei) + (sigz) * (sigz));
  if ((r2 > (rad2o(h, absZ) + R(kHalfTolerance))) && absZ > R(0) && absZ < R(h.dz))
    return approx_dist_outside(r, absZ, R(h.rmax), R(h.tOut));
  if ((r2 < (rad2i(h, absZ) - R(kHalfTolerance))) && absZ > R(0) && absZ < R(h.dz))
    return approx_dist_inside(r, absZ, R(h.rmin), R(h.tIn2));
  return R(0);
}

// HypeImplementation.h:297-330
template <class R>
VGMI_DEV R SafetyToOut(HypeData const &h, Vec3<R> const &p)
{
  const R r    = vsqrt(p.x * p.x + p.y * p.y);
  const R absZ = vabs(p.z);
  bool cin, cout;
  contains_inside(h, p, true, cin, cout);
  if (!cin && !cout) return R(0);
  if (cout) return R(-1);
  const R distZ = vabs(vabs(p.z) - R(h.dz));
  R distInner   = R(0);
  if (h.inner && h.stIn != 0.) distInner = approx_dist_outside(r, absZ, R(h.rmin), R(h.tIn));
  if (h.inner && h.stIn == 0.) distInner = r - R(h.rmin);
  if (!h.inner && h.stIn == 0.) distInner = InfLen<R>();
  const R distOuter = approx_dist_inside(r, absZ, R(h.rmax), R(h.tOut2));
  return vmin(vmin(distInner, distOuter), distZ);
}

} // namespace hype
} // namespace vgmi
