// CutTube kernels (VecGeom/volumes/kernel/CutTubeImplementation.h:84-322 with CutPlanes.h:123-173 and
// Plane.h:100-191): the unbounded tube's kernels combined with the two cut planes; templated on the
// arithmetic type.
#pragma once
#include "dev_common.h"
#include "dev_tube.h"
#include "shape_data.h"

namespace vgmi {
namespace ctub {

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

template <class R>
VGMI_DEV R plane(CutTubeData const &s, int i, Vec3<R> const &p)
{
  const double *n = i ? s.n1 : s.n0;
  return p.x * R(n[0]) + p.y * R(n[1]) + p.z * R(n[2]) + R(i ? s.d1 : s.d0);
}

template <class R>
VGMI_DEV R ndot(CutTubeData const &s, int i, Vec3<R> const &d)
{
  const double *n = i ? s.n1 : s.n0;
  return d.x * R(n[0]) + d.y * R(n[1]) + d.z * R(n[2]);
}

// CutPlanes::Inside (CutPlanes.h:133-144)
template <class R>
VGMI_DEV int cut_inside(CutTubeData const &s, Vec3<R> const &p)
{
  const R d0 = plane(s, 0, p), d1 = plane(s, 1, p);
  int in = (d0 < R(0.) && d1 < R(0.)) ? kInside : kOutside;
  if (vabs(d0) < R(kTolerance) || vabs(d1) < R(kTolerance)) in = kSurface;
  return in;
}

// CutPlanes::DistanceToOut (Plane.h:149-171 per plane, the smaller wins)
template <class R>
VGMI_DEV R cut_dist_out(CutTubeData const &s, Vec3<R> const &p, Vec3<R> const &d)
{
  R distance = InfLen<R>();
  for (int i = 0; i < 2; ++i) {
    const R ndd = NonZero(ndot(s, i, d));
    const R saf = plane(s, i, p);
    R di        = InfLen<R>();
    if (saf > R(kTolerance)) di = -InfLen<R>();
    if (ndd > R(0.) && saf < R(kTolerance)) di = -saf / ndd;
    distance = i == 0 ? di : Mn(distance, di);
  }
  return distance;
}

// :84-99
template <class R>
VGMI_DEV bool Contains(CutTubeData const &s, Vec3<R> const &p)
{
  if (!(plane(s, 0, p) < R(0.) && plane(s, 1, p) < R(0.))) return false;
  return tube::Contains(s.tube, p);
}

// :102-120
template <class R>
VGMI_DEV int Inside(CutTubeData const &s, Vec3<R> const &p)
{
  const int ic = cut_inside(s, p);
  if (ic == kOutside) return kOutside;
  int in = tube::Inside(s.tube, p);
  if (ic == kSurface && in != kOutside) in = ic;
  return in;
}

// :148-244 on the scalar backend
template <class R>
VGMI_DEV R DistanceToInKernel(CutTubeData const &s, Vec3<R> const &p, Vec3<R> const &d, R stepMax)
{
  R distance  = InfLen<R>();
  int ic      = cut_inside(s, p);
  int instart = tube::Inside(s.tube, p);
  if (ic == kOutside || (ic == kSurface && instart != kOutside)) instart = ic;
  if (instart == kInside) distance = R(-1.);
  R dpl[2];
  for (int i = 0; i < 2; ++i) {
    const R dn  = ndot(s, i, d);
    const R ndd = NonZero(dn);
    const R saf = plane(s, i, p);
    dpl[i]      = (ndd < R(0.) && saf > -R(kTolerance)) ? -saf / ndd : -InfLen<R>();
    if (dn > R(0.)) dpl[i] = -InfLen<R>();
  }
  R dplanes            = Mx(dpl[0], dpl[1]);
  const R safplanes    = Mx(plane(s, 0, p), plane(s, 1, p));
  const bool hitplanes = vabs(dplanes) < stepMax && safplanes > -R(kTolerance);
  if (ic != kInside && !hitplanes) return distance;
  Vec3<R> propagated = p;
  if (!hitplanes) dplanes = R(0.);
  if (hitplanes) {
    propagated = propagated + d * dplanes;
    ic         = cut_inside(s, propagated);
  }
  bool done = ic == kOutside || instart == kInside;
  if (done) return distance;
  instart           = tube::Inside(s.tube, propagated);
  const bool inside = hitplanes && instart != kOutside;
  if (inside) {
    distance = dplanes;
    return vabs(distance) < R(kTolerance) ? R(0.) : distance;
  }
  const R dexit = cut_dist_out(s, propagated, d);
  R dtube       = tube::DistanceToInKernel(s.tube, propagated, d);
  if (dtube < R(0.)) dtube = R(0.);
  if (dexit < dtube) dtube = InfLen<R>();
  if ((dtube + dplanes) < stepMax) distance = dtube + dplanes;
  return distance;
}

// :123-145: a point beyond 100 max(rmax, dz) is first moved along the direction to that radius
template <class R>
VGMI_DEV R DistanceToInStep(CutTubeData const &s, Vec3<R> const &pt, Vec3<R> const &d, R stepMax)
{
  Vec3<R> p      = pt;
  const R ptDist = vsqrt(p.Mag2());
  R move         = R(0.);
  const R far    = R(100.) * R(s.maxVal);
  if (ptDist > far) {
    move = ptDist - far;
    p    = p + d * move;
  }
  return DistanceToInKernel(s, p, d, stepMax) + move;
}

template <class R>
VGMI_DEV R DistanceToIn(CutTubeData const &s, Vec3<R> const &p, Vec3<R> const &d)
{
  return DistanceToInStep(s, p, d, InfLen<R>());
}

// :248-266
template <class R>
VGMI_DEV R DistanceToOut(CutTubeData const &s, Vec3<R> const &p, Vec3<R> const &d)
{
  R distance    = cut_dist_out(s, p, d);
  const R dtube = tube::DistanceToOut(s.tube, p, d);
  if (dtube < distance) distance = dtube;
  return distance;
}

// :269-282
template <class R>
VGMI_DEV R SafetyToIn(CutTubeData const &s, Vec3<R> const &p)
{
  R safety      = Mx(plane(s, 0, p), plane(s, 1, p));
  const R stube = tube::SafetyToIn(s.tube, p);
  if (stube > safety) safety = stube;
  return safety;
}

// :285-298
template <class R>
VGMI_DEV R SafetyToOut(CutTubeData const &s, Vec3<R> const &p)
{
  R safety      = Mn(-plane(s, 0, p), -plane(s, 1, p));
  const R stube = tube::SafetyToOut(s.tube, p);
  if (stube < safety) safety = stube;
  return safety;
}

} // namespace ctub
} // namespace vgmi
