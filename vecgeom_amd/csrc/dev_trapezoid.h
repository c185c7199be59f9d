// Trapezoid kernels (VecGeom/volumes/kernel/TrapezoidImplementation.h:98-400)
// over a four-plane shell (VecGeom/base/PlaneShell.h:105-270).
#pragma once
#include "dev_common.h"
#include "shape_data.h"

namespace vgmi {
namespace trap {

template <class R>
VGMI_DEV R plane_dist(TrapezoidData const &t, int i, Vec3<R> const &p)
{
  return R(t.a[i]) * p.x + R(t.b[i]) * p.y + R(t.c[i]) * p.z + R(t.d[i]);
}
template <class R>
VGMI_DEV R plane_proj(TrapezoidData const &t, int i, Vec3<R> const &d)
{
  return R(t.a[i]) * d.x + R(t.b[i]) * d.y + R(t.c[i]) * d.z;
}

template <class R>
VGMI_DEV void ContainsInside(TrapezoidData const &t, Vec3<R> const &p, bool forInside, bool &cin, bool &cout)
{
  const R dz = R(t.dz);
  cout       = vabs(p.z) > (forInside ? dz + R(kHalfTolerance) : dz);
  cin        = forInside && vabs(p.z) < dz - R(kHalfTolerance);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    R dist = plane_dist(t, i, p);
    cout   = cout || (dist > (forInside ? R(kHalfTolerance) : R(0)));
    if (forInside) cin = cin && (dist < R(-kHalfTolerance));
  }
}

template <class R>
VGMI_DEV bool Contains(TrapezoidData const &t, Vec3<R> const &p)
{
  bool ci, co;
  ContainsInside(t, p, false, ci, co);
  return !co;
}

template <class R>
VGMI_DEV int Inside(TrapezoidData const &t, Vec3<R> const &p)
{
  bool ci, co;
  ContainsInside(t, p, true, ci, co);
  return ci ? kInside : (co ? kOutside : kSurface);
}

template <class R>
VGMI_DEV R DistanceToIn(TrapezoidData const &t, Vec3<R> const &p, Vec3<R> const &d)
{
  const R dz = R(t.dz);
  R signZ    = Sign(d.z);
  R mx       = signZ * dz - p.z;
  if (signZ * mx < R(kHalfTolerance)) return InfLen<R>();
  R invdir = R(1) / NonZero(d.z);
  R smax   = mx * invdir;
  R smin   = -(signZ * dz + p.z) * invdir;
  R pdist[4], proj[4], vdist[4];
  bool done = false;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    pdist[i] = plane_dist(t, i, p);
    proj[i]  = plane_proj(t, i, d);
    vdist[i] = -pdist[i] / NonZero(proj[i]);
    done     = done || (pdist[i] > R(kHalfTolerance) && proj[i] >= R(0)) ||
           (pdist[i] > R(-kHalfTolerance) && proj[i] > R(0));
  }
  if (done) return InfLen<R>();
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    bool posPoint = pdist[i] > R(-kHalfTolerance);
    bool posDir   = proj[i] > R(0);
    bool fromIn   = !posPoint && posDir;
    bool fromOut  = posPoint && !posDir;
    if ((fromIn && vdist[i] < smin) || (fromOut && vdist[i] > smax)) return InfLen<R>();
    if (fromOut && vdist[i] > smin) smin = vdist[i];
    if (fromIn && vdist[i] < smax) smax = vdist[i];
  }
  return smin <= smax ? smin : InfLen<R>();
}

template <class R>
VGMI_DEV R DistanceToOut(TrapezoidData const &t, Vec3<R> const &p, Vec3<R> const &d)
{
  const R dz = R(t.dz);
  if (vabs(p.z) > dz + R(kHalfTolerance)) return R(-1);
  R dist = InfLen<R>();
  if (d.z != R(0)) dist = (Sign(d.z) * dz - p.z) / NonZero(d.z);
  R out = InfLen<R>();
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    R pd = plane_dist(t, i, p);
    R pr = plane_proj(t, i, d);
    R vd = -pd / NonZero(pr);
    if (pd > R(kHalfTolerance)) out = R(-1);
    if (pr > R(0) && vd < out) out = vd;
  }
  return out < dist ? out : dist;
}

template <class R>
VGMI_DEV R SafetyToIn(TrapezoidData const &t, Vec3<R> const &p)
{
  R safety = vabs(p.z) - R(t.dz);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    R dd = plane_dist(t, i, p);
    if (dd > safety) safety = dd;
  }
  return safety;
}

template <class R>
VGMI_DEV R SafetyToOut(TrapezoidData const &t, Vec3<R> const &p)
{
  R safety = R(t.dz) - vabs(p.z);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    R dd = -plane_dist(t, i, p);
    if (dd < safety) safety = dd;
  }
  return safety;
}

} // namespace trap
} // namespace vgmi
