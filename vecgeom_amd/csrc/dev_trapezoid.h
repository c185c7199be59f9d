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

// Certified FP32 paths (see dev_common.h). A plane distance n.p + D of FP32 copies is within
// 4u (|a x| + |b y| + |c z| + |D|) of its FP64 value, a projection n.d within 4u (|a dx| + |b dy| +
// |c dz|), and their quotient -pd / pr within (epd + |v| epr) / |pr| + u |v|. Sign tests whose
// operand lies within its bound, and interval comparisons that overlap, answer NaN.
struct TrapPlaneF {
  float pd, epd, pr, epr;
};
VGMI_DEV TrapPlaneF trap_plane_f32(TrapezoidData const &t, int i, Vec3<float> const &p, Vec3<float> const &d)
{
  const float a = float(t.a[i]), b = float(t.b[i]), c = float(t.c[i]), dd = float(t.d[i]);
  TrapPlaneF r;
  r.pd  = a * p.x + b * p.y + c * p.z + dd;
  r.epd = 4.f * kU32 * (vabs(a * p.x) + vabs(b * p.y) + vabs(c * p.z) + vabs(dd));
  r.pr  = a * d.x + b * d.y + c * d.z;
  r.epr = 4.f * kU32 * (vabs(a * d.x) + vabs(b * d.y) + vabs(c * d.z));
  return r;
}

VGMI_DEV FInt trap_vdist_f32(TrapPlaneF const &q)
{
  const float v = -q.pd / NonZero(q.pr);
  return fint(v, (q.epd + vabs(v) * q.epr) / vabs(q.pr) + 2.f * kU32 * vabs(v));
}

VGMI_DEV float DistanceToInF32(TrapezoidData const &t, Vec3<float> const &p, Vec3<float> const &d)
{
  const float dz = float(t.dz), ht = float(kHalfTolerance);
  const float signZ = Sign(d.z);
  const float mx    = signZ * dz - p.z;
  const float ez    = 3.f * kU32 * (dz + vabs(p.z));
  if (signZ * mx + ez < ht) return InfLen<float>();
  if (signZ * mx - ez < ht) return Undecided();
  const float invdir = 1.f / NonZero(d.z);
  const float a0 = mx * invdir, a1 = -(signZ * dz + p.z) * invdir;
  FInt smax = fint(a0, ez * vabs(invdir) + 2.f * kU32 * vabs(a0));
  FInt smin = fint(a1, ez * vabs(invdir) + 2.f * kU32 * vabs(a1));
  float vsmin = a1;
  TrapPlaneF q[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    q[i] = trap_plane_f32(t, i, p, d);
    // done = (pd > ht && pr >= 0) || (pd > -ht && pr > 0): the sign tests must be decided
    if (vabs(q[i].pd) <= q[i].epd + ht || vabs(q[i].pr) <= q[i].epr) return Undecided();
    if (q[i].pd > 0.f && q[i].pr > 0.f) return InfLen<float>();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const bool posPoint = q[i].pd > 0.f, posDir = q[i].pr > 0.f;
    const bool fromIn = !posPoint && posDir, fromOut = posPoint && !posDir;
    if (!fromIn && !fromOut) continue;
    const FInt v = trap_vdist_f32(q[i]);
    if (fromIn) {
      if (v.hi < smin.lo) return InfLen<float>();
      if (v.lo < smin.hi) return Undecided();
      smax = fmini(smax, v);
    } else {
      if (v.lo > smax.hi) return InfLen<float>();
      if (v.hi > smax.lo) return Undecided();
      vsmin = vmax(vsmin, -q[i].pd / NonZero(q[i].pr)); // the estimate stays inside the interval
      smin  = fmaxi(smin, v);
    }
  }
  if (smin.hi <= smax.lo) return smin.hi - smin.lo <= kMixDistErr ? vsmin : Undecided();
  if (smin.lo > smax.hi) return InfLen<float>();
  return Undecided();
}

VGMI_DEV float DistanceToOutF32(TrapezoidData const &t, Vec3<float> const &p, Vec3<float> const &d)
{
  // wrong side: beyond the z planes the reference answers -1 at once. Outside a side plane it sets
  // -1 at every such plane and then keeps the smaller vd of the later planes with pr > 0, so the
  // answer is min(-1, vd of those planes): negative, which is all the reference's convention gives a
  // wrong-side result; its FP32 magnitude is returned once every plane's side is decided.
  const float dz = float(t.dz), ht = float(kHalfTolerance);
  const float mz = vabs(p.z) - dz, ez = 3.f * kU32 * (dz + vabs(p.z));
  if (mz - ez > ht) return -1.f;
  if (mz + ez > ht) return Undecided();
  {
    float neg = 0.f;
    bool outside = false;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const TrapPlaneF q = trap_plane_f32(t, i, p, d);
      if (vabs(q.pd - ht) <= q.epd) return Undecided();
      if (q.pd > ht) {
        outside = true;
        neg     = -1.f;
      }
      if (outside && q.pr > 0.f) neg = vmin(neg, -q.pd / NonZero(q.pr));
    }
    if (outside) return neg;
  }
  FInt out{FLT_MAX, FLT_MAX};
  float vout = FLT_MAX;
  if (d.z != 0.f) {
    const float inv = 1.f / NonZero(d.z);
    const float v   = (Sign(d.z) * dz - p.z) * inv;
    out             = fint(v, 3.f * kU32 * (dz + vabs(p.z)) * vabs(inv) + 2.f * kU32 * vabs(v));
    vout            = v;
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const TrapPlaneF q = trap_plane_f32(t, i, p, d);
    if (q.pd + q.epd > float(-kHalfTolerance)) return Undecided(); // not certainly inside this plane
    if (q.pr + q.epr <= 0.f) continue;                              // certainly not exiting through it
    if (q.pr - q.epr <= 0.f) {
      // exit direction undecided: only harmless when even the nearest possible crossing is beyond
      const float lo = (-q.pd - q.epd) / (vabs(q.pr) + q.epr);
      if (lo > out.hi) continue;
      return Undecided();
    }
    const FInt v = trap_vdist_f32(q);
    vout         = vmin(vout, -q.pd / NonZero(q.pr));
    out          = fmini(out, v);
  }
  return out.hi - out.lo <= kMixDistErr ? vout : Undecided();
}

template <class R>
VGMI_DEV R DistanceToIn(TrapezoidData const &t, Vec3<R> const &p, Vec3<R> const &d)
{
  if constexpr (sizeof(R) == sizeof(float)) return DistanceToInF32(t, p, d);
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
  if constexpr (sizeof(R) == sizeof(float)) return DistanceToOutF32(t, p, d);
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
