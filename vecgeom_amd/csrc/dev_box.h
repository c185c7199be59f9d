// Box kernels (VecGeom/volumes/kernel/BoxImplementation.h:76-167).
#pragma once
#include "dev_common.h"
#include "shape_data.h"

namespace vgmi {
namespace box {

template <class R>
VGMI_DEV Vec3<R> half(BoxData const &b)
{
  return Vec3<R>(R(b.h[0]), R(b.h[1]), R(b.h[2]));
}

template <class R>
VGMI_DEV bool Contains(BoxData const &b, Vec3<R> const &p)
{
  return Max3(Abs3(p) - half<R>(b)) < R(0);
}

template <class R>
VGMI_DEV int Inside(BoxData const &b, Vec3<R> const &p)
{
  R dist = Max3(Abs3(p) - half<R>(b));
  int in = dist < R(0) ? kInside : kOutside;
  return vabs(dist) < R(kHalfTolerance) ? kSurface : in;
}

// Certified FP32 DistanceToIn (see dev_common.h): the slab parameters as intervals. With FP32
// copies of h, p, d (each within u), t = (+-h - p) * (1 / d) is within 6u (|h| + |p|) / |d| of its
// FP64 value; orthog (the distance to an exit plane) within 3u (|h| + |p|).
VGMI_DEV float DistanceToInF32(BoxData const &b, Vec3<float> const &p, Vec3<float> const &d)
{
  const float h[3] = {float(b.h[0]), float(b.h[1]), float(b.h[2])};
  const float pp[3] = {p.x, p.y, p.z}, dd[3] = {d.x, d.y, d.z};
  FInt tin{-FLT_MAX, -FLT_MAX}, tout{FLT_MAX, FLT_MAX};
  float dist = -FLT_MAX, dout = FLT_MAX, orthog = FLT_MAX, eorth = 0.f;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float inv = 1.f / NonZero(dd[i]), sg = Sign(dd[i]);
    const float a = (-sg * h[i] - pp[i]) * inv, c = (sg * h[i] - pp[i]) * inv;
    const float e = 6.f * kU32 * (h[i] + vabs(pp[i])) * vabs(inv);
    tin  = fmaxi(tin, fint(a, e + 2.f * kU32 * vabs(a)));
    tout = fmini(tout, fint(c, e + 2.f * kU32 * vabs(c)));
    dist = vmax(dist, a);
    dout = vmin(dout, c);
    const float o = sg * h[i] - pp[i];
    if (sg * o < orthog) {
      orthog = sg * o;
      eorth  = 3.f * kU32 * (h[i] + vabs(pp[i]));
    }
  }
  orthog = vabs(orthog);
  // miss = dist >= dout || dout <= kHalfTolerance || orthog <= kHalfTolerance, each decided or NaN
  const float ht = float(kHalfTolerance);
  if (tout.hi <= ht || orthog + eorth <= ht) return InfLen<float>();
  if (tout.lo <= ht || orthog - eorth <= ht) return Undecided();
  if (tin.lo >= tout.hi) return InfLen<float>();
  if (tin.hi >= tout.lo) return Undecided();
  return tin.hi - tin.lo <= kMixDistErr ? dist : Undecided();
}

template <class R>
VGMI_DEV R DistanceToIn(BoxData const &b, Vec3<R> const &p, Vec3<R> const &d)
{
  if constexpr (sizeof(R) == sizeof(float)) return DistanceToInF32(b, p, d);
  Vec3<R> h = half<R>(b);
  Vec3<R> inv(R(1) / NonZero(d.x), R(1) / NonZero(d.y), R(1) / NonZero(d.z));
  Vec3<R> sgn(Sign(d.x), Sign(d.y), Sign(d.z));
  Vec3<R> tin  = -sgn * h - p;
  Vec3<R> tout = sgn * h - p;
  R orthog     = vabs(Min3(sgn * tout));
  R dout       = Min3(tout * inv);
  R dist       = Max3(tin * inv);
  bool miss    = dist >= dout || dout <= R(kHalfTolerance) || orthog <= R(kHalfTolerance);
  return miss ? InfLen<R>() : dist;
}

// Certified FP32 DistanceToOut: the exit parameters as intervals (bounds as in DistanceToInF32);
// an outside point is answered -1 as in the reference.
VGMI_DEV float DistanceToOutF32(BoxData const &b, Vec3<float> const &p, Vec3<float> const &d)
{
  const float h[3] = {float(b.h[0]), float(b.h[1]), float(b.h[2])};
  const float pp[3] = {p.x, p.y, p.z}, dd[3] = {d.x, d.y, d.z};
  int outside = kNo;
  FInt tout{FLT_MAX, FLT_MAX};
  float dout = FLT_MAX;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    outside = or3(outside, cmpGT(vabs(pp[i]) - h[i], float(kHalfTolerance), 3.f * kU32 * (h[i] + vabs(pp[i]))));
    const float inv = 1.f / NonZero(dd[i]);
    const float c   = (Sign(dd[i]) * h[i] - pp[i]) * inv;
    tout = fmini(tout, fint(c, 6.f * kU32 * (h[i] + vabs(pp[i])) * vabs(inv) + 2.f * kU32 * vabs(c)));
    dout = vmin(dout, c);
  }
  if (outside == kYes) return -1.f;
  if (outside == kUnsure) return Undecided();
  return tout.hi - tout.lo <= kMixDistErr ? dout : Undecided();
}

template <class R>
VGMI_DEV R DistanceToOut(BoxData const &b, Vec3<R> const &p, Vec3<R> const &d)
{
  if constexpr (sizeof(R) == sizeof(float)) return DistanceToOutF32(b, p, d);
  Vec3<R> h = half<R>(b);
  Vec3<R> inv(R(1) / NonZero(d.x), R(1) / NonZero(d.y), R(1) / NonZero(d.z));
  Vec3<R> sgn(Sign(d.x), Sign(d.y), Sign(d.z));
  R safeIn     = Max3(Abs3(p) - h);
  Vec3<R> tout = sgn * h - p;
  R dist       = Min3(tout * inv);
  return safeIn > R(kHalfTolerance) ? R(-1) : dist;
}

template <class R>
VGMI_DEV R SafetyToIn(BoxData const &b, Vec3<R> const &p)
{
  return Max3(Abs3(p) - half<R>(b));
}

template <class R>
VGMI_DEV R SafetyToOut(BoxData const &b, Vec3<R> const &p)
{
  return Min3(half<R>(b) - Abs3(p));
}

} // namespace box
} // namespace vgmi
