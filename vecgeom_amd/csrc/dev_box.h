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

template <class R>
VGMI_DEV R DistanceToIn(BoxData const &b, Vec3<R> const &p, Vec3<R> const &d)
{
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

template <class R>
VGMI_DEV R DistanceToOut(BoxData const &b, Vec3<R> const &p, Vec3<R> const &d)
{
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
