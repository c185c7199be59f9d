// Parallelepiped (VecGeom/volumes/kernel/ParallelepipedImplementation.h:60-183) and Tet
// (VecGeom/volumes/kernel/TetImplementation.h:61-185) kernels: planar solids, templated on the
// arithmetic type (FP64, or FP32 for the mixed-precision point methods).
#pragma once
#include "dev_common.h"
#include "shape_data.h"

namespace vgmi {
namespace ppd {

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

// ParallelepipedImplementation::Transform (:60-66): to the oblique frame
template <class R>
VGMI_DEV Vec3<R> oblique(ParallelepipedData const &s, Vec3<R> p)
{
  p.y -= R(s.tanThetaSinPhi) * p.z;
  p.x -= R(s.tanThetaCosPhi) * p.z + R(s.tanAlpha) * p.y;
  return p;
}

// ::SafetyVector (:68-76)
template <class R>
VGMI_DEV Vec3<R> safety_vector(ParallelepipedData const &s, Vec3<R> const &lp)
{
  Vec3<R> v(vabs(lp.x) - R(s.dim[0]), vabs(lp.y) - R(s.dim[1]), vabs(lp.z) - R(s.dim[2]));
  v.x *= R(s.ctx);
  v.y *= R(s.cty);
  return v;
}

template <class R>
VGMI_DEV R max_safety(ParallelepipedData const &s, Vec3<R> const &p)
{
  const Vec3<R> v = safety_vector(s, oblique(s, p));
  return Mx(Mx(v.x, v.y), v.z);
}

template <class R>
VGMI_DEV bool Contains(ParallelepipedData const &s, Vec3<R> const &p)
{
  return max_safety(s, p) < R(0.0);
}

template <class R>
VGMI_DEV int Inside(ParallelepipedData const &s, Vec3<R> const &p)
{
  const R safety = max_safety(s, p);
  return vabs(safety) < R(kHalfTolerance) ? kSurface : (safety < R(0.0) ? kInside : kOutside);
}

template <class R>
VGMI_DEV R SafetyToIn(ParallelepipedData const &s, Vec3<R> const &p)
{
  const R safety = max_safety(s, p);
  return vabs(safety) < R(kHalfTolerance) ? R(0.0) : safety;
}

template <class R>
VGMI_DEV R SafetyToOut(ParallelepipedData const &s, Vec3<R> const &p)
{
  const R safety = -max_safety(s, p);
  return vabs(safety) < R(kHalfTolerance) ? R(0.0) : safety;
}

// :130-162
template <class R>
VGMI_DEV R DistanceToIn(ParallelepipedData const &s, Vec3<R> const &point, Vec3<R> const &direction)
{
  const Vec3<R> p = oblique(s, point), v = oblique(s, direction), sv = safety_vector(s, p);
  const R hT         = R(kHalfTolerance);
  const bool leaving = (sv.x >= -hT && p.x * v.x >= R(0.)) || (sv.y >= -hT && p.y * v.y >= R(0.)) ||
                       (sv.z >= -hT && p.z * v.z >= R(0.));
  const Vec3<R> inv(R(1.) / NonZero(v.x), R(1.) / NonZero(v.y), R(1.) / NonZero(v.z));
  const Vec3<R> t(Sign(inv.x) * R(s.dim[0]), Sign(inv.y) * R(s.dim[1]), Sign(inv.z) * R(s.dim[2]));
  const R distIn  = Mx(Mx((-t.x - p.x) * inv.x, (-t.y - p.y) * inv.y), (-t.z - p.z) * inv.z);
  const R distOut = Mn(Mn((t.x - p.x) * inv.x, (t.y - p.y) * inv.y), (t.z - p.z) * inv.z);
  return !leaving && distOut > distIn + hT ? distIn : InfLen<R>();
}

// :164-183
template <class R>
VGMI_DEV R DistanceToOut(ParallelepipedData const &s, Vec3<R> const &point, Vec3<R> const &direction)
{
  const Vec3<R> p = oblique(s, point), v = oblique(s, direction), sv = safety_vector(s, p);
  const Vec3<R> dir(NonZero(v.x), NonZero(v.y), NonZero(v.z));
  const R dist = Mn(Mn((Sign(dir.x) * R(s.dim[0]) - p.x) / dir.x, (Sign(dir.y) * R(s.dim[1]) - p.y) / dir.y),
                    (Sign(dir.z) * R(s.dim[2]) - p.z) / dir.z);
  return Mx(Mx(sv.x, sv.y), sv.z) > R(kHalfTolerance) ? R(-1.0) : dist;
}

} // namespace ppd

namespace tet {

using ppd::Mn;
using ppd::Mx;

template <class R>
VGMI_DEV Vec3<R> normal(TetData const &s, int i)
{
  return Vec3<R>(R(s.n[i][0]), R(s.n[i][1]), R(s.n[i][2]));
}

// GenericKernelForContainsAndInside (:85-107): the largest plane distance n.p + d
template <class R>
VGMI_DEV R max_plane(TetData const &s, Vec3<R> const &p)
{
  R dist[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) dist[i] = normal<R>(s, i).Dot(p) + R(s.d[i]);
  return Mx(Mx(Mx(dist[0], dist[1]), dist[2]), dist[3]);
}

template <class R>
VGMI_DEV bool Contains(TetData const &s, Vec3<R> const &p)
{
  return !(max_plane(s, p) > R(kHalfTolerance));
}

template <class R>
VGMI_DEV int Inside(TetData const &s, Vec3<R> const &p)
{
  const R safety = max_plane(s, p);
  return safety <= -R(kHalfTolerance) ? kInside : (safety > R(kHalfTolerance) ? kOutside : kSurface);
}

// :109-135 (absSafe is a minimum of absolute values: its "<= -kHalfTolerance" test never holds and is
// left out)
template <class R>
VGMI_DEV R DistanceToIn(TetData const &s, Vec3<R> const &p, Vec3<R> const &d)
{
  R distance = -InfLen<R>(), distanceOut = InfLen<R>();
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const Vec3<R> n = normal<R>(s, i);
    const R cosa    = NonZero(n.Dot(d));
    const R safe    = n.Dot(p) + R(s.d[i]);
    const R dist    = -safe / cosa;
    if (cosa < R(0.)) distance = Mx(distance, dist);
    if (cosa > R(0.)) distanceOut = Mn(distanceOut, dist);
  }
  return distance >= distanceOut || distanceOut <= R(kHalfTolerance) ? InfLen<R>() : distance;
}

// :137-157
template <class R>
VGMI_DEV R DistanceToOut(TetData const &s, Vec3<R> const &p, Vec3<R> const &d)
{
  R distance = InfLen<R>(), safety = -InfLen<R>();
  R cosa[4], safe[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const Vec3<R> n = normal<R>(s, i);
    cosa[i]         = NonZero(n.Dot(d));
    safe[i]         = n.Dot(p) + R(s.d[i]);
    safety          = Mx(safety, safe[i]);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (cosa[i] > R(0.)) distance = Mn(distance, -safe[i] / cosa[i]);
  return safety > R(kHalfTolerance) ? R(-1.) : distance;
}

// :159-185
template <class R>
VGMI_DEV R SafetyToIn(TetData const &s, Vec3<R> const &p)
{
  const R safety = max_plane(s, p);
  return vabs(safety) <= R(kHalfTolerance) ? R(0.) : safety;
}

template <class R>
VGMI_DEV R SafetyToOut(TetData const &s, Vec3<R> const &p)
{
  const R safety = -max_plane(s, p);
  return vabs(safety) <= R(kHalfTolerance) ? R(0.) : safety;
}

} // namespace tet
} // namespace vgmi
