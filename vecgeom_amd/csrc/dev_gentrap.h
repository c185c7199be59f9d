// GenTrap kernels for planar lateral faces: the reference's scalar CPU path
// (VecGeom/volumes/kernel/GenTrapImplementation.h:330-608), which hands a planar trap to its
// TessellatedSection helper (VecGeom/volumes/TessellatedSection.h: z slab and one convex quadrilateral
// plane per cluster; TessellatedCluster.h DistanceToInConvex / DistanceToOutConvex / DistPlanes).
// Templated on the arithmetic type.
#pragma once
#include "dev_common.h"
#include "shape_data.h"

namespace vgmi {
namespace gtra {

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
VGMI_DEV Vec3<R> fnormal(GenTrapData const &s, int i)
{
  return Vec3<R>(R(s.fnormal[i][0]), R(s.fnormal[i][1]), R(s.fnormal[i][2]));
}
// TessellatedCluster::DistPlanes: point . normal + distance
template <class R>
VGMI_DEV R dist_plane(GenTrapData const &s, int i, Vec3<R> const &p)
{
  return p.Dot(fnormal<R>(s, i)) + R(s.fdist[i]);
}

template <class R>
VGMI_DEV bool Contains(GenTrapData const &s, Vec3<R> const &p)
{
  if (vabs(p.z) > R(s.dz)) return false;
  for (int i = 0; i < s.nfacets; ++i)
    if (dist_plane(s, i, p) > R(0)) return false;
  return true;
}

template <class R>
VGMI_DEV int Inside(GenTrapData const &s, Vec3<R> const &p)
{
  const R tol  = R(1e-9);
  const R safZ = vabs(p.z) - R(s.dz);
  if (safZ > tol) return kOutside;
  bool in = true, out = false;
  for (int i = 0; i < s.nfacets; ++i) {
    const R d = dist_plane(s, i, p);
    out       = out || d > tol;
    in        = in && d < -tol;
  }
  const int inside = out ? kOutside : (in ? kInside : kSurface);
  return safZ < -tol || inside == kOutside ? inside : kSurface;
}

// DistanceToIn with the step limit (TessellatedSection::DistanceToIn<false>): the z slab entry, raised by
// every face the ray enters and cut by every face it leaves through, up to min(z slab exit, stepMax)
template <class R>
VGMI_DEV R DistanceToInStep(GenTrapData const &s, Vec3<R> const &p, Vec3<R> const &d, R stepMax)
{
  const R tol     = R(1e-9);
  const R invdirz = R(1.) / NonZero(d.z);
  const R pz      = p.z;
  if ((vabs(pz) - R(s.dz)) > -tol && pz * d.z >= R(0)) return InfLen<R>();
  const R ddz = vcopysign(R(s.dz), invdirz);
  R distance  = -(pz + ddz) * invdirz;
  R limit     = Mn(-(pz - ddz) * invdirz, stepMax);
  for (int i = 0; i < s.nfacets; ++i) {
    const R proj            = NonZero(d.Dot(fnormal<R>(s, i)));
    const bool moving_away  = proj > -tol;
    const R pdist           = dist_plane(s, i, p);
    const bool side_correct = pdist > -tol;
    if (side_correct && moving_away) return InfLen<R>();
    const R dist = -pdist / NonZero(proj);
    distance     = Mx(distance, side_correct && !moving_away ? dist : -InfLen<R>());
    limit        = Mn(limit, !side_correct && moving_away ? dist : InfLen<R>());
    if (!(distance < limit - tol)) return InfLen<R>();
  }
  return distance;
}

template <class R>
VGMI_DEV R DistanceToIn(GenTrapData const &s, Vec3<R> const &p, Vec3<R> const &d)
{
  return DistanceToInStep(s, p, d, InfLen<R>());
}

// TessellatedSection::DistanceToOut: -kTolerance from outside the z slab or a face's plane
template <class R>
VGMI_DEV R DistanceToOut(GenTrapData const &s, Vec3<R> const &p, Vec3<R> const &d)
{
  const R tol = R(1e-9);
  if (vabs(p.z) - R(s.dz) > tol) return -tol;
  R distance = (vcopysign(R(s.dz), d.z) - p.z) / NonZero(d.z);
  for (int i = 0; i < s.nfacets; ++i) {
    const R saf = dist_plane(s, i, p);
    if (!(saf < tol)) return -tol;
    const R ndd  = NonZero(d.Dot(fnormal<R>(s, i)));
    const R dist = ndd > R(0.) ? -saf / ndd : InfLen<R>();
    if (dist < distance) distance = dist;
  }
  return distance;
}

// TessellatedSection::SafetyToIn / SafetyToOut (fZ = 0): the largest of the z-plane and face-plane distances
template <class R>
VGMI_DEV R max_plane(GenTrapData const &s, Vec3<R> const &p)
{
  R safety = Mx(R(0.) - R(s.dz) - p.z, p.z - R(0.) - R(s.dz));
  for (int i = 0; i < s.nfacets; ++i) {
    const R v = dist_plane(s, i, p);
    if (v > safety) safety = v;
  }
  return safety;
}

template <class R>
VGMI_DEV R SafetyToIn(GenTrapData const &s, Vec3<R> const &p)
{
  return max_plane(s, p);
}

template <class R>
VGMI_DEV R SafetyToOut(GenTrapData const &s, Vec3<R> const &p)
{
  return -max_plane(s, p);
}

} // namespace gtra
} // namespace vgmi
