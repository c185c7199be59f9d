// Trd kernels (VecGeom/volumes/kernel/TrdImplementation.h:25-518) for TrdTypes::UniversalTrd
// (VecGeom/volumes/kernel/shapetypes/TrdTypes.h:46-76): the y faces count as varying when dy1 != dy2,
// and the Inside/Contains y test always takes the cross product.
#pragma once
#include "dev_common.h"
#include "shape_data.h"

namespace vgmi {
namespace trd {

// TrdUtilities::UnplacedInside (TrdImplementation.h:171-210): each lateral test is the sign of the
// cross product of (|x| - dx1, z + dz) with the side's direction (dx2 - dx1, 2 dz).
template <class R>
VGMI_DEV void flags(TrdData const &t, Vec3<R> const &p, bool surface, bool &cin, bool &cout)
{
  const R pz  = p.z + R(t.dz);
  const R hT  = R(kHalfTolerance);
  cout        = vabs(p.z) > (surface ? R(t.dz) + hT : R(t.dz));
  cin         = surface && vabs(p.z) < R(t.dz) - hT;
  const R cx  = R(t.x2minusx1) * pz - R(2.0) * R(t.dz) * (vabs(p.x) - R(t.dx1));
  if (surface) {
    cout = cout || cx < -R(t.tolx);
    cin  = cin && cx > R(t.tolx);
  } else {
    cout = cout || cx < R(0);
  }
  if (t.plainY) { // TrdTypes::Trd1 (HasVaryingY kNo)
    cout = cout || vabs(p.y) > (surface ? R(t.dy1) + hT : R(t.dy1));
    if (surface) cin = cin && vabs(p.y) < R(t.dy1) - hT;
    return;
  }
  const R cy = R(t.y2minusy1) * pz - R(2.0) * R(t.dz) * (vabs(p.y) - R(t.dy1));
  if (surface) {
    cout = cout || cy < -R(t.toly);
    cin  = cin && cy > R(t.toly);
  } else {
    cout = cout || cy < R(0);
  }
}

template <class R>
VGMI_DEV bool Contains(TrdData const &t, Vec3<R> const &p)
{
  bool ci, co;
  flags(t, p, false, ci, co);
  return !co;
}

template <class R>
VGMI_DEV int Inside(TrdData const &t, Vec3<R> const &p)
{
  bool ci, co;
  flags(t, p, true, ci, co);
  return co ? kOutside : (ci ? kInside : kSurface);
}

// TrdUtilities::FaceTrajectoryIntersection (TrdImplementation.h:75-130): the ray against face +x (forY
// false) or +y, or the mirror face, entering (toIn) or leaving; dist is written once the face's
// orientation admits a crossing.
template <class R>
VGMI_DEV bool face(TrdData const &t, bool forY, bool mirrored, bool toIn, Vec3<R> const &p, Vec3<R> const &d, R &dist)
{
  R alongV, v1, posV, posK, dirV, dirK, fK, fV, halfKplus;
  if (forY) {
    alongV = R(t.y2minusy1), v1 = R(t.dy1), posV = p.y, posK = p.x, dirV = d.y, dirK = d.x, fK = R(t.fx);
    fV = R(t.fy), halfKplus = R(t.halfx1plusx2);
  } else {
    alongV = R(t.x2minusx1), v1 = R(t.dx1), posV = p.x, posK = p.y, dirV = d.x, dirK = d.y, fK = R(t.fy);
    fV = R(t.fx), halfKplus = R(t.halfy1plusy2);
  }
  if (mirrored) posV *= R(-1.), dirV *= R(-1.);
  const R ndotv = dirV + fV * d.z;
  if (!(toIn ? ndotv < R(0.) : ndotv > R(0.))) return false;
  const R alongZ = R(2.0) * R(t.dz);
  dist = (alongZ * (posV - v1) - alongV * (p.z + R(t.dz))) / (d.z * alongV - dirV * alongZ + R(kTiny));
  if (!(dist > R(0.) - R(kHalfTolerance))) return false;
  const R hitz = p.z + dist * d.z;
  const R hitk = posK + dist * dirK;
  const bool ok = vabs(hitz) <= R(t.dz) && vabs(hitk) <= halfKplus - fK * hitz;
  if (ok && vabs(dist) < R(kHalfTolerance)) dist = R(0.0);
  return ok;
}

template <class R>
VGMI_DEV R finish(R dist)
{
  return vabs(dist) < R(kHalfTolerance) ? R(0.0) : dist;
}

// TrdImplementation.h:299-385
template <class R>
VGMI_DEV R DistanceToIn(TrdData const &t, Vec3<R> const &p, Vec3<R> const &d)
{
  const R hT   = R(kHalfTolerance);
  R distance   = InfLen<R>();
  const bool inz = vabs(p.z) < R(t.dz) - hT;
  R distx        = R(t.halfx1plusx2) - R(t.fx) * p.z;
  const bool inx = (distx - vabs(p.x)) * R(t.calfx) > R(0.) + hT;
  R disty;
  bool iny;
  if (t.varyingY) {
    disty = R(t.halfy1plusy2) - R(t.fy) * p.z;
    iny   = (disty - vabs(p.y)) * R(t.calfy) > R(0.) + hT;
  } else {
    disty = vabs(p.y) - R(t.dy1);
    iny   = disty < R(0.) - hT;
  }
  if (inx && iny && inz) return R(-1.);
  if (p.z * d.z < R(0.) && !inz) {
    const R distz = (vabs(p.z) - R(t.dz)) / vabs(d.z);
    const R hitx  = vabs(p.x + distz * d.x);
    const R hity  = vabs(p.y + distz * d.y);
    const bool top = p.z > (R(t.dz) - hT) && hitx <= R(t.dx2) && hity <= R(t.dy2);
    const bool bot = p.z < (-R(t.dz) + hT) && hitx <= R(t.dx1) && hity <= R(t.dy1);
    if (top || bot) return finish(distz);
  }
  if (!inx) {
    // the second face decides whether the search stops here (the reference keeps its mask only)
    if (face(t, false, false, true, p, d, distx)) distance = distx;
    if (face(t, false, true, true, p, d, distx)) return finish(distx);
  }
  if (t.varyingY) {
    if (!iny) {
      if (face(t, true, false, true, p, d, disty)) distance = disty;
      if (face(t, true, true, true, p, d, disty)) distance = disty;
    }
  } else if (!iny) {
    disty /= vabs(d.y);
    const R zhit = p.z + disty * d.z;
    const R xhit = p.x + disty * d.x;
    const R dx   = R(t.halfx1plusx2) - R(t.fx) * zhit;
    if (p.y * d.y < R(0) && disty > -hT && vabs(xhit) < dx && vabs(zhit) < R(t.dz)) distance = disty;
  }
  return finish(distance);
}

// TrdImplementation.h:387-495 (scalar early returns)
template <class R>
VGMI_DEV R DistanceToOut(TrdData const &t, Vec3<R> const &p, Vec3<R> const &d)
{
  const R hT     = R(kHalfTolerance);
  const R invdir = R(1.) / vabs(d.z + R(kTiny));
  bool out       = R(t.dz) - vabs(p.z) < R(0.) - hT;
  R distx        = R(t.halfx1plusx2) - R(t.fx) * p.z;
  out            = out || (distx - vabs(p.x)) * R(t.calfx) < R(0.) - hT;
  R disty;
  if (t.varyingY) {
    disty = R(t.halfy1plusy2) - R(t.fy) * p.z;
    out   = out || (disty - vabs(p.y)) * R(t.calfy) < R(0.) - hT;
  } else {
    disty = R(t.dy1) - vabs(p.y);
    out   = out || disty < R(0.) - hT;
  }
  if (out) return R(-1.);
  if (d.z > R(0.)) {
    const R distz = (R(t.dz) - p.z) * invdir;
    if (vabs(p.x + distz * d.x) <= R(t.dx2) && vabs(p.y + distz * d.y) <= R(t.dy2)) return finish(distz);
  }
  if (d.z < R(0.)) {
    const R distz = (p.z + R(t.dz)) * invdir;
    if (vabs(p.x + distz * d.x) <= R(t.dx1) && vabs(p.y + distz * d.y) <= R(t.dy1)) return finish(distz);
  }
  if (face(t, false, false, false, p, d, distx)) return finish(distx);
  if (face(t, false, true, false, p, d, distx)) return finish(distx);
  R distance = R(0.0);
  if (t.varyingY) {
    if (face(t, true, false, false, p, d, disty)) return finish(disty);
    if (face(t, true, true, false, p, d, disty)) distance = disty;
  } else {
    const R plane = d.y < R(0.) ? -R(t.dy1) : R(t.dy1);
    disty         = (plane - p.y) / d.y;
    const R zhit  = p.z + disty * d.z;
    const R xhit  = p.x + disty * d.x;
    const R dx    = R(t.halfx1plusx2) - R(t.fx) * zhit;
    if (vabs(xhit) < dx && vabs(zhit) < R(t.dz)) distance = disty;
  }
  return finish(distance);
}

// TrdUtilities::Safety (TrdImplementation.h:132-165)
template <class R>
VGMI_DEV R safety(TrdData const &t, Vec3<R> const &p, bool inside)
{
  R dist       = R(t.dz) - vabs(p.z);
  const R dxz  = R(t.halfx1plusx2) - R(t.fx) * p.z;
  const R safx = (dxz - vabs(p.x)) * R(t.calfx);
  if (dxz >= R(0) && safx < dist) dist = safx;
  if (t.varyingY) {
    const R dyz  = R(t.halfy1plusy2) - R(t.fy) * p.z;
    const R safy = (dyz - vabs(p.y)) * R(t.calfy);
    if (dyz >= R(0) && safy < dist) dist = safy;
  } else {
    const R safy = R(t.dy1) - vabs(p.y);
    if (safy < dist) dist = safy;
  }
  return inside ? dist : -dist;
}

template <class R>
VGMI_DEV R SafetyToIn(TrdData const &t, Vec3<R> const &p)
{
  return safety(t, p, false);
}
template <class R>
VGMI_DEV R SafetyToOut(TrdData const &t, Vec3<R> const &p)
{
  return safety(t, p, true);
}

} // namespace trd
} // namespace vgmi
