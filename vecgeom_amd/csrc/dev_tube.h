// Tube kernels: the run-time-specialised "UniversalTube" of
// VecGeom/volumes/kernel/TubeImplementation.h:389-964 (rmin and phi treatment
// are record flags, VecGeom/volumes/kernel/shapetypes/TubeTypes.h:92-144),
// phi helpers TubeImplementation.h:58-308 and the evolution wedge
// VecGeom/volumes/Wedge_Evolution.h:305-392.
#pragma once
#include "dev_common.h"
#include "shape_data.h"

namespace vgmi {
namespace tube {

// PointInCyclicalSector (onSurface = false)
template <class R>
VGMI_DEV bool InSector(TubeData const &t, R x, R y, bool includeSurface)
{
  R startCheck = (-x * R(t.a1y) + y * R(t.a1x));
  R endCheck   = (-R(t.a2x) * y + R(t.a2y) * x);
  R tol        = includeSurface ? R(-kHalfTolerance) : R(kHalfTolerance);
  if (t.phiLePi) return (startCheck >= tol) && (endCheck >= tol);
  return (startCheck >= tol) || (endCheck >= tol);
}

// Wedge::IsOnSurfaceGeneric
template <class R>
VGMI_DEV bool OnWedgePlane(R ax, R ay, R nx, R ny, Vec3<R> const &p)
{
  if (!(ax * p.x + ay * p.y >= R(0))) return false;
  return vabs(nx * p.x + ny * p.y) < R(kTolerance);
}

// Wedge::GenericKernelForContainsAndInside
template <class R>
VGMI_DEV void WedgeContainsInside(TubeData const &t, Vec3<R> const &p, bool forInside, bool &cin, bool &cout)
{
  R startCheck = (-p.x * R(t.a1y) + p.y * R(t.a1x));
  R endCheck   = (-R(t.a2x) * p.y + R(t.a2y) * p.x);
  cout         = startCheck < R(0);
  if (t.wedgeLtPi)
    cout = cout || (endCheck < R(0));
  else
    cout = cout && (endCheck < R(0));
  if (forInside) {
    bool on = OnWedgePlane(R(t.a1x), R(t.a1y), R(t.n1x), R(t.n1y), p) ||
              OnWedgePlane(R(t.a2x), R(t.a2y), R(t.n2x), R(t.n2y), p);
    cout = cout && !on;
    cin  = !on && !cout;
  }
}

// CircleTrajectoryIntersection
template <class R>
VGMI_DEV bool CircleHit(TubeData const &t, R b, R c, Vec3<R> const &p, Vec3<R> const &d, bool largest,
                        bool sectorCheck, R &dist)
{
  R delta = b * b - c;
  bool ok = delta > R(0);
  if (largest) ok = ok || (delta == R(0));
  if (!ok) delta = R(0);
  delta = vsqrt(delta);
  if (!largest) delta = -delta;
  dist = -b + delta;
  ok   = ok && (dist >= R(-2 * kTolerance));
  if (!ok) return false;
  if (sectorCheck) {
    R hitz = p.z + dist * d.z;
    if (!(vabs(hitz) <= R(t.z))) return false;
    if (t.hasPhi) {
      R hitx = p.x + dist * d.x;
      R hity = p.y + dist * d.y;
      ok     = InSector(t, hitx, hity, true);
    }
  }
  return ok;
}

// PhiPlaneTrajectoryIntersection with PositiveDirectionOfPhiVector = true
template <class R>
VGMI_DEV bool PhiPlaneHit(TubeData const &t, R ax, R ay, R nx, R ny, Vec3<R> const &p, Vec3<R> const &d, bool toIn,
                          R &dist)
{
  R dirDotNorm = d.x * nx + d.y * ny;
  bool ok      = toIn ? (dirDotNorm > R(0)) : (dirDotNorm < R(0));
  R dirDotXY   = (d.y * ax - d.x * ay);
  dist         = (ay * p.x - ax * p.y) / NonZero(dirDotXY);
  ok           = ok && ((dist * vabs(dirDotNorm)) > R(-kHalfTolerance));
  R hitx       = p.x + dist * d.x;
  R hity       = p.y + dist * d.y;
  if (toIn) {
    R hitz = p.z + dist * d.z;
    R r2   = hitx * hitx + hity * hity;
    ok     = ok && vabs(hitz) <= R(t.z + kHalfTolerance) && (r2 >= R(t.tolOrmin2)) && (r2 <= R(t.tolOrmax2));
    ok     = ok && (hitx * ax + hity * ay) > R(0);
  } else {
    ok = ok && (hitx * ax + hity * ay) >= R(0);
  }
  return ok;
}

// GenericKernelForContainsAndInside with MakePlus/MinusTolerant(Square)
// (VecGeom/volumes/kernel/GenericKernels.h:26-82)
template <class R>
VGMI_DEV void ContainsInside(TubeData const &t, Vec3<R> const &p, bool forInside, bool &cin, bool &cout)
{
  const R z = R(t.z), rmax = R(t.rmax), rmin = R(t.rmin);
  R absz    = vabs(p.z);
  cout      = absz > (forInside ? z + R(kHalfTolerance) : z);
  cin       = forInside && absz < z - R(kHalfTolerance);
  R r2      = p.x * p.x + p.y * p.y;
  cout      = cout || (r2 > (forInside ? rmax * (rmax + R(2.0 * kTolerance)) : rmax * rmax));
  if (forInside) cin = cin && (r2 < vabs(rmax * (rmax - R(2.0 * kTolerance))));
  if (t.hasRmin) {
    cout = cout || (r2 <= (forInside ? vabs(rmin * (rmin - R(2.0 * kTolerance))) : rmin * rmin));
    if (forInside) cin = cin && (r2 > rmin * (rmin + R(2.0 * kTolerance)));
  }
  if (t.hasPhi) {
    bool wi = false, wo = false;
    WedgeContainsInside(t, p, forInside, wi, wo);
    cout = cout || wo;
    if (forInside) cin = cin && wi;
  }
}

template <class R>
VGMI_DEV bool Contains(TubeData const &t, Vec3<R> const &p)
{
  bool ci, co;
  ContainsInside(t, p, false, ci, co);
  return !co;
}

template <class R>
VGMI_DEV int Inside(TubeData const &t, Vec3<R> const &p)
{
  bool ci, co;
  ContainsInside(t, p, true, ci, co);
  return ci ? kInside : (co ? kOutside : kSurface);
}

// DistanceToInKernel (TubeImplementation.h:504-681)
template <class R>
VGMI_DEV R DistanceToInKernel(TubeData const &t, Vec3<R> const &p, Vec3<R> const &d)
{
  const R z = R(t.z);
  R distz   = vabs(p.z) - z;
  if (distz > R(kHalfTolerance) && p.z * d.z >= R(0)) return InfLen<R>();
  R rsq   = p.x * p.x + p.y * p.y;
  R rdotn = p.x * d.x + p.y * d.y;
  const R tolIrmax2 = R(t.tolIrmax2);
  if (rsq > tolIrmax2 && rdotn >= R(0)) return InfLen<R>();

  bool inside = distz < R(-kHalfTolerance) && rsq < tolIrmax2;
  if (t.hasRmin) inside = inside && rsq > R(t.tolIrmin2);
  if (t.hasPhi && inside) inside = InSector(t, p.x, p.y, false);
  if (inside) return R(-1);

  // entry through a z plane
  distz /= NonZeroAbs(d.z);
  {
    R hitx   = p.x + distz * d.x;
    R hity   = p.y + distz * d.y;
    R r2     = hitx * hitx + hity * hity;
    bool okz = distz > R(-kHalfTolerance) && (p.z * d.z < R(0)) && (r2 <= R(t.rmax2));
    if (t.hasRmin) okz = okz && (R(t.rmin2) <= r2);
    if (t.hasPhi && okz) okz = InSector(t, hitx, hity, true);
    if (okz) return distz;
  }

  // on a cylindrical surface and moving inwards (IsMovingInsideTubeSurface)
  {
    R rho     = rsq;
    bool zok  = vabs(p.z) < z + R(kTolerance);
    bool onO  = (rho >= tolIrmax2) && (rho <= R(t.tolOrmax2)) && zok && (d.x * p.x + d.y * p.y + d.z * R(0) <= R(0));
    bool surf = onO;
    if (t.hasRmin) {
      bool onI = (rho >= R(t.tolOrmin2)) && (rho <= R(t.tolIrmin2)) && zok &&
                 (d.x * -p.x + d.y * -p.y + d.z * R(0) <= R(0));
      surf = surf || onI;
    }
    if (t.hasPhi) surf = surf && InSector(t, p.x, p.y, true);
    if (surf) return R(0);
  }

  R invnsq = R(1) / NonZero(R(1) - d.z * d.z);
  R b      = invnsq * rdotn;
  R dist   = InfLen<R>();
  {
    R crmax = invnsq * (rsq - R(t.rmax2));
    R dr;
    if (CircleHit(t, b, crmax, p, d, false, true, dr) && dr < dist) return dr;
  }
  if (t.hasRmin) {
    R crmin = invnsq * (rsq - R(t.rmin2));
    R dr;
    if (CircleHit(t, b, crmin, p, d, true, true, dr) && dr < dist) dist = dr;
  }
  if (t.hasPhi) {
    R dp;
    if (PhiPlaneHit(t, R(t.a1x), R(t.a1y), R(t.n1x), R(t.n1y), p, d, true, dp) && dp < dist) dist = dp;
    if (PhiPlaneHit(t, R(t.a2x), R(t.a2y), R(t.n2x), R(t.n2y), p, d, true, dp) && dp < dist) dist = dp;
  }
  return dist;
}

// ---------------------------------------------------------------- certified FP32 paths
// (see dev_common.h) The same decisions as DistanceToInKernel / DistanceToOut, each taken only when
// its margin exceeds the rounding bound of the FP32 operands; ep bounds the error of the FP32 point.
// FP32 forms 1 - dz^2 as dx^2 + dy^2 (the same value for a unit direction) so that rays close to
// the axis keep a small relative error.

// PointInCyclicalSector of a computed point (x, y), position error ex
VGMI_DEV int InSectorF(TubeData const &t, float x, float y, float ex, bool includeSurface)
{
  const float a1x = float(t.a1x), a1y = float(t.a1y), a2x = float(t.a2x), a2y = float(t.a2y);
  const float s1 = -x * a1y + y * a1x, s2 = -a2x * y + a2y * x;
  const float e1 = 1.5f * ex + 4.f * kU32 * (vabs(x * a1y) + vabs(y * a1x));
  const float e2 = 1.5f * ex + 4.f * kU32 * (vabs(a2x * y) + vabs(a2y * x));
  const float tol = includeSurface ? float(-kHalfTolerance) : float(kHalfTolerance);
  const int c1 = cmpGT(s1, tol, e1), c2 = cmpGT(s2, tol, e2); // >= and > agree once decided
  return t.phiLePi ? and3(c1, c2) : or3(c1, c2);
}

// CircleTrajectoryIntersection with b, c carrying errors eb, ec
VGMI_DEV int CircleHitF(TubeData const &t, float b, float eb, float c, float ec, Vec3<float> const &p,
                        Vec3<float> const &d, float ep, bool largest, bool sectorCheck, float &dist, float &edist)
{
  const float delta = b * b - c;
  const float ed    = 2.f * vabs(b) * eb + ec + 3.f * kU32 * (b * b + vabs(c));
  if (delta < -ed) return kNo;
  if (delta <= ed) return kUnsure; // tangent within rounding (largest also accepts delta == 0)
  const float sq  = vsqrt(delta);
  const float esq = ed / (sq + vsqrt(vmax(delta - ed, 0.f))) + kU32 * sq;
  dist            = largest ? -b + sq : -b - sq;
  edist           = eb + esq + 2.f * kU32 * vabs(dist);
  const int ok    = cmpGT(dist, float(-2 * kTolerance), edist);
  if (ok != kYes) return ok;
  if (!sectorCheck) return kYes;
  const float hitz = p.z + dist * d.z;
  const float eh   = edist + 3.f * kU32 * (vabs(p.x) + vabs(p.y) + vabs(p.z) + vabs(dist)) + ep;
  const int zok    = cmpGT(float(t.z), vabs(hitz), eh + kU32 * float(t.z));
  if (zok != kYes || !t.hasPhi) return zok;
  return InSectorF(t, p.x + dist * d.x, p.y + dist * d.y, eh, true);
}

// PhiPlaneTrajectoryIntersection (PositiveDirectionOfPhiVector)
VGMI_DEV int PhiPlaneHitF(TubeData const &t, float ax, float ay, float nx, float ny, Vec3<float> const &p,
                          Vec3<float> const &d, float ep, bool toIn, float &dist, float &edist)
{
  const float dn  = d.x * nx + d.y * ny;
  const float edn = 4.f * kU32 * (vabs(d.x * nx) + vabs(d.y * ny));
  const int dirok = toIn ? cmpGT(dn, 0.f, edn) : cmpGT(0.f, dn, edn);
  if (dirok != kYes) return dirok;
  const float dxy  = d.y * ax - d.x * ay;
  const float edxy = 4.f * kU32 * (vabs(d.y * ax) + vabs(d.x * ay));
  if (vabs(dxy) <= edxy) return kUnsure;
  const float num  = ay * p.x - ax * p.y;
  const float enm  = 4.f * kU32 * (vabs(ay * p.x) + vabs(ax * p.y)) + 1.5f * ep;
  dist             = num / NonZero(dxy);
  edist            = (enm + vabs(dist) * edxy) / (vabs(dxy) - edxy) + 2.f * kU32 * vabs(dist);
  int ok           = cmpGT(dist * vabs(dn), float(-kHalfTolerance), edist * vabs(dn) + vabs(dist) * edn);
  if (ok != kYes) return ok;
  const float hx = p.x + dist * d.x, hy = p.y + dist * d.y;
  const float eh = edist + 3.f * kU32 * (vabs(p.x) + vabs(p.y) + vabs(p.z) + vabs(dist)) + ep;
  if (toIn) {
    const float hz = p.z + dist * d.z;
    ok             = cmpGT(float(t.z + kHalfTolerance), vabs(hz), eh + kU32 * float(t.z));
    const float r2 = hx * hx + hy * hy;
    const float er = 2.f * vsqrt(r2) * eh + eh * eh + 4.f * kU32 * r2;
    ok             = and3(ok, cmpGT(r2, float(t.tolOrmin2), er + kU32 * float(t.tolOrmin2)));
    ok             = and3(ok, cmpGT(float(t.tolOrmax2), r2, er + kU32 * float(t.tolOrmax2)));
  }
  const float ha = hx * ax + hy * ay;
  return and3(ok, cmpGT(ha, 0.f, 1.5f * eh + 3.f * kU32 * (vabs(hx * ax) + vabs(hy * ay))));
}

VGMI_DEV float DistanceToInF32(TubeData const &t, Vec3<float> const &pt, Vec3<float> const &d)
{
  Vec3<float> p = pt;
  float ep      = 2.f * kU32 * (vabs(p.x) + vabs(p.y) + vabs(p.z));
  float move = 0.f, emove = 0.f;
  const float ptDist = vsqrt(p.Mag2());
  if (ptDist > 100.f * float(t.maxVal)) { // the far-point shift (TubeImplementation.h:475-502)
    move  = ptDist - 100.f * float(t.maxVal);
    p     = p + move * d;
    emove = 4.f * kU32 * ptDist;
    ep += 8.f * kU32 * ptDist;
  }
  const float z = float(t.z), ht = float(kHalfTolerance);
  const float distz = vabs(p.z) - z;
  const float edz   = 2.f * kU32 * (vabs(p.z) + z) + ep;
  if (p.z * d.z >= 0.f) {
    const int c = cmpGT(distz, ht, edz);
    if (c == kYes) return InfLen<float>();
    if (c == kUnsure) return Undecided();
  }
  const float rsq   = p.x * p.x + p.y * p.y;
  const float ersq  = 4.f * kU32 * rsq + 2.f * vsqrt(rsq) * ep + ep * ep;
  const float rdotn = p.x * d.x + p.y * d.y;
  const float erd   = 4.f * kU32 * (vabs(p.x * d.x) + vabs(p.y * d.y)) + ep;
  const float irmax2 = float(t.tolIrmax2);
  {
    int c = cmpGT(rsq, irmax2, ersq + kU32 * irmax2);
    if (c != kNo) c = and3(c, not3(cmpGT(0.f, rdotn, erd)));
    if (c == kYes) return InfLen<float>();
    if (c == kUnsure) return Undecided();
  }
  {
    int in = and3(cmpGT(-ht, distz, edz), cmpGT(irmax2, rsq, ersq + kU32 * irmax2));
    if (t.hasRmin) in = and3(in, cmpGT(rsq, float(t.tolIrmin2), ersq + kU32 * float(t.tolIrmin2)));
    if (t.hasPhi && in != kNo) in = and3(in, InSectorF(t, p.x, p.y, ep, false));
    if (in == kYes) return -1.f;
    if (in == kUnsure) return Undecided();
  }
  // entry through a z plane
  {
    const float adz = NonZeroAbs(d.z);
    const float tz  = distz / adz;
    const float etz = edz / adz + 2.f * kU32 * vabs(tz);
    int okz         = (p.z * d.z < 0.f) ? cmpGT(tz, -ht, etz) : kNo;
    if (okz != kNo) {
      const float hx = p.x + tz * d.x, hy = p.y + tz * d.y;
      const float eh = etz + 3.f * kU32 * (vabs(p.x) + vabs(p.y) + vabs(tz)) + ep;
      const float r2 = hx * hx + hy * hy;
      const float er = 2.f * vsqrt(r2) * eh + eh * eh + 4.f * kU32 * r2;
      okz            = and3(okz, cmpGT(float(t.rmax2), r2, er + kU32 * float(t.rmax2)));
      if (t.hasRmin) okz = and3(okz, cmpGT(r2, float(t.rmin2), er + kU32 * float(t.rmin2)));
      if (t.hasPhi && okz != kNo) okz = and3(okz, InSectorF(t, hx, hy, eh, true));
    }
    if (okz == kUnsure) return Undecided();
    if (okz == kYes) return etz + emove <= kMixDistErr ? tz + move : Undecided();
  }
  // (a start point on a cylindrical surface is within kMixedBand of the surface: never here)
  const float n2   = d.x * d.x + d.y * d.y;
  const float inv  = 1.f / NonZero(n2);
  const float rel  = 6.f * kU32;
  const float b    = inv * rdotn;
  const float eb   = vabs(b) * rel + erd * vabs(inv);
  FInt best{FLT_MAX, FLT_MAX};
  float vbest = FLT_MAX;
  {
    const float cr = rsq - float(t.rmax2);
    const float c  = inv * cr;
    const float ec = vabs(c) * rel + (ersq + 2.f * kU32 * (rsq + float(t.rmax2))) * vabs(inv);
    float dr, edr;
    const int h = CircleHitF(t, b, eb, c, ec, p, d, ep, false, true, dr, edr);
    if (h == kUnsure) return Undecided();
    if (h == kYes) return edr + emove <= kMixDistErr ? dr + move : Undecided();
  }
  if (t.hasRmin) {
    const float cr = rsq - float(t.rmin2);
    const float c  = inv * cr;
    const float ec = vabs(c) * rel + (ersq + 2.f * kU32 * (rsq + float(t.rmin2))) * vabs(inv);
    float dr, edr;
    const int h = CircleHitF(t, b, eb, c, ec, p, d, ep, true, true, dr, edr);
    if (h == kUnsure) return Undecided();
    if (h == kYes) {
      best  = fmini(best, fint(dr, edr));
      vbest = vmin(vbest, dr);
    }
  }
  if (t.hasPhi) {
    float dp, edp;
    int h = PhiPlaneHitF(t, float(t.a1x), float(t.a1y), float(t.n1x), float(t.n1y), p, d, ep, true, dp, edp);
    if (h == kUnsure) return Undecided();
    if (h == kYes) {
      best  = fmini(best, fint(dp, edp));
      vbest = vmin(vbest, dp);
    }
    h = PhiPlaneHitF(t, float(t.a2x), float(t.a2y), float(t.n2x), float(t.n2y), p, d, ep, true, dp, edp);
    if (h == kUnsure) return Undecided();
    if (h == kYes) {
      best  = fmini(best, fint(dp, edp));
      vbest = vmin(vbest, dp);
    }
  }
  if (vbest >= FLT_MAX) return InfLen<float>();
  return best.hi - best.lo + 2.f * emove <= kMixDistErr ? vbest + move : Undecided();
}

VGMI_DEV float DistanceToOutF32(TubeData const &t, Vec3<float> const &p, Vec3<float> const &d)
{
  const float ep = 2.f * kU32 * (vabs(p.x) + vabs(p.y) + vabs(p.z));
  const float z = float(t.z), ht = float(kHalfTolerance);
  // wrong side: every outside point gets -1
  {
    const float distz = z - vabs(p.z);
    int out = cmpGT(-ht, distz, 2.f * kU32 * (vabs(p.z) + z) + ep);
    const float rsq  = p.x * p.x + p.y * p.y;
    const float ersq = 4.f * kU32 * rsq + 2.f * vsqrt(rsq) * ep + ep * ep;
    out = or3(out, cmpGT(rsq - float(t.rmax2), float(2.0 * kTolerance * t.rmax),
                         ersq + 2.f * kU32 * (rsq + float(t.rmax2))));
    if (t.hasRmin)
      out = or3(out, cmpGT(float(-2.0 * kTolerance * t.rmin), rsq - float(t.rmin2),
                           ersq + 2.f * kU32 * (rsq + float(t.rmin2))));
    if (t.hasPhi) {
      const float sc = -p.x * float(t.a1y) + p.y * float(t.a1x), ec = -float(t.a2x) * p.y + float(t.a2y) * p.x;
      const float e1 = 1.5f * ep + 4.f * kU32 * (vabs(p.x) + vabs(p.y));
      const int o1 = cmpGT(0.f, sc, e1), o2 = cmpGT(0.f, ec, e1);
      int wo = t.wedgeLtPi ? or3(o1, o2) : and3(o1, o2);
      // on a phi plane (|n.p| < kTolerance on the plane's half) is within kMixedBand of the surface
      // for a point inside the z and r ranges: such a point never reaches the FP32 pass
      out = or3(out, wo);
    }
    if (out == kYes) return -1.f;
    if (out == kUnsure) return Undecided();
  }
  const float rsq   = p.x * p.x + p.y * p.y;
  const float ersq  = 4.f * kU32 * rsq + 2.f * vsqrt(rsq) * ep + ep * ep;
  const float rdotn = d.x * p.x + d.y * p.y;
  const float erd   = 4.f * kU32 * (vabs(p.x * d.x) + vabs(p.y * d.y)) + ep;
  FInt best{FLT_MAX, FLT_MAX};
  float vbest = FLT_MAX;
  if (d.z != 0.f) {
    const float inv = 1.f / NonZero(d.z);
    const float v   = (d.z < 0.f ? (-z - p.z) : (z - p.z)) * inv;
    best            = fint(v, (2.f * kU32 * (z + vabs(p.z)) + ep) * vabs(inv) + 2.f * kU32 * vabs(v));
    vbest           = v;
  }
  const float n2  = d.x * d.x + d.y * d.y;
  const float inv = 1.f / NonZero(n2);
  const float rel = 6.f * kU32;
  const float b   = inv * rdotn;
  const float eb  = vabs(b) * rel + erd * vabs(inv);
  if (t.hasRmin) {
    const float c  = inv * (rsq - float(t.rmin2));
    const float ec = vabs(c) * rel + (ersq + 2.f * kU32 * (rsq + float(t.rmin2))) * vabs(inv);
    float dr, edr;
    const int h = CircleHitF(t, b, eb, c, ec, p, d, ep, false, false, dr, edr);
    if (h == kUnsure) return Undecided();
    if (h == kYes) {
      best  = fmini(best, fint(dr, edr));
      vbest = vmin(vbest, dr);
    }
  }
  {
    const float c  = inv * (rsq - float(t.rmax2));
    const float ec = vabs(c) * rel + (ersq + 2.f * kU32 * (rsq + float(t.rmax2))) * vabs(inv);
    float dr, edr;
    const int h = CircleHitF(t, b, eb, c, ec, p, d, ep, true, false, dr, edr);
    if (h == kUnsure) return Undecided();
    if (h == kYes) {
      best  = fmini(best, fint(dr, edr));
      vbest = vmin(vbest, dr);
    }
  }
  if (t.hasPhi) {
    float dp, edp;
    int h = PhiPlaneHitF(t, float(t.a1x), float(t.a1y), float(t.n1x), float(t.n1y), p, d, ep, false, dp, edp);
    if (h == kUnsure) return Undecided();
    if (h == kYes) {
      best  = fmini(best, fint(dp, edp));
      vbest = vmin(vbest, dp);
    }
    h = PhiPlaneHitF(t, float(t.a2x), float(t.a2y), float(t.n2x), float(t.n2y), p, d, ep, false, dp, edp);
    if (h == kUnsure) return Undecided();
    if (h == kYes) {
      best  = fmini(best, fint(dp, edp));
      vbest = vmin(vbest, dp);
    }
  }
  if (vbest >= FLT_MAX) return InfLen<float>();
  return best.hi - best.lo <= kMixDistErr ? vbest : Undecided();
}

// DistanceToIn with the far-point shift (TubeImplementation.h:475-502)
template <class R>
VGMI_DEV R DistanceToIn(TubeData const &t, Vec3<R> const &pt, Vec3<R> const &d)
{
  if constexpr (sizeof(R) == sizeof(float)) return DistanceToInF32(t, pt, d);
  Vec3<R> p = pt;
  R ptDist  = vsqrt(p.Mag2());
  R move    = R(0);
  if (ptDist > R(100.) * R(t.maxVal)) {
    move = ptDist - R(100.) * R(t.maxVal);
    p    = p + move * d;
  }
  return DistanceToInKernel(t, p, d) + move;
}

// DistanceToOut (TubeImplementation.h:686-821, UniversalTube phi branch)
template <class R>
VGMI_DEV R DistanceToOut(TubeData const &t, Vec3<R> const &p, Vec3<R> const &d)
{
  if constexpr (sizeof(R) == sizeof(float)) return DistanceToOutF32(t, p, d);
  const R z = R(t.z);
  R distz   = z - vabs(p.z);
  if (distz < R(-kHalfTolerance)) return R(-1);
  R rsq   = p.x * p.x + p.y * p.y;
  R rdotn = d.x * p.x + d.y * p.y;
  R crmax = rsq - R(t.rmax2);
  R crmin = rsq;
  if (crmax > R(2.0 * kTolerance * t.rmax)) return R(-1);
  if (t.hasRmin) {
    crmin -= R(t.rmin2);
    if (crmin < R(-2.0 * kTolerance * t.rmin)) return R(-1);
  }
  if (t.hasPhi) {
    bool wi, wo;
    WedgeContainsInside(t, p, true, wi, wo);
    if (wo) return R(-1);
  }
  R dist    = InfLen<R>();
  R invdirz = R(1) / NonZero(d.z);
  distz     = d.z < R(0) ? (-z - p.z) * invdirz : (z - p.z) * invdirz;
  if (d.z != R(0) && distz < dist) dist = distz;

  R invnsq = R(1) / NonZero(R(1) - d.z * d.z);
  R b      = invnsq * rdotn;
  if (t.hasRmin) {
    R dr;
    if (CircleHit(t, b, crmin * invnsq, p, d, false, false, dr) && dr < dist) dist = dr;
  }
  {
    R dr;
    if (CircleHit(t, b, crmax * invnsq, p, d, true, false, dr) && dr < dist) dist = dr;
  }
  if (t.hasPhi) {
    R dp;
    if (PhiPlaneHit(t, R(t.a1x), R(t.a1y), R(t.n1x), R(t.n1y), p, d, false, dp) && dp < dist) dist = dp;
    if (PhiPlaneHit(t, R(t.a2x), R(t.a2y), R(t.n2x), R(t.n2y), p, d, false, dp) && dp < dist) dist = dp;
  }
  return dist;
}

// SafetyToInOld (TubeImplementation.h:909-941) with PhiPlaneSafety<inside=false>
template <class R>
VGMI_DEV R SafetyToIn(TubeData const &t, Vec3<R> const &p)
{
  R safety = vabs(p.z) - R(t.z);
  R r      = vsqrt(p.x * p.x + p.y * p.y);
  R sr     = r - R(t.rmax);
  if (sr > safety) safety = sr;
  if (t.hasRmin) {
    R s = R(t.rmin) - r;
    if (s > safety) safety = s;
  }
  if (t.hasPhi) {
    if (InSector(t, p.x, p.y, false)) return safety;
    R sp   = t.phiGtPi ? r : InfLen<R>();
    R phi1 = p.x * R(t.a1y) - p.y * R(t.a1x);
    if (phi1 > R(-kHalfTolerance) && phi1 < sp) sp = phi1;
    R phi2 = -(p.x * R(t.a2y) - p.y * R(t.a2x));
    if (phi2 > R(-kHalfTolerance) && phi2 < sp) sp = phi2;
    if (sp < InfLen<R>() && sp > safety) safety = sp;
  }
  return safety;
}

// SafetyToOutOld (TubeImplementation.h:943-964) with Wedge::SafetyToOut
template <class R>
VGMI_DEV R SafetyToOut(TubeData const &t, Vec3<R> const &p)
{
  R safety = R(t.z) - vabs(p.z);
  R r      = vsqrt(p.x * p.x + p.y * p.y);
  R sr     = R(t.rmax) - r;
  if (sr < safety) safety = sr;
  if (t.hasRmin) {
    R s = r - R(t.rmin);
    if (s < safety) safety = s;
  }
  if (t.hasPhi) {
    R d1 = p.x * R(t.n1x) + p.y * R(t.n1y);
    R d2 = p.x * R(t.n2x) + p.y * R(t.n2y);
    R s  = t.wedgeLtPi ? vmin(d1, d2) : vmax(d1, d2);
    if (s < safety) safety = s;
  }
  return safety;
}

} // namespace tube
} // namespace vgmi
