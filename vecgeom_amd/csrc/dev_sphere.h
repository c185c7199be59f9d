// Sphere kernels: rmin/rmax shells, phi wedge and theta cones
// (VecGeom/volumes/kernel/SphereImplementation.h:65-474, VecGeom/volumes/SphereUtilities.h:55-195,
// VecGeom/volumes/ThetaCone_Evolution.h:150-788, VecGeom/volumes/Wedge_Evolution.h:235-436).
//
// Branch order, tolerance bands and early returns follow the scalar backend so FP64
// results match it; the same code instantiated with R = float is the bulk pass of
// the mixed-precision path. The wedge and theta data come from the raw constructor
// angles, the shell flags from the normalised ones, as in SphereStruct.
#pragma once
#include "dev_common.h"
#include "shape_data.h"

namespace vgmi {
namespace sph {

template <class R>
VGMI_DEV R PlusTolSq(R x)
{
  return x * (x + R(2.0 * kTolerance));
}
template <class R>
VGMI_DEV R MinusTolSq(R x)
{
  return vabs(x * (x - R(2.0 * kTolerance)));
}

// ---------------------------------------------------------------- wedge
template <class R>
VGMI_DEV bool WOnPlane(R ax, R ay, R nx, R ny, Vec3<R> const &p, R band = R(kTolerance))
{
  if (!(ax * p.x + ay * p.y >= R(0))) return false;
  return vabs(nx * p.x + ny * p.y) < band;
}

template <class R>
VGMI_DEV void WContainsInside(SphereData const &s, Vec3<R> const &p, bool forInside, bool &ci, bool &co)
{
  const R startCheck = (-p.x * R(s.w1y) + p.y * R(s.w1x));
  const R endCheck   = (-R(s.w2x) * p.y + R(s.w2y) * p.x);
  co                 = startCheck < R(0);
  if (s.wedgeLtPi)
    co = co || (endCheck < R(0));
  else
    co = co && (endCheck < R(0));
  if (forInside) {
    const bool on = WOnPlane(R(s.w1x), R(s.w1y), R(s.n1x), R(s.n1y), p) ||
                    WOnPlane(R(s.w2x), R(s.w2y), R(s.n2x), R(s.n2y), p);
    co = co && !on;
    ci = !on && !co;
  }
}

template <class R>
VGMI_DEV bool WContains(SphereData const &s, Vec3<R> const &p)
{
  bool ci = false, co = false;
  WContainsInside(s, p, false, ci, co);
  return !co;
}

template <class R>
VGMI_DEV bool WOnSurfaceMoving(SphereData const &s, bool start, bool out, Vec3<R> const &p, Vec3<R> const &d)
{
  const R ax = R(start ? s.w1x : s.w2x), ay = R(start ? s.w1y : s.w2y);
  const R nx = R(start ? s.n1x : s.n2x), ny = R(start ? s.n1y : s.n2y);
  if (!WOnPlane(ax, ay, nx, ny, p)) return false;
  const R c = d.x * -nx + d.y * -ny + d.z * -R(0);
  return out ? (c > R(0.005 * kHalfTolerance)) : (c < R(0.005 * kHalfTolerance));
}

template <class R>
VGMI_DEV void WDistanceToIn(SphereData const &s, Vec3<R> const &p, Vec3<R> const &d, R &d1, R &d2)
{
  d1             = InfLen<R>();
  d2             = InfLen<R>();
  const R comp1  = d.x * R(s.n1x) + d.y * R(s.n1y);
  const R comp2  = d.x * R(s.n2x) + d.y * R(s.n2y);
  if (comp1 > R(0)) {
    const R t = -(p.x * R(s.n1x) + p.y * R(s.n1y)) / comp1;
    if (t > R(0)) d1 = t;
  }
  if (comp2 > R(0)) {
    const R t = -(p.x * R(s.n2x) + p.y * R(s.n2y)) / comp2;
    if (t > R(0)) d2 = t;
  }
}

template <class R>
VGMI_DEV void WDistanceToOut(SphereData const &s, Vec3<R> const &p, Vec3<R> const &d, R &d1, R &d2)
{
  const R comp1 = d.x * R(s.n1x) + d.y * R(s.n1y);
  const R comp2 = d.x * R(s.n2x) + d.y * R(s.n2y);
  d1            = InfLen<R>();
  d2            = InfLen<R>();
  if (comp1 < R(0)) {
    const R t = -(p.x * R(s.n1x) + p.y * R(s.n1y)) / comp1;
    if (t > R(0)) d1 = t;
  }
  if (comp2 < R(0)) {
    const R t = -(p.x * R(s.n2x) + p.y * R(s.n2y)) / comp2;
    if (t > R(0)) d2 = t;
  }
}

template <class R>
VGMI_DEV R WSafetyToIn(SphereData const &s, Vec3<R> const &p)
{
  const R d1 = p.x * R(s.n1x) + p.y * R(s.n1y);
  const R d2 = p.x * R(s.n2x) + p.y * R(s.n2y);
  return s.wedgeLtPi ? vmax(-1 * d1, -1 * d2) : vmin(-1 * d1, -1 * d2);
}

template <class R>
VGMI_DEV R WSafetyToOut(SphereData const &s, Vec3<R> const &p)
{
  const R d1 = p.x * R(s.n1x) + p.y * R(s.n1y);
  const R d2 = p.x * R(s.n2x) + p.y * R(s.n2y);
  return s.wedgeLtPi ? vmin(d1, d2) : vmax(d1, d2);
}

// ---------------------------------------------------------------- theta cones
constexpr double kHP   = 0.5 * kPi;
constexpr double kHAT  = 0.5 * kTolerance; // half angular tolerance of the theta cones
constexpr double kTol20 = 2 * kTolerance * 10.;

template <class R>
VGMI_DEV bool TOnSurface(SphereData const &s, bool start, Vec3<R> const &p)
{
  const R rhs  = start ? vabs(R(s.tanS) * p.z) : vabs(R(s.tanE) * p.z);
  const R rho2 = p.Perp2();
  return rho2 >= MinusTolSq(rhs) && rho2 <= PlusTolSq(rhs);
}

template <class R>
VGMI_DEV bool TOnSurfaceMoving(SphereData const &s, bool start, bool out, Vec3<R> const &p, Vec3<R> const &d)
{
  if (!TOnSurface(s, start, p)) return false;
  Vec3<R> n;
  if (start) {
    n = Vec3<R>(p.x, p.y, -R(s.tanS2) * p.z);
    if (s.tS <= kHP) n = -n;
  } else {
    n = Vec3<R>(p.x, p.y, -R(s.tanE2) * p.z);
    if (!(s.tE <= kHP)) n = -n;
  }
  const R c = d.Dot(n);
  return out ? c > R(0) : c < R(0);
}

template <class R>
VGMI_DEV bool TCompletelyInside(SphereData const &s, Vec3<R> const &p)
{
  const double sT = s.tS, eT = s.tE;
  const R rho     = vsqrt(p.Mag2() - (p.z * p.z));
  const R c1r     = vabs(p.z * R(s.tanS));
  const R c2r     = vabs(p.z * R(s.tanE));
  const bool onZ  = p.z != R(0) && p.x == R(0) && p.y == R(0);
  const bool onXY = p.z == R(0) && (p.x != R(0) || p.y != R(0));
  bool ci = onZ && ((sT == 0. && eT == kPi) || (p.z > R(0) && sT == 0.) || (p.z < R(0) && eT == kPi));
  ci      = ci || (onXY && (sT < kHP && eT > kHP && (kHP - sT) > kTolerance && (eT - kHP) > kTolerance));
  if (sT < kHP + kHAT) {
    if (eT < kHP + kHAT && sT < eT) {
      const R mn = c1r + R(kTol20), mx = c2r - R(kTol20);
      ci = ci || (((rho <= mx) && (rho >= mn) && (p.z > R(0)) && (sT != 0.)) ||
                  ((rho <= mx) && (sT == 0.) && (p.z > R(0))));
    }
    if (eT > kHP + kHAT && sT < eT) {
      const R mn = c1r + R(kTol20), mx = c2r + R(kTol20);
      ci = ci || (((rho >= mn) && (p.z > R(0))) || ((rho >= mx) && (p.z < R(0))));
    }
    if (eT >= kHP - kHAT && eT <= kHP + kHAT) ci = ci && !(vabs(p.z) < R(kHAT));
  }
  if (eT > kHP + kHAT) {
    if (sT >= kHP - kHAT && sT <= kHP + kHAT) ci = ci && !(vabs(p.z) < R(kHAT));
    if (sT > kHP + kHAT && sT < eT) {
      const R mn = c1r - R(kTol20), mx = c2r + R(kTol20);
      ci = ci || (((rho <= mn) && (rho >= mx) && (p.z < R(0)) && (eT != kPi)) ||
                  ((rho <= mn) && (p.z < R(0)) && (eT == kPi)));
    }
  }
  return ci;
}

template <class R>
VGMI_DEV bool TCompletelyOutside(SphereData const &s, Vec3<R> const &p)
{
  const double sT = s.tS, eT = s.tE;
  const R rho     = vsqrt(p.Perp2());
  const R c1r     = vabs(p.z * R(s.tanS));
  const R c2r     = vabs(p.z * R(s.tanE));
  const bool onZ  = p.z != R(0) && p.x == R(0) && p.y == R(0);
  const bool onXY = p.z == R(0) && (p.x != R(0) || p.y != R(0));
  bool co = onZ && ((sT != 0. && eT != kPi) || (p.z > R(0) && sT != 0.) || (p.z < R(0) && eT != kPi));
  co = co || (onXY && (((sT < kHP) && (eT < kHP) && ((kHP - sT) > kTolerance) && ((kHP - eT) > kTolerance)) ||
                       ((sT > kHP && eT > kHP) && ((sT - kHP) > kTolerance) && ((eT - kHP) > kTolerance))));
  if (sT < kHP + kHAT) {
    if (eT < kHP + kHAT && sT < eT) {
      const R mn = c1r - R(kTol20), mx = c2r + R(kTol20);
      co = co || ((rho < mn) || (rho > mx) || (p.z < R(0)));
    }
    if (eT > kHP + kHAT && sT < eT) {
      const R mn = c1r - R(kTol20), mx = c2r - R(kTol20);
      co = co || (((rho < mn) && (p.z > R(0))) || ((rho < mx) && (p.z < R(0))));
    }
    if (eT >= kHP - kHAT && eT <= kHP + kHAT) co = co && !(vabs(p.z) < R(kHAT));
  }
  if (eT > kHP + kHAT) {
    if (sT >= kHP - kHAT && sT <= kHP + kHAT) co = co && !(vabs(p.z) < R(kHAT));
    if (sT > kHP + kHAT && sT < eT) {
      const R mn = c1r + R(kTol20), mx = c2r - R(kTol20);
      co = co || ((rho < mx) || (rho > mn) || (p.z > R(0)));
    }
  }
  return co;
}

template <class R>
VGMI_DEV bool TContains(SphereData const &s, Vec3<R> const &p)
{
  return !TCompletelyOutside(s, p);
}

template <class R>
VGMI_DEV R TDistanceToLine(double slope, R x, R y)
{
  const R dist = (y - R(slope) * x) / vsqrt(R(1.) + R(slope) * R(slope));
  return vabs(dist);
}

template <class R>
VGMI_DEV R TSafetyToIn(SphereData const &s, Vec3<R> const &p)
{
  const double sT = s.tS, eT = s.tE;
  const R rad     = vsqrt(p.x * p.x + p.y * p.y);
  const R s1      = TDistanceToLine(s.slope1, rad, p.z);
  const R s2      = TDistanceToLine(s.slope2, rad, p.z);
  R safe          = vmin(s1, s2);
  if (TContains(s, p)) return R(0);
  if (sT < kPi / 2 + kHAT) {
    if (eT < kPi / 2 + kHAT && sT < eT && p.z < R(0)) safe = s2;
    if (eT > kPi / 2 + kHAT && sT < eT) {
      if (p.z > R(0)) safe = s1;
      if (p.z < R(0)) safe = s2;
    }
  }
  if (eT > kPi / 2 + kHAT && sT > kPi / 2 + kHAT && sT < eT && p.z > R(0)) safe = s1;
  return safe;
}

template <class R>
VGMI_DEV R TSafetyToOut(SphereData const &s, Vec3<R> const &p)
{
  const double sT = s.tS, eT = s.tE;
  const R rad     = vsqrt(p.x * p.x + p.y * p.y);
  const R bis     = vabs(p.z * R(s.tanBis));
  bool cond       = false;
  const R s1      = TDistanceToLine(s.slope1, rad, p.z);
  const R s2      = TDistanceToLine(s.slope2, rad, p.z);
  if (sT < kPi / 2 + kHAT) {
    if (eT < kPi / 2 + kHAT && sT < eT) cond = (rad < bis) && (sT != 0.);
    if (eT > kPi / 2 + kHAT && sT < eT) cond = s1 < s2;
  }
  if (eT > kPi / 2 + kHAT && sT > kPi / 2 + kHAT && sT < eT) cond = !((rad < bis) && (eT != kPi));
  return cond ? s1 : s2;
}

template <class R>
VGMI_DEV void TDistanceToIn(SphereData const &s, Vec3<R> const &p, Vec3<R> const &d, R &d1, R &d2, bool &i1,
                            bool &i2)
{
  const double sT = s.tS, eT = s.tE;
  const R inf     = InfLen<R>();
  const R t3      = R(3. * kTolerance);
  bool done       = false;
  R r1 = inf, r2 = inf;
  const R pv   = p.x * d.x + p.y * d.y;
  const R rho2 = p.x * p.x + p.y * p.y;
  const R dr2  = d.Perp2();
  const R b    = pv - p.z * d.z * R(s.tanS2);
  const R a    = dr2 - d.z * d.z * R(s.tanS2);
  const R c    = rho2 - p.z * p.z * R(s.tanS2);
  const R dd   = b * b - a * c;
  const R ai   = R(1.) / NonZero(a);
  if (dd > R(0)) r1 = QuadRoot(b, dd, ai, c, +1);
  done = done || (vabs(r1) < t3);
  if (vabs(r1) < t3) r1 = R(0);
  if (!done && r1 < R(0)) r1 = inf;
  const R b2  = pv - p.z * d.z * R(s.tanE2);
  const R a2  = dr2 - d.z * d.z * R(s.tanE2);
  const R c2  = rho2 - p.z * p.z * R(s.tanE2);
  const R dd2 = b2 * b2 - a2 * c2;
  const R a2i = R(1.) / NonZero(a2);
  if (dd2 > R(0)) r2 = QuadRoot(b2, dd2, a2i, c2, -1);
  if (!done && vabs(r2) < t3) r2 = R(0);
  done = done || (vabs(r2) < t3);
  if (!done && r2 < R(0)) r2 = inf;

  if (sT < kHP + kHAT) {
    if (eT < kHP + kHAT && sT < eT) {
      d1         = r1;
      d2         = r2;
      const R z1 = p.z + d1 * d.z;
      const R z2 = p.z + d2 * d.z;
      i1         = (dd > R(0)) && (z1 > R(0));
      i2         = (dd2 > R(0)) && (z2 > R(0));
    }
    if (eT >= kHP - kHAT && eT <= kHP + kHAT) {
      if (d.z > R(0)) d2 = -p.z / d.z;
      const R z2 = p.z + d2 * d.z;
      i2         = (d2 != inf) && (vabs(z2) < HitTol<R>(kHAT));
    }
    if (eT > kHP + kHAT && sT < eT) {
      d1 = r1;
      if (dd2 > R(0)) r2 = QuadRoot(b2, dd2, a2i, c2, +1);
      done = vabs(r2) < t3;
      if (vabs(r2) < t3) r2 = R(0);
      if (!done && r2 < R(0)) r2 = inf;
      d2         = r2;
      const R z1 = p.z + d1 * d.z;
      const R z2 = p.z + d2 * d.z;
      i1         = (dd > R(0)) && (d1 != inf) && (z1 > R(0));
      i2         = (dd2 > R(0)) && (d2 != inf) && (z2 < R(0));
    }
  }
  if (sT >= kHP - kHAT && eT > kHP + kHAT && sT < eT) {
    if (dd > R(0)) r1 = QuadRoot(b, dd, ai, c, -1);
    done = vabs(r1) < t3;
    if (vabs(r1) < t3) r1 = R(0);
    if (!done && r1 < R(0)) r1 = inf;
    d1 = r1;
    if (dd2 > R(0)) r2 = QuadRoot(b2, dd2, a2i, c2, +1);
    done = vabs(r2) < t3;
    if (vabs(r2) < t3) r2 = R(0);
    if (!done && r2 < R(0)) r2 = inf;
    d2         = r2;
    const R z1 = p.z + d1 * d.z;
    const R z2 = p.z + d2 * d.z;
    i1         = (dd > R(0)) && (d1 != inf) && (z1 < R(0));
    i2         = (dd2 > R(0)) && (d2 != inf) && (z2 < R(0));
  }
  if (sT >= kHP - kHAT && sT <= kHP + kHAT) {
    if (d.z < R(0)) d1 = -p.z / d.z;
    const R z1 = p.z + d1 * d.z;
    i1         = (d1 != inf) && (vabs(z1) < HitTol<R>(kHAT));
  }
}

template <class R>
VGMI_DEV void TDistanceToOut(SphereData const &s, Vec3<R> const &p, Vec3<R> const &d, R &d1, R &d2, bool &i1,
                             bool &i2)
{
  const double sT = s.tS, eT = s.tE;
  const R inf     = InfLen<R>();
  const R hT      = R(kHalfTolerance);
  R r1 = inf, r2 = inf;
  const R pv   = p.x * d.x + p.y * d.y;
  const R rho2 = p.x * p.x + p.y * p.y;
  const R b    = pv - p.z * d.z * R(s.tanS2);
  const R a    = d.x * d.x + d.y * d.y - d.z * d.z * R(s.tanS2);
  const R c    = rho2 - p.z * p.z * R(s.tanS2);
  R dd         = b * b - a * c;
  if (dd < R(0) && vabs(dd) < hT) dd = R(0);
  if ((dd >= R(0)) && b >= R(0) && a != R(0)) r1 = (-b - vsqrt(vabs(dd))) / NonZero(a);
  if ((dd >= R(0)) && b < R(0)) r1 = c / NonZero(-b + vsqrt(vabs(dd)));
  if (r1 < R(0)) r1 = inf;
  const R b2 = p.x * d.x + p.y * d.y - p.z * d.z * R(s.tanE2);
  const R a2 = d.x * d.x + d.y * d.y - d.z * d.z * R(s.tanE2);
  const R c2 = p.x * p.x + p.y * p.y - p.z * p.z * R(s.tanE2);
  R dd2      = (b2 * b2) - (a2 * c2);
  if (dd2 < R(0) && vabs(dd2) < hT) dd2 = R(0);
  if ((dd2 >= R(0)) && b2 >= R(0)) r2 = c2 / NonZero(-b2 - vsqrt(vabs(dd2)));
  if ((dd2 >= R(0)) && b2 < R(0) && a2 != R(0)) r2 = (-b2 + vsqrt(vabs(dd2))) / NonZero(a2);
  if (r2 < R(0) && vabs(r2) > R(kTolerance)) r2 = inf;
  if (vabs(r2) < R(kTolerance)) r2 = R(0);

  if (sT < kPi / 2 + kHAT) {
    if (eT < kPi / 2 + kHAT && sT < eT) {
      d1         = r1;
      d2         = r2;
      const R z1 = p.z + d1 * d.z;
      const R z2 = p.z + d2 * d.z;
      i1         = (dd > R(0)) && (d1 != inf) && (z1 > -hT);
      i2         = (dd2 > R(0)) && (d2 != inf) && (z2 > -hT);
      const R dr2 = d.x * d.x + d.y * d.y;
      R zs = inf, ze = inf;
      if (sT != 0.) zs = dr2 / R(s.tanS);
      if (eT != 0.) ze = dr2 / R(s.tanE);
      const bool cond = (p.x == R(0) && p.y == R(0) && p.z == R(0) && d.z < zs && d.z < ze);
      if (cond) d1 = R(0), d2 = R(0);
      i1 = i1 || cond;
      i2 = i2 || cond;
    }
    if (eT >= kPi / 2 - kHAT && eT <= kPi / 2 + kHAT) {
      d1 = r1;
      d2 = inf;
      if (d.z < R(0)) d2 = R(-1.) * p.z / d.z;
      const R z2 = p.z + d2 * d.z;
      const R z1 = p.z + d1 * d.z;
      i2         = (d2 != inf) && (vabs(z2) < HitTol<R>(kHalfTolerance)) && !(d.z == R(0));
      i1         = (dd >= R(0)) && (d1 != inf) && (vabs(z1) < hT) && !(d.z == R(0));
    }
    if (eT > kPi / 2 + kHAT && sT < eT) {
      d1 = r1;
      if ((dd2 >= R(0)) && b2 > R(0) && a2 != R(0)) r2 = (-b2 - vsqrt(vabs(dd2))) / NonZero(a2);
      if ((dd2 >= R(0)) && b2 <= R(0)) r2 = c2 / NonZero(-b2 + vsqrt(vabs(dd2)));
      if (r2 < R(0)) r2 = inf;
      d2         = r2;
      const R z1 = p.z + d1 * d.z;
      const R z2 = p.z + d2 * d.z;
      i1         = (dd >= R(0)) && (d1 != inf) && (z1 > -hT);
      i2         = (dd2 >= R(0)) && (d2 != inf) && (z2 < hT);
    }
  }
  if (eT > kPi / 2 + kHAT && sT < eT) {
    r2 = inf;
    if ((dd2 >= R(0)) && b2 > R(0) && a2 != R(0)) r2 = (-b2 - vsqrt(vabs(dd2))) / NonZero(a2);
    if ((dd2 >= R(0)) && b2 <= R(0)) r2 = c2 / NonZero(-b2 + vsqrt(vabs(dd2)));
    d2 = r2;
    if (sT > kPi / 2 + kHAT) {
      if ((dd >= R(0)) && b > R(0)) r1 = c / NonZero(-b - vsqrt(vabs(dd)));
      if ((dd >= R(0)) && b <= R(0) && a != R(0)) r1 = (-b + vsqrt(vabs(dd))) / NonZero(a);
      if (r1 < R(0)) r1 = inf;
      d1          = r1;
      const R z1  = p.z + d1 * d.z;
      i1          = (dd > R(0)) && (d1 != inf) && (z1 < hT);
      const R z2  = p.z + d2 * d.z;
      i2          = (dd2 > R(0)) && (d2 != inf) && (z2 < hT);
      const R dr2 = d.x * d.x + d.y * d.y;
      R zs = -inf, ze = -inf;
      if (s.tanS != 0.) zs = -dr2 / R(s.tanS);
      if (s.tanE != 0.) ze = -dr2 / R(s.tanE);
      const bool cond = (p.x == R(0) && p.y == R(0) && p.z == R(0) && d.z > zs && d.z > ze);
      if (cond) d1 = R(0), d2 = R(0);
      i1 = i1 || cond;
      i2 = i2 || cond;
    }
  }
  if (sT >= kPi / 2 - kHAT && sT <= kPi / 2 + kHAT) {
    d2 = r2;
    d1 = inf;
    if (d.z > R(0)) d1 = R(-1.) * p.z / NonZero(d.z);
    const R z1 = p.z + d1 * d.z;
    const R z2 = p.z + d2 * d.z;
    i1         = (d1 != inf) && (vabs(z1) < HitTol<R>(kHalfTolerance)) && (d.z != R(0));
    i2         = (dd2 >= R(0)) && (d2 != inf) && (vabs(z2) < hT) && (d.z != R(0));
  }
}

// ---------------------------------------------------------------- sphere
template <class R>
VGMI_DEV void ContainsInside(SphereData const &s, Vec3<R> const &p, bool forInside, bool &ci, bool &co)
{
  const R r2   = p.Mag2();
  const R rmin = R(s.rmin), rmax = R(s.rmax);
  if (s.rmin != 0.)
    ci = r2 <= MinusTolSq(rmax) && r2 >= PlusTolSq(rmin);
  else
    ci = r2 <= MinusTolSq(rmax);
  if (s.rmin != 0.)
    co = r2 >= PlusTolSq(rmax) || r2 <= MinusTolSq(rmin);
  else
    co = r2 >= PlusTolSq(rmax);
  if (!s.fullPhi) {
    bool cip = false, cop = false;
    WContainsInside(s, p, forInside, cip, cop);
    co = co || cop;
    if (forInside) ci = ci && cip;
  }
  if (!s.fullTheta) {
    const bool cit = forInside ? TCompletelyInside(s, p) : false;
    const bool cot = TCompletelyOutside(s, p);
    co             = co || cot;
    if (forInside) ci = ci && cit;
  }
}

template <class R>
VGMI_DEV bool CompletelyOutside(SphereData const &s, Vec3<R> const &p)
{
  const R rad     = vsqrt(p.Mag2());
  const bool outR = (rad > (R(s.rmax) + R(kTolerance))) || (rad < (R(s.rmin) - R(kTolerance)));
  bool inPhi = false, outPhi = false;
  WContainsInside(s, p, true, inPhi, outPhi);
  return outR || outPhi || TCompletelyOutside(s, p);
}

template <class R>
VGMI_DEV bool CompletelyInside(SphereData const &s, Vec3<R> const &p)
{
  const R rad    = vsqrt(p.Mag2());
  const bool inR = (rad < (R(s.rmax) - R(kTolerance))) && (rad > (R(s.rmin) + R(kTolerance)));
  bool inPhi = false, outPhi = false;
  WContainsInside(s, p, true, inPhi, outPhi);
  return inR && inPhi && TCompletelyInside(s, p);
}

template <class R>
VGMI_DEV bool OnSurfaceMoving(SphereData const &s, bool out, Vec3<R> const &p, Vec3<R> const &d)
{
  const R rmin = R(s.rmin), rmax = R(s.rmax);
  const R m2   = p.Mag2();
  const R sqt  = R(3.1622777e-5); // kSqrtTolerance, base/Math.h:29
  const bool onOuter = m2 <= PlusTolSq(rmax) && m2 >= MinusTolSq(rmax);
  const R dp         = d.Dot(p);
  bool outerR        = onOuter && (out ? (dp > sqt) : (dp < -sqt));
  bool innerR = false, sPhi = false, ePhi = false, sTh = false, eTh = false;
  if (s.rmin != 0.) {
    const bool onInner = m2 <= PlusTolSq(rmin) && m2 >= MinusTolSq(rmin);
    const R dmp        = d.Dot(-p);
    innerR             = onInner && (out ? (dmp > sqt) : (dmp < -sqt));
  }
  if (s.dphi < (kTwoPi - kHalfTolerance)) {
    sPhi = WOnSurfaceMoving(s, true, out, p, d);
    ePhi = WOnSurfaceMoving(s, false, out, p, d);
  }
  if (s.dtheta < (kPi - kHalfTolerance)) {
    sTh = TOnSurfaceMoving(s, true, out, p, d);
    eTh = TOnSurfaceMoving(s, false, out, p, d);
  }
  return ((outerR || innerR) && WContains(s, p) && TContains(s, p)) ||
         ((sPhi || ePhi) && (m2 >= rmin * rmin) && (m2 <= rmax * rmax) && TContains(s, p)) ||
         ((sTh || eTh) && (m2 >= rmin * rmin) && (m2 <= rmax * rmax) && WContains(s, p));
}

template <class R>
VGMI_DEV void MinDistFromPhi(SphereData const &s, bool toIn, Vec3<R> const &p, Vec3<R> const &d, R &distance)
{
  const R rmin = R(s.rmin), rmax = R(s.rmax);
  R dp1, dp2;
  if (toIn)
    WDistanceToIn(s, p, d, dp1, dp2);
  else
    WDistanceToOut(s, p, d, dp1, dp2);
  R dist    = vmin(dp1, dp2);
  Vec3<R> t = p + d * dist;
  R r2      = t.Mag2();
  const R band = HitTol<R>(kTolerance);
  bool tc      = ((dist == dp1) && WOnPlane(R(s.w1x), R(s.w1y), R(s.n1x), R(s.n1y), t, band)) ||
            ((dist == dp2) && WOnPlane(R(s.w2x), R(s.w2y), R(s.n2x), R(s.n2y), t, band));
  const bool c1 = tc && (r2 > rmin * rmin) && (r2 < rmax * rmax) && TContains(s, t);
  if (c1) distance = vmin(dist, distance);
  dist = vmax(dp1, dp2);
  t    = p + d * dist;
  r2   = t.Mag2();
  tc   = ((dist == dp1) && WOnPlane(R(s.w1x), R(s.w1y), R(s.n1x), R(s.n1y), t, band)) ||
       ((dist == dp2) && WOnPlane(R(s.w2x), R(s.w2y), R(s.n2x), R(s.n2y), t, band));
  const bool c2 = tc && (r2 > rmin * rmin) && (r2 < rmax * rmax) && TContains(s, t);
  if (!c1 && c2) distance = vmin(dist, distance);
}

template <class R>
VGMI_DEV bool Contains(SphereData const &s, Vec3<R> const &p)
{
  bool ci = false, co = false;
  ContainsInside(s, p, false, ci, co);
  return !co;
}

template <class R>
VGMI_DEV int Inside(SphereData const &s, Vec3<R> const &p)
{
  bool ci = false, co = false;
  ContainsInside(s, p, true, ci, co);
  int r = kSurface;
  if (co) r = kOutside;
  if (ci) r = kInside;
  return r;
}

template <class R>
VGMI_DEV R DistanceToIn(SphereData const &s, Vec3<R> const &p, Vec3<R> const &d)
{
  const R inf  = InfLen<R>();
  const R rmin = R(s.rmin), rmax = R(s.rmax);
  const R rad2 = p.Mag2();
  const R pv   = p.Dot(d);
  R c          = rad2 - rmax * rmax;
  if (CompletelyInside(s, p)) return R(-1);
  if (OnSurfaceMoving(s, false, p, d)) return R(0);
  R sd1 = inf, sd2 = inf;
  R d2  = (pv * pv - c);
  if (d2 < R(0) || ((c > R(0)) && (pv > R(0)))) return inf;
  if (d2 >= R(0)) sd1 = (-pv - vsqrt(vabs(d2)));
  R outerDist = inf, innerDist = inf;
  if (s.full) {
    if (sd1 >= R(0)) outerDist = sd1;
  } else {
    const Vec3<R> t = p + d * sd1;
    if (WContains(s, t) && TContains(s, t) && (sd1 >= R(0))) outerDist = sd1;
  }
  if (s.rmin != 0.) {
    c  = rad2 - rmin * rmin;
    d2 = pv * pv - c;
    if (d2 >= R(0)) sd2 = QuadRoot(pv, d2, R(1), c, +1);
    if (s.full) {
      if (sd2 >= R(0)) innerDist = sd2;
    } else {
      const Vec3<R> t = p + d * sd2;
      if ((sd2 >= R(0)) && WContains(s, t) && TContains(s, t)) innerDist = sd2;
    }
  }
  R distance = vmin(outerDist, innerDist);
  if (!s.fullPhi) MinDistFromPhi(s, true, p, d, distance);
  R dThetaMin = inf;
  if (!s.fullTheta) {
    bool i1 = false, i2 = false;
    R dt1 = inf, dt2 = inf;
    TDistanceToIn(s, p, d, dt1, dt2, i1, i2);
    const Vec3<R> q1 = p + d * dt1;
    const R m1       = q1.Mag2();
    const Vec3<R> q2 = p + d * dt2;
    const R m2       = q2.Mag2();
    bool v1          = (m1 >= rmin * rmin && m1 <= rmax * rmax) && i1;
    bool v2          = (m2 >= rmin * rmin && m2 <= rmax * rmax) && i2;
    if (!s.fullPhi) {
      v1 = v1 && WContains(s, q1);
      v2 = v2 && WContains(s, q2);
    }
    if (v2 && !v1) dThetaMin = dt2;
    if (v1 && !v2) dThetaMin = dt1;
    if (v1 && v2) dThetaMin = vmin(dt1, dt2);
  }
  distance         = vmin(dThetaMin, distance);
  const Vec3<R> dd = Vec3<R>(R(0), R(0), R(0)) - p;
  const R m2       = dd.Mag2();
  const R nd       = vsqrt(m2);
  const R inv      = vsqrt(m2 + R(DBL_MIN)); // Vector3D::Unit, base/Vector3D.h:284-290
  const R ux = dd.x / inv, uy = dd.y / inv, uz = dd.z / inv;
  const R ht = R(kHalfTolerance);
  if ((s.stheta > kHalfTolerance || s.etheta < (kPi - kHalfTolerance)) && (vabs(ux - d.x) < ht) &&
      (vabs(uy - d.y) < ht) && (vabs(uz - d.z) < ht))
    distance = vmin(distance, nd);
  return distance;
}

template <class R>
VGMI_DEV R DistanceToOut(SphereData const &s, Vec3<R> const &p, Vec3<R> const &d)
{
  const R inf  = InfLen<R>();
  const R rmin = R(s.rmin), rmax = R(s.rmax);
  const R pv   = p.Dot(d);
  const R rad2 = p.Mag2();
  R c          = rad2 - rmax * rmax;
  R sd1 = inf, sd2 = inf;
  if (CompletelyOutside(s, p)) return R(-1);
  if (OnSurfaceMoving(s, true, p, d)) return R(0);
  R d2 = (pv * pv - c);
  if (d2 >= R(0)) sd1 = QuadRoot(pv, d2, R(1), c, +1);
  if (s.rmin != 0.) {
    c  = rad2 - rmin * rmin;
    d2 = (pv * pv - c);
    if ((d2 >= R(0)) && (pv < R(0))) sd2 = (-pv - vsqrt(vabs(d2)));
  }
  const R snxt    = vmin(sd1, sd2);
  const bool semi = ((s.stheta == 0. && s.etheta == kPi / 2.) && d.z >= R(0)) ||
                    ((s.stheta == kPi / 2. && s.etheta == kPi) && d.z <= R(0));
  if (semi) return snxt;
  R dThetaMin = inf;
  if (!s.fullTheta) {
    bool i1 = false, i2 = false;
    R dt1 = inf, dt2 = inf;
    TDistanceToOut(s, p, d, dt1, dt2, i1, i2);
    if (i2 && !i1) dThetaMin = dt2;
    if (!i2 && i1) dThetaMin = dt1;
    if (i2 && i1) dThetaMin = vmin(dt1, dt2);
  }
  R distance = vmin(dThetaMin, snxt);
  if (!s.fullPhi) {
    if (s.dphi <= kPi) {
      R dp1, dp2;
      WDistanceToOut(s, p, d, dp1, dp2);
      distance = vmin(vmin(dp1, dp2), distance);
    } else {
      MinDistFromPhi(s, false, p, d, distance);
    }
  }
  return distance;
}

template <class R>
VGMI_DEV R SafetyToIn(SphereData const &s, Vec3<R> const &p)
{
  const R rad = vsqrt(p.Mag2());
  bool ci = false, co = false;
  ContainsInside(s, p, true, ci, co);
  if (ci) return R(-1);
  if (!ci && !co) return R(0);
  R safety;
  if (s.rmin != 0.) {
    const R sMin = R(s.rmin) - rad, sMax = rad - R(s.rmax);
    safety = (sMin > sMax) ? sMin : sMax;
  } else {
    safety = rad - R(s.rmax);
  }
  if (!s.fullPhi) safety = vmax(WSafetyToIn(s, p), safety);
  if (!s.fullTheta) safety = vmax(TSafetyToIn(s, p), safety);
  return safety;
}

template <class R>
VGMI_DEV R SafetyToOut(SphereData const &s, Vec3<R> const &p)
{
  const R rad = vsqrt(p.Mag2());
  bool ci = false, co = false;
  ContainsInside(s, p, true, ci, co);
  if (co) return R(-1);
  if (!ci && !co) return R(0);
  R safety;
  if (s.rmin != 0.) {
    const R sMin = rad - R(s.rmin), sMax = R(s.rmax) - rad;
    safety = (sMin < sMax) ? sMin : sMax;
  } else {
    safety = R(s.rmax) - rad;
  }
  if (!s.fullPhi) safety = vmin(WSafetyToOut(s, p), safety);
  if (!s.fullTheta) safety = vmin(TSafetyToOut(s, p), safety);
  return safety;
}

} // namespace sph
} // namespace vgmi
