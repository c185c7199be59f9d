// Torus2 kernels (VecGeom/volumes/kernel/TorusImplementation2.h:14-660): quartic roots via the
// resolvent cubic, Newton polishing to kTolerance, bounding-tube pre-step with the UniversalTube
// kernels (the VECGEOM_NO_SPECIALIZATION branch of DistanceToIn).
//
// In the mixed path the FP32 pass answers only DistanceToOut's early exits (point outside) and
// defers every quartic solve, and all of DistanceToIn, to the FP64 evaluation.
//
// The cubic uses the device math library's cbrt, pow, atan2 and cos where the scalar backend
// uses the host C library's; the Newton polishing that follows makes both converge to the same
// root within kTolerance, but the last bits may differ from a host evaluation.
#pragma once
#include "dev_common.h"
#include "dev_tube.h"
#include "shape_data.h"

namespace vgmi {
namespace tor {

template <class R>
VGMI_DEV unsigned SolveCubic(R a, R b, R c, R *x)
{
  const R ott     = R(1. / 3.);
  const R sq3     = vsqrt(R(3.));
  const R inv6sq3 = R(1.) / (R(6.) * sq3);
  unsigned ireal  = 1;
  R p             = b - a * a * ott;
  R q             = c - a * b * ott + R(2.) * a * a * a * ott * ott * ott;
  R delta         = R(4) * p * p * p + R(27.) * q * q;
  R t, u;
  if (delta >= R(0)) {
    delta = vsqrt(delta);
    t     = (R(-3) * q * sq3 + delta) * inv6sq3;
    u     = (R(3) * q * sq3 + delta) * inv6sq3;
    x[0]  = vcopysign(R(1.), t) * vcbrt(vabs(t)) - vcopysign(R(1.), u) * vcbrt(vabs(u)) - a * ott;
  } else {
    delta = vsqrt(-delta);
    t     = R(-0.5) * q;
    u     = delta * inv6sq3;
    x[0]  = R(2.) * vpow(t * t + u * u, R(0.5) * ott) * vcos(ott * vatan2(u, t));
    x[0] -= a * ott;
  }
  t     = x[0] * x[0] + a * x[0] + b;
  u     = a + x[0];
  delta = u * u - R(4.) * t;
  if (delta >= R(0)) {
    ireal = 3;
    delta = vsqrt(delta);
    x[1]  = R(0.5) * (-u - delta);
    x[2]  = R(0.5) * (-u + delta);
  }
  return ireal;
}

template <class R>
VGMI_DEV void CmpAndSwap(R *a, int i, int j)
{
  if (a[i] > a[j]) {
    const R c = a[j];
    a[j]      = a[i];
    a[i]      = c;
  }
}
template <class R>
VGMI_DEV void Sort4(R *a)
{
  CmpAndSwap(a, 0, 2);
  CmpAndSwap(a, 1, 3);
  CmpAndSwap(a, 0, 1);
  CmpAndSwap(a, 2, 3);
  CmpAndSwap(a, 1, 2);
}

template <class R>
VGMI_DEV int SolveQuartic(R a, R b, R c, R d, R *x)
{
  R e     = b - R(3.) * a * a / R(8.);
  R f     = c + a * a * a / R(8.) - R(0.5) * a * b;
  R g     = d - R(3.) * a * a * a * a / R(256.) + a * a * b / R(16.) - a * c / R(4.);
  R xx[4] = {InfLen<R>(), InfLen<R>(), InfLen<R>(), InfLen<R>()};
  R delta;
  R h            = R(0.);
  unsigned ireal = 0;
  if (vabs(f) < R(1E-6)) {
    delta = e * e - R(4.) * g;
    if (delta < R(0.)) return 0;
    delta = vsqrt(delta);
    h     = R(0.5) * (-e - delta);
    if (h >= R(0)) {
      h          = vsqrt(h);
      x[ireal++] = -h - R(0.25) * a;
      x[ireal++] = h - R(0.25) * a;
    }
    h = R(0.5) * (-e + delta);
    if (h >= R(0)) {
      h          = vsqrt(h);
      x[ireal++] = -h - R(0.25) * a;
      x[ireal++] = h - R(0.25) * a;
    }
    Sort4(x);
    return ireal;
  }
  if (vabs(g) < R(1E-6)) {
    x[ireal++]     = R(-0.25) * a;
    unsigned ncube = SolveCubic(R(0), e, f, xx);
    for (unsigned i = 0; i < ncube; i++) x[ireal++] = xx[i] - R(0.25) * a;
    Sort4(x);
    return ireal;
  }
  ireal = SolveCubic(R(2.) * e, e * e - R(4.) * g, -f * f, xx);
  if (ireal == 1) {
    if (xx[0] <= R(0)) return 0;
    h = vsqrt(xx[0]);
  } else {
    for (unsigned i = 0; i < 3; i++) {
      h = xx[i];
      if (h >= R(0)) break;
    }
    if (h <= R(0)) return 0;
    h = vsqrt(h);
  }
  R j   = R(0.5) * (e + h * h - f / h);
  ireal = 0;
  delta = h * h - R(4.) * j;
  if (delta >= R(0)) {
    delta      = vsqrt(delta);
    x[ireal++] = R(0.5) * (-h - delta) - R(0.25) * a;
    x[ireal++] = R(0.5) * (-h + delta) - R(0.25) * a;
  }
  delta = h * h - R(4.) * g / j;
  if (delta >= R(0)) {
    delta      = vsqrt(delta);
    x[ireal++] = R(0.5) * (h - delta) - R(0.25) * a;
    x[ireal++] = R(0.5) * (h + delta) - R(0.25) * a;
  }
  Sort4(x);
  return ireal;
}

// Wedge::GenericKernelForContainsAndInside on the torus wedge
template <class R>
VGMI_DEV void WedgeContainsInside(TorusData const &t, Vec3<R> const &p, bool forInside, bool &ci, bool &co)
{
  const R startCheck = (-p.x * R(t.w1y) + p.y * R(t.w1x));
  const R endCheck   = (-R(t.w2x) * p.y + R(t.w2y) * p.x);
  co                 = startCheck < R(0);
  if (t.wedgeLtPi)
    co = co || (endCheck < R(0));
  else
    co = co && (endCheck < R(0));
  if (forInside) {
    const bool on = tube::OnWedgePlane(R(t.w1x), R(t.w1y), R(t.n1x), R(t.n1y), p) ||
                    tube::OnWedgePlane(R(t.w2x), R(t.w2y), R(t.n2x), R(t.n2y), p);
    co = co && !on;
    ci = !on && !co;
  }
}

// TorusImplementation2.h:317-365
template <class R>
VGMI_DEV void ContainsInside(TorusData const &t, Vec3<R> const &p, bool forInside, bool notForDisk, bool &ci, bool &co)
{
  const R tol   = R(100. * kTolerance);
  const R rtor  = R(t.rtor);
  const R rxy   = vsqrt(p.x * p.x + p.y * p.y);
  const R radsq = (rxy - rtor) * (rxy - rtor) + p.z * p.z;
  if (forInside) {
    co = radsq > (tol * R(t.rmax) + R(t.rmax2));
    ci = radsq < (-tol * R(t.rmax) + R(t.rmax2));
  } else {
    co = radsq > R(t.rmax2);
  }
  if (co) return;
  if (forInside) {
    co = co || radsq < (-tol * R(t.rmin) + R(t.rmin2));
    ci = ci && radsq > (tol * R(t.rmin) + R(t.rmin2));
  } else {
    co = co || radsq < R(t.rmin2);
  }
  if (co) return;
  if (t.hasPhi && notForDisk) {
    bool cip = false, cop = false;
    WedgeContainsInside(t, p, forInside, cip, cop);
    co = co || cop;
    if (forInside) ci = ci && cip;
  }
}

template <class R>
VGMI_DEV bool Contains(TorusData const &t, Vec3<R> const &p)
{
  bool ci = false, co = false;
  ContainsInside(t, p, true, false, ci, co);
  return !co;
}

template <class R>
VGMI_DEV int Inside(TorusData const &t, Vec3<R> const &p)
{
  bool ci = false, co = false;
  ContainsInside(t, p, true, true, ci, co);
  int r = kSurface;
  if (co) r = kOutside;
  if (ci) r = kInside;
  return r;
}

// TorusImplementation2.h:367-493. Out of line: the quartic solve with the math library's cbrt,
// pow, cos and atan2 needs far more registers than any other shape kernel, and inlined at its
// four call sites it would set the register budget of every kernel that also holds the torus.
template <class R>
__device__ __noinline__ R ToBoundary(TorusData const &t, Vec3<R> const &pt, Vec3<R> const &dir, double radiusD, bool out)
{
  const R radius = R(radiusD);
  const R rtor2  = R(t.rtor2);
  R s            = InfLen<R>();
  const R tol    = R(100. * kTolerance);
  const R r0sq   = pt.x * pt.x + pt.y * pt.y + pt.z * pt.z;
  const R rdn    = pt.x * dir.x + pt.y * dir.y + pt.z * dir.z;
  const R rsum   = rtor2 + radius * radius;
  const R a      = R(4.) * rdn;
  const R b      = R(2.) * (r0sq + R(2.) * rdn * rdn - rsum + R(2.) * rtor2 * dir.z * dir.z);
  const R c      = R(4.) * (r0sq * rdn - rsum * rdn + R(2.) * rtor2 * pt.z * dir.z);
  const R d      = r0sq * r0sq - R(2.) * r0sq * rsum + R(4.) * rtor2 * pt.z * pt.z +
              (rtor2 - radius * radius) * (rtor2 - radius * radius);
  R x[4]   = {InfLen<R>(), InfLen<R>(), InfLen<R>(), InfLen<R>()};
  int nsol = 0;
  if (vabs(dir.z) < R(1E-3) && vabs(pt.z) < R(0.1) * radius) {
    R r0              = R(t.rtor) - vsqrt((radius - pt.z) * (radius + pt.z));
    const R invdirxy2 = R(1.) / (R(1) - dir.z * dir.z);
    const R b0        = (pt.x * dir.x + pt.y * dir.y) * invdirxy2;
    R c0              = (pt.x * pt.x + (pt.y - r0) * (pt.y + r0)) * invdirxy2;
    R delta           = b0 * b0 - c0;
    if (delta > R(0)) {
      x[nsol] = -b0 - vsqrt(delta);
      if (x[nsol] > -tol) nsol++;
      x[nsol] = -b0 + vsqrt(delta);
      if (x[nsol] > -tol) nsol++;
    }
    r0    = R(t.rtor) + vsqrt((radius - pt.z) * (radius + pt.z));
    c0    = (pt.x * pt.x + (pt.y - r0) * (pt.y + r0)) * invdirxy2;
    delta = b0 * b0 - c0;
    if (delta > R(0)) {
      x[nsol] = -b0 - vsqrt(delta);
      if (x[nsol] > -tol) nsol++;
      x[nsol] = -b0 + vsqrt(delta);
      if (x[nsol] > -tol) nsol++;
    }
    if (nsol) Sort4(x);
  } else {
    nsol = SolveQuartic(a, b, c, d, x);
  }
  if (!nsol) return InfLen<R>();
  const bool inner = vabs(radiusD - t.rmin) < kTolerance;
  for (int i = 0; i < nsol; i++) {
    if (x[i] < R(-10)) continue;
    Vec3<R> r0   = pt + dir * x[i];
    Vec3<R> norm = r0;
    r0.z         = R(0.);
    r0           = r0 * (R(1.) / vsqrt(r0.Mag2()));
    r0           = r0 * R(t.rtor);
    norm         = norm - r0;
    const R ndotd = norm.Dot(dir);
    if (inner ^ out) {
      if (ndotd < R(0)) continue;
    } else {
      if (ndotd > R(0)) continue;
    }
    if (t.hasPhi) {
      bool ci = false, co = false;
      WedgeContainsInside(t, r0, true, ci, co);
      if (co) continue;
    }
    s          = x[i];
    R eps      = InfLen<R>();
    R delta    = s * s * s * s + a * s * s * s + b * s * s + c * s + d;
    R eps0     = -delta / (R(4.) * s * s * s + R(3.) * a * s * s + R(2.) * b * s + c);
    int ntry   = 0;
    while (vabs(eps) > R(kTolerance)) {
      if (vabs(eps0) > R(100)) break;
      s += eps0;
      if (vabs(s + eps0) < R(kTolerance)) break;
      delta = s * s * s * s + a * s * s * s + b * s * s + c * s + d;
      eps   = -delta / (R(4.) * s * s * s + R(3.) * a * s * s + R(2.) * b * s + c);
      if (vabs(eps) >= vabs(eps0)) break;
      ntry++;
      if (ntry > 100) break;
      eps0 = eps;
    }
    if (s < -tol) continue;
    return vmax(R(0.), s);
  }
  return InfLen<R>();
}

// TorusImplementation2.h:199-299
template <class R>
VGMI_DEV R DistanceToOut(TorusData const &t, Vec3<R> const &p, Vec3<R> const &d)
{
  const R hT = R(kHalfTolerance);
  bool done  = false;
  done       = done || (vabs(p.z) - R(t.rmax)) > hT;
  const R rsq = p.x * p.x + p.y * p.y;
  const R outerExcl = R(t.rtor + t.rmax + kHalfTolerance);
  done              = done || rsq > outerExcl * outerExcl;
  const R innerExcl = R(t.rtor - t.rmax - kHalfTolerance);
  done              = done || rsq < innerExcl * innerExcl;
  if (done) return R(-1.);
  if (Inside(t, p) == kOutside) return R(-1.);
  // FP32 pass of the mixed path: the quartic is too ill-conditioned in FP32 (coefficients
  // scale with rtor^4); the root solve and its Newton polishing are left to FP64 (NaN defers).
  if (sizeof(R) == sizeof(float)) return R(__builtin_nanf(""));
  const R dout = ToBoundary(t, p, d, t.rmax, true);
  R din        = InfLen<R>();
  if (t.rmin > 0) din = ToBoundary(t, p, d, t.rmin, true);
  R distance = vmin(dout, din);
  if (t.hasPhi) {
    // Wedge::DistanceToOut
    const R comp1 = d.x * R(t.n1x) + d.y * R(t.n1y);
    const R comp2 = d.x * R(t.n2x) + d.y * R(t.n2y);
    R dp1 = InfLen<R>(), dp2 = InfLen<R>();
    if (comp1 < R(0)) {
      const R tt = -(p.x * R(t.n1x) + p.y * R(t.n1y)) / comp1;
      if (tt > R(0)) dp1 = tt;
    }
    if (comp2 < R(0)) {
      const R tt = -(p.x * R(t.n2x) + p.y * R(t.n2y)) / comp2;
      if (tt > R(0)) dp2 = tt;
    }
    bool smaller = dp1 < distance;
    if (smaller) {
      const Vec3<R> ip = p + d * dp1;
      bool ci = false, co = false;
      ContainsInside(t, ip, false, false, ci, co);
      if (!co) {
        const R diri = ip.x * R(t.w1x) + ip.y * R(t.w1y);
        if (diri >= R(0) && smaller) distance = dp1;
      }
    }
    smaller = dp2 < distance;
    if (smaller) {
      const Vec3<R> ip = p + d * dp2;
      bool ci = false, co = false;
      ContainsInside(t, ip, false, false, ci, co);
      if (!co) {
        const R diri2 = ip.x * R(t.w2x) + ip.y * R(t.w2y);
        if (diri2 >= R(0) && (dp2 < distance) && smaller) distance = dp2;
      }
    }
  }
  if (distance >= InfLen<R>()) distance = R(-1.);
  return distance;
}

template <class R>
VGMI_DEV R DistSqrToTorusR(TorusData const &t, Vec3<R> const &point, Vec3<R> const &dir, R dist)
{
  const Vec3<R> p = point + dir * dist;
  const R rxy     = vsqrt(p.x * p.x + p.y * p.y);
  return (rxy - R(t.rtor)) * (rxy - R(t.rtor)) + p.z * p.z;
}

// TorusImplementation2.h:538-628
template <class R>
VGMI_DEV R DistanceToIn(TorusData const &t, Vec3<R> const &point, Vec3<R> const &dir)
{
  // FP32 pass of the mixed path: the bounding tube is tangent to the torus, so its FP32
  // pre-test can miss where FP64 hits; the whole query is deferred to FP64 (NaN).
  if (sizeof(R) == sizeof(float)) return R(__builtin_nanf(""));
  Vec3<R> lp          = point;
  R distance          = InfLen<R>();
  R tubeDist          = InfLen<R>();
  const bool inBounds = tube::Contains(t.bounding, lp);
  if (!inBounds)
    tubeDist = tube::DistanceToIn(t.bounding, lp, dir);
  else
    tubeDist = R(0.);
  if (inBounds) {
    if (Inside(t, point) == kInside) return R(-1.);
  } else {
    if (tubeDist == InfLen<R>()) return InfLen<R>();
  }
  lp = lp + dir * tubeDist;
  if (t.hasPhi) {
    // Wedge::DistanceToIn
    const R comp1 = dir.x * R(t.n1x) + dir.y * R(t.n1y);
    const R comp2 = dir.x * R(t.n2x) + dir.y * R(t.n2y);
    R d1 = InfLen<R>(), d2 = InfLen<R>();
    if (comp1 > R(0)) {
      const R tt = -(lp.x * R(t.n1x) + lp.y * R(t.n1y)) / comp1;
      if (tt > R(0)) d1 = tt;
    }
    if (comp2 > R(0)) {
      const R tt = -(lp.x * R(t.n2x) + lp.y * R(t.n2y)) / comp2;
      if (tt > R(0)) d2 = tt;
    }
    if (d1 != InfLen<R>()) {
      const R daxis = DistSqrToTorusR(t, lp, dir, d1);
      if (daxis >= R(t.rmin2) && daxis < R(t.rmax2)) {
        distance = d1;
        if (vabs(distance) < R(kTolerance)) return distance + tubeDist;
      }
    }
    if (d2 != InfLen<R>()) {
      const R daxis = DistSqrToTorusR(t, lp, dir, d2);
      if (daxis >= R(t.rmin2) && daxis < R(t.rmax2)) {
        distance = vmin(d2, distance);
        if (vabs(distance) < R(kTolerance)) return distance + tubeDist;
      }
    }
    distance = InfLen<R>();
  }
  R dd = ToBoundary(t, lp, dir, t.rmax, false);
  if (t.rmin > 0.) dd = vmin(dd, ToBoundary(t, lp, dir, t.rmin, false));
  distance = vmin(distance, dd);
  distance += tubeDist;
  if (vabs(distance) > InfLen<R>()) distance = InfLen<R>();
  return distance;
}

template <class R>
VGMI_DEV R SafetyToIn(TorusData const &t, Vec3<R> const &p)
{
  const R rxy = vsqrt(p.x * p.x + p.y * p.y);
  const R rad = vsqrt((rxy - R(t.rtor)) * (rxy - R(t.rtor)) + p.z * p.z);
  R safety    = rad - R(t.rmax);
  if (t.rmin != 0.) safety = vmax(R(t.rmin) - rad, rad - R(t.rmax));
  if (t.hasPhi && rxy != R(0.)) {
    const R d1 = p.x * R(t.n1x) + p.y * R(t.n1y);
    const R d2 = p.x * R(t.n2x) + p.y * R(t.n2y);
    const R sp = t.wedgeLtPi ? vmax(-1 * d1, -1 * d2) : vmin(-1 * d1, -1 * d2);
    safety     = vmax(sp, safety);
  }
  return safety;
}

template <class R>
VGMI_DEV R SafetyToOut(TorusData const &t, Vec3<R> const &p)
{
  const R rxy = vsqrt(p.x * p.x + p.y * p.y);
  const R rad = vsqrt((rxy - R(t.rtor)) * (rxy - R(t.rtor)) + p.z * p.z);
  R safety    = R(t.rmax) - rad;
  if (t.rmin != 0.) safety = vmin(rad - R(t.rmin), R(t.rmax) - rad);
  if (t.hasPhi) {
    const R d1 = p.x * R(t.n1x) + p.y * R(t.n1y);
    const R d2 = p.x * R(t.n2x) + p.y * R(t.n2y);
    const R sp = t.wedgeLtPi ? vmin(d1, d2) : vmax(d1, d2);
    safety     = vmin(sp, safety);
  }
  return safety;
}

} // namespace tor
} // namespace vgmi
