// Surface normals: VUnplacedVolume::Normal (VecGeom/volumes/UnplacedVolume.h:252) of every shape, in
// FP64 with the reference's scalar operation order (the library is built with -ffp-contract=off).
//
// Each function returns the validity flag and writes the normal to n, setting `set`; where the
// reference leaves its normal argument unassigned (a tube or hype point near none of the surfaces
// its kernel tests, a torus point on no surface) `set` stays false and the caller's vector is kept.
// Vector3D conventions (VecGeom/base/Vector3D.h): Normalize()/Normalized() multiply by 1 / Mag()
// (:198-202), Unit() divides by Sqrt(Mag2() + kMinimum) (:284-290).
#pragma once
#include "dev_shapes.h"

namespace vgmi {
namespace nrm {

using V3d = Vec3<double>;

VGMI_DEV V3d Unit(V3d const &v)
{
  const double m = vsqrt(v.Mag2() + DBL_MIN);
  return V3d(v.x / m, v.y / m, v.z / m);
}
VGMI_DEV V3d Scale(V3d const &v, double s) { return V3d(v.x * s, v.y * s, v.z * s); }
VGMI_DEV V3d Div(V3d const &v, double s) { return V3d(v.x / s, v.y / s, v.z / s); }
// vecCore Min / Max of two values (only compared with == afterwards, so ties do not matter)
VGMI_DEV double Min2(double a, double b) { return a < b ? a : b; }
VGMI_DEV double Max2(double a, double b) { return a > b ? a : b; }

// BoxImplementation::NormalKernel (VecGeom/volumes/kernel/BoxImplementation.h:172-191)
VGMI_DEV bool Normal(BoxData const &b, V3d const &p, V3d &n, bool &set)
{
  const V3d h(b.h[0], b.h[1], b.h[2]);
  const V3d safety    = Abs3(Abs3(p) - h);
  const double safmin = Min3(safety);
  V3d normal(0., 0., 0.);
  if (safety.x - safmin < kHalfTolerance) normal.x = Sign(p.x);
  if (safety.y - safmin < kHalfTolerance) normal.y = Sign(p.y);
  if (safety.z - safmin < kHalfTolerance) normal.z = Sign(p.z);
  if (normal.Mag2() > 1.0) normal = Scale(normal, 1. / vsqrt(normal.Mag2()));
  n   = normal;
  set = true;
  return safmin < kHalfTolerance;
}

// OrbImplementation::NormalKernel (VecGeom/volumes/kernel/OrbImplementation.h:227-240)
VGMI_DEV bool Normal(OrbData const &o, V3d const &p, V3d &n, bool &set)
{
  const double rad2 = p.Mag2();
  n                 = Scale(p, 1. / vsqrt(rad2));
  set               = true;
  const double tolO = o.r + kTolerance, tolI = o.r - kTolerance;
  return rad2 <= tolO * tolO && rad2 >= tolI * tolI;
}

// TrapezoidImplementation::NormalKernel (VecGeom/volumes/kernel/TrapezoidImplementation.h:406-459) with
// PlaneShell<4>::NormalKernel (VecGeom/base/PlaneShell.h:342-370)
VGMI_DEV bool Normal(TrapezoidData const &t, V3d const &p, V3d &n, bool &set)
{
  V3d normal(0., 0., 0.);
  double safety = -kInfLength;
  double dist[4];
  for (int i = 0; i < 4; ++i) dist[i] = t.a[i] * p.x + t.b[i] * p.y + t.c[i] * p.z + t.d[i];
  for (int i = 0; i < 4; ++i) {
    const double saf_i = dist[i] - safety;
    const V3d ni(t.a[i], t.b[i], t.c[i]);
    if (vabs(saf_i) < kHalfTolerance && dist[i] >= -kHalfTolerance) normal = normal + ni;
    if (saf_i > 0.0) {
      normal = ni;
      safety = dist[i];
    }
  }
  const double safz = Sign(p.z) * p.z - t.dz;
  if (vabs(safz - safety) < kHalfTolerance && safz >= -kHalfTolerance) normal = normal + V3d(0, 0, Sign(p.z));
  if (safz > safety && safz >= -kHalfTolerance) normal = V3d(0, 0, Sign(p.z));
  if (safz > safety) safety = safz;
  if (normal.Mag2() > 1.0) normal = Scale(normal, 1. / vsqrt(normal.Mag2()));
  n   = normal;
  set = true;
  return vabs(safety) <= 1000. * kTolerance;
}

// UnplacedTrd::Normal (source/UnplacedTrd.cpp:276-353)
VGMI_DEV bool Normal(TrdData const &t, V3d const &p, V3d &n, bool &set)
{
  const double dx1 = t.dx1, dx2 = t.dx2, dy1 = t.dy1, dy2 = t.dy2, dz = t.dz;
  int noSurfaces     = 0;
  V3d sumnorm(0., 0., 0.);
  const double distz = vabs(vabs(p.z) - dz);
  const double xnorm = 1.0 / vsqrt(4 * dz * dz + (dx2 - dx1) * (dx2 - dx1));
  const double ynorm = 1.0 / vsqrt(4 * dz * dz + (dy2 - dy1) * (dy2 - dy1));
  double distmx      = -2.0 * dz * p.x - (dx2 - dx1) * p.z - dz * (dx1 + dx2);
  distmx *= xnorm;
  double distpx = 2.0 * dz * p.x - (dx2 - dx1) * p.z - dz * (dx1 + dx2);
  distpx *= xnorm;
  double distmy = -2.0 * dz * p.y - (dy2 - dy1) * p.z - dz * (dy1 + dy2);
  distmy *= ynorm;
  double distpy = 2.0 * dz * p.y - (dy2 - dy1) * p.z - dz * (dy1 + dy2);
  distpy *= ynorm;
  double safmin = Min2(vabs(distmx), vabs(distpx));
  safmin        = Min2(safmin, Min2(vabs(distmy), vabs(distpy)));
  safmin        = Min2(safmin, vabs(distz));
  if (vabs(distmx) - safmin <= kHalfTolerance) {
    noSurfaces++;
    sumnorm = sumnorm + Scale(V3d(-2.0 * dz, 0.0, -(dx2 - dx1)), xnorm);
  }
  if (vabs(distpx) - safmin <= kHalfTolerance) {
    noSurfaces++;
    sumnorm = sumnorm + Scale(V3d(2.0 * dz, 0.0, -(dx2 - dx1)), xnorm);
  }
  if (vabs(distpy) - safmin <= kHalfTolerance) {
    noSurfaces++;
    sumnorm = sumnorm + Scale(V3d(0.0, 2.0 * dz, -(dy2 - dy1)), ynorm);
  }
  if (vabs(distmy) - safmin <= kHalfTolerance) {
    noSurfaces++;
    sumnorm = sumnorm + Scale(V3d(0.0, -2.0 * dz, -(dy2 - dy1)), ynorm);
  }
  if (vabs(distz) - safmin <= kHalfTolerance) {
    noSurfaces++;
    sumnorm = sumnorm + (p.z >= 0. ? V3d(0., 0., 1.) : V3d(0., 0., -1.));
  }
  if (noSurfaces == 0)
    n = V3d(0., 0., 1.);
  else if (noSurfaces == 1)
    n = sumnorm;
  else
    n = Unit(sumnorm);
  set = true;
  return safmin <= kHalfTolerance;
}

// ParaboloidImplementation::NormalKernel (VecGeom/volumes/kernel/ParaboloidImplementation.h:428-487).
// The normal computed at :441-449 goes to a block-local vector that shadows the one used afterwards,
// so the parabolic surface contributes -(0, 0, Sign(-a)); kept as the reference computes it.
VGMI_DEV bool Normal(ParaboloidData const &q, V3d const &p, V3d &n, bool &set)
{
  V3d normal(0., 0., 0.);
  int nsurf = 0;
  const V3d normParabolic(0., 0., Sign(-q.a));
  const double r   = vsqrt(p.Perp2());
  const bool top   = para::on_z_plane(q, true, p);
  const bool bot   = para::on_z_plane(q, false, p);
  const double val = q.a * p.Perp2() + q.b - p.z;
  const bool onPar = val > -kTolerance && val < kTolerance;
  if (top || bot) nsurf++;
  if (top) normal.z = 1.;
  if (bot) normal.z = -1.;
  if (onPar) {
    nsurf++;
    normal = normal - normParabolic;
  }
  const bool valid = nsurf > 0;
  if (!valid) {
    V3d norm(0., 0., 0.);
    if (p.z > 0.) norm.z = 1.;
    if (p.z < 0.) norm.z = -1.;
    const double safz = q.dz - vabs(p.z);
    const double safr = vabs(r - vsqrt((p.z - q.b) * q.invA));
    if (safz >= 0. && safr < safz) norm = normParabolic;
    normal = norm;
  }
  n   = Scale(normal, 1. / vsqrt(normal.Mag2()));
  set = true;
  return valid;
}

// UnplacedHype::Normal (VecGeom/volumes/UnplacedHype.h:276-303): always valid; PointOnZSurface tests
// the +z cap only; a point matching none of the three tests keeps the caller's vector.
VGMI_DEV bool Normal(HypeData const &h, V3d const &p, V3d &n, bool &set)
{
  const double absZ       = vabs(p.z);
  const double distZ      = absZ - h.dz;
  const double dist2Z     = distZ * distZ;
  const double xR2        = p.Perp2();
  const double dist2Outer = vabs(xR2 - (h.rmax2 + h.tOut2 * absZ * absZ));
  const double dist2Inner = vabs(xR2 - (h.rmin2 + h.tIn2 * absZ * absZ));
  const bool onZ          = p.z > (h.dz - h.zTol) && p.z < (h.dz + h.zTol);
  const double outR2 = h.rmax2 + h.tOut2 * p.z * p.z, inR2 = h.rmin2 + h.tIn2 * p.z * p.z;
  const bool onOuter = xR2 > (outR2 - h.outerRadTol) && xR2 < (outR2 + h.outerRadTol);
  const bool onInner = xR2 > (inR2 - h.outerRadTol) && xR2 < (inR2 + h.outerRadTol);
  if (onZ || ((dist2Z < dist2Inner) && (dist2Z < dist2Outer))) {
    n   = V3d(0.0, 0.0, p.z < 0 ? -1.0 : 1.0);
    set = true;
  }
  if (onOuter || ((dist2Outer < dist2Inner) && (dist2Outer < dist2Z))) {
    n   = Unit(V3d(p.x, p.y, -p.z * h.tOut2));
    set = true;
  }
  if (onInner || ((dist2Inner < dist2Outer) && (dist2Inner < dist2Z))) {
    n   = Unit(V3d(-p.x, -p.y, p.z * h.tIn2));
    set = true;
  }
  return true;
}

// TubeImplementation::ApproxSurfaceNormalKernel (VecGeom/volumes/kernel/TubeImplementation.h:969-1016)
VGMI_DEV V3d TubeApprox(TubeData const &t, V3d const &p)
{
  V3d norm(0., 0., 0.);
  const double radius = vsqrt(p.Perp2());
  const double distRMax = vabs(radius - t.rmax);
  double distRMin       = kInfLength;
  if (t.rmin) distRMin = vabs(t.rmin - radius);
  double distMin  = Min2(distRMin, distRMax);
  double distPhi1 = kInfLength, distPhi2 = kInfLength;
  if (t.dphi != kTwoPi) {
    distPhi1 = p.x * t.n1x + p.y * t.n1y;
    distPhi2 = p.x * t.n2x + p.y * t.n2y;
    if (distPhi1 < 0.) distPhi1 = kInfLength;
    if (distPhi2 < 0.) distPhi2 = kInfLength;
    distMin = Min2(distMin, Min2(distPhi1, distPhi2));
  }
  double distZ = kInfLength;
  if (p.z < 0.) distZ = vabs(p.z + t.z);
  if (p.z >= 0.) distZ = vabs(p.z - t.z);
  distMin = Min2(distMin, distZ);
  if (t.dphi) {
    if (distMin == distPhi1) norm = -V3d(t.n1x, t.n1y, 0.);
    if (distMin == distPhi2) norm = -V3d(t.n2x, t.n2y, 0.);
  }
  if (distMin == distZ && p.z < 0.) norm = V3d(0., 0., -1.);
  if (distMin == distZ && p.z >= 0.) norm = V3d(0., 0., 1.);
  if (vabs(p.z) < (t.z + kTolerance)) {
    const V3d temp(p.x, p.y, 0.);
    if (distMin == distRMax) norm = Unit(temp);
    if (t.rmin && distMin == distRMin) norm = -Unit(temp);
  }
  return norm;
}

// TubeImplementation::NormalKernel (VecGeom/volumes/kernel/TubeImplementation.h:1021-1097) via
// UnplacedTube::Normal (source/UnplacedTube.cpp:181-186)
VGMI_DEV bool Normal(TubeData const &t, V3d const &p, V3d &norm, bool &set)
{
  bool ci = false, co = false;
  tube::ContainsInside(t, p, true, ci, co);
  if (ci || co) {
    norm = TubeApprox(t, p);
    set  = true;
    return false;
  }
  int nosurface     = 0;
  const double x2y2 = vsqrt(p.x * p.x + p.y * p.y);
  const bool inZ    = (p.z < t.z + kTolerance) && (p.z > -t.z - kTolerance);
  const bool inR    = (x2y2 >= t.rmin - kTolerance) && (x2y2 <= t.rmax + kTolerance);
  if (inR && vabs(p.z - t.z) <= kTolerance) {
    norm = V3d(0., 0., 1.);
    nosurface++;
  }
  if (inR && vabs(p.z + t.z) <= kTolerance) {
    if (nosurface > 0)
      norm.z += -1.;
    else
      norm = V3d(0., 0., -1.);
    nosurface++;
  }
  if (t.rmin > 0.) {
    if (inZ && vabs(x2y2 - t.rmin) <= kTolerance) {
      const double inv = 1. / x2y2;
      if (nosurface == 0) {
        norm = V3d(-p.x * inv, -p.y * inv, 0.);
      } else {
        norm.x += -p.x * inv;
        norm.y += -p.y * inv;
      }
      nosurface++;
    }
  }
  if (inZ && vabs(x2y2 - t.rmax) <= kTolerance) {
    const double inv = 1. / x2y2;
    if (nosurface > 0) {
      norm.x += p.x * inv;
      norm.y += p.y * inv;
    } else {
      norm = V3d(p.x * inv, p.y * inv, 0.);
    }
    nosurface++;
  }
  if (t.dphi < kTwoPi) {
    const V3d mn1 = -V3d(t.n1x, t.n1y, 0.), mn2 = -V3d(t.n2x, t.n2y, 0.);
    if (inR && tube::OnWedgePlane(t.a1x, t.a1y, t.n1x, t.n1y, p)) {
      norm = nosurface == 0 ? mn1 : norm + mn1;
      nosurface++;
    }
    if (inR && tube::OnWedgePlane(t.a2x, t.a2y, t.n2x, t.n2y, p)) {
      norm = nosurface == 0 ? mn2 : norm + mn2;
      nosurface++;
    }
  }
  if (nosurface > 1) norm = Div(norm, vsqrt(1. * nosurface));
  set = nosurface > 0;
  return nosurface != 0;
}

// ConeStruct::Normal (VecGeom/volumes/ConeStruct.h:177-306) via UnplacedCone::Normal
// (source/UnplacedCone.cpp:216-219). The point's azimuth comes from the device library's atan2; it
// only enters comparisons with the phi tolerances.
VGMI_DEV bool Normal(ConeData const &c, V3d const &p, V3d &norm, bool &set)
{
  int noSurfaces  = 0;
  double distSPhi = kInfLength, distEPhi = kInfLength;
  V3d sumnorm(0., 0., 0.);
  const V3d nZ(0., 0., 1.);
  V3d nR(0., 0., 0.), nr(0., 0., 0.), nPs(0., 0., 0.), nPe(0., 0., 0.);
  double distZ       = vabs(p.z) - c.dz;
  const double rho   = vsqrt(p.x * p.x + p.y * p.y);
  const double pRMin = rho - p.z * c.tanRMin, widRMin = c.rmin2 - c.dz * c.tanRMin;
  double distRMin;
  if (vabs(c.origRmin1 - c.origRmin2) < c.innerTol)
    distRMin = rho - c.origRmin2;
  else
    distRMin = (pRMin - widRMin) / c.secRMin;
  const double pRMax = rho - p.z * c.tanRMax, widRMax = c.rmax2 - c.dz * c.tanRMax;
  double distRMax;
  if (vabs(c.origRmax1 - c.origRmax2) < c.outerTol)
    distRMax = rho - c.origRmax2;
  else
    distRMax = (pRMax - widRMax) / c.secRMax;
  const bool hasR = c.rmin1 || c.rmin2;
  bool inside     = distZ < kTolerance && distRMax < c.outerTol;
  if (hasR) inside &= distRMin > -c.innerTol;
  distZ    = vabs(distZ);
  distRMax = vabs(distRMax);
  distRMin = vabs(distRMin);
  double distNearest = distZ;
  V3d normNearest    = nZ;
  if (p.z < 0.) normNearest = V3d(0, 0, -1.);
  const bool fullPhi = c.dphi == kTwoPi;
  if (!fullPhi) {
    if (rho) {
      double pPhi = vatan2(p.y, p.x);
      if (pPhi < c.sphi - kHalfTolerance)
        pPhi += 2 * kPi;
      else if (pPhi > c.sphi + c.dphi + kHalfTolerance)
        pPhi -= 2 * kPi;
      distSPhi = rho * (pPhi - c.sphi);
      distEPhi = rho * (pPhi - c.sphi - c.dphi);
      inside   = inside && (distSPhi > -kTolerance) && (distEPhi < kTolerance);
      distSPhi = vabs(distSPhi);
      distEPhi = vabs(distEPhi);
    } else if (!(c.rmin1) || !(c.rmin2)) {
      distSPhi = 0.;
      distEPhi = 0.;
    }
    nPs = V3d(c.a1y, -c.a1x, 0);
    nPe = V3d(-c.a2y, c.a2x, 0);
  }
  if (rho > kHalfTolerance) {
    nR = V3d(p.x / rho / c.secRMax, p.y / rho / c.secRMax, -c.tanRMax / c.secRMax);
    if (hasR) nr = V3d(-p.x / rho / c.secRMin, -p.y / rho / c.secRMin, c.tanRMin / c.secRMin);
  }
  if (inside && distZ <= kHalfTolerance) {
    noSurfaces++;
    if (p.z >= 0.)
      sumnorm = sumnorm + nZ;
    else
      sumnorm = V3d(0, 0, -1.);
  }
  if (inside && distRMax <= c.outerTol) {
    noSurfaces++;
    sumnorm = sumnorm + nR;
  } else if (noSurfaces == 0 && distRMax < distNearest) {
    distNearest = distRMax;
    normNearest = nR;
  }
  if (hasR) {
    if (inside && distRMin <= c.innerTol) {
      noSurfaces++;
      sumnorm = sumnorm + nr;
    } else if (noSurfaces == 0 && distRMin < distNearest) {
      distNearest = distRMin;
      normNearest = nr;
    }
  }
  if (!fullPhi) {
    if (inside && distSPhi <= kHalfTolerance) {
      noSurfaces++;
      sumnorm = sumnorm + nPs;
    } else if (noSurfaces == 0 && distSPhi < distNearest) {
      distNearest = distSPhi;
      normNearest = nPs;
    }
    if (inside && distEPhi <= kHalfTolerance) {
      noSurfaces++;
      sumnorm = sumnorm + nPe;
    } else if (noSurfaces == 0 && distEPhi < distNearest) {
      normNearest = nPe;
    }
  }
  if (noSurfaces == 0)
    norm = normNearest;
  else if (noSurfaces == 1)
    norm = sumnorm;
  else
    norm = Unit(sumnorm);
  set = true;
  return noSurfaces != 0 && noSurfaces <= 2;
}

// UnplacedPolycone::Normal (source/UnplacedPolycone.cpp:520-565)
VGMI_DEV bool Normal(PolyconeData const &P, V3d const &p, V3d &norm, bool &set)
{
  set             = true;
  const int index = pcon::SectionIndex(P, p.z - kTolerance);
  if (index < 0) {
    norm = V3d(0., 0., index == -1 ? -1. : 1.);
    return false;
  }
  bool s1, s2;
  bool valid = Normal(pcon::sec(P, index), pcon::local(P, index, p), norm, s1);
  if (index + 1 < P.nsec && vabs(p.z - pcon::zb(P, index + 1)) < kTolerance) {
    V3d norm2(0., 0., 0.);
    const bool valid2 = Normal(pcon::sec(P, index + 1), pcon::local(P, index + 1, p), norm2, s2);
    if (!valid && valid2) {
      norm  = norm2;
      valid = valid2;
    }
    if (valid && valid2) {
      const V3d q   = p + (kTolerance * 10) * norm2;
      const bool c2 = cone::Contains(pcon::sec(P, index + 1), pcon::local(P, index + 1, q));
      if (c2) {
        norm = norm2;
      } else {
        norm = norm + norm2;
        if (norm.Mag2() < kTolerance) norm = V3d(p.x, p.y, 0);
      }
    }
  }
  if (valid) norm = Div(norm, vsqrt(norm.Mag2()));
  return valid;
}

// ThetaCone::GetNormal1 / GetNormal2 (VecGeom/volumes/ThetaCone_Evolution.h:108-138)
VGMI_DEV V3d ThetaNormal(SphereData const &s, bool start, V3d const &p)
{
  if (start) {
    const V3d v(2. * p.x, 2. * p.y, -2. * s.tanS2 * p.z);
    return s.tS <= kPi / 2. ? -v : v;
  }
  const V3d v(2 * p.x, 2 * p.y, -2 * s.tanE2 * p.z);
  return s.tE <= kPi / 2. ? v : -v;
}

// SphereImplementation::ApproxSurfaceNormalKernel (VecGeom/volumes/kernel/SphereImplementation.h:489-533)
VGMI_DEV V3d SphereApprox(SphereData const &s, V3d const &p)
{
  V3d norm(0., 0., 0.);
  const double radius   = vsqrt(p.Mag2());
  const double distRMax = vabs(radius - s.rmax);
  double distRMin = kInfLength, distPhi1 = kInfLength, distPhi2 = kInfLength, distTheta1 = kInfLength,
         distTheta2 = kInfLength;
  double distMin = distRMax;
  if (s.rmin > 0.) {
    distRMin = vabs(s.rmin - radius);
    distMin  = Min2(distRMin, distRMax);
  }
  if (!s.fullPhi) {
    distPhi1 = vabs(p.x * s.n1x + p.y * s.n1y);
    distPhi2 = vabs(p.x * s.n2x + p.y * s.n2y);
    distMin  = Min2(distMin, Min2(distPhi1, distPhi2));
  }
  if (!s.fullTheta) {
    const double rho = vsqrt(p.Perp2());
    distTheta1       = sph::TDistanceToLine(s.slope1, rho, p.z);
    distTheta2       = sph::TDistanceToLine(s.slope2, rho, p.z);
    distMin          = Min2(distMin, Min2(distTheta1, distTheta2));
  }
  if (distMin == distRMax) norm = Unit(p);
  if (distMin == distRMin) norm = -Unit(p);
  if (distMin == distPhi1) norm = -V3d(s.n1x, s.n1y, 0.);
  if (distMin == distPhi2) norm = -V3d(s.n2x, s.n2y, 0.);
  if (distMin == distTheta1) norm = norm + ThetaNormal(s, true, p);
  if (distMin == distTheta2) norm = norm + ThetaNormal(s, false, p);
  return norm;
}

// SphereImplementation::Normal (VecGeom/volumes/kernel/SphereImplementation.h:537-637)
VGMI_DEV bool Normal(SphereData const &s, V3d const &p, V3d &n, bool &set)
{
  V3d normal(1e-30, 1e-30, 1e-30);
  bool ci = false, co = false;
  sph::ContainsInside(s, p, true, ci, co);
  if (ci || co) normal = SphereApprox(s, p);
  int noSurfaces     = 0;
  const double m2    = p.Mag2();
  const bool onOuter = m2 <= sph::PlusTolSq(s.rmax) && m2 >= sph::MinusTolSq(s.rmax);
  if (onOuter) noSurfaces++;
  if (!co && onOuter) normal = normal + Unit(p);
  if (s.rmin) {
    const bool onInner = m2 <= sph::PlusTolSq(s.rmin) && m2 >= sph::MinusTolSq(s.rmin);
    if (onInner) noSurfaces++;
    if (!co && onInner) normal = normal - Unit(p);
  }
  if (!s.fullPhi) {
    const bool onS = sph::WOnPlane(s.w1x, s.w1y, s.n1x, s.n1y, p);
    const bool onE = sph::WOnPlane(s.w2x, s.w2y, s.n2x, s.n2y, p);
    if (onS) noSurfaces++;
    if (onE) noSurfaces++;
    if (!co && onS) normal = normal - V3d(s.n1x, s.n1y, 0.);
    if (!co && onE) normal = normal - V3d(s.n2x, s.n2y, 0.);
  }
  if (!s.fullTheta) {
    const bool onS = sph::TOnSurface(s, true, p);
    const bool onE = sph::TOnSurface(s, false, p);
    if (onS) noSurfaces++;
    if (!co && onS) normal = normal + ThetaNormal(s, true, p);
    if (onE) noSurfaces++;
    if (!co && onE) normal = normal + ThetaNormal(s, false, p);
    if (!co && onS && onE && (s.etheta <= kPi / 2.)) normal = V3d(0., 0., -1.);
    if (!co && onS && onE && (s.stheta >= kPi / 2.)) normal = V3d(0., 0., 1.);
  }
  n   = Scale(normal, 1. / vsqrt(normal.Mag2()));
  set = true;
  return noSurfaces > 0;
}

// PolyhedronImplementation::ScalarNormalKernel (VecGeom/volumes/kernel/PolyhedronImplementation.h:998-1067)
VGMI_DEV bool Normal(PolyhedronData const &P, V3d const &p, V3d &normal, bool &set)
{
  set              = true;
  double safety    = kInfLength;
  const int zMax   = P.nseg;
  const int zIndex = phed::FindZSegment(P, p.z);
  if (zIndex < 0) {
    normal = V3d(0, 0, -1);
    return true;
  }
  if (zIndex >= zMax) {
    normal = V3d(0, 0, 1);
    return true;
  }
  int iSeg = zIndex, iSurf = -1, iz = 0;
  int phiIndex = phed::FindPhiSegment(P, p);
  for (int z = zIndex; z < zMax;) {
    int iSurfCrt           = -1;
    const double safetySeg = phed::SafetyToZSegmentSquared(P, z, phiIndex, p, true, &iSurfCrt);
    if (safetySeg < safety) {
      safety = safetySeg;
      iSeg   = z;
      iSurf  = iSurfCrt;
    }
    ++z;
    const double dz = phed::zp(P, z) - p.z;
    if (dz * dz > safety) break;
  }
  for (int z = zIndex - 1; z >= 0; --z) {
    int iSurfCrt           = -1;
    const double safetySeg = phed::SafetyToZSegmentSquared(P, z, phiIndex, p, true, &iSurfCrt);
    if (safetySeg < safety) {
      safety = safetySeg;
      iSeg   = z;
      iSurf  = iSurfCrt;
    }
    const double dz = p.z - phed::zp(P, z);
    if (dz * dz > safety) break;
  }
  phed::SafetyToEndcapsSquared(P, p, safety, &iz);
  if (iz != 0) {
    normal = V3d(0, 0, iz);
    return true;
  }
  // the reference indexes the quadrilateral sets without a check; an index out of range (not
  // reached for the solids the constructor builds) gives the zero vector instead of a stray read
  const int which = (iSurf >= 0 && iSurf < 2) ? 1 : (iSurf == 2 ? 2 : 0);
  const int i     = which == 1 ? iSurf : phiIndex;
  const phed::QSet q = phed::qset(P, iSeg, which);
  const V3d qn       = (i >= 0 && i < q.n) ? phed::qn<double>(q.q + kQuadDoubles * i) : V3d(0, 0, 0);
  normal             = which == 2 ? Scale(qn, -1.) : qn;
  return true;
}

// UnplacedTorus2::Normal (source/UnplacedTorus2.cpp:252-330). The azimuth comes from the device
// library's atan2 and only enters comparisons with the angular tolerance.
VGMI_DEV bool Normal(TorusData const &t, V3d const &p, V3d &norm, bool &set)
{
  int noSurfaces      = 0;
  double distRMin     = kInfLength;
  double distSPhi     = kInfLength, distEPhi = kInfLength;
  const double delta  = Max2(10.0 * kTolerance, 1.0e-8 * (t.rtor + t.rmax));
  const double dAngle = 10.0 * kTolerance;
  V3d nR(0., 0., 0.), nPs(0., 0., 0.), nPe(0., 0., 0.), sumnorm(0., 0., 0.);
  const double rho2 = p.x * p.x + p.y * p.y;
  const double rho  = vsqrt(rho2);
  double pt2        = rho2 + p.z * p.z + t.rtor * (t.rtor - 2 * rho);
  pt2               = Max2(pt2, 0.);
  const double pt   = vsqrt(pt2);
  const double distRMax = vabs(pt - t.rmax);
  if (t.rmin) distRMin = vabs(pt - t.rmin);
  if (rho > delta && pt != 0.) {
    const double redFactor = (rho - t.rtor) / rho;
    nR                     = Scale(V3d(p.x * redFactor, p.y * redFactor, p.z), 1. / pt);
  }
  if (t.dphi < kTwoPi) {
    if (rho) {
      double pPhi = vatan2(p.y, p.x);
      if (pPhi < t.sphi - delta)
        pPhi += kTwoPi;
      else if (pPhi > t.sphi + t.dphi + delta)
        pPhi -= kTwoPi;
      distSPhi = vabs(pPhi - t.sphi);
      distEPhi = vabs(pPhi - t.sphi - t.dphi);
    }
    nPs = V3d(-t.n1x, -t.n1y, 0); // (sin sphi, -cos sphi, 0)
    nPe = V3d(-t.n2x, -t.n2y, 0); // (-sin(sphi + dphi), cos(sphi + dphi), 0)
  }
  if (distRMax <= delta) {
    noSurfaces++;
    sumnorm = sumnorm + nR;
  } else if (t.rmin && (distRMin <= delta)) {
    noSurfaces++;
    sumnorm = sumnorm - nR;
  }
  if ((t.dphi < kTwoPi) && (t.rmin - delta <= pt) && (pt <= (t.rmax + delta))) {
    if (distSPhi <= dAngle) {
      noSurfaces++;
      sumnorm = sumnorm + nPs;
    }
    if (distEPhi <= dAngle) {
      noSurfaces++;
      sumnorm = sumnorm + nPe;
    }
  }
  if (noSurfaces == 0) return false;
  norm = noSurfaces == 1 ? sumnorm : Unit(sumnorm);
  set  = true;
  return true;
}

// TessellatedImplementation::NormalKernel (VecGeom/volumes/kernel/TessellatedImplementation.h:162-172):
// the normal of the facet nearest by SafetySq<false>. Always valid.
VGMI_DEV bool Normal(TessellatedData const &s, V3d const &p, V3d &n, bool &set)
{
  const int slot = tess::nearest_facet_slot(s, p, false);
  n              = slot >= 0 ? tess::at3(tess::facet(s, slot) + 3) : V3d(0, 0, 0);
  set            = true;
  return true;
}

// ParallelepipedImplementation::NormalKernel (VecGeom/volumes/kernel/ParallelepipedImplementation.h:185-232)
VGMI_DEV bool Normal(ParallelepipedData const &s, V3d const &point, V3d &n, bool &set)
{
  const V3d p = ppd::oblique(s, point), sv = ppd::safety_vector(s, p);
  const V3d signs(Sign(p.x), Sign(p.y), Sign(p.z));
  const V3d n0(s.n0[0], s.n0[1], s.n0[2]), n1(s.n1[0], s.n1[1], s.n1[2]);
  V3d normal(0., 0., 0.);
  if (vabs(sv.z) <= kHalfTolerance) normal = V3d(0., 0., signs.z);
  if (vabs(sv.y) <= kHalfTolerance) normal = normal + signs.y * n1;
  if (vabs(sv.x) <= kHalfTolerance) normal = normal + signs.x * n0;
  const double mag2 = normal.Mag2();
  if (mag2 > 1.) normal = Unit(normal);
  set = true;
  if (mag2 > 0.) {
    n = normal;
    return true;
  }
  const double safety = Max2(Max2(sv.x, sv.y), sv.z);
  normal              = signs.x * n0;
  if (sv.y == safety) normal = signs.y * n1;
  if (sv.z == safety) normal = V3d(0., 0., signs.z);
  n = normal;
  return false;
}

// TetImplementation::NormalKernel (VecGeom/volumes/kernel/TetImplementation.h:187-218)
VGMI_DEV bool Normal(TetData const &s, V3d const &p, V3d &n, bool &set)
{
  V3d normal(0., 0., 0.);
  double dist[4];
  for (int i = 0; i < 4; ++i) {
    const V3d ni = tet::normal<double>(s, i);
    dist[i]      = ni.Dot(p) + s.d[i];
    if (vabs(dist[i]) <= kHalfTolerance) normal = normal + ni;
  }
  if (normal.Mag2() > 1) normal = Unit(normal);
  set = true;
  if (normal.Mag2() > 0.) {
    n = normal;
    return true;
  }
  double safety = -kInfLength;
  for (int i = 0; i < 4; ++i)
    if (dist[i] > safety) normal = tet::normal<double>(s, i), safety = dist[i];
  n = normal;
  return false;
}

// EllipticalTubeImplementation::NormalKernel (VecGeom/volumes/kernel/EllipticalTubeImplementation.h:233-268)
VGMI_DEV bool Normal(EllipticalTubeData const &s, V3d const &point, V3d &n, bool &set)
{
  V3d normal(0., 0., 0.);
  const double x = point.x * s.sx, y = point.y * s.sy;
  double distR = s.q1 * (x * x + y * y) - s.q2;
  if (vabs(distR) <= kHalfTolerance) normal = Unit(V3d(point.x * s.ddy, point.y * s.ddx, 0.));
  const double distZ = vabs(point.z) - s.dz;
  if (vabs(distZ) <= kHalfTolerance) normal.z = Sign(point.z);
  if (normal.Mag2() > 1.) normal = Unit(normal);
  set = true;
  if (normal.Mag2() > 0.) {
    n = normal;
    return true;
  }
  normal.z = Sign(point.z);
  distR    = vsqrt(x * x + y * y) - s.r;
  if (distR > distZ && (x * x + y * y) > 0.) normal = Unit(V3d(point.x * s.ddy, point.y * s.ddx, 0.));
  n = normal;
  return false;
}

// EllipsoidImplementation::NormalKernel (VecGeom/volumes/kernel/EllipsoidImplementation.h:245-287)
VGMI_DEV bool Normal(EllipsoidData const &s, V3d const &point, V3d &n, bool &set)
{
  V3d normal(0., 0., 0.);
  const double px = point.x, py = point.y, pz = point.z;
  const double A = s.dx, B = s.dy, C = s.dz;
  const double x = px * s.sx, y = py * s.sy, z = pz * s.sz;
  const double mag2 = x * x + y * y + z * z;
  double distR      = s.q1 * mag2 - s.q2;
  if (vabs(distR) <= kHalfTolerance) normal = Unit(V3d(px / (A * A), py / (B * B), pz / (C * C)));
  const double distZ = vabs(z - s.scZMidCut) - s.scZDimCut;
  if (vabs(distZ) <= kHalfTolerance) normal.z = normal.z + Sign(z - s.scZMidCut);
  if (normal.Mag2() > 1.) normal = Unit(normal);
  set = true;
  if (normal.Mag2() > 0.) {
    n = normal;
    return true;
  }
  normal.z = Sign(z - s.scZMidCut);
  distR    = vsqrt(mag2) - s.r;
  if (distR > distZ && mag2 > 0.) normal = Unit(V3d(px / (A * A), py / (B * B), pz / (C * C)));
  n = normal;
  return false;
}

// GenTrapImplementation::NormalKernel and GetClosestEdge (VecGeom/volumes/kernel/GenTrapImplementation.h:
// 610-735), planar faces: +-z within 10 kTolerance of a z plane, else the shell normal of the section edge
// closest to the point (the section from the reverted vertices, the normals in the order given)
VGMI_DEV bool Normal(GenTrapData const &s, V3d const &p, V3d &n, bool &set)
{
  set               = true;
  const double safz = vabs(s.dz - vabs(p.z));
  if (safz < 10. * kTolerance) {
    n = V3d(0., 0., p.z > 0. ? 1. : (p.z < 0. ? -1. : 0.));
    return true;
  }
  const double cf = s.halfInverseDz * (s.dz - p.z);
  double vertexX[4], vertexY[4];
  for (int i = 0; i < 4; i++) {
    vertexX[i] = s.vx[i + 4] + cf * s.connX[i];
    vertexY[i] = s.vy[i + 4] + cf * s.connY[i];
  }
  int iseg      = 0;
  double safe   = kInfLength;
  for (int i = 0; i < 4; ++i) {
    const int j     = (i + 1) % 4;
    const double dx = vertexX[j] - vertexX[i], dy = vertexY[j] - vertexY[i];
    double dpx = p.x - vertexX[i], dpy = p.y - vertexY[i];
    const double lsq = dx * dx + dy * dy;
    if (lsq < kTolerance) {
      const double ssq = dpx * dpx + dpy * dpy;
      if (ssq < safe) iseg = i, safe = ssq;
      continue;
    }
    const double u = (dpx * dx + dpy * dy) / NonZero(lsq);
    if (u > 1) dpx = p.x - vertexX[j], dpy = p.y - vertexY[j];
    if (u >= 0 && u <= 1) dpx = dpx - u * dx, dpy = dpy - u * dy;
    const double ssq = dpx * dpx + dpy * dpy;
    if (ssq < safe) iseg = i, safe = ssq;
  }
  n = V3d(s.shellNormal[iseg][0], s.shellNormal[iseg][1], s.shellNormal[iseg][2]);
  return true;
}

// EllipticalConeImplementation::NormalKernel (VecGeom/volumes/kernel/EllipticalConeImplementation.h:307-333):
// the z plane and/or the lateral surface within half a tolerance, summed and made unit; elsewhere
// invalid, the z direction or the lateral normal, whichever surface is nearer
VGMI_DEV bool Normal(EllipticalConeData const &s, V3d const &p, V3d &n, bool &set)
{
  set = true;
  V3d normal(0., 0., 0.);
  bool valid     = true;
  const double z = vabs(p.z) - s.zCut;
  if (vabs(z) <= kHalfTolerance) normal.z = vcopysign(1., p.z);
  const double nx = p.x * s.invDx * s.invDx, ny = p.y * s.invDy * s.invDy;
  double nz       = vsqrt(p.x * nx + p.y * ny);
  if ((nx * nx + ny * ny) == 0.) nz = 1.;
  const V3d nside(nx, ny, nz);
  const double ds = (nz + p.z - s.dz) * s.cosAxisMin;
  if (vabs(ds) <= kHalfTolerance) normal = Unit(normal + Unit(nside));
  if (!(normal.Mag2() > 0.)) {
    valid    = false;
    normal.z = vcopysign(1., p.z);
    if (ds > z) normal = Unit(nside);
  }
  n = normal;
  return valid;
}

// CutTubeImplementation::NormalKernel (VecGeom/volumes/kernel/CutTubeImplementation.h:301-322): the
// unbounded tube's normal where its surface is nearer than the cut planes, else the cut plane of the
// point's z half
VGMI_DEV bool Normal(CutTubeData const &s, V3d const &p, V3d &n, bool &set)
{
  const double safcut  = ctub::Mn(-ctub::plane(s, 0, p), -ctub::plane(s, 1, p));
  const double saftube = tube::SafetyToOut(s.tube, p);
  if (vabs(saftube) < vabs(safcut)) return Normal(s.tube, p, n, set);
  set = true;
  n   = p.z < 0 ? V3d(s.n0[0], s.n0[1], s.n0[2]) : V3d(s.n1[0], s.n1[1], s.n1[2]);
  return true;
}

} // namespace nrm

} // namespace vgmi
