// Oracle: VUnplacedVolume::Normal (VecGeom/volumes/UnplacedVolume.h:252) for every shape.
// TEST INFRASTRUCTURE ONLY (see common.h).
//
// Each function restates the reference's scalar Normal of that shape, operation for operation.
// `n` is the caller's vector: where the reference leaves its normal argument unassigned (a tube or
// hype point near none of the surfaces the kernel tests), it is left as the caller set it.
// Vector3D conventions (VecGeom/base/Vector3D.h): Normalize() and Normalized() multiply by
// 1 / Mag() (:198-202), Unit() divides by Sqrt(Mag2() + kMinimum) (:284-290), v / s divides each
// component by s (:329-339).
#include "shapes.h"

namespace vgo {
namespace {
inline V3 Unit(V3 const &v)
{
  const double m = std::sqrt(v.Mag2() + kMinimum);
  return V3(v.x / m, v.y / m, v.z / m);
}
inline V3 Normalized(V3 const &v) { return v * (1. / v.Mag()); }
inline V3 Div(V3 const &v, double s) { return V3(v.x / s, v.y / s, v.z / s); }
// vecCore Min/Max of two values: a < b ? a : b (only compared with == below, so ties do not matter)
inline double Min2(double a, double b) { return a < b ? a : b; }
inline double Max2(double a, double b) { return a > b ? a : b; }
} // namespace

static V3 SphereThetaNormal(Sphere const &s, bool start, V3 const &p);

// BoxImplementation::NormalKernel (VecGeom/volumes/kernel/BoxImplementation.h:172-191): the unit
// average of the face normals within kHalfTolerance of the nearest face.
bool Box::Normal(V3 const &p, V3 &n) const
{
  const V3 safety((p.Abs() - h).Abs());
  const double safmin = safety.Min();
  const bool valid    = safmin < kHalfTolerance;
  V3 normal(0., 0., 0.);
  if (safety.x - safmin < kHalfTolerance) normal.x = Sign(p.x);
  if (safety.y - safmin < kHalfTolerance) normal.y = Sign(p.y);
  if (safety.z - safmin < kHalfTolerance) normal.z = Sign(p.z);
  if (normal.Mag2() > 1.0) normal = normal * (1. / normal.Mag());
  n = normal;
  return valid;
}

// OrbImplementation::NormalKernel (VecGeom/volumes/kernel/OrbImplementation.h:227-240)
bool Orb::Normal(V3 const &p, V3 &n) const
{
  const double rad2      = p.Mag2();
  const double invRadius = 1. / std::sqrt(rad2);
  n                      = p * invRadius;
  const double tolO = r + kTolerance, tolI = r - kTolerance;
  return rad2 <= tolO * tolO && rad2 >= tolI * tolI;
}

// TrapezoidImplementation::NormalKernel (VecGeom/volumes/kernel/TrapezoidImplementation.h:406-459) with
// PlaneShell<4>::NormalKernel (VecGeom/base/PlaneShell.h:342-370). A plane within kHalfTolerance of
// the farthest one so far adds its normal, a farther one replaces it (so an addition survives only
// when later planes are not farther); then the z planes, valid within 1000 kTolerance.
bool Trapezoid::Normal(V3 const &p, V3 &n) const
{
  const double delta = 1000. * kTolerance;
  V3 normal(0., 0., 0.);
  double safety = -kInfLength;
  double dist[4];
  for (int i = 0; i < 4; ++i) dist[i] = A[i] * p.x + B[i] * p.y + C[i] * p.z + D[i];
  for (int i = 0; i < 4; ++i) {
    const double saf_i = dist[i] - safety;
    if (std::fabs(saf_i) < kHalfTolerance && dist[i] >= -kHalfTolerance) normal = normal + V3(A[i], B[i], C[i]);
    if (saf_i > 0.0) {
      normal = V3(A[i], B[i], C[i]);
      safety = dist[i];
    }
  }
  const double safz = Sign(p.z) * p.z - dz;
  if (std::fabs(safz - safety) < kHalfTolerance && safz >= -kHalfTolerance) normal = normal + V3(0, 0, Sign(p.z));
  if (safz > safety && safz >= -kHalfTolerance) normal = V3(0, 0, Sign(p.z));
  if (safz > safety) safety = safz;
  const bool valid = std::fabs(safety) <= delta;
  if (normal.Mag2() > 1.0) normal = normal * (1. / normal.Mag());
  n = normal;
  return valid;
}

// UnplacedTrd::Normal (source/UnplacedTrd.cpp:276-353): every face within kHalfTolerance of the
// nearest adds its unit normal; the sum of several is made a unit vector.
bool Trd::Normal(V3 const &p, V3 &n) const
{
  int noSurfaces = 0;
  V3 sumnorm(0., 0., 0.), vecnorm(0., 0., 0.);
  const double distz = std::fabs(std::fabs(p.z) - dz);
  const double xnorm = 1.0 / std::sqrt(4 * dz * dz + (dx2 - dx1) * (dx2 - dx1));
  const double ynorm = 1.0 / std::sqrt(4 * dz * dz + (dy2 - dy1) * (dy2 - dy1));
  double distmx      = -2.0 * dz * p.x - (dx2 - dx1) * p.z - dz * (dx1 + dx2);
  distmx *= xnorm;
  double distpx = 2.0 * dz * p.x - (dx2 - dx1) * p.z - dz * (dx1 + dx2);
  distpx *= xnorm;
  double distmy = -2.0 * dz * p.y - (dy2 - dy1) * p.z - dz * (dy1 + dy2);
  distmy *= ynorm;
  double distpy = 2.0 * dz * p.y - (dy2 - dy1) * p.z - dz * (dy1 + dy2);
  distpy *= ynorm;
  double safmin    = Min2(std::fabs(distmx), std::fabs(distpx));
  safmin           = Min2(safmin, Min2(std::fabs(distmy), std::fabs(distpy)));
  safmin           = Min2(safmin, std::fabs(distz));
  const bool valid = safmin <= kHalfTolerance;
  if (std::fabs(distmx) - safmin <= kHalfTolerance) {
    noSurfaces++;
    sumnorm = sumnorm + V3(-2.0 * dz, 0.0, -(dx2 - dx1)) * xnorm;
  }
  if (std::fabs(distpx) - safmin <= kHalfTolerance) {
    noSurfaces++;
    sumnorm = sumnorm + V3(2.0 * dz, 0.0, -(dx2 - dx1)) * xnorm;
  }
  if (std::fabs(distpy) - safmin <= kHalfTolerance) {
    noSurfaces++;
    sumnorm = sumnorm + V3(0.0, 2.0 * dz, -(dy2 - dy1)) * ynorm;
  }
  if (std::fabs(distmy) - safmin <= kHalfTolerance) {
    noSurfaces++;
    sumnorm = sumnorm + V3(0.0, -2.0 * dz, -(dy2 - dy1)) * ynorm;
  }
  if (std::fabs(distz) - safmin <= kHalfTolerance) {
    noSurfaces++;
    sumnorm = sumnorm + (p.z >= 0. ? V3(0., 0., 1.) : V3(0., 0., -1.));
  }
  if (noSurfaces == 0)
    vecnorm = V3(0., 0., 1.);
  else if (noSurfaces == 1)
    vecnorm = sumnorm;
  else
    vecnorm = Unit(sumnorm);
  n = vecnorm;
  return valid;
}

// ParaboloidImplementation::NormalKernel (VecGeom/volumes/kernel/ParaboloidImplementation.h:428-487).
// The paraboloid's own normal is computed into a block-local vector (:441-449) that shadows the one
// used afterwards, so the parabolic surface contributes (0, 0, Sign(-a)) negated: the reference's
// result, kept as it is.
bool Paraboloid::Normal(V3 const &p, V3 &n) const
{
  V3 normal(0., 0., 0.);
  int nsurf = 0;
  const V3 normParabolic(0., 0., Sign(-a));
  const double r   = p.Perp();
  const bool top   = OnZPlane(true, p);
  const bool bot   = OnZPlane(false, p);
  const double val = a * p.Perp2() + b - p.z; // IsOnParabolicSurface (:136-144)
  const bool onPar = val > -kTolerance && val < kTolerance;
  if (top || bot) nsurf++;
  if (top) normal.z = 1.;
  if (bot) normal.z = -1.;
  if (onPar) {
    nsurf++;
    normal.x = normal.x - normParabolic.x;
    normal.y = normal.y - normParabolic.y;
    normal.z = normal.z - normParabolic.z;
  }
  const bool valid = nsurf > 0;
  if (!valid) {
    V3 norm(0., 0., 0.);
    if (p.z > 0.) norm.z = 1.;
    if (p.z < 0.) norm.z = -1.;
    const double safz = dz - std::fabs(p.z);
    const double safr = std::fabs(r - std::sqrt((p.z - b) * invA));
    if (safz >= 0. && safr < safz) norm = normParabolic;
    normal = norm;
  }
  n = Normalized(normal);
  return valid;
}

// UnplacedHype::Normal (VecGeom/volumes/UnplacedHype.h:276-303): the end cap, outer or inner
// hyperboloid that the point is on (PointOnZSurface tests the +z cap only) or nearest to; later
// matches override earlier ones, and a point matching none leaves n as it was. Always valid.
bool Hype::Normal(V3 const &p, V3 &n) const
{
  const double absZ       = std::fabs(p.z);
  const double distZ      = absZ - dz;
  const double dist2Z     = distZ * distZ;
  const double xR2        = p.Perp2();
  const double dist2Outer = std::fabs(xR2 - (rmax2 + tOut2 * absZ * absZ));
  const double dist2Inner = std::fabs(xR2 - (rmin2 + tIn2 * absZ * absZ));
  const bool onZ          = p.z > (dz - zTol) && p.z < (dz + zTol);
  const double outR2 = rmax2 + tOut2 * p.z * p.z, inR2 = rmin2 + tIn2 * p.z * p.z;
  const bool onOuter = xR2 > (outR2 - outerRadTol) && xR2 < (outR2 + outerRadTol);
  const bool onInner = xR2 > (inR2 - outerRadTol) && xR2 < (inR2 + outerRadTol);
  if (onZ || ((dist2Z < dist2Inner) && (dist2Z < dist2Outer))) n = V3(0.0, 0.0, p.z < 0 ? -1.0 : 1.0);
  if (onOuter || ((dist2Outer < dist2Inner) && (dist2Outer < dist2Z))) n = Unit(V3(p.x, p.y, -p.z * tOut2));
  if (onInner || ((dist2Inner < dist2Outer) && (dist2Inner < dist2Z))) n = Unit(V3(-p.x, -p.y, p.z * tIn2));
  return true;
}

// TubeImplementation::ApproxSurfaceNormalKernel (VecGeom/volumes/kernel/TubeImplementation.h:969-1016):
// the normal of the nearest surface, for points off the surface.
static V3 TubeApproxNormal(Tube const &t, V3 const &p)
{
  V3 norm(0., 0., 0.);
  const double radius = p.Perp();
  double distRMax     = std::fabs(radius - t.rmax);
  double distRMin     = kInfLength;
  if (t.rmin) distRMin = std::fabs(t.rmin - radius);
  double distMin  = Min2(distRMin, distRMax);
  double distPhi1 = kInfLength, distPhi2 = kInfLength;
  if (t.dphi != kTwoPi) {
    distPhi1 = p.x * t.wedge.normal1.x + p.y * t.wedge.normal1.y;
    distPhi2 = p.x * t.wedge.normal2.x + p.y * t.wedge.normal2.y;
    if (distPhi1 < 0.) distPhi1 = kInfLength;
    if (distPhi2 < 0.) distPhi2 = kInfLength;
    distMin = Min2(distMin, Min2(distPhi1, distPhi2));
  }
  double distZ = kInfLength;
  if (p.z < 0.) distZ = std::fabs(p.z + t.z);
  if (p.z >= 0.) distZ = std::fabs(p.z - t.z);
  distMin = Min2(distMin, distZ);
  if (t.dphi) {
    if (distMin == distPhi1) norm = -t.wedge.normal1;
    if (distMin == distPhi2) norm = -t.wedge.normal2;
  }
  if (distMin == distZ && p.z < 0.) norm = V3(0., 0., -1.);
  if (distMin == distZ && p.z >= 0.) norm = V3(0., 0., 1.);
  if (std::fabs(p.z) < (t.z + kTolerance)) {
    const V3 temp(p.x, p.y, 0.);
    if (distMin == distRMax) norm = Unit(temp);
    if (t.rmin && distMin == distRMin) norm = -Unit(temp);
  }
  return norm;
}

// TubeImplementation::NormalKernel (VecGeom/volumes/kernel/TubeImplementation.h:1021-1097) via
// UnplacedTube::Normal (source/UnplacedTube.cpp:181-186): points inside or outside get the nearest
// surface's normal and valid = false; surface points sum the normals of the surfaces they are on.
bool Tube::Normal(V3 const &p, V3 &norm) const
{
  bool ci = false, co = false;
  ContainsInside(p, true, ci, co);
  if (ci || co) {
    norm = TubeApproxNormal(*this, p);
    return false;
  }
  int nosurface     = 0;
  const double x2y2 = std::sqrt(p.x * p.x + p.y * p.y);
  const bool inZ    = (p.z < z + kTolerance) && (p.z > -z - kTolerance);
  const bool inR    = (x2y2 >= rmin - kTolerance) && (x2y2 <= rmax + kTolerance);
  if (inR && std::fabs(p.z - z) <= kTolerance) {
    norm = V3(0., 0., 1.);
    nosurface++;
  }
  if (inR && std::fabs(p.z + z) <= kTolerance) {
    if (nosurface > 0)
      norm.z += -1.;
    else
      norm = V3(0., 0., -1.);
    nosurface++;
  }
  if (rmin > 0.) {
    if (inZ && std::fabs(x2y2 - rmin) <= kTolerance) {
      const double inv = 1. / x2y2;
      if (nosurface == 0) {
        norm = V3(-p.x * inv, -p.y * inv, 0.);
      } else {
        norm.x += -p.x * inv;
        norm.y += -p.y * inv;
      }
      nosurface++;
    }
  }
  if (inZ && std::fabs(x2y2 - rmax) <= kTolerance) {
    const double inv = 1. / x2y2;
    if (nosurface > 0) {
      norm.x += p.x * inv;
      norm.y += p.y * inv;
    } else {
      norm = V3(p.x * inv, p.y * inv, 0.);
    }
    nosurface++;
  }
  if (dphi < kTwoPi) {
    if (inR && wedge.IsOnSurfaceGeneric(wedge.along1, wedge.normal1, p)) {
      norm = nosurface == 0 ? -wedge.normal1 : norm + (-wedge.normal1);
      nosurface++;
    }
    if (inR && wedge.IsOnSurfaceGeneric(wedge.along2, wedge.normal2, p)) {
      norm = nosurface == 0 ? -wedge.normal2 : norm + (-wedge.normal2);
      nosurface++;
    }
  }
  if (nosurface > 1) norm = Div(norm, std::sqrt(1. * nosurface));
  return nosurface != 0;
}

// ConeStruct::Normal (VecGeom/volumes/ConeStruct.h:177-306) via UnplacedCone::Normal
// (source/UnplacedCone.cpp:216-219): the surfaces within tolerance of an inside-or-on point add
// their normals (unit sum for several); otherwise the nearest surface's normal. Valid for one or
// two surfaces.
bool Cone::Normal(V3 const &p, V3 &norm) const
{
  int noSurfaces = 0;
  double pPhi;
  double distSPhi = kInfLength, distEPhi = kInfLength;
  V3 sumnorm(0., 0., 0.), nZ(0., 0., 1.);
  V3 nR, nr(0., 0., 0.), nPs, nPe;
  norm = sumnorm;

  double distZ      = std::fabs(p.z) - dz;
  const double rho  = std::sqrt(p.x * p.x + p.y * p.y);
  const double pRMin = rho - p.z * tanRMin, widRMin = rmin2 - dz * tanRMin;
  double distRMin;
  if (std::fabs(origRmin1 - origRmin2) < innerTol)
    distRMin = rho - origRmin2;
  else
    distRMin = (pRMin - widRMin) / secRMin;
  const double pRMax = rho - p.z * tanRMax, widRMax = rmax2 - dz * tanRMax;
  double distRMax;
  if (std::fabs(origRmax1 - origRmax2) < outerTol)
    distRMax = rho - origRmax2;
  else
    distRMax = (pRMax - widRMax) / secRMax;

  bool inside = distZ < kTolerance && distRMax < outerTol;
  if (rmin1 || rmin2) inside &= distRMin > -innerTol;
  distZ    = std::fabs(distZ);
  distRMax = std::fabs(distRMax);
  distRMin = std::fabs(distRMin);

  double distNearest = distZ;
  V3 normNearest     = nZ;
  if (p.z < 0.) normNearest = V3(0, 0, -1.);

  const bool fullPhi = dphi == kTwoPi;
  if (!fullPhi) {
    if (rho) {
      pPhi = std::atan2(p.y, p.x);
      if (pPhi < sphi - kHalfTolerance)
        pPhi += 2 * kPi;
      else if (pPhi > sphi + dphi + kHalfTolerance)
        pPhi -= 2 * kPi;
      distSPhi = rho * (pPhi - sphi);
      distEPhi = rho * (pPhi - sphi - dphi);
      inside   = inside && (distSPhi > -kTolerance) && (distEPhi < kTolerance);
      distSPhi = std::fabs(distSPhi);
      distEPhi = std::fabs(distEPhi);
    } else if (!(rmin1) || !(rmin2)) {
      distSPhi = 0.;
      distEPhi = 0.;
    }
    nPs = V3(a1y, -a1x, 0);  // (sin sphi, -cos sphi, 0)
    nPe = V3(-a2y, a2x, 0);  // (-sin(sphi + dphi), cos(sphi + dphi), 0)
  }
  if (rho > kHalfTolerance) {
    nR = V3(p.x / rho / secRMax, p.y / rho / secRMax, -tanRMax / secRMax);
    if (rmin1 || rmin2) nr = V3(-p.x / rho / secRMin, -p.y / rho / secRMin, tanRMin / secRMin);
  }
  if (inside && distZ <= kHalfTolerance) {
    noSurfaces++;
    if (p.z >= 0.)
      sumnorm = sumnorm + nZ;
    else
      sumnorm = V3(0, 0, -1.);
  }
  if (inside && distRMax <= outerTol) {
    noSurfaces++;
    sumnorm = sumnorm + nR;
  } else if (noSurfaces == 0 && distRMax < distNearest) {
    distNearest = distRMax;
    normNearest = nR;
  }
  if (rmin1 || rmin2) {
    if (inside && distRMin <= innerTol) {
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
  return noSurfaces != 0 && noSurfaces <= 2;
}

// UnplacedPolycone::Normal (source/UnplacedPolycone.cpp:520-565): the normal of the section holding
// z - kTolerance; on the plane shared with the next section, that section's normal is taken when
// only it is valid or when stepping 10 kTolerance along it stays inside it, else the two are added.
bool Polycone::Normal(V3 const &p, V3 &norm) const
{
  const int index = SectionIndex(p.z - kTolerance);
  if (index < 0) {
    if (index == -1) norm = V3(0., 0., -1.);
    if (index == -2) norm = V3(0., 0., 1.);
    return false;
  }
  bool valid = secs[index].Normal(p - V3(0, 0, shift[index]), norm);
  if (index + 1 < NSec() && std::fabs(p.z - zs[index + 1]) < kTolerance) {
    V3 norm2;
    const bool valid2 = secs[index + 1].Normal(p - V3(0, 0, shift[index + 1]), norm2);
    if (!valid && valid2) {
      norm  = norm2;
      valid = valid2;
    }
    if (valid && valid2) {
      const bool c2 = secs[index + 1].Contains(p + kTolerance * 10 * norm2 - V3(0, 0, shift[index + 1]));
      if (c2) {
        norm = norm2;
      } else {
        norm = norm + norm2;
        if (norm.Mag2() < kTolerance) norm = V3(p.x, p.y, 0);
      }
    }
  }
  if (valid) norm = Div(norm, norm.Mag());
  return valid;
}

// SphereImplementation::ApproxSurfaceNormalKernel (VecGeom/volumes/kernel/SphereImplementation.h:489-533)
static V3 SphereApproxNormal(Sphere const &s, V3 const &p)
{
  V3 norm(0., 0., 0.);
  const double radius = p.Mag();
  const double distRMax = std::fabs(radius - s.rmax);
  double distRMin = kInfLength, distPhi1 = kInfLength, distPhi2 = kInfLength, distTheta1 = kInfLength,
         distTheta2 = kInfLength;
  double distMin = distRMax;
  if (s.rmin > 0.) {
    distRMin = std::fabs(s.rmin - radius);
    distMin  = Min2(distRMin, distRMax);
  }
  if (!s.fullPhi) {
    distPhi1 = std::fabs(p.x * s.wedge.normal1.x + p.y * s.wedge.normal1.y);
    distPhi2 = std::fabs(p.x * s.wedge.normal2.x + p.y * s.wedge.normal2.y);
    distMin  = Min2(distMin, Min2(distPhi1, distPhi2));
  }
  if (!s.fullTheta) {
    const double rho = p.Perp();
    distTheta1       = s.theta.DistanceToLine(s.theta.slope1, rho, p.z);
    distTheta2       = s.theta.DistanceToLine(s.theta.slope2, rho, p.z);
    distMin          = Min2(distMin, Min2(distTheta1, distTheta2));
  }
  if (distMin == distRMax) norm = Unit(p);
  if (distMin == distRMin) norm = -Unit(p);
  if (distMin == distPhi1) norm = -s.wedge.normal1;
  if (distMin == distPhi2) norm = -s.wedge.normal2;
  if (distMin == distTheta1) norm = norm + SphereThetaNormal(s, true, p);
  if (distMin == distTheta2) norm = norm + SphereThetaNormal(s, false, p);
  return norm;
}

// ThetaCone::GetNormal1 / GetNormal2 (VecGeom/volumes/ThetaCone_Evolution.h:108-138): the gradient
// of the theta cone, towards the outside of the theta section (not unit length).
static V3 SphereThetaNormal(Sphere const &s, bool start, V3 const &p)
{
  if (start) {
    const V3 n(2. * p.x, 2. * p.y, -2. * s.theta.tanS2 * p.z);
    return s.theta.sTheta <= kPi / 2. ? -n : n;
  }
  const V3 n(2 * p.x, 2 * p.y, -2 * s.theta.tanE2 * p.z);
  return s.theta.eTheta <= kPi / 2. ? n : -n;
}

// SphereImplementation::Normal (VecGeom/volumes/kernel/SphereImplementation.h:537-637) via
// UnplacedSphere::Normal (VecGeom/volumes/UnplacedSphere.h:260-265): start from (1e-30, 1e-30, 1e-30)
// (the nearest surface's normal for points off the surface), add the normal of every surface the
// point is on unless it is outside, then normalise; valid when on at least one surface.
bool Sphere::Normal(V3 const &p, V3 &n) const
{
  V3 normal(1e-30, 1e-30, 1e-30);
  bool ci = false, co = false;
  ContainsInside(p, true, ci, co);
  if (ci || co) normal = SphereApproxNormal(*this, p);
  int noSurfaces  = 0;
  const double m2 = p.Mag2();
  const bool onOuter = m2 <= PlusTolSq(rmax) && m2 >= MinusTolSq(rmax); // SphereUtilities.h:65-70
  if (onOuter) noSurfaces++;
  if (!co && onOuter) normal = normal + Unit(p);
  if (rmin) {
    const bool onInner = m2 <= PlusTolSq(rmin) && m2 >= MinusTolSq(rmin);
    if (onInner) noSurfaces++;
    if (!co && onInner) normal = normal - Unit(p);
  }
  if (!fullPhi) {
    const bool onS = wedge.IsOnSurfaceGeneric(wedge.along1, wedge.normal1, p);
    const bool onE = wedge.IsOnSurfaceGeneric(wedge.along2, wedge.normal2, p);
    if (onS) noSurfaces++;
    if (onE) noSurfaces++;
    if (!co && onS) normal = normal - wedge.normal1;
    if (!co && onE) normal = normal - wedge.normal2;
  }
  if (!fullTheta) {
    const bool onS = theta.IsOnSurface(true, p);
    const bool onE = theta.IsOnSurface(false, p);
    if (onS) noSurfaces++;
    if (!co && onS) normal = normal + SphereThetaNormal(*this, true, p);
    if (onE) noSurfaces++;
    if (!co && onE) normal = normal + SphereThetaNormal(*this, false, p);
    if (!co && onS && onE && (etheta <= kPi / 2.)) normal = V3(0., 0., -1.);
    if (!co && onS && onE && (stheta >= kPi / 2.)) normal = V3(0., 0., 1.);
  }
  n = normal * (1. / normal.Mag());
  return noSurfaces > 0;
}

// PolyhedronImplementation::ScalarNormalKernel (VecGeom/volumes/kernel/PolyhedronImplementation.h:
// 998-1067) via UnplacedPolyhedron::Normal: the normal of the nearest surface found by the safety
// search (z segments outwards from the point's, then the end caps). Always valid.
bool Polyhedron::Normal(V3 const &p, V3 &normal) const
{
  double safety    = kInfLength;
  const int zMax   = NSeg();
  const int zIndex = FindZSegment(p.z);
  if (zIndex < 0) {
    normal = V3(0, 0, -1);
    return true;
  }
  if (zIndex >= zMax) {
    normal = V3(0, 0, 1);
    return true;
  }
  int iSeg = zIndex, iSurf = -1, iz = 0;
  int phiIndex = FindPhiSegment(p);
  for (int z = zIndex; z < zMax;) {
    int iSurfCrt           = -1;
    const double safetySeg = SafetyToZSegmentSquared(z, phiIndex, p, true, &iSurfCrt);
    if (safetySeg < safety) {
      safety = safetySeg;
      iSeg   = z;
      iSurf  = iSurfCrt;
    }
    ++z;
    const double dz = zs[z] - p.z;
    if (dz * dz > safety) break;
  }
  for (int z = zIndex - 1; z >= 0; --z) {
    int iSurfCrt           = -1;
    const double safetySeg = SafetyToZSegmentSquared(z, phiIndex, p, true, &iSurfCrt);
    if (safetySeg < safety) {
      safety = safetySeg;
      iSeg   = z;
      iSurf  = iSurfCrt;
    }
    const double dz = p.z - zs[z];
    if (dz * dz > safety) break;
  }
  SafetyToEndcapsSquared(p, safety, &iz);
  if (iz != 0) {
    normal = V3(0, 0, iz);
    return true;
  }
  ZSeg const &S = segs[iSeg];
  // the reference indexes the quadrilateral sets without a check; an index that is out of range
  // (an unreachable combination for the solids the constructor builds) gives the zero vector here
  auto quadNormal = [](Quads const &q, int i) { return (i >= 0 && i < q.size()) ? q.q[i].n : V3(0, 0, 0); };
  if (iSurf >= 0 && iSurf < 2)
    normal = quadNormal(S.phi, iSurf);
  else if (iSurf == 2)
    normal = -1. * quadNormal(S.inner, phiIndex);
  else
    normal = quadNormal(S.outer, phiIndex);
  return true;
}

// UnplacedTorus2::Normal (source/UnplacedTorus2.cpp:252-330)
bool Torus2::Normal(V3 const &p, V3 &norm) const
{
  int noSurfaces    = 0;
  bool valid        = true;
  double distRMin   = kInfLength;
  double distSPhi   = kInfLength, distEPhi = kInfLength;
  const double delta  = Max2(10.0 * kTolerance, 1.0e-8 * (rtor + rmax));
  const double dAngle = 10.0 * kTolerance;
  V3 nR, nPs, nPe, sumnorm(0., 0., 0.);
  const double rho2 = p.x * p.x + p.y * p.y;
  const double rho  = std::sqrt(rho2);
  double pt2        = rho2 + p.z * p.z + rtor * (rtor - 2 * rho);
  pt2               = Max2(pt2, 0.);
  const double pt   = std::sqrt(pt2);
  const double distRMax = std::fabs(pt - rmax);
  if (rmin) distRMin = std::fabs(pt - rmin);
  if (rho > delta && pt != 0.) {
    const double redFactor = (rho - rtor) / rho;
    nR                     = V3(p.x * redFactor, p.y * redFactor, p.z);
    nR                     = nR * (1. / pt);
  }
  if (dphi < kTwoPi) {
    if (rho) {
      double pPhi = std::atan2(p.y, p.x);
      if (pPhi < sphi - delta)
        pPhi += kTwoPi;
      else if (pPhi > sphi + dphi + delta)
        pPhi -= kTwoPi;
      distSPhi = std::fabs(pPhi - sphi);
      distEPhi = std::fabs(pPhi - sphi - dphi);
    }
    nPs = V3(std::sin(sphi), -std::cos(sphi), 0);
    nPe = V3(-std::sin(sphi + dphi), std::cos(sphi + dphi), 0);
  }
  if (distRMax <= delta) {
    noSurfaces++;
    sumnorm = sumnorm + nR;
  } else if (rmin && (distRMin <= delta)) {
    noSurfaces++;
    sumnorm = sumnorm - nR;
  }
  if ((dphi < kTwoPi) && (rmin - delta <= pt) && (pt <= (rmax + delta))) {
    if (distSPhi <= dAngle) {
      noSurfaces++;
      sumnorm = sumnorm + nPs;
    }
    if (distEPhi <= dAngle) {
      noSurfaces++;
      sumnorm = sumnorm + nPe;
    }
  }
  if (noSurfaces == 0)
    valid = false;
  else if (noSurfaces == 1)
    norm = sumnorm;
  else
    norm = Unit(sumnorm);
  return valid;
}

// TessellatedImplementation::NormalKernel (VecGeom/volumes/kernel/TessellatedImplementation.h:162-172)
// via UnplacedTessellated::Normal (source/UnplacedTessellated.cpp:83-89): the normal of the facet that
// SafetySq<false> finds nearest. Always valid.
bool Tessellated::Normal(V3 const &p, V3 &n) const
{
  int isurf = -1;
  SafetySq(false, p, isurf);
  n = isurf >= 0 ? f[isurf].n : V3(0, 0, 0);
  return true;
}

} // namespace vgo
