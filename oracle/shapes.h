// Oracle shapes: scalar FP64 restatement of VecGeom's unplaced-shape kernels.
// TEST INFRASTRUCTURE ONLY (see common.h).
//
// Each shape mirrors the VUnplacedVolume virtual interface
// (VecGeom/volumes/UnplacedVolume.h:88-223) as reached through
// CommonUnplacedVolumeImplHelper (VecGeom/volumes/UnplacedVolumeImplHelper.h:77-180):
// the output is initialised to kInfLength and the static kernel is called.
#pragma once
#include "common.h"
#include <vector>

namespace vgo {

class Shape {
public:
  virtual ~Shape() {}
  virtual bool Contains(V3 const &p) const = 0;
  virtual int Inside(V3 const &p) const = 0;
  virtual double DistanceToIn(V3 const &p, V3 const &d, double step_max) const = 0;
  virtual double DistanceToOut(V3 const &p, V3 const &d, double step_max) const = 0;
  virtual double SafetyToIn(V3 const &p) const = 0;
  virtual double SafetyToOut(V3 const &p) const = 0;
  virtual void Extent(V3 &lo, V3 &hi) const = 0;
  // VUnplacedVolume::Normal (VecGeom/volumes/UnplacedVolume.h:252), restated in normal.cpp: n is the
  // caller's vector (left as it was where the reference leaves it unassigned); returns valid
  virtual bool Normal(V3 const &p, V3 &n) const = 0;
};

// Phi wedge: VecGeom/volumes/Wedge_Evolution.h, ctor source/Wedge_Evolution.cpp:17-128
struct Wedge;
// Box reduced to the phi range of a wedge: source/UnplacedTube.cpp:220-262 (the same rule is used by
// the cone, polycone and sphere extents). C, D at rOut and A, B at rIn, at the start and end angles.
void PhiExtent(Wedge const &w, double rOut, double rIn, double sphi, double dphi, V3 &lo, V3 &hi);

struct Wedge {
  double dphi = 0;
  V3 along1, along2, normal1, normal2;
  void Set(double angle, double zeroangle);
  bool IsOnSurfaceGeneric(V3 const &along, V3 const &normal, V3 const &p) const;
  void ContainsInside(V3 const &p, bool forInside, bool &cin, bool &cout) const;
  double SafetyToOut(V3 const &p) const;
  double SafetyToIn(V3 const &p) const;
  // Wedge_Evolution.h:293-303 (Contains), :235-247 (IsPointOnSurfaceAndMovingOut), :394-436 (DistanceToIn/Out)
  bool Contains(V3 const &p) const
  {
    bool ci = false, co = false;
    ContainsInside(p, false, ci, co);
    return !co;
  }
  bool OnSurfaceMoving(bool start, bool movingOut, V3 const &p, V3 const &d) const;
  void DistanceToIn(V3 const &p, V3 const &d, double &d1, double &d2) const;
  void DistanceToOut(V3 const &p, V3 const &d, double &d1, double &d2) const;
};

// VecGeom/volumes/ThetaCone_Evolution.h:42-820 (theta section of the sphere)
struct ThetaCone {
  double sTheta = 0, dTheta = 0, eTheta = 0, angTol = 0, halfAngTol = 0;
  double tanS = 0, tanE = 0, tanBis = 0, slope1 = 0, slope2 = 0, tanS2 = 0, tanE2 = 0;
  void Set(double s, double d);
  V3 Normal(bool start, V3 const &p) const;
  bool IsOnSurface(bool start, V3 const &p) const;
  bool OnSurfaceMoving(bool start, bool movingOut, V3 const &p, V3 const &d) const;
  bool IsCompletelyInside(V3 const &p) const;
  bool IsCompletelyOutside(V3 const &p) const;
  bool Contains(V3 const &p) const { return !IsCompletelyOutside(p); }
  double DistanceToLine(double slope, double x, double y) const;
  double SafetyToIn(V3 const &p) const;
  double SafetyToOut(V3 const &p) const;
  void DistanceToIn(V3 const &p, V3 const &d, double &d1, double &d2, bool &i1, bool &i2) const;
  void DistanceToOut(V3 const &p, V3 const &d, double &d1, double &d2, bool &i1, bool &i2) const;
};

// ---------------------------------------------------------------- Box
// VecGeom/volumes/kernel/BoxImplementation.h:76-167
class Box : public Shape {
public:
  V3 h;
  Box(double dx, double dy, double dz) : h(dx, dy, dz) {}
  bool Contains(V3 const &p) const override;
  int Inside(V3 const &p) const override;
  double DistanceToIn(V3 const &p, V3 const &d, double step_max) const override;
  double DistanceToOut(V3 const &p, V3 const &d, double step_max) const override;
  double SafetyToIn(V3 const &p) const override;
  double SafetyToOut(V3 const &p) const override;
  bool Normal(V3 const &p, V3 &n) const override;
  void Extent(V3 &lo, V3 &hi) const override { lo = -h; hi = h; }
};

// ---------------------------------------------------------------- Tube
// VecGeom/volumes/TubeStruct.h:15-176, VecGeom/volumes/kernel/TubeImplementation.h:389-964
// (UniversalTube: rmin and phi treatment decided at run time, TubeTypes.h:92-144)
class Tube : public Shape {
public:
  double rmin, rmax, z, sphi, dphi;
  double rmin2, rmax2, a1x, a1y, a2x, a2y;
  double tolIrmin2, tolOrmin2, tolIrmax2, tolOrmax2, tolIz, tolOz;
  double tolIrmin, tolOrmin, tolIrmax, tolOrmax, maxVal;
  Wedge wedge;
  Tube(double rmin, double rmax, double dz, double sphi, double dphi);
  bool hasRmin() const { return rmin > 0; }
  bool hasPhi() const { return dphi < kTwoPi; }
  bool Contains(V3 const &p) const override;
  int Inside(V3 const &p) const override;
  double DistanceToIn(V3 const &p, V3 const &d, double step_max) const override;
  double DistanceToOut(V3 const &p, V3 const &d, double step_max) const override;
  double SafetyToIn(V3 const &p) const override;
  double SafetyToOut(V3 const &p) const override;
  bool Normal(V3 const &p, V3 &n) const override;
  void Extent(V3 &lo, V3 &hi) const override;
  // internals shared with other shapes
  void ContainsInside(V3 const &p, bool forInside, bool &cin, bool &cout) const;
  double DistanceToInKernel(V3 const &p, V3 const &d) const;
  bool InSector(double x, double y, bool includeSurface) const;
  bool CircleHit(double b, double c, V3 const &p, V3 const &d, bool largest, bool sectorCheck, double &dist) const;
  bool PhiPlaneHit(double ax, double ay, double nx, double ny, V3 const &p, V3 const &d, bool posDir, bool toIn,
                   double &dist) const;
  double PhiPlaneSafetyOutside(V3 const &p) const;
};

// Factory used by the C ABI. Parameter meaning per type is listed in capi.cpp.
Shape *MakeShape(int type, const double *params, int nparams);

} // namespace vgo

namespace vgo {
// ---------------------------------------------------------------- Trapezoid
// VecGeom/volumes/TrapezoidStruct.h, source/UnplacedTrapezoid.cpp:447-675 (planes from corners),
// VecGeom/volumes/kernel/TrapezoidImplementation.h:98-400 with PlaneShell<4> (VecGeom/base/PlaneShell.h:35-270)
class Trapezoid : public Shape {
public:
  double dz, theta, phi, dy1, dx1, dx2, tanAlpha1, dy2, dx3, dx4, tanAlpha2, tthetaCphi, tthetaSphi;
  double A[4], B[4], C[4], D[4];
  Trapezoid(double dz, double theta, double phi, double dy1, double dx1, double dx2, double alpha1, double dy2,
            double dx3, double dx4, double alpha2);
  explicit Trapezoid(V3 const pt[8]);
  void SetPlanes(V3 const pt[8]);
  bool Contains(V3 const &p) const override;
  int Inside(V3 const &p) const override;
  double DistanceToIn(V3 const &p, V3 const &d, double step_max) const override;
  double DistanceToOut(V3 const &p, V3 const &d, double step_max) const override;
  double SafetyToIn(V3 const &p) const override;
  double SafetyToOut(V3 const &p) const override;
  bool Normal(V3 const &p, V3 &n) const override;
  void Extent(V3 &lo, V3 &hi) const override;
  void ContainsInside(V3 const &p, bool forInside, bool &cin, bool &cout) const;
};
} // namespace vgo

namespace vgo {
// ---------------------------------------------------------------- Orb
// VecGeom/volumes/kernel/OrbImplementation.h:65-220
class Orb : public Shape {
public:
  double r;
  explicit Orb(double r_) : r(r_) {}
  bool Contains(V3 const &p) const override;
  int Inside(V3 const &p) const override;
  double DistanceToIn(V3 const &p, V3 const &d, double step_max) const override;
  double DistanceToOut(V3 const &p, V3 const &d, double step_max) const override;
  double SafetyToIn(V3 const &p) const override;
  double SafetyToOut(V3 const &p) const override;
  bool Normal(V3 const &p, V3 &n) const override;
  void Extent(V3 &lo, V3 &hi) const override { lo = V3(-r, -r, -r), hi = V3(r, r, r); }
};

// ---------------------------------------------------------------- Hype (UniversalHype)
// VecGeom/volumes/HypeStruct.h:16-146 (cached members), VecGeom/volumes/kernel/HypeImplementation.h:40-330,
// VecGeom/volumes/HypeUtilities.h:25-438 (scalar HypeHelpers<Precision> specialisation :374-436)
class Hype : public Shape {
public:
  double rmin, rmax, stIn, stOut, dz;
  double tIn, tOut, tIn2, tOut2, rmin2, rmax2, dz2;
  double endInnerRadius2, endOuterRadius2, endInnerRadius, endOuterRadius;
  double zTol, innerRadTol, outerRadTol;
  Hype(double rmin, double rmax, double stIn, double stOut, double dz);
  bool inner() const { return rmin > 0. || stIn != 0.; } // InnerSurfaceExists (HypeStruct.h:87)
  bool Contains(V3 const &p) const override;
  int Inside(V3 const &p) const override;
  double DistanceToIn(V3 const &p, V3 const &d, double step_max) const override;
  double DistanceToOut(V3 const &p, V3 const &d, double step_max) const override;
  double SafetyToIn(V3 const &p) const override;
  double SafetyToOut(V3 const &p) const override;
  bool Normal(V3 const &p, V3 &n) const override;
  void Extent(V3 &lo, V3 &hi) const override
  {
    lo = V3(-endOuterRadius, -endOuterRadius, -dz), hi = V3(endOuterRadius, endOuterRadius, dz);
  }
  void ContainsInside(V3 const &p, bool forInside, bool &cin, bool &cout) const;
  bool CompletelyInside(V3 const &p) const;
  bool CompletelyOutside(V3 const &p) const;
  bool OnSurfaceMovingIn(V3 const &p, V3 const &d) const;
  bool OnSurfaceMovingOut(V3 const &p, V3 const &d) const;
  bool OnOuterMovingOut(V3 const &p, V3 const &d) const;
  bool OnInnerMovingOut(V3 const &p, V3 const &d) const;
  bool ZPlane(V3 const &p, V3 const &d, bool toIn, double &zDist) const;
  bool Hyperbolic(V3 const &p, V3 const &d, bool toIn, bool innerSurf, double &dist) const;
};

// ---------------------------------------------------------------- Trd (UniversalTrd)
// VecGeom/volumes/TrdStruct.h:18-121 (cached members), VecGeom/volumes/kernel/TrdImplementation.h:25-518
// with TrdTypes::UniversalTrd (VecGeom/volumes/kernel/shapetypes/TrdTypes.h:46-76): the y faces are
// taken as varying whenever dy1 != dy2, and the Inside/Contains y test always uses the cross product.
class Trd : public Shape {
public:
  double dx1, dx2, dy1, dy2, dz;
  double x2minusx1, y2minusy1, halfx1plusx2, halfy1plusy2, fx, fy, secxz, secyz, calfx, calfy, tolx, toly;
  bool plainY = false; // TrdTypes::Trd1 (HasVaryingY kNo): Inside/Contains test |y| against dy1
  Trd(double x1, double x2, double y1, double y2, double z, bool trd1 = false);
  bool varyingY() const { return dy1 != dy2; }
  bool Contains(V3 const &p) const override;
  int Inside(V3 const &p) const override;
  double DistanceToIn(V3 const &p, V3 const &d, double step_max) const override;
  double DistanceToOut(V3 const &p, V3 const &d, double step_max) const override;
  double SafetyToIn(V3 const &p) const override;
  double SafetyToOut(V3 const &p) const override;
  bool Normal(V3 const &p, V3 &n) const override;
  void Extent(V3 &lo, V3 &hi) const override
  {
    const double mx = dx1 > dx2 ? dx1 : dx2, my = dy1 > dy2 ? dy1 : dy2;
    lo = V3(-mx, -my, -dz), hi = V3(mx, my, dz);
  }
  void InsideFlags(V3 const &p, bool surface, bool &cin, bool &cout) const;
  bool Face(bool forY, bool mirrored, bool toIn, V3 const &p, V3 const &d, double &dist) const;
  double Safety(V3 const &p, bool inside) const;
};

// ---------------------------------------------------------------- Paraboloid
// VecGeom/volumes/ParaboloidStruct.h:17-100 (cached members), VecGeom/volumes/kernel/ParaboloidImplementation.h:70-420
class Paraboloid : public Shape {
public:
  double rlo, rhi, dz;
  double dx, dy, a, invA, a2, b, b2, invB, k1, k2, rlo2, rhi2;
  Paraboloid(double rlo, double rhi, double dz);
  bool Contains(V3 const &p) const override;
  int Inside(V3 const &p) const override;
  double DistanceToIn(V3 const &p, V3 const &d, double step_max) const override;
  double DistanceToOut(V3 const &p, V3 const &d, double step_max) const override;
  double SafetyToIn(V3 const &p) const override;
  double SafetyToOut(V3 const &p) const override;
  bool Normal(V3 const &p, V3 &n) const override;
  void Extent(V3 &lo, V3 &hi) const override { lo = V3(-dx, -dy, -dz), hi = V3(dx, dy, dz); }
  void Flags(V3 const &p, bool forInside, bool &cin, bool &cout) const;
  bool OnZPlane(bool top, V3 const &p) const;
};

// ---------------------------------------------------------------- Parallelepiped
// VecGeom/volumes/ParallelepipedStruct.h:80-124 (cached tangents, normals, safety scale factors),
// VecGeom/volumes/kernel/ParallelepipedImplementation.h:60-236, source/UnplacedParallelepiped.cpp:79-89
class Parallelepiped : public Shape {
public:
  V3 dim;
  double tanAlpha, tanThetaSinPhi, tanThetaCosPhi, cosTheta, ctx, cty;
  V3 normals[3];
  Parallelepiped(double dx, double dy, double dz, double alpha, double theta, double phi);
  bool Contains(V3 const &p) const override;
  int Inside(V3 const &p) const override;
  double DistanceToIn(V3 const &p, V3 const &d, double step_max) const override;
  double DistanceToOut(V3 const &p, V3 const &d, double step_max) const override;
  double SafetyToIn(V3 const &p) const override;
  double SafetyToOut(V3 const &p) const override;
  bool Normal(V3 const &p, V3 &n) const override;
  void Extent(V3 &lo, V3 &hi) const override;
  V3 Oblique(V3 p) const;           // ParallelepipedImplementation::Transform
  V3 SafetyVector(V3 const &lp) const;
};

// ---------------------------------------------------------------- Tet
// VecGeom/volumes/TetStruct.h:57-90 (face planes from the corners, outward),
// VecGeom/volumes/kernel/TetImplementation.h:61-221, source/UnplacedTet.cpp:34-50
class Tet : public Shape {
public:
  V3 vertex[4];
  V3 n[4];
  double d[4];
  Tet(V3 const &p0, V3 const &p1, V3 const &p2, V3 const &p3);
  bool Contains(V3 const &p) const override;
  int Inside(V3 const &p) const override;
  double DistanceToIn(V3 const &p, V3 const &d, double step_max) const override;
  double DistanceToOut(V3 const &p, V3 const &d, double step_max) const override;
  double SafetyToIn(V3 const &p) const override;
  double SafetyToOut(V3 const &p) const override;
  bool Normal(V3 const &p, V3 &nrm) const override;
  void Extent(V3 &lo, V3 &hi) const override;
  double MaxPlaneDistance(V3 const &p) const;
};

// ---------------------------------------------------------------- EllipticalTube
// source/UnplacedEllipticalTube.cpp:28-71 (CheckParameters: cached scale factors and coefficients),
// VecGeom/volumes/kernel/EllipticalTubeImplementation.h:61-300
class EllipticalTube : public Shape {
public:
  double dx, dy, dz, rsph, ddx, ddy, sx, sy, r, q1, q2, scratch;
  EllipticalTube(double dx, double dy, double dz);
  bool Contains(V3 const &p) const override;
  int Inside(V3 const &p) const override;
  double DistanceToIn(V3 const &p, V3 const &d, double step_max) const override;
  double DistanceToOut(V3 const &p, V3 const &d, double step_max) const override;
  double SafetyToIn(V3 const &p) const override;
  double SafetyToOut(V3 const &p) const override;
  bool Normal(V3 const &p, V3 &n) const override;
  void Extent(V3 &lo, V3 &hi) const override { lo = V3(-dx, -dy, -dz), hi = V3(dx, dy, dz); }
  double Safety(V3 const &p) const;
};

// ---------------------------------------------------------------- Ellipsoid
// source/UnplacedEllipsoid.cpp:27-116 (CheckParameters), VecGeom/volumes/kernel/EllipsoidImplementation.h:61-318
class Ellipsoid : public Shape {
public:
  double dx, dy, dz, zBottomCut, zTopCut, xmax, ymax, rsph, r, sx, sy, sz, scZMidCut, scZDimCut, q1, q2;
  Ellipsoid(double dx, double dy, double dz, double zBottomCut, double zTopCut);
  bool Contains(V3 const &p) const override;
  int Inside(V3 const &p) const override;
  double DistanceToIn(V3 const &p, V3 const &d, double step_max) const override;
  double DistanceToOut(V3 const &p, V3 const &d, double step_max) const override;
  double SafetyToIn(V3 const &p) const override;
  double SafetyToOut(V3 const &p) const override;
  bool Normal(V3 const &p, V3 &n) const override;
  void Extent(V3 &lo, V3 &hi) const override { lo = V3(-xmax, -ymax, zBottomCut), hi = V3(xmax, ymax, zTopCut); }
  double Safety(V3 const &p) const;
};

// ---------------------------------------------------------------- GenTrap (planar)
// VecGeom/volumes/GenTrapStruct.h:57-190 (vertex checks, clockwise reversal, twist), the scalar CPU path of
// VecGeom/volumes/kernel/GenTrapImplementation.h:330-608, which hands planar traps to the
// TessellatedSection helper (VecGeom/volumes/TessellatedSection.h, TessellatedCluster.h, Tile.h with one
// facet per cluster), and NormalKernel (:610-735) with SecondOrderSurfaceShell's face normals
// (VecGeom/volumes/SecondOrderSurfaceShell.h:65-122). Twisted traps are not built (ok = false).
class GenTrap : public Shape {
public:
  double dz, halfInverseDz;
  V3 vertices[8];
  double vx[8], vy[8], connX[4], connY[4];
  bool twisted = false, ok = false;
  int nfacets = 0;
  V3 fnormal[4];
  double fdist[4];
  V3 shellNormal[4];
  GenTrap(const double *x, const double *y, double h);
  bool Contains(V3 const &p) const override;
  int Inside(V3 const &p) const override;
  double DistanceToIn(V3 const &p, V3 const &d, double step_max) const override;
  double DistanceToOut(V3 const &p, V3 const &d, double step_max) const override;
  double SafetyToIn(V3 const &p) const override;
  double SafetyToOut(V3 const &p) const override;
  bool Normal(V3 const &p, V3 &n) const override;
  void Extent(V3 &lo, V3 &hi) const override;
  double DistPlane(int i, V3 const &p) const { return p.Dot(fnormal[i]) + fdist[i]; }
};

// ---------------------------------------------------------------- EllipticalCone
// VecGeom/volumes/EllipticalConeStruct.h, source/UnplacedEllipticalCone.cpp:28-75 (parameter clamps,
// precalculated values, extent), VecGeom/volumes/kernel/EllipticalConeImplementation.h:63-386
class EllipticalCone : public Shape {
public:
  double dx, dy, dz, zCut, rsph, invDx, invDy, cosAxisMin, dApex;
  EllipticalCone(double a, double b, double h, double zcut);
  double Safety(V3 const &p) const;
  bool Contains(V3 const &p) const override;
  int Inside(V3 const &p) const override;
  double DistanceToIn(V3 const &p, V3 const &d, double step_max) const override;
  double DistanceToOut(V3 const &p, V3 const &d, double step_max) const override;
  double SafetyToIn(V3 const &p) const override;
  double SafetyToOut(V3 const &p) const override;
  bool Normal(V3 const &p, V3 &n) const override;
  void Extent(V3 &lo, V3 &hi) const override;
};

// ---------------------------------------------------------------- CutTube
// VecGeom/volumes/CutTubeStruct.h:36-47 (a tube of infinite half-length and two cut planes through
// (0, 0, -dz) and (0, 0, dz)), VecGeom/volumes/CutPlanes.h:123-173 with Plane.h:44-191,
// VecGeom/volumes/kernel/CutTubeImplementation.h:84-370, source/UnplacedCutTube.cpp:128-192 (extent,
// normal), VecGeom/volumes/UnplacedCutTube.h:36-50 (constructor check).
class CutTube : public Shape {
public:
  Tube tube; // rmin, rmax, kInfLength, sphi, dphi
  double dz, maxVal, cosPhi1, sinPhi1, cosPhi2, sinPhi2;
  V3 n[2];        // bottom, top unit normals (outward)
  double dist[2]; // plane offsets: DistPlane = p.n + dist
  bool ok = false;
  CutTube(double rmin, double rmax, double z, double sphi, double dphi, V3 const &bottom, V3 const &top);
  double DistPlane(int i, V3 const &p) const { return p.Dot(n[i]) + dist[i]; }
  int CutInside(V3 const &p) const;
  bool Contains(V3 const &p) const override;
  int Inside(V3 const &p) const override;
  double DistanceToIn(V3 const &p, V3 const &d, double step_max) const override;
  double DistanceToOut(V3 const &p, V3 const &d, double step_max) const override;
  double SafetyToIn(V3 const &p) const override;
  double SafetyToOut(V3 const &p) const override;
  bool Normal(V3 const &p, V3 &n) const override;
  void Extent(V3 &lo, V3 &hi) const override;
  double DistanceToInKernel(V3 const &p, V3 const &d, double step_max) const;
};

// ---------------------------------------------------------------- Cone (UniversalCone)
// VecGeom/volumes/ConeStruct.h:15-431, VecGeom/volumes/kernel/ConeImplementation.h:95-521 and the scalar
// ConeHelpers<Precision> specialisation in VecGeom/volumes/ConeUtilities.h:580-820
class Cone : public Shape {
public:
  double rmin1, rmax1, rmin2, rmax2, dz, sphi, dphi;
  double origRmax1, origRmax2, origRmin1, origRmin2; // constructor values (_frmax1.. in ConeStruct)
  double a1x, a1y, a2x, a2y;
  Wedge wedge;
  double innerSlope, outerSlope, innerOffset, outerOffset, innerTol, outerTol;
  double sqRmin1, sqRmin2, sqRmax1, sqRmax2, tolIz, tolOz;
  double innerApex, tanInnerApex, outerApex, tanOuterApex;
  double secRMin, secRMax, invSecRMin, invSecRMax, tanRMin, tanRMax, zNormInner, zNormOuter;
  Cone(double rmin1, double rmax1, double rmin2, double rmax2, double dz, double sphi, double dphi);
  bool hasRmin() const { return rmin1 > 0 || rmin2 > 0; }
  bool hasPhi() const { return dphi < kTwoPi; }
  bool Contains(V3 const &p) const override;
  int Inside(V3 const &p) const override;
  double DistanceToIn(V3 const &p, V3 const &d, double step_max) const override;
  double DistanceToOut(V3 const &p, V3 const &d, double step_max) const override;
  double SafetyToIn(V3 const &p) const override;
  double SafetyToOut(V3 const &p) const override;
  bool Normal(V3 const &p, V3 &n) const override;
  void Extent(V3 &lo, V3 &hi) const override;
  // helpers shared with the polycone
  double RadiusAt(bool inner, double z) const;
  bool InSector(double x, double y, bool includeSurface) const;
  void ContainsInside(V3 const &p, bool forInside, bool &cin, bool &cout) const;
  bool ConicalHit(bool toIn, bool inner, V3 const &p, V3 const &d, double &dist) const;
  bool PhiPlaneHit(double ax, double ay, double nx, double ny, V3 const &p, V3 const &d, bool toIn,
                   double &dist) const;
  V3 Normal(bool inner, V3 const &p) const;
  double SafeDistConical(bool inner, V3 const &p) const;
};
} // namespace vgo

namespace vgo {
// ---------------------------------------------------------------- Polycone
// VecGeom/volumes/PolyconeStruct.h:144-290 (sections = cones between consecutive distinct z planes,
// shifted by their mid z), VecGeom/volumes/kernel/PolyconeImplementation.h:70-520
class Polycone : public Shape {
public:
  double sphi, dphi;
  std::vector<double> zs;        // section boundaries
  std::vector<double> shift;     // per section
  std::vector<Cone> secs;
  double rmaxExtent = 0, zlo = 0, zhi = 0;
  double startPhi = 0, deltaPhi = 0; // fStartPhi, fDeltaPhi (PolyconeStruct.h:152-184)
  Polycone(double sphi, double dphi, int nz, const double *z, const double *rmin, const double *rmax);
  int NSec() const { return int(secs.size()); }
  int SectionIndex(double z) const;
  void SectionInOut(int isec, V3 const &p, bool forInside, bool &cin, bool &cout) const;
  void ContainsInside(V3 const &p, bool forInside, bool &cin, bool &cout) const;
  bool Contains(V3 const &p) const override;
  int Inside(V3 const &p) const override;
  double DistanceToIn(V3 const &p, V3 const &d, double step_max) const override;
  double DistanceToOut(V3 const &p, V3 const &d, double step_max) const override;
  double SafetyToIn(V3 const &p) const override;
  double SafetyToOut(V3 const &p) const override;
  bool Normal(V3 const &p, V3 &n) const override;
  void Extent(V3 &lo, V3 &hi) const override;
};
} // namespace vgo

namespace vgo {
// ---------------------------------------------------------------- Sphere
// VecGeom/volumes/SphereStruct.h:25-300, VecGeom/volumes/kernel/SphereImplementation.h:65-417,
// VecGeom/volumes/SphereUtilities.h:55-195, extent source/UnplacedSphere.cpp:76-165
class Sphere : public Shape {
public:
  double rmin, rmax, sphi, dphi, stheta, dtheta, etheta;
  bool fullPhi = true, fullTheta = true, full = true;
  Wedge wedge;     // built from the raw (sphi, dphi) parameters
  ThetaCone theta; // built from the raw (stheta, dtheta) parameters
  Sphere(double rmin, double rmax, double sphi, double dphi, double stheta, double dtheta);
  void ContainsInside(V3 const &p, bool forInside, bool &cin, bool &cout) const;
  bool CompletelyInside(V3 const &p) const;
  bool CompletelyOutside(V3 const &p) const;
  bool OnSurfaceMoving(bool movingOut, V3 const &p, V3 const &d) const;
  void MinDistFromPhi(bool toIn, V3 const &p, V3 const &d, double &distance) const;
  bool Contains(V3 const &p) const override;
  int Inside(V3 const &p) const override;
  double DistanceToIn(V3 const &p, V3 const &d, double step_max) const override;
  double DistanceToOut(V3 const &p, V3 const &d, double step_max) const override;
  double SafetyToIn(V3 const &p) const override;
  double SafetyToOut(V3 const &p) const override;
  bool Normal(V3 const &p, V3 &n) const override;
  void Extent(V3 &lo, V3 &hi) const override;
};
} // namespace vgo

namespace vgo {
// ---------------------------------------------------------------- Polyhedron
// VecGeom/volumes/Quadrilaterals.h:105-630 with source/Quadrilaterals.cpp:36-118 and
// source/Planes.cpp:30-52 (plane sets); VecGeom/volumes/PolyhedronStruct.h:91-430 (segments);
// VecGeom/volumes/kernel/PolyhedronImplementation.h:298-1200 (kernels, generic
// EInnerRadii/EPhiCutout specialisation); extent source/UnplacedPolyhedron.cpp:273-303.
struct Quad {
  V3 n;
  double d = 0;
  V3 sn[4];
  double sd[4] = {0, 0, 0, 0};
  V3 c[4];
};

struct Quads {
  std::vector<Quad> q;
  int size() const { return int(q.size()); }
  void Set(int i, V3 const &c0, V3 const &c1, V3 const &c2, V3 const &c3);
  void FlipSign(int i) { q[i].n = -q[i].n, q[i].d = -q[i].d; }
  bool Contains(V3 const &p) const;
  int Inside(V3 const &p) const;
  int InsideAt(V3 const &p, int i) const;
  bool RayHits(int i, V3 const &x) const;
  double DistanceToIn(bool behind, V3 const &p, V3 const &d) const;
  double DistanceToOut(V3 const &p, V3 const &d, double zMin, double zMax) const;
  double DistanceSquared(int i, V3 const &p) const;
};

struct ZSeg {
  Quads outer, phi, inner;
  bool hasInner() const { return inner.size() > 0; }
};

class Polyhedron : public Shape {
public:
  int sides = 0, nz = 0;
  bool hasInnerRadii = false, hasPhiCutout = false, hasLargePhiCutout = false;
  double phiStart = 0, phiDelta = 0;
  Wedge wedge;
  std::vector<ZSeg> segs;
  std::vector<double> zs, rmin, rmax;
  std::vector<char> sameZ;
  std::vector<V3> phiSections;
  Tube bounding{0, 1, 1, 0, kTwoPi};
  double boundingOffset = 0;
  Polyhedron(double phiStart, double phiDelta, int sides, int nz, const double *z, const double *rmin,
             const double *rmax);
  int NSeg() const { return int(segs.size()); }
  int FindZSegment(double z) const;
  int FindPhiSegment(V3 const &p) const;
  double DistanceToInZSegment(int i, V3 const &p, V3 const &d) const;
  double DistanceToOutZSegment(int i, double zMin, double zMax, V3 const &p, V3 const &d) const;
  double SafetyToZSegmentSquared(int i, int &phiIndex, V3 const &p, bool ptInside, int *iSurf = nullptr) const;
  void DistanceToEndcaps(bool inside, V3 const &p, V3 const &d, double &distance) const;
  void SafetyToEndcapsSquared(V3 const &p, double &dsq, int *iz = nullptr) const;
  bool InPhiCutoutWedge(ZSeg const &s, V3 const &p) const;
  int InsideSegBorder(V3 const &p, int zIndex) const;
  int InsideSegPhi(V3 const &p, int zIndex, int phiIndex) const;
  double Safety(V3 const &p, bool ptInside) const;
  int WedgeInside(V3 const &p) const;
  bool Contains(V3 const &p) const override;
  int Inside(V3 const &p) const override;
  double DistanceToIn(V3 const &p, V3 const &d, double step_max) const override;
  double DistanceToOut(V3 const &p, V3 const &d, double step_max) const override;
  double SafetyToIn(V3 const &p) const override { return Safety(p, false); }
  double SafetyToOut(V3 const &p) const override { return Safety(p, true); }
  bool Normal(V3 const &p, V3 &n) const override;
  void Extent(V3 &lo, V3 &hi) const override;
};
} // namespace vgo

namespace vgo {
// ---------------------------------------------------------------- Torus2
// VecGeom/volumes/TorusStruct2.h:22-45, VecGeom/volumes/kernel/TorusImplementation2.h:14-660 (quartic
// via the resolvent cubic, Newton polishing), bounding tube with UniversalTube kernels (the
// VECGEOM_NO_SPECIALIZATION branch of DistanceToIn, :560-575).
class Torus2 : public Shape {
public:
  double rmin, rmax, rtor, sphi, dphi, rmin2, rmax2, rtor2;
  Wedge wedge;
  Tube bounding{0, 1, 1, 0, kTwoPi};
  Torus2(double rmin, double rmax, double rtor, double sphi, double dphi);
  void ContainsInside(V3 const &p, bool forInside, bool notForDisk, bool &ci, bool &co) const;
  double ToBoundary(V3 const &pt, V3 const &dir, double radius, bool out) const;
  double DistSqrToTorusR(V3 const &p, V3 const &d, double dist) const;
  bool Contains(V3 const &p) const override;
  int Inside(V3 const &p) const override;
  double DistanceToIn(V3 const &p, V3 const &d, double step_max) const override;
  double DistanceToOut(V3 const &p, V3 const &d, double step_max) const override;
  double SafetyToIn(V3 const &p) const override;
  double SafetyToOut(V3 const &p) const override;
  bool Normal(V3 const &p, V3 &n) const override;
  void Extent(V3 &lo, V3 &hi) const override { bounding.Extent(lo, hi); }
};
// TorusImplementation2.h:14-190
unsigned SolveCubic(double a, double b, double c, double *x);
int SolveQuartic(double a, double b, double c, double d, double *x);
} // namespace vgo

namespace vgo {
// ---------------------------------------------------------------- Tessellated
// Triangular facets (VecGeom/volumes/Tile.h:58-160, TessellatedStruct<3,double>), queried through
// the scalar backend's one-facet clusters (VecGeom/volumes/TessellatedCluster.h:183-300,:470-540) and
// TessellatedImplementation.h:54-300. The reference loops over the clusters in the order of their
// bounding-box hit distance (HybridNavigator2.h:111-190) and keeps the first of equal distances;
// the restatement visits every facet that passes the same single-precision box test and keeps,
// among equal distances, the lowest facet index.
struct TriFacet {
  V3 v[3];     // vertices, anticlockwise seen from outside
  V3 side[3];  // cluster side vectors: normal x (v[i+1] - v[i]), normalised
  V3 n;        // outward unit normal
  double dist; // plane offset: n.p + dist = signed distance
  float blo[3], bhi[3]; // bounding box of the facet in single precision (ABBoxManager::ABBox_s)
};
class Tessellated : public Shape {
public:
  std::vector<TriFacet> f;
  V3 minExt, maxExt, testDir;
  bool ok = false;
  // tri: 9 doubles per triangle (v0, v1, v2)
  Tessellated(const double *tri, int ntri);
  void DistanceToSolid(bool toIn, V3 const &p, V3 const &d, double stepMax, double &distance, int &isurf,
                       double &distother, int &isurfother) const;
  double SafetySq(bool toIn, V3 const &p, int &isurf) const;
  bool Contains(V3 const &p) const override;
  int Inside(V3 const &p) const override;
  double DistanceToIn(V3 const &p, V3 const &d, double step_max) const override;
  double DistanceToOut(V3 const &p, V3 const &d, double step_max) const override;
  double SafetyToIn(V3 const &p) const override;
  double SafetyToOut(V3 const &p) const override;
  bool Normal(V3 const &p, V3 &n) const override;
  void Extent(V3 &lo, V3 &hi) const override
  {
    lo = minExt;
    hi = maxExt;
  }
};
// Direction used by Contains/Inside: the first of a fixed list with |n.dir| >= 1e-8 for every
// facet (the reference draws a random direction once, TessellatedStruct.h:380-390).
V3 TessellatedTestDir(std::vector<TriFacet> const &f);
} // namespace vgo
