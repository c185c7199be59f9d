// Flat shape-parameter records shared by the host builder and the device kernels.
//
// Each record holds the shape's defining parameters plus the cached members the
// reference precomputes in its Unplaced*/…Struct constructors, so kernels never
// recompute trigonometry. Records are POD so they can be passed by value as
// kernel arguments (scalar registers) for single-shape launches, or stored in a
// device table indexed by shape id for navigation and mixed batches.
#pragma once
#include <stdint.h>

namespace vgmi {

enum ShapeType : int32_t {
  kBox = 0,
  kTube = 1,
  kCone = 2,
  kPolycone = 3,
  kPolyhedron = 4,
  kTrapezoid = 5,
  kSphere = 6,
  kTorus2 = 7,
  kOrb = 8,
  kTessellated = 9,
  kHype = 10,
  kTrd = 11,
  kParaboloid = 12,
  kParallelepiped = 13,
  kTet = 14,
  kEllipticalTube = 15,
  kEllipsoid = 16,
  kGenTrap = 17,
  kCutTube = 18,
  kEllipticalCone = 19,
  kNumShapeTypes = 20
};

enum QueryOp : int32_t {
  kOpContains = 0,
  kOpInside = 1,
  kOpDistIn = 2,
  kOpDistOut = 3,
  kOpSafetyIn = 4,
  kOpSafetyOut = 5,
  kNumOps = 6
};

// Box: half lengths (VecGeom/volumes/BoxStruct.h:20)
struct BoxData {
  double h[3];
};

// Tube (VecGeom/volumes/TubeStruct.h:15-43): defining parameters, cached
// squares and tolerance bands, phi-sector unit vectors and wedge normals.
struct TubeData {
  double rmin, rmax, z, sphi, dphi;
  double rmin2, rmax2;
  double a1x, a1y, a2x, a2y;     // along vectors of the two phi planes
  double n1x, n1y, n2x, n2y;     // inward wedge normals
  double tolIrmin2, tolOrmin2, tolIrmax2, tolOrmax2;
  double maxVal;
  int32_t hasRmin, hasPhi;       // rmin > 0, dphi < 2 pi
  int32_t phiGtPi;               // dphi > pi   (PhiPlaneSafety)
  int32_t phiLePi;               // dphi <= pi  (PointInCyclicalSector)
  int32_t wedgeLtPi;             // dphi < pi   (Wedge kernels)
  int32_t pad_;
};

// Trapezoid (VecGeom/volumes/TrapezoidStruct.h): half length in z and the four
// side planes a*x + b*y + c*z + d = 0 with outward unit normals (PlaneShell<4>).
struct TrapezoidData {
  double dz;
  double a[4], b[4], c[4], d[4];
  double params[11]; // defining parameters, alphas already converted to tan(alpha)
};

// Orb (VecGeom/volumes/OrbStruct.h)
struct OrbData {
  double r;
};

// Cone (VecGeom/volumes/ConeStruct.h:15-140): defining parameters plus the cached
// slopes, offsets, apex data and tolerances of CalculateCached().
struct ConeData {
  double rmin1, rmax1, rmin2, rmax2, dz, sphi, dphi;
  double origRmax1, origRmax2, origRmin1, origRmin2; // constructor radii (extent)
  double a1x, a1y, a2x, a2y;             // phi-plane vectors of the normalised angles
  double w1x, w1y, w2x, w2y;             // wedge along vectors (raw input angles)
  double n1x, n1y, n2x, n2y;             // wedge inward normals (raw input angles)
  double innerSlope, outerSlope, innerOffset, outerOffset, innerTol, outerTol;
  double sqRmin1, sqRmin2, sqRmax1, sqRmax2, tolIz;
  double innerApex, tanInnerApex, outerApex, tanOuterApex;
  double secRMin, secRMax, invSecRMin, invSecRMax, tanRMin, tanRMax, zNormInner, zNormOuter;
  int32_t hasRmin, hasPhi, phiLePi, wedgeLtPi;
};

// Polycone (VecGeom/volumes/PolyconeStruct.h:144-290): nsec cone sections; the
// section records, their z shifts and the nsec+1 section boundaries live in the
// shape's variable-length block (doubles), addressed from `base` once uploaded.
struct PolyconeData {
  int32_t nsec;
  int32_t secOff;   // ConeData[nsec]
  int32_t shiftOff; // double[nsec]
  int32_t zsOff;    // double[nsec + 1]
  const double *base;
  double ext[6]; // extent lo[3], hi[3]
};

// Sphere (VecGeom/volumes/SphereStruct.h:25-266): shell radii, the normalised phi
// range and clamped theta range, plus the wedge vectors and theta-cone cache built
// from the raw constructor angles (Wedge_Evolution.cpp:17-108, ThetaCone_Evolution.h:62-83).
struct SphereData {
  double rmin, rmax, sphi, dphi, stheta, dtheta, etheta;
  double w1x, w1y, w2x, w2y;      // wedge along vectors
  double n1x, n1y, n2x, n2y;      // wedge inward normals
  double tS, tE;                  // theta-cone start and end angles (raw)
  double tanS, tanE, tanS2, tanE2, tanBis, slope1, slope2;
  int32_t fullPhi, fullTheta, full;
  int32_t wedgeLtPi;              // raw dphi < pi (Wedge kernels)
};

// Polyhedron (VecGeom/volumes/PolyhedronStruct.h:45-60): z segments of quadrilateral
// plane sets. The z planes, radii, same-z flags, phi-section vectors, a 6-entry header per
// segment (outer, phi, inner: offset and count) and 32 doubles per quadrilateral (plane
// normal and distance, four side planes, four corners) live in the variable-length block.
struct PolyhedronData {
  int32_t sides, nz, nseg;
  int32_t hasPhiCutout, hasLargePhiCutout, wedgeLtPi;
  int32_t zsOff, rminOff, rmaxOff, sameZOff, phiSecOff, segOff;
  const double *base;
  double phiStart, phiDelta, boundingOffset;
  double w1x, w1y, w2x, w2y, n1x, n1y, n2x, n2y; // phi wedge
  TubeData bounding;                             // bounding tube (UniversalTube kernels)
  double ext[6];                                 // extent lo[3], hi[3]
};
constexpr int kQuadDoubles = 32;

// Torus2 (VecGeom/volumes/TorusStruct2.h:22-45): radii and squares, the phi wedge of the
// raw angles and the bounding tube (rtor -/+ (rmax + kTolerance), half length rmax).
struct TorusData {
  double rmin, rmax, rtor, sphi, dphi, rmin2, rmax2, rtor2;
  double w1x, w1y, w2x, w2y, n1x, n1y, n2x, n2y;
  int32_t hasPhi, wedgeLtPi;
  TubeData bounding;
};

// Hype (VecGeom/volumes/HypeStruct.h:16-83): defining parameters and the cached tangents,
// squares, end radii and tolerance levels of CalculateCached.
struct HypeData {
  double rmin, rmax, stIn, stOut, dz;
  double tIn, tOut, tIn2, tOut2, rmin2, rmax2, dz2;
  double endInnerRadius2, endOuterRadius2, endInnerRadius, endOuterRadius;
  double zTol, innerRadTol, outerRadTol;
  int32_t inner; // InnerSurfaceExists: rmin > 0 or stIn != 0
  int32_t pad_;
};

// Trd (VecGeom/volumes/TrdStruct.h:18-121): half-lengths and the cached differences, half sums,
// slopes, cosines and lateral tolerances of CalculateCached (the y tolerance from the x slope, as
// there). varyingY: dy1 != dy2 (TrdTypes::checkVaryingY for UniversalTrd). plainY: the TrdTypes::Trd1
// specialisation GeoManager::MakeInstance picks when dy1 == dy2 (source/UnplacedTrd.cpp:60-63), whose
// Inside/Contains test |y| against dy1 instead of the cross product.
struct TrdData {
  double dx1, dx2, dy1, dy2, dz;
  double x2minusx1, y2minusy1, halfx1plusx2, halfy1plusy2, fx, fy, calfx, calfy, tolx, toly;
  int32_t varyingY;
  int32_t plainY;
};

// Paraboloid (VecGeom/volumes/ParaboloidStruct.h:17-70): rho^2 = k1 z + k2 between z = -dz (radius rlo)
// and z = +dz (radius rhi), i.e. z = a rho^2 + b; cached coefficients and squares.
struct ParaboloidData {
  double rlo, rhi, dz, dx, a, invA, b, k1, k2, rlo2, rhi2;
};

// Parallelepiped (VecGeom/volumes/ParallelepipedStruct.h:80-124): half-lengths in the oblique frame,
// the cached tangents, the x and y face normals (the z face's is +z) and the safety scale factors.
struct ParallelepipedData {
  double dim[3];
  double tanAlpha, tanThetaSinPhi, tanThetaCosPhi;
  double n0[3], n1[3];
  double ctx, cty;
};

// Tet (VecGeom/volumes/TetStruct.h:57-90): the corners and the four outward face planes n.p + d.
struct TetData {
  double v[4][3];
  double n[4][3];
  double d[4];
};

// EllipticalTube (VecGeom/volumes/EllipticalTubeStruct.h; cached by source/UnplacedEllipticalTube.cpp:28-62):
// the ellipse scaled to a circle of radius r = min(dx, dy); dist ~ q1 (x^2 + y^2) - q2.
struct EllipticalTubeData {
  double dx, dy, dz, rsph, ddx, ddy, sx, sy, r, q1, q2, scratch;
};

// Ellipsoid (VecGeom/volumes/EllipsoidStruct.h; cached by source/UnplacedEllipsoid.cpp:40-116): scaled to a
// sphere of radius r = min(dx, dy, dz), z cuts (also scaled), x-y extent.
struct EllipsoidData {
  double dx, dy, dz, zBottomCut, zTopCut, xmax, ymax, rsph, r, sx, sy, sz, scZMidCut, scZDimCut, q1, q2;
};

// GenTrap with planar lateral faces (VecGeom/volumes/GenTrapStruct.h:57-190): the vertices after the
// clockwise reversal, bottom-to-top connecting components, the lateral faces that the reference's
// TessellatedSection keeps (outward unit normal, offset: n.p + d), and SecondOrderSurfaceShell's face
// normals in the vertex order as given (used by Normal).
struct GenTrapData {
  double dz, halfInverseDz;
  double vx[8], vy[8], connX[4], connY[4];
  double fnormal[4][3], fdist[4];
  double shellNormal[4][3];
  int32_t nfacets, pad_;
};

// EllipticalCone (VecGeom/volumes/EllipticalConeStruct.h): sqrt((x dx^-1)^2 + (y dy^-1)^2) + z <= dz,
// |z| <= zCut, with the precalculated bounding radius, inverse slopes and apex offset.
struct EllipticalConeData {
  double dx, dy, dz, zCut, rsph, invDx, invDy, cosAxisMin, dApex;
};

// CutTube (VecGeom/volumes/CutTubeStruct.h): a tube of infinite half-length (its z fields are
// kInfLength) cut by two planes through (0, 0, -dz) and (0, 0, dz); n0/n1 are the bottom/top unit
// outward normals and d0/d1 their offsets (signed distance = p.n + d).
struct CutTubeData {
  TubeData tube;
  double n0[3], d0, n1[3], d1;
  double dz, maxVal, cosPhi1, sinPhi1, cosPhi2, sinPhi2;
};

// Tessellated solid of triangular facets (VecGeom/volumes/TessellatedStruct.h, Tile.h). The facets
// and a threaded BVH over them live in the variable-length pool:
//   facet k (kFacetDoubles doubles): single-precision box lo[3] hi[3] (3 doubles), normal[3],
//     plane offset, v0, side0, v1, side1, v2, side2, facet index (as a double);
//   node j (kNodeDoubles doubles): box lo[3] hi[3] (enlarged by a margin covering the
//     single-precision facet boxes), then int32 first facet, facet count (0: internal node),
//     escape node (next node in depth-first order after this subtree), unused.
// The pool also holds an image of what the walks read most, laid out for a workgroup to copy to
// LDS as one block (imageDoubles doubles at imageOff; 0 when the solid has 65536 facets or more):
// the nodes as TessNode32 (box rounded outwards to single precision), then every facet's plane
// (normal[3], offset), then every facet's single-precision box (lo[3], hi[3]).
struct TessellatedData {
  int32_t nfacets, nnodes;
  int32_t facetOff, nodeOff;
  const double *base;
  double minExt[3], maxExt[3];
  double testDir[3]; // direction of the Contains/Inside ray
  int32_t imageOff, imageDoubles;
};
constexpr int kFacetDoubles   = 26;
constexpr int kNodeDoubles    = 8;
constexpr int kTessLeafFacets = 4; // facets per BVH leaf, at most

struct TessNode32 {
  float lo[3], hi[3];
  uint16_t first, count;
  int32_t escape;
};
static_assert(sizeof(TessNode32) == 32, "TessNode32 is four doubles");
constexpr int kImageNodeDoubles  = 4;
constexpr int kImagePlaneDoubles = 4;
constexpr int kImageBoxFloats    = 6;

// Union of all records: the device shape table stores one per unplaced shape.
// Variable-length data (polycone sections, tessellated facets, ...) lives in a
// separate double pool addressed by an offset inside the record.
// 16-byte aligned, with the parameters from offset 16, so adjacent double pairs load as dwordx4.
struct alignas(16) ShapeRecord {
  int32_t type;
  int32_t pad_[3];
  union alignas(16) {
    BoxData box;
    TubeData tube;
    TrapezoidData trap;
    OrbData orb;
    ConeData cone;
    PolyconeData pcon;
    SphereData sph;
    PolyhedronData phed;
    TorusData tor;
    TessellatedData tess;
    HypeData hype;
    TrdData trd;
    ParaboloidData para;
    ParallelepipedData ppd;
    TetData tet;
    EllipticalTubeData eltu;
    EllipsoidData ells;
    GenTrapData gtra;
    CutTubeData ctub;
    EllipticalConeData elco;
    double raw[64];
  } u;
};
static_assert(sizeof(PolyhedronData) <= 64 * sizeof(double), "shape records must fit the 512-byte union");
static_assert(sizeof(SphereData) <= 64 * sizeof(double), "shape records must fit the 512-byte union");
static_assert(sizeof(TorusData) <= 64 * sizeof(double), "shape records must fit the 512-byte union");
static_assert(sizeof(ConeData) <= 64 * sizeof(double), "shape records must fit the 512-byte union");
static_assert(sizeof(GenTrapData) <= 64 * sizeof(double), "shape records must fit the 512-byte union");

} // namespace vgmi
