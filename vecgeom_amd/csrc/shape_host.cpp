// Host-side construction of the flat shape records (cached members as computed
// by the reference constructors). Runs once per shape at geometry build time.
#include "shape_host.h"
#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>

namespace vgmi {

static constexpr double hTol  = 1e-9;
static constexpr double hHTol = 0.5e-9;
static constexpr double hPi   = 3.14159265358979323846;

// Wedge::Contains for a wedge given by its along vectors (Wedge_Evolution.h:293-336, no surface test)
static bool WedgeContainsXY(double a1x, double a1y, double a2x, double a2y, bool ltPi, double x, double y)
{
  const double startCheck = (-x * a1y + y * a1x);
  const double endCheck   = (-a2x * y + a2y * x);
  bool out                = startCheck < 0.;
  if (ltPi)
    out = out || (endCheck < 0.);
  else
    out = out && (endCheck < 0.);
  return !out;
}

// Box reduced to the phi range (source/UnplacedTube.cpp:220-262; the cone, polycone and sphere
// extents use the same rule): C, D at rOut and A, B at rIn, at the start and end angles.
static void PhiExtent(double a1x, double a1y, double a2x, double a2y, bool ltPi, double rOut, double rIn, double sphi,
                      double dphi, double lo[3], double hi[3])
{
  const double Rin = 0.5 * (rOut + rIn);
  const bool o0    = !WedgeContainsXY(a1x, a1y, a2x, a2y, ltPi, Rin, 0);
  const bool o90   = !WedgeContainsXY(a1x, a1y, a2x, a2y, ltPi, 0, Rin);
  const bool o180  = !WedgeContainsXY(a1x, a1y, a2x, a2y, ltPi, -Rin, 0);
  const bool o270  = !WedgeContainsXY(a1x, a1y, a2x, a2y, ltPi, 0, -Rin);
  if (!(o0 || o90 || o180 || o270)) return;
  const double Cx = rOut * std::cos(sphi), Dx = rOut * std::cos(sphi + dphi);
  const double Cy = rOut * std::sin(sphi), Dy = rOut * std::sin(sphi + dphi);
  if (o0) hi[0] = std::max(Cx, Dx);
  if (o90) hi[1] = std::max(Cy, Dy);
  if (o180) lo[0] = std::min(Cx, Dx);
  if (o270) lo[1] = std::min(Cy, Dy);
  if (dphi >= hPi) return;
  const double Ax = rIn * std::cos(sphi), Bx = rIn * std::cos(sphi + dphi);
  const double Ay = rIn * std::sin(sphi), By = rIn * std::sin(sphi + dphi);
  double t = std::max(Ax, Bx);
  hi[0]    = t > hi[0] ? t : hi[0];
  t        = std::max(Ay, By);
  hi[1]    = t > hi[1] ? t : hi[1];
  t        = std::min(Ax, Bx);
  lo[0]    = t < lo[0] ? t : lo[0];
  t        = std::min(Ay, By);
  lo[1]    = t < lo[1] ? t : lo[1];
}

// source/UnplacedCone.cpp:149-213
static void ConeExtent(ConeData const &c, double lo[3], double hi[3])
{
  const double mx = c.origRmax1 > c.origRmax2 ? c.origRmax1 : c.origRmax2;
  const double mn = c.origRmin1 > c.origRmin2 ? c.origRmin2 : c.origRmin1;
  lo[0] = lo[1] = -mx;
  hi[0] = hi[1] = mx;
  lo[2]         = -c.dz;
  hi[2]         = c.dz;
  if (c.dphi == 2. * hPi) return;
  PhiExtent(c.w1x, c.w1y, c.w2x, c.w2y, c.wedgeLtPi != 0, mx, mn, c.sphi, c.dphi, lo, hi);
}

// VecGeom/volumes/BoxStruct.h:26
int MakeBox(double const *a, int n, ShapeRecord &r, std::string &err)
{
  if (n < 3) return err = "box needs 3 parameters (dx, dy, dz)", -1;
  if (!(a[0] > 0 && a[1] > 0 && a[2] > 0)) return err = "box half-lengths must be positive", -1;
  r      = ShapeRecord();
  r.type = kBox;
  for (int i = 0; i < 3; ++i) r.u.box.h[i] = a[i];
  return 0;
}

// VecGeom/volumes/TubeStruct.h:44-175: the constructor keeps sphi/dphi as given and
// caches squares, tolerance bands and the phi-plane vectors; the wedge vectors follow
// source/Wedge_Evolution.cpp:118-127.
int MakeTube(double const *a, int n, ShapeRecord &r, std::string &err)
{
  if (n < 5) return err = "tube needs 5 parameters (rmin, rmax, dz, sphi, dphi)", -1;
  if (!(a[1] > 0 && a[2] > 0 && a[4] > 0)) return err = "tube rmax, dz and dphi must be positive", -1;
  r          = ShapeRecord();
  r.type     = kTube;
  TubeData &t = r.u.tube;
  t.rmin      = a[0] < 0.0 ? 0.0 : a[0];
  t.rmax      = a[1];
  t.z         = a[2];
  t.sphi      = a[3];
  t.dphi      = a[4];
  t.rmin2     = t.rmin * t.rmin;
  t.rmax2     = t.rmax * t.rmax;
  double tolOrmin = t.rmin - hHTol, tolIrmin = t.rmin + hHTol;
  double tolOrmax = t.rmax + hHTol, tolIrmax = t.rmax - hHTol;
  t.tolOrmin2 = tolOrmin * tolOrmin;
  t.tolIrmin2 = tolIrmin * tolIrmin;
  t.tolOrmax2 = tolOrmax * tolOrmax;
  t.tolIrmax2 = tolIrmax * tolIrmax;
  t.a1x       = std::cos(t.sphi);
  t.a1y       = std::sin(t.sphi);
  t.a2x       = std::cos(t.sphi + t.dphi);
  t.a2y       = std::sin(t.sphi + t.dphi);
  t.n1x       = -std::sin(t.sphi);
  t.n1y       = std::cos(t.sphi);
  t.n2x       = std::sin(t.sphi + t.dphi);
  t.n2y       = -std::cos(t.sphi + t.dphi);
  t.maxVal    = std::max(t.rmax, t.z);
  t.hasRmin   = t.rmin > 0;
  t.hasPhi    = t.dphi < 2. * hPi;
  t.phiGtPi   = t.dphi > M_PI;
  t.phiLePi   = t.dphi <= M_PI;
  t.wedgeLtPi = t.dphi < hPi;
  (void)hTol;
  return 0;
}

// Trapezoid: volumes/UnplacedTrapezoid.h:56-63 (tan(alpha)), TrapezoidStruct::CalculateCached,
// corners source/UnplacedTrapezoid.cpp:447-465, side planes from corner diagonals :575-619
// (outward unit normal (p3-p1)x(p4-p2), offset through the corner centroid).
// Parameter counts select the reference constructor: 11 = the full list (dz theta phi dy1 dx1 dx2 alpha1
// dy2 dx3 dx4 alpha2); 3 = xbox ybox zbox; 4 = dx1 dx2 dy dz; 5 = dx1 dx2 dy1 dy2 dz; 6 = dx dy dz alpha
// theta phi (source/UnplacedTrapezoid.cpp:148-190, each the full form with zero or shared angles);
// 24 = eight corners in TrapCorners order (:129-134, parameters recovered by fromCornersToParameters
// :496-523, planes through the given corners).
int MakeTrapezoid(double const *a, int n, ShapeRecord &r, std::string &err)
{
  struct P3 {
    double x, y, z;
  };
  double f[11];
  P3 c[8];
  double prm[11];
  if (n == 24) {
    for (int i = 0; i < 8; ++i) c[i] = {a[3 * i], a[3 * i + 1], a[3 * i + 2]};
    const double dz = c[7].z, dzRecip = 1.0 / dz;
    const double dy1 = 0.50 * (c[2].y - c[0].y), dx1 = 0.50 * (c[1].x - c[0].x), dx2 = 0.50 * (c[3].x - c[2].x);
    const double ta1 = 0.25 * (c[2].x + c[3].x - c[1].x - c[0].x) / dy1;
    const double dy2 = 0.50 * (c[6].y - c[4].y), dx3 = 0.50 * (c[5].x - c[4].x), dx4 = 0.50 * (c[7].x - c[6].x);
    const double ta2 = 0.25 * (c[6].x + c[7].x - c[5].x - c[4].x) / dy2;
    const double tc = (c[4].x + dy2 * ta2 + dx3) * dzRecip, ts = (c[4].y + dy2) * dzRecip;
    if (!(dz > 0 && dy1 > 0 && dx1 > 0 && dx2 > 0 && dy2 > 0 && dx3 > 0 && dx4 > 0))
      return err = "trapezoid corners must give positive half-lengths (TrapCorners order)", -1;
    double g[11] = {dz, std::atan(std::sqrt(ts * ts + tc * tc)), std::atan2(ts, tc), dy1, dx1, dx2, ta1, dy2,
                    dx3, dx4, ta2};
    for (int i = 0; i < 11; ++i) prm[i] = g[i];
  } else {
    if (n == 11) {
      for (int i = 0; i < 11; ++i) f[i] = a[i];
    } else if (n == 3) {
      double g[11] = {a[2], 0., 0., a[1], a[0], a[0], 0., a[1], a[0], a[0], 0.};
      for (int i = 0; i < 11; ++i) f[i] = g[i];
    } else if (n == 4) {
      double g[11] = {a[3], 0., 0., a[2], a[0], a[0], 0., a[2], a[1], a[1], 0.};
      for (int i = 0; i < 11; ++i) f[i] = g[i];
    } else if (n == 5) {
      double g[11] = {a[4], 0., 0., a[2], a[0], a[0], 0., a[3], a[1], a[1], 0.};
      for (int i = 0; i < 11; ++i) f[i] = g[i];
    } else if (n == 6) {
      double g[11] = {a[2], a[4], a[5], a[1], a[0], a[0], a[3], a[1], a[0], a[0], a[3]};
      for (int i = 0; i < 11; ++i) f[i] = g[i];
    } else {
      return err = "trapezoid takes 11 parameters (dz theta phi dy1 dx1 dx2 alpha1 dy2 dx3 dx4 alpha2), "
                   "3, 4, 5 or 6 (reduced forms) or 24 (eight corners)",
             -1;
    }
    if (!(f[0] > 0 && f[3] > 0 && f[4] > 0 && f[5] > 0 && f[7] > 0 && f[8] > 0 && f[9] > 0))
      return err = "trapezoid half-lengths must be positive", -1;
    const double dz = f[0], theta = f[1], phi = f[2], dy1 = f[3], dx1 = f[4], dx2 = f[5];
    const double ta1 = std::tan(f[6]), dy2 = f[7], dx3 = f[8], dx4 = f[9], ta2 = std::tan(f[10]);
    const double tthetaCphi = std::tan(theta) * std::cos(phi);
    const double tthetaSphi = std::tan(theta) * std::sin(phi);
    double g[11] = {dz, theta, phi, dy1, dx1, dx2, ta1, dy2, dx3, dx4, ta2};
    for (int i = 0; i < 11; ++i) prm[i] = g[i];
    const double dxdyDy1 = ta1 * dy1, dxdyDy2 = ta2 * dy2, dxdzDz = tthetaCphi * dz, dydzDz = tthetaSphi * dz;
    P3 cc[8] = {{-dxdzDz - dxdyDy1 - dx1, -dydzDz - dy1, -dz}, {-dxdzDz - dxdyDy1 + dx1, -dydzDz - dy1, -dz},
                {-dxdzDz + dxdyDy1 - dx2, -dydzDz + dy1, -dz}, {-dxdzDz + dxdyDy1 + dx2, -dydzDz + dy1, -dz},
                {+dxdzDz - dxdyDy2 - dx3, +dydzDz - dy2, +dz}, {+dxdzDz - dxdyDy2 + dx3, +dydzDz - dy2, +dz},
                {+dxdzDz + dxdyDy2 - dx4, +dydzDz + dy2, +dz}, {+dxdzDz + dxdyDy2 + dx4, +dydzDz + dy2, +dz}};
    for (int i = 0; i < 8; ++i) c[i] = cc[i];
  }
  r                = ShapeRecord();
  r.type           = kTrapezoid;
  TrapezoidData &t = r.u.trap;
  t.dz             = prm[0];
  for (int i = 0; i < 11; ++i) t.params[i] = prm[i];
  auto plane = [&](int i, int i1, int i2, int i3, int i4) {
    P3 const &p1 = c[i1], &p2 = c[i2], &p3 = c[i3], &p4 = c[i4];
    double ux = p3.x - p1.x, uy = p3.y - p1.y, uz = p3.z - p1.z;
    double vx = p4.x - p2.x, vy = p4.y - p2.y, vz = p4.z - p2.z;
    double nx = uy * vz - uz * vy, ny = uz * vx - ux * vz, nz = ux * vy - uy * vx;
    double inv = 1. / std::sqrt(nx * nx + ny * ny + nz * nz);
    nx *= inv;
    ny *= inv;
    nz *= inv;
    double cx = 0.25 * (((p1.x + p2.x) + p3.x) + p4.x);
    double cy = 0.25 * (((p1.y + p2.y) + p3.y) + p4.y);
    double cz = 0.25 * (((p1.z + p2.z) + p3.z) + p4.z);
    t.a[i] = nx;
    t.b[i] = ny;
    t.c[i] = nz;
    t.d[i] = -(nx * cx + ny * cy + nz * cz);
  };
  plane(0, 0, 1, 5, 4); // ~ -y
  plane(1, 2, 6, 7, 3); // ~ +y
  plane(2, 0, 4, 6, 2); // ~ -x
  plane(3, 1, 3, 7, 5); // ~ +x
  return 0;
}

// Extent from the corners recomputed from the planes (source/UnplacedTrapezoid.cpp:298-337, :467-494).
static void TrapezoidExtent(TrapezoidData const &t, double lo[3], double hi[3])
{
  const double dz = t.dz;
  const double *A = t.a, *B = t.b, *C = t.c, *D = t.d;
  double y[8], x[8];
  y[0] = y[1] = -(D[0] - dz * C[0]) / B[0];
  y[2] = y[3] = -(D[1] - dz * C[1]) / B[1];
  y[4] = y[5] = -(D[0] + dz * C[0]) / B[0];
  y[6] = y[7] = -(D[1] + dz * C[1]) / B[1];
  x[0] = ((B[2] / B[0]) * (D[0] - dz * C[0]) - (D[2] - dz * C[2])) / A[2];
  x[1] = ((B[3] / B[0]) * (D[0] - dz * C[0]) - (D[3] - dz * C[3])) / A[3];
  x[2] = ((B[2] / B[1]) * (D[1] - dz * C[1]) - (D[2] - dz * C[2])) / A[2];
  x[3] = ((B[3] / B[1]) * (D[1] - dz * C[1]) - (D[3] - dz * C[3])) / A[3];
  x[4] = ((B[2] / B[0]) * (D[0] + dz * C[0]) - (D[2] + dz * C[2])) / A[2];
  x[5] = ((B[3] / B[0]) * (D[0] + dz * C[0]) - (D[3] + dz * C[3])) / A[3];
  x[6] = ((B[2] / B[1]) * (D[1] + dz * C[1]) - (D[2] + dz * C[2])) / A[2];
  x[7] = ((B[3] / B[1]) * (D[1] + dz * C[1]) - (D[3] + dz * C[3])) / A[3];
  auto mx = [](double a, double b) { return a < b ? b : a; };
  auto mn = [](double a, double b) { return b < a ? b : a; };
  hi[0] = mx(mx(mx(x[0], x[1]), mx(x[2], x[3])), mx(mx(x[4], x[5]), mx(x[6], x[7])));
  lo[0] = mn(mn(mn(x[0], x[1]), mn(x[2], x[3])), mn(mn(x[4], x[5]), mn(x[6], x[7])));
  hi[1] = mx(mx(mx(y[0], y[1]), mx(y[2], y[3])), mx(mx(y[4], y[5]), mx(y[6], y[7])));
  lo[1] = mn(mn(mn(y[0], y[1]), mn(y[2], y[3])), mn(mn(y[4], y[5]), mn(y[6], y[7])));
  lo[2] = -dz;
  hi[2] = dz;
}

// Orb: VecGeom/volumes/OrbStruct.h
int MakeOrb(double const *a, int n, ShapeRecord &r, std::string &err)
{
  if (n < 1 || !(a[0] > 0)) return err = "orb needs a positive radius", -1;
  r          = ShapeRecord();
  r.type     = kOrb;
  r.u.orb.r  = a[0];
  return 0;
}

// Hype: VecGeom/volumes/HypeStruct.h:54-83 (CalculateCached); UniversalHype (the inner surface
// is treated when rmin > 0 or stIn != 0, HypeStruct.h:87 and kernel/shapetypes/HypeTypes.h:63-70)
int MakeHype(double const *a, int n, ShapeRecord &r, std::string &err)
{
  if (n < 5) return err = "hype needs 5 parameters (rmin rmax stIn stOut dz)", -1;
  if (!(a[1] > 0 && a[4] > 0 && a[0] >= 0 && a[0] < a[1])) return err = "hype needs 0 <= rmin < rmax and dz > 0", -1;
  r           = ShapeRecord();
  r.type      = kHype;
  HypeData &h = r.u.hype;
  h.rmin = a[0], h.rmax = a[1], h.stIn = a[2], h.stOut = a[3], h.dz = a[4];
  h.tIn             = std::tan(h.stIn);
  h.tOut            = std::tan(h.stOut);
  h.tIn2            = h.tIn * h.tIn;
  h.tOut2           = h.tOut * h.tOut;
  h.rmin2           = h.rmin * h.rmin;
  h.rmax2           = h.rmax * h.rmax;
  h.dz2             = h.dz * h.dz;
  h.endInnerRadius2 = h.tIn2 * h.dz2 + h.rmin2;
  h.endOuterRadius2 = h.tOut2 * h.dz2 + h.rmax2;
  h.endInnerRadius  = std::sqrt(h.endInnerRadius2);
  h.endOuterRadius  = std::sqrt(h.endOuterRadius2);
  h.zTol            = 1e-9 * h.dz;
  h.innerRadTol     = 1e-9 * h.endInnerRadius;
  h.outerRadTol     = 1e-9 * h.endOuterRadius;
  h.inner           = (h.rmin > 0. || h.stIn != 0.) ? 1 : 0;
  return 0;
}

// Trd: the UnplacedTrd constructors (VecGeom/volumes/UnplacedTrd.h:51-80), with an optional sixth value 1
// for the Trd1 kind of GeoManager::MakeInstance (source/UnplacedTrd.cpp:60-63), and TrdStruct::CalculateCached
// (VecGeom/volumes/TrdStruct.h:102-121; the y tolerance is computed from the x slope, as there).
int MakeTrd(double const *a, int n, ShapeRecord &r, std::string &err)
{
  double x1, x2, y1, y2, z;
  if (n == 3)
    x1 = x2 = a[0], y1 = y2 = a[1], z = a[2];
  else if (n == 4)
    x1 = a[0], x2 = a[1], y1 = y2 = a[2], z = a[3];
  else if (n == 5 || n == 6)
    x1 = a[0], x2 = a[1], y1 = a[2], y2 = a[3], z = a[4];
  else
    return err = "trd takes x1 x2 y1 y2 z, x1 x2 y z or x y z (and a sixth value 1 for the Trd1 kind)", -1;
  const bool plainY = n == 6 && a[5] == 1.0;
  if (n == 6 && !(a[5] == 0.0 || (plainY && y1 == y2)))
    return err = "trd kind must be 0 (UniversalTrd) or 1 (Trd1, which needs y1 == y2)", -1;
  if (!(x1 >= 0 && x2 >= 0 && y1 >= 0 && y2 >= 0 && z > 0 && (x1 > 0 || x2 > 0) && (y1 > 0 || y2 > 0)))
    return err = "trd half-lengths must be non-negative (not both zero at the two ends) and z > 0", -1;
  r          = ShapeRecord();
  r.type     = kTrd;
  TrdData &t = r.u.trd;
  t.dx1 = x1, t.dx2 = x2, t.dy1 = y1, t.dy2 = y2, t.dz = z;
  t.x2minusx1    = t.dx2 - t.dx1;
  t.y2minusy1    = t.dy2 - t.dy1;
  t.halfx1plusx2 = 0.5 * (t.dx1 + t.dx2);
  t.halfy1plusy2 = 0.5 * (t.dy1 + t.dy2);
  t.fx           = 0.5 * (t.dx1 - t.dx2) / t.dz;
  t.fy           = 0.5 * (t.dy1 - t.dy2) / t.dz;
  t.calfx        = 1. / std::sqrt(1.0 + t.fx * t.fx);
  t.calfy        = 1. / std::sqrt(1.0 + t.fy * t.fy);
  t.tolx         = 1e-9 * std::sqrt(t.x2minusx1 * t.x2minusx1 + 4 * t.dz * t.dz);
  t.toly         = 1e-9 * std::sqrt(t.x2minusx1 * t.x2minusx1 + 4 * t.dz * t.dz);
  t.varyingY     = t.dy1 != t.dy2 ? 1 : 0;
  t.plainY       = plainY ? 1 : 0;
  return 0;
}

// Paraboloid: UnplacedParaboloid(rlo, rhi, dz) (VecGeom/volumes/UnplacedParaboloid.h:60) and
// ParaboloidStruct::CalculateCached (VecGeom/volumes/ParaboloidStruct.h:48-70).
int MakeParaboloid(double const *a, int n, ShapeRecord &r, std::string &err)
{
  if (n != 3) return err = "paraboloid takes rlo rhi dz", -1;
  if (!(a[0] >= 0 && a[1] >= 0 && a[2] > 0 && a[0] != a[1]))
    return err = "paraboloid needs rlo >= 0, rhi >= 0, rlo != rhi and dz > 0", -1;
  r                 = ShapeRecord();
  r.type            = kParaboloid;
  ParaboloidData &q = r.u.para;
  q.rlo = a[0], q.rhi = a[1], q.dz = a[2];
  q.rlo2          = q.rlo * q.rlo;
  q.rhi2          = q.rhi * q.rhi;
  const double dd = 1. / (q.rhi2 - q.rlo2);
  q.a             = 2. * q.dz * dd;
  q.b             = -q.dz * (q.rlo2 + q.rhi2) * dd;
  q.k1            = (q.rhi2 - q.rlo2) * 0.5 / q.dz;
  q.k2            = (q.rhi2 + q.rlo2) * 0.5;
  q.invA          = 1 / q.a;
  q.dx            = q.rhi < q.rlo ? q.rlo : q.rhi;
  return 0;
}

// Parallelepiped: UnplacedParallelepiped(dx, dy, dz, alpha, theta, phi) (VecGeom/volumes/
// UnplacedParallelepiped.h:65-71) and ParallelepipedStruct's SetAlpha, SetThetaAndPhi and ComputeNormals
// (VecGeom/volumes/ParallelepipedStruct.h:62-124): tangents, face normals, safety scale factors.
int MakeParallelepiped(double const *a, int n, ShapeRecord &r, std::string &err)
{
  if (n != 6) return err = "parallelepiped takes dx dy dz alpha theta phi", -1;
  if (!(a[0] > 0 && a[1] > 0 && a[2] > 0) || !std::isfinite(a[3]) || !std::isfinite(a[4]) || !std::isfinite(a[5]))
    return err = "parallelepiped needs positive half-lengths and finite angles", -1;
  r                     = ShapeRecord();
  r.type                = kParallelepiped;
  ParallelepipedData &q = r.u.ppd;
  q.dim[0] = a[0], q.dim[1] = a[1], q.dim[2] = a[2];
  q.tanAlpha       = std::tan(a[3]);
  q.tanThetaCosPhi = std::tan(a[4]) * std::cos(a[5]);
  q.tanThetaSinPhi = std::tan(a[4]) * std::sin(a[5]);
  // n0 = (tanAlpha, 1, 0) x (tcp, tsp, 1), n1 = (tcp, tsp, 1) x (1, 0, 0); each times 1 / |n|
  const double vy[3] = {q.tanAlpha, 1., 0.}, vz[3] = {q.tanThetaCosPhi, q.tanThetaSinPhi, 1.}, vx[3] = {1., 0., 0.};
  auto cross = [](const double *l, const double *rr, double *o) {
    o[0] = l[1] * rr[2] - l[2] * rr[1];
    o[1] = l[2] * rr[0] - l[0] * rr[2];
    o[2] = l[0] * rr[1] - l[1] * rr[0];
  };
  cross(vy, vz, q.n0);
  cross(vz, vx, q.n1);
  for (double *v : {q.n0, q.n1}) {
    const double f = 1. / std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] *= f, v[1] *= f, v[2] *= f;
  }
  q.ctx = std::fabs(q.n0[0]);
  q.cty = std::fabs(q.n1[1]);
  return 0;
}

// Tet: UnplacedTet(p0, p1, p2, p3) (VecGeom/volumes/UnplacedTet.h:48-49) and TetStruct::CalculateCached
// (VecGeom/volumes/TetStruct.h:57-90): unit face normals (Vector3D::Unit), flipped to point away from
// the opposite corner, and the plane offsets.
int MakeTet(double const *a, int n, ShapeRecord &r, std::string &err)
{
  if (n != 12) return err = "tet takes its four corners (12 values)", -1;
  for (int i = 0; i < 12; ++i)
    if (!std::isfinite(a[i])) return err = "tet corners must be finite", -1;
  r          = ShapeRecord();
  r.type     = kTet;
  TetData &t = r.u.tet;
  for (int k = 0; k < 4; ++k)
    for (int c = 0; c < 3; ++c) t.v[k][c] = a[3 * k + c];
  for (int k = 0; k < 4; ++k) {
    const double *v0 = t.v[k], *v1 = t.v[(k + 1) % 4], *v2 = t.v[(k + 2) % 4], *v3 = t.v[(k + 3) % 4];
    const double e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
    const double e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
    double c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    const double m = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2] + DBL_MIN);
    c[0] /= m, c[1] /= m, c[2] /= m;
    const double o[3] = {v3[0] - v0[0], v3[1] - v0[1], v3[2] - v0[2]};
    if (c[0] * o[0] + c[1] * o[1] + c[2] * o[2] > 0) c[0] = -c[0], c[1] = -c[1], c[2] = -c[2];
    for (int i = 0; i < 3; ++i) t.n[k][i] = c[i];
    t.d[k] = -(c[0] * v0[0] + c[1] * v0[1] + c[2] * v0[2]);
  }
  return 0;
}

// EllipticalTube: UnplacedEllipticalTube(dx, dy, dz) -> SetParameters -> CheckParameters
// (source/UnplacedEllipticalTube.cpp:20-62). Semi-axes below 2 kTolerance reset all three to it, as there
// (whose test reads dy twice and never dz).
int MakeEllipticalTube(double const *a, int n, ShapeRecord &r, std::string &err)
{
  if (n != 3) return err = "elliptical tube takes dx dy dz", -1;
  if (!std::isfinite(a[0]) || !std::isfinite(a[1]) || !std::isfinite(a[2]))
    return err = "elliptical tube semi-axes must be finite", -1;
  r                     = ShapeRecord();
  r.type                = kEllipticalTube;
  EllipticalTubeData &e = r.u.eltu;
  e.dx = a[0], e.dy = a[1], e.dz = a[2];
  const double tol = 2. * 1e-9;
  if (e.dx < tol || e.dy < tol || e.dy < tol) e.dx = e.dy = e.dz = tol;
  const double R = e.dx < e.dy ? e.dx : e.dy;
  e.rsph    = std::sqrt(e.dx * e.dx + e.dy * e.dy + e.dz * e.dz);
  e.ddx     = e.dx * e.dx;
  e.ddy     = e.dy * e.dy;
  e.sx      = R / e.dx;
  e.sy      = R / e.dy;
  e.r       = R;
  e.q1      = 0.5 / R;
  e.q2      = 0.5 * (R + 0.5e-9 * 0.5e-9 / R);
  e.scratch = 2. * R * R * DBL_EPSILON;
  return 0;
}

// Ellipsoid: UnplacedEllipsoid(dx, dy, dz, zBottomCut = 0, zTopCut = 0) -> CheckParameters
// (source/UnplacedEllipsoid.cpp:27-116): semi-axes below 2 kTolerance reset as for the elliptical tube;
// cuts both 0 or invalid become -dz, +dz, then clamped to [-dz, dz]; the x-y extent shrinks when both
// cuts lie on one side of z = 0.
int MakeEllipsoid(double const *a, int n, ShapeRecord &r, std::string &err)
{
  if (n != 3 && n != 5) return err = "ellipsoid takes dx dy dz [zBottomCut zTopCut]", -1;
  for (int i = 0; i < n; ++i)
    if (!std::isfinite(a[i])) return err = "ellipsoid parameters must be finite", -1;
  r                = ShapeRecord();
  r.type           = kEllipsoid;
  EllipsoidData &e = r.u.ells;
  e.dx = a[0], e.dy = a[1], e.dz = a[2];
  e.zBottomCut = n == 5 ? a[3] : 0., e.zTopCut = n == 5 ? a[4] : 0.;
  const double tol = 2. * 1e-9;
  if (e.dx < tol || e.dy < tol || e.dy < tol) e.dx = e.dy = e.dz = tol;
  const double A = e.dx, B = e.dy, C = e.dz;
  if (e.zBottomCut == 0 && e.zTopCut == 0) e.zBottomCut = -C, e.zTopCut = C;
  if (e.zBottomCut >= C || e.zTopCut <= -C || e.zBottomCut >= e.zTopCut) e.zBottomCut = -C, e.zTopCut = C;
  e.zBottomCut = e.zBottomCut > -e.dz ? e.zBottomCut : -e.dz;
  e.zTopCut    = e.zTopCut < e.dz ? e.zTopCut : e.dz;
  e.xmax = A, e.ymax = B;
  if (e.zBottomCut > 0.) {
    const double ratio = e.zBottomCut / C, scale = std::sqrt((1. - ratio) * (1 + ratio));
    e.xmax = A * scale, e.ymax = B * scale;
  }
  if (e.zTopCut < 0.) {
    const double ratio = e.zTopCut / C, scale = std::sqrt((1. - ratio) * (1 + ratio));
    e.xmax = A * scale, e.ymax = B * scale;
  }
  e.rsph = std::max(std::max(A, B), C);
  e.r    = std::min(std::min(A, B), C);
  e.sx   = e.r / A;
  e.sy   = e.r / B;
  e.sz   = e.r / C;
  const double bot = e.zBottomCut * e.sz, top = e.zTopCut * e.sz;
  e.scZMidCut = 0.5 * (top + bot);
  e.scZDimCut = 0.5 * (top - bot);
  e.q1        = 0.5 / e.r;
  e.q2        = 0.5 * (e.r + 0.5e-9 * 0.5e-9 / e.r);
  return 0;
}

// EllipticalCone (source/UnplacedEllipticalCone.cpp:28-63): values below 2 kTolerance raised to it, zcut
// at most h, then the bounding radius, inverse slopes, min axis cosine and apex offset.
int MakeEllipticalCone(double const *a, int n, ShapeRecord &r, std::string &err)
{
  if (n < 4) return err = "elliptical cone needs 4 parameters (a, b, h, zcut)", -1;
  r                     = ShapeRecord();
  r.type                = kEllipticalCone;
  EllipticalConeData &c = r.u.elco;
  const double tol      = 2. * hTol;
  c.dx                  = a[0] < tol ? tol : a[0];
  c.dy                  = a[1] < tol ? tol : a[1];
  c.dz                  = a[2] < tol ? tol : a[2];
  c.zCut                = a[3] < tol ? tol : a[3];
  c.zCut                = std::min(c.dz, c.zCut);
  const double h2 = c.dz + c.zCut, xmax = c.dx * h2, ymax = c.dy * h2, m = std::min(c.dx, c.dy);
  c.rsph       = std::sqrt(xmax * xmax + ymax * ymax + c.zCut * c.zCut);
  c.invDx      = 1. / c.dx;
  c.invDy      = 1. / c.dy;
  c.cosAxisMin = m / std::sqrt(1. + m * m);
  c.dApex      = hHTol / c.cosAxisMin;
  return 0;
}

// CutTube (VecGeom/volumes/UnplacedCutTube.h:54-66, CutTubeStruct.h:36-47): the tube of infinite
// half-length, the normals made unit (Vector3D::Unit: / sqrt(mag2 + DBL_MIN)) and set as planes
// through (0, 0, -dz) and (0, 0, dz) (Plane::Set, Plane.h:44-51: components below kTolerance zeroed,
// normalised again, offset -n.origin). A bottom normal with z >= 0 or a top one with z <= 0 is refused.
int MakeCutTube(double const *a, int n, ShapeRecord &r, std::string &err)
{
  if (n < 11) return err = "cut tube needs 11 parameters (rmin, rmax, dz, sphi, dphi, bottom normal, top normal)", -1;
  if (!(a[2] > 0)) return err = "cut tube dz must be positive", -1;
  if (a[7] >= 0 || a[10] <= 0) return err = "Illegal normal direction for cut planes", -1;
  const double tp[5] = {a[0], a[1], DBL_MAX, a[3], a[4]}; // kInfLength
  ShapeRecord t;
  if (MakeTube(tp, 5, t, err) != 0) return -1;
  r            = ShapeRecord();
  r.type       = kCutTube;
  CutTubeData &c = r.u.ctub;
  c.tube       = t.u.tube;
  c.dz         = a[2];
  for (int i = 0; i < 2; ++i) {
    const double *v = a + 5 + 3 * i;
    const double m  = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + DBL_MIN);
    double f[3]     = {v[0] / m, v[1] / m, v[2] / m};
    for (double &x : f)
      if (std::fabs(x) < hTol) x = 0.0;
    const double inv = 1. / std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    double *nn       = i ? c.n1 : c.n0;
    for (int k = 0; k < 3; ++k) nn[k] = f[k] * inv;
    const double oz = i ? c.dz : -c.dz;
    (i ? c.d1 : c.d0) = inv * -(f[0] * 0. + f[1] * 0. + f[2] * oz);
  }
  c.cosPhi1 = std::cos(a[3]);
  c.sinPhi1 = std::sin(a[3]);
  c.cosPhi2 = std::cos(a[3] + a[4]);
  c.sinPhi2 = std::sin(a[3] + a[4]);
  c.maxVal  = std::max(c.tube.rmax, c.dz);
  return 0;
}

// GenTrap: UnplacedGenTrap(x[8], y[8], dz) (VecGeom/volumes/UnplacedGenTrap.h:46) and GenTrapStruct::
// Initialize (VecGeom/volumes/GenTrapStruct.h:57-188): SecondOrderSurfaceShell's face normals from the
// vertices as given; anticlockwise polygons reverted to clockwise; opposite orientations, crossing opposite
// segments and non-convex polygons refused (the reference prints an error and leaves the solid half made);
// the lateral faces of the TessellatedSection (Tile<4>::AddVertex, VecGeom/volumes/Tile.h:70-130: a face
// with fewer than three distinct corners is left out). Twisted traps (non-parallel bottom and top edges of a
// face) are refused: only the planar path is built.
int MakeGenTrap(double const *a, int n, ShapeRecord &r, std::string &err)
{
  if (n != 17) return err = "gentrap takes x[8] y[8] dz", -1;
  for (int i = 0; i < 17; ++i)
    if (!std::isfinite(a[i])) return err = "gentrap parameters must be finite", -1;
  const double h = a[16], tol = 1e-9;
  if (!(h > 0)) return err = "gentrap half-length must be positive", -1;
  const double *x = a, *y = a + 8;
  using V = std::array<double, 3>;
  auto sub3  = [](V const &p, V const &q) { return V{p[0] - q[0], p[1] - q[1], p[2] - q[2]}; };
  auto cross = [](V const &l, V const &rr) {
    return V{l[1] * rr[2] - l[2] * rr[1], l[2] * rr[0] - l[0] * rr[2], l[0] * rr[1] - l[1] * rr[0]};
  };
  auto mag2 = [](V const &v) { return v[0] * v[0] + v[1] * v[1] + v[2] * v[2]; };
  auto normalized = [&](V v) {
    const double f = 1. / std::sqrt(mag2(v));
    return V{v[0] * f, v[1] * f, v[2] * f};
  };
  auto dot = [](V const &l, V const &rr) { return l[0] * rr[0] + l[1] * rr[1] + l[2] * rr[2]; };
  r              = ShapeRecord();
  r.type         = kGenTrap;
  GenTrapData &g = r.u.gtra;
  g.dz = h, g.halfInverseDz = 0.5 / h;
  for (int i = 0; i < 4; ++i) { // SecondOrderSurfaceShell::Initialize (SecondOrderSurfaceShell.h:65-105)
    const int j = (i + 1) % 4;
    const V va{x[i], y[i], -h}, vb{x[i + 4], y[i + 4], h}, vc{x[j], y[j], -h}, vd{x[j + 4], y[j + 4], h};
    V nn = cross(sub3(vb, va), sub3(vc, va));
    if (mag2(nn) < tol) {
      nn = cross(sub3(vb, va), sub3(vd, vb));
      if (mag2(nn) < tol) nn = V{0., 0., 1.};
    }
    nn = normalized(nn);
    for (int c = 0; c < 3; ++c) g.shellNormal[i][c] = nn[c];
  }
  V v[8];
  for (int i = 0; i < 8; ++i) v[i] = V{x[i], y[i], i < 4 ? -h : h};
  double sum1 = 0., sum2 = 0.;
  for (int i = 0; i < 4; ++i) {
    const int j = (i + 1) % 4;
    sum1 += v[i][0] * v[j][1] - v[j][0] * v[i][1];
    sum2 += v[i + 4][0] * v[j + 4][1] - v[j + 4][0] * v[i + 4][1];
  }
  if (sum1 * sum2 < -tol) return err = "gentrap bottom and top polygons have opposite orientations", -1;
  if (sum1 > tol) std::swap(v[1], v[3]), std::swap(v[5], v[7]);
  for (int i = 0; i < 4; ++i) g.connX[i] = sub3(v[i], v[i + 4])[0], g.connY[i] = sub3(v[i], v[i + 4])[1];
  for (int i = 0; i < 8; ++i) g.vx[i] = v[i][0], g.vy[i] = v[i][1];
  auto crossing = [&](V const &p, V const &p1, V const &q, V const &q1) { // GenTrapStruct.h:296-317
    const V rr = sub3(p1, p), ss = sub3(q1, q), rxs = cross(rr, ss);
    if (mag2(rxs) < tol) return false;
    const double t = cross(sub3(q, p), ss)[2] / rxs[2];
    if (t < tol || t > 1. - tol) return false;
    const double u = cross(sub3(q, p), rr)[2] / rxs[2];
    return !(u < tol || u > 1. - tol);
  };
  if (crossing(v[0], v[1], v[3], v[2]) || crossing(v[1], v[2], v[0], v[3]) || crossing(v[4], v[5], v[7], v[6]) ||
      crossing(v[5], v[6], v[4], v[7]))
    return err = "gentrap polygons have crossing opposite segments", -1;
  double dX[8], dY[8];
  for (int i = 0; i < 4; ++i) {
    const int j = (i + 1) % 4;
    dX[i] = g.vx[j] - g.vx[i], dX[i + 4] = g.vx[j + 4] - g.vx[i + 4];
    dY[i] = g.vy[j] - g.vy[i], dY[i + 4] = g.vy[j + 4] - g.vy[i + 4];
  }
  for (int i = 0; i < 4; ++i) {
    const int j = (i + 1) % 4;
    if (dX[i] * dY[j] - dY[i] * dX[j] > tol || dX[i + 4] * dY[j + 4] - dY[i + 4] * dX[j + 4] > tol)
      return err = "gentrap bottom or top polygon is not convex", -1;
  }
  for (int i = 0; i < 4; ++i) { // ComputeIsTwisted (GenTrapStruct.h:246-276)
    const double dx1 = v[(i + 1) % 4][0] - v[i][0], dy1 = v[(i + 1) % 4][1] - v[i][1];
    const double dx2 = v[4 + (i + 1) % 4][0] - v[4 + i][0], dy2 = v[4 + (i + 1) % 4][1] - v[4 + i][1];
    if ((dx1 == 0 && dy1 == 0) || (dx2 == 0 && dy2 == 0)) continue;
    if (std::fabs(dy1 * dx2 - dx1 * dy2) >= tol) return err = "twisted gentrap faces are not supported", -1;
  }
  const int quad[4][4] = {{0, 4, 5, 1}, {1, 5, 6, 2}, {2, 6, 7, 3}, {3, 7, 4, 0}};
  g.nfacets = 0;
  for (auto const &q : quad) {
    const V c[4] = {v[q[0]], v[q[1]], v[q[2]], v[q[3]]};
    int nvert = 4;
    for (int i = 0; i < 4; ++i)
      if (mag2(sub3(c[(i + 1) % 4], c[i])) < tol) nvert--;
    if (nvert < 3) continue;
    bool found = false;
    V nn{0., 0., 0.};
    for (int i = 0; i < 3 && !found; ++i) {
      const V e1 = sub3(c[i + 1], c[i]);
      if (mag2(e1) < tol) continue;
      for (int j = i + 1; j < 4; ++j) {
        const V e2 = sub3(c[(j + 1) % 4], c[j]);
        if (mag2(e2) < tol) continue;
        nn = cross(e1, e2);
        if (mag2(nn) < tol) continue;
        nn = normalized(nn), found = true;
        break;
      }
    }
    if (!found) continue;
    double dd = 0.;
    for (int i = 0; i < 4; ++i)
      if (!(mag2(sub3(c[(i + 1) % 4], c[i])) < tol)) {
        dd = -dot(nn, c[i]);
        break;
      }
    for (int k = 0; k < 3; ++k) g.fnormal[g.nfacets][k] = nn[k];
    g.fdist[g.nfacets++] = dd;
  }
  return 0;
}

// Cone: VecGeom/volumes/ConeStruct.h:415-426 (constructor), :309-346 (angle normalisation: the
// wedge keeps the raw angles, the phi-plane vectors use the normalised ones) and :76-140 (cache).
int MakeCone(double const *a, int n, ShapeRecord &r, std::string &err)
{
  if (n < 7) return err = "cone needs 7 parameters (rmin1 rmax1 rmin2 rmax2 dz sphi dphi)", -1;
  if (!(a[4] > 0 && a[6] > 0 && a[1] >= 0 && a[3] >= 0)) return err = "cone dz and dphi must be positive", -1;
  const double kConeTol = 1e-7, kAngTol = 1e-9;
  r           = ShapeRecord();
  r.type      = kCone;
  ConeData &c = r.u.cone;
  c.rmin1 = a[0] < 0.0 ? 0.0 : a[0];
  c.rmax1 = a[1];
  c.rmin2 = a[2] < 0.0 ? 0.0 : a[2];
  c.rmax2 = a[3];
  c.dz    = a[4];
  const double sphi0 = a[5], dphi0 = a[6];
  c.w1x = std::cos(sphi0);
  c.w1y = std::sin(sphi0);
  c.w2x = std::cos(sphi0 + dphi0);
  c.w2y = std::sin(sphi0 + dphi0);
  c.n1x = -std::sin(sphi0);
  c.n1y = std::cos(sphi0);
  c.n2x = std::sin(sphi0 + dphi0);
  c.n2y = -std::cos(sphi0 + dphi0);
  c.dphi = dphi0;
  c.sphi = sphi0;
  if (dphi0 >= 2. * hPi - 0.5 * kAngTol) {
    c.dphi = 2. * hPi;
    c.sphi = 0;
  }
  c.sphi = sphi0 < 0 ? 2. * hPi - std::fmod(std::fabs(sphi0), 2. * hPi) : std::fmod(sphi0, 2. * hPi);
  if (c.sphi + c.dphi > 2. * hPi) c.sphi -= 2. * hPi;
  c.a1x = std::cos(c.sphi);
  c.a1y = std::sin(c.sphi);
  c.a2x = std::cos(c.sphi + c.dphi);
  c.a2y = std::sin(c.sphi + c.dphi);
  c.origRmax1 = c.rmax1;
  c.origRmax2 = c.rmax2;
  c.origRmin1 = a[0];
  c.origRmin2 = a[2];
  if (c.rmin1 == c.rmax1) c.rmax1 += kConeTol;
  if (c.rmin2 == c.rmax2) c.rmax2 += kConeTol;
  c.sqRmin1     = c.rmin1 * c.rmin1;
  c.sqRmax1     = c.rmax1 * c.rmax1;
  c.sqRmin2     = c.rmin2 * c.rmin2;
  c.sqRmax2     = c.rmax2 * c.rmax2;
  c.tanRMin     = (c.rmin2 - c.rmin1) * 0.5 / c.dz;
  c.secRMin     = std::sqrt(1.0 + c.tanRMin * c.tanRMin);
  c.invSecRMin  = 1. / (c.secRMin + std::copysign(1e-30, c.secRMin));
  c.tanRMax     = (c.rmax2 - c.rmax1) * 0.5 / c.dz;
  c.secRMax     = std::sqrt(1.0 + c.tanRMax * c.tanRMax);
  c.invSecRMax  = 1. / (c.secRMax + std::copysign(1e-30, c.secRMax));
  c.innerSlope  = -(c.rmin1 - c.rmin2) / (2. * c.dz);
  c.outerSlope  = -(c.rmax1 - c.rmax2) / (2. * c.dz);
  c.innerOffset = c.rmin2 - c.innerSlope * c.dz;
  c.outerOffset = c.rmax2 - c.outerSlope * c.dz;
  c.innerTol    = kConeTol * c.secRMin;
  c.outerTol    = kConeTol * c.secRMax;
  auto nz       = [](double x) { return x + std::copysign(1e-30, x); };
  if (c.rmin2 > c.rmin1) {
    c.innerApex    = 2 * c.dz * c.rmin1 / (c.rmin2 - c.rmin1);
    c.tanInnerApex = c.rmin2 / (2 * c.dz + c.innerApex);
  } else {
    c.innerApex    = 2 * c.dz * c.rmin2 / nz(c.rmin1 - c.rmin2);
    c.tanInnerApex = c.rmin1 / (2 * c.dz + c.innerApex);
  }
  if (c.rmin1 == 0. || c.rmin2 == 0.) c.innerApex = 0.;
  if (c.rmin1 == 0.) c.tanInnerApex = c.rmin2 / (2 * c.dz);
  if (c.rmin2 == 0.) c.tanInnerApex = c.rmin1 / (2 * c.dz);
  if (c.rmax2 > c.rmax1) {
    c.outerApex    = 2 * c.dz * c.rmax1 / (c.rmax2 - c.rmax1);
    c.tanOuterApex = c.rmax2 / (2 * c.dz + c.outerApex);
  } else {
    c.outerApex    = 2 * c.dz * c.rmax2 / nz(c.rmax1 - c.rmax2);
    c.tanOuterApex = c.rmax1 / (2 * c.dz + c.outerApex);
  }
  if (c.rmax1 == 0. || c.rmax2 == 0.) c.outerApex = 0.;
  if (c.rmax1 == 0.) c.tanOuterApex = c.rmax2 / (2 * c.dz);
  if (c.rmax2 == 0.) c.tanOuterApex = c.rmax1 / (2 * c.dz);
  c.zNormInner = c.tanRMin / nz(c.secRMin);
  c.zNormOuter = -c.tanRMax / nz(c.secRMax);
  c.tolIz      = c.dz - hHTol;
  c.hasRmin    = c.rmin1 > 0 || c.rmin2 > 0;
  c.hasPhi     = c.dphi < 2. * hPi;
  c.phiLePi    = c.dphi <= M_PI;
  c.wedgeLtPi  = c.dphi < hPi;
  return 0;
}

// Polycone: PolyconeStruct::Init (VecGeom/volumes/PolyconeStruct.h:144-290). Planes are
// put in increasing z; every pair of consecutive planes more than kTolerance apart
// becomes a cone section (built with the raw phi angles) centred at its mid z.
// Parameters: sphi dphi nz z[nz] rmin[nz] rmax[nz].
int MakePolycone(double const *a, int n, ShapeRecord &r, std::vector<double> &extra, std::string &err)
{
  if (n < 3) return err = "polycone needs sphi dphi nz z[nz] rmin[nz] rmax[nz]", -1;
  const int nz = int(a[2]);
  if (nz < 2 || n < 3 + 3 * nz) return err = "polycone needs at least 2 z planes and 3*nz values", -1;
  std::vector<double> Z(a + 3, a + 3 + nz), Rmin(a + 3 + nz, a + 3 + 2 * nz), Rmax(a + 3 + 2 * nz, a + 3 + 3 * nz);
  if (Z[0] > Z[nz - 1]) {
    std::reverse(Z.begin(), Z.end());
    std::reverse(Rmin.begin(), Rmin.end());
    std::reverse(Rmax.begin(), Rmax.end());
  }
  std::vector<ShapeRecord> cones;
  std::vector<double> zs, shifts;
  double prevZ = Z[0], prevRmin = 0, prevRmax = 0, rmaxAll = 0;
  const int dirZ = Z[1] < Z[0] ? -1 : 1;
  for (int i = 0; i < nz; ++i) {
    rmaxAll = std::max(rmaxAll, Rmax[i]);
    if (i == 0) {
      zs.push_back(Z[0]);
    } else if ((Z[i] > prevZ + hTol && dirZ > 0) || (Z[i] < prevZ - hTol && dirZ < 0)) {
      double cp[7] = {prevRmin, prevRmax, Rmin[i], Rmax[i], (Z[i] - prevZ) / 2, a[0], a[1]};
      ShapeRecord c;
      std::string e;
      if (MakeCone(cp, 7, c, e) != 0) return err = "polycone section: " + e, -1;
      cones.push_back(c);
      zs.push_back(Z[i]);
      const size_t zi = zs.size() - 1;
      shifts.push_back(zs[zi - 1] + 0.5 * (zs[zi] - zs[zi - 1]));
    }
    prevZ    = Z[i];
    prevRmin = Rmin[i];
    prevRmax = Rmax[i];
  }
  if (cones.empty()) return err = "polycone has no section of non-zero length", -1;
  r              = ShapeRecord();
  r.type         = kPolycone;
  PolyconeData &P = r.u.pcon;
  P.nsec         = int(cones.size());
  static_assert(sizeof(ConeData) % sizeof(double) == 0, "cone records must be a whole number of doubles");
  const size_t coneWords = sizeof(ConeData) / sizeof(double);
  extra.clear();
  P.secOff = 0;
  for (auto const &c : cones) {
    const double *w = reinterpret_cast<const double *>(&c.u.cone);
    extra.insert(extra.end(), w, w + coneWords);
  }
  P.shiftOff = int(extra.size());
  extra.insert(extra.end(), shifts.begin(), shifts.end());
  P.zsOff = int(extra.size());
  extra.insert(extra.end(), zs.begin(), zs.end());
  P.base = nullptr;
  // extent (source/UnplacedPolycone.cpp:865-898): z from the planes; x/y from the cone
  // (minR, maxR, minR, maxR, 1, fStartPhi, fDeltaPhi) over the sections' constructor radii,
  // with fStartPhi/fDeltaPhi as PolyconeStruct.h:152-184 leaves them
  double maxR = 0, minR = 1.7976931348623157e308;
  for (auto const &c : cones) {
    maxR = std::max(maxR, std::max(c.u.cone.origRmax1, c.u.cone.origRmax2));
    minR = std::min(minR, std::min(c.u.cone.origRmin1, c.u.cone.origRmin2));
  }
  const double sphi0 = a[0], dphi0 = a[1];
  double startPhi, deltaPhi = dphi0;
  if (dphi0 <= 0. || dphi0 > 2. * hPi - hTol) {
    startPhi = 0;
    deltaPhi = 2. * hPi;
  } else {
    if (dphi0 >= 2. * hPi - 0.5 * 1e-9) deltaPhi = 2. * hPi;
    startPhi = sphi0;
    while (startPhi < 0) startPhi += 2. * hPi;
  }
  const double tp[7] = {minR, maxR, minR, maxR, 1, startPhi, deltaPhi};
  ShapeRecord tmp;
  std::string e;
  if (MakeCone(tp, 7, tmp, e) != 0) return err = "polycone extent: " + e, -1;
  ConeExtent(tmp.u.cone, P.ext, P.ext + 3);
  P.ext[2] = zs.front();
  P.ext[5] = zs.back();
  (void)rmaxAll;
  return 0;
}

// Sphere: SphereStruct.h:245-266 (constructor), CheckPhiAngles :128-150 with :180-199,
// CheckThetaAngles :100-126; wedge from the raw (sphi, dphi), theta cones from the raw
// (stheta, dtheta) (ThetaCone_Evolution.h:62-83).
int MakeSphere(double const *a, int n, ShapeRecord &r, std::string &err)
{
  if (n < 6) return err = "sphere needs rmin rmax sphi dphi stheta dtheta", -1;
  if (!(a[1] > 0) || a[0] < 0 || !(a[0] < a[1])) return err = "sphere needs 0 <= rmin < rmax", -1;
  r              = ShapeRecord();
  r.type         = kSphere;
  SphereData &s  = r.u.sph;
  const double sphi = a[2], dphi = a[3], sth = a[4], dth = a[5];
  s.rmin         = a[0];
  s.rmax         = a[1];
  s.sphi         = sphi;
  s.dphi         = dphi;
  s.stheta       = sth;
  s.dtheta       = dth;
  // wedge (Wedge_Evolution.cpp:100-108)
  s.w1x = std::cos(sphi);
  s.w1y = std::sin(sphi);
  s.w2x = std::cos(sphi + dphi);
  s.w2y = std::sin(sphi + dphi);
  s.n1x = -std::sin(sphi);
  s.n1y = std::cos(sphi);
  s.n2x = std::sin(sphi + dphi);
  s.n2y = -std::cos(sphi + dphi);
  s.wedgeLtPi = dphi < hPi;
  // theta cones
  s.tS      = sth;
  s.tE      = sth + dth;
  double ts = s.tS, te = s.tE;
  if (s.tS > hPi / 2) ts = hPi - s.tS;
  if (s.tE > hPi / 2) te = hPi - s.tE;
  s.tanS   = std::tan(ts);
  s.tanS2  = s.tanS * s.tanS;
  s.tanE   = std::tan(te);
  s.tanE2  = s.tanE * s.tanE;
  s.tanBis = std::tan(ts + (dth / 2));
  if (s.tS > hPi / 2 && s.tE > hPi / 2) s.tanBis = std::tan(ts - (dth / 2));
  s.slope1 = std::tan(hPi / 2 - s.tS);
  s.slope2 = std::tan(hPi / 2 - s.tE);
  // phi range checks
  bool fullPhi;
  if (dphi >= 2 * hPi - 1e-9 * 0.5) {
    s.dphi  = 2 * hPi;
    s.sphi  = 0;
    fullPhi = true;
  } else {
    fullPhi = false;
    if (dphi > 0) s.dphi = dphi;
    if (sphi < 0)
      s.sphi = 2 * hPi - std::fmod(std::fabs(sphi), 2 * hPi);
    else
      s.sphi = std::fmod(sphi, 2 * hPi);
    if (s.sphi + s.dphi > 2 * hPi) s.sphi -= 2 * hPi;
  }
  // theta range checks
  if (dth + sth >= hPi)
    s.dtheta = hPi - sth;
  else if (dth > 0)
    s.dtheta = dth;
  const bool fullTheta = !(s.dtheta - s.stheta < hPi);
  s.fullPhi   = fullPhi;
  s.fullTheta = fullTheta;
  s.full      = fullPhi && fullTheta;
  s.etheta    = s.stheta + s.dtheta;
  return 0;
}

static bool SphereWedgeContains(SphereData const &s, double x, double y)
{
  // Wedge::Contains (Wedge_Evolution.h:293-336 without the surface test)
  const double startCheck = (-x * s.w1y + y * s.w1x);
  const double endCheck   = (-s.w2x * y + s.w2y * x);
  bool out                = startCheck < 0.;
  if (s.wedgeLtPi)
    out = out || (endCheck < 0.);
  else
    out = out && (endCheck < 0.);
  return !out;
}

// source/UnplacedSphere.cpp:76-165
static void SphereExtent(SphereData const &s, double lo[3], double hi[3])
{
  for (int i = 0; i < 3; ++i) lo[i] = -s.rmax, hi[i] = s.rmax;
  if (s.full) return;
  double st1 = 0., st2 = 0.;
  if (!s.fullTheta) {
    hi[2] = ((s.stheta <= hPi / 2.) ? s.rmax : s.rmin) * std::cos(s.stheta);
    lo[2] = ((s.etheta <= hPi / 2.) ? s.rmin : s.rmax) * std::cos(s.etheta);
    st1   = std::sin(s.stheta);
    st2   = std::sin(s.etheta);
    double sx = 0;
    bool set  = false;
    if (s.stheta <= hPi / 2.) {
      if (s.etheta < hPi / 2.) sx = s.rmax * st2, set = true;
    } else {
      sx = s.rmax * st1, set = true;
    }
    if (set) hi[0] = sx, lo[0] = -sx, hi[1] = sx, lo[1] = -sx;
  }
  if (!s.fullPhi) {
    double Rmax = s.rmax;
    if (s.stheta > hPi / 2.) Rmax *= st1;
    if (s.etheta < hPi / 2.) Rmax *= st2;
    const double Rmin = s.rmin * std::min(st1, st2);
    const double Rin  = 0.5 * (Rmax + Rmin);
    const bool o0 = !SphereWedgeContains(s, Rin, 0), o90 = !SphereWedgeContains(s, 0, Rin);
    const bool o180 = !SphereWedgeContains(s, -Rin, 0), o270 = !SphereWedgeContains(s, 0, -Rin);
    if (!(o0 || o90 || o180 || o270)) return;
    const double Cx = Rmax * std::cos(s.sphi), Dx = Rmax * std::cos(s.sphi + s.dphi);
    const double Cy = Rmax * std::sin(s.sphi), Dy = Rmax * std::sin(s.sphi + s.dphi);
    if (o0) hi[0] = std::max(Cx, Dx);
    if (o90) hi[1] = std::max(Cy, Dy);
    if (o180) lo[0] = std::min(Cx, Dx);
    if (o270) lo[1] = std::min(Cy, Dy);
    if (s.dphi >= hPi) return;
    const double Ax = Rmin * std::cos(s.sphi), Bx = Rmin * std::cos(s.sphi + s.dphi);
    const double Ay = Rmin * std::sin(s.sphi), By = Rmin * std::sin(s.sphi + s.dphi);
    hi[0] = std::max(std::max(Ax, Bx), hi[0]);
    hi[1] = std::max(std::max(Ay, By), hi[1]);
    lo[0] = std::min(std::min(Ax, Bx), lo[0]);
    lo[1] = std::min(std::min(Ay, By), lo[1]);
  }
}

// ---------------------------------------------------------------- Polyhedron
// VecGeom/volumes/PolyhedronStruct.h:91-103 and :248-429 (segments, phi sections, bounding
// tube, vertices), source/Quadrilaterals.cpp:36-110 (plane of each quadrilateral from its
// first three distinct corners, side planes), source/Planes.cpp:30-52.
namespace {
struct HV {
  double v[3];
  double &operator[](int i) { return v[i]; }
  double operator[](int i) const { return v[i]; }
};
inline HV hv(double x, double y, double z) { return HV{{x, y, z}}; }
inline double hdot(HV const &a, HV const &b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline HV hsub(HV const &a, HV const &b) { return hv(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
inline HV hscale(HV const &a, double s) { return hv(a[0] * s, a[1] * s, a[2] * s); }
inline HV hcross(HV const &l, HV const &r)
{
  return hv(l[1] * r[2] - l[2] * r[1], l[2] * r[0] - l[0] * r[2], l[0] * r[1] - l[1] * r[0]);
}
inline double hmag(HV const &a) { return std::sqrt(hdot(a, a)); }
inline HV hnormalized(HV const &a) { return hscale(a, 1. / hmag(a)); }
inline HV hfix(HV a)
{
  for (int i = 0; i < 3; ++i)
    if (std::fabs(a[i]) < hTol) a[i] = 0.0;
  return a;
}
inline bool htolEqual(HV const &a, HV const &b)
{
  return std::fabs(a[0] - b[0]) < hTol && std::fabs(a[1] - b[1]) < hTol && std::fabs(a[2] - b[2]) < hTol;
}
inline double hnormAngle(double a) { return a + 2. * hPi * (int(a < 0) - int(a / (2. * hPi))); }

// One quadrilateral into 32 doubles: normal, distance, 4 x (side normal, side distance), 4 corners.
void PutQuad(double *q, HV const &c0, HV const &c1, HV const &c2, HV const &c3)
{
  HV const cs[4] = {c0, c1, c2, c3};
  HV ch[4];
  ch[0]        = cs[0];
  int assigned = 1;
  for (int t = 1; t < 4; ++t) {
    bool chosen = true;
    for (int j = 0; j < assigned; ++j)
      if (htolEqual(ch[j], cs[t])) {
        chosen = false;
        break;
      }
    if (chosen) ch[assigned++] = cs[t];
  }
  double a = ch[0][1] * (ch[1][2] - ch[2][2]) + ch[1][1] * (ch[2][2] - ch[0][2]) + ch[2][1] * (ch[0][2] - ch[1][2]);
  double b = ch[0][2] * (ch[1][0] - ch[2][0]) + ch[1][2] * (ch[2][0] - ch[0][0]) + ch[2][2] * (ch[0][0] - ch[1][0]);
  double c = ch[0][0] * (ch[1][1] - ch[2][1]) + ch[1][0] * (ch[2][1] - ch[0][1]) + ch[2][0] * (ch[0][1] - ch[1][1]);
  double d = -ch[0][0] * (ch[1][1] * ch[2][2] - ch[2][1] * ch[1][2]) -
             ch[1][0] * (ch[2][1] * ch[0][2] - ch[0][1] * ch[2][2]) -
             ch[2][0] * (ch[0][1] * ch[1][2] - ch[1][1] * ch[0][2]);
  HV n             = hv(a, b, c);
  const double inv = 1. / hmag(n);
  n                = hscale(n, inv);
  d *= inv;
  n = hfix(n);
  q[0] = n[0], q[1] = n[1], q[2] = n[2], q[3] = d;
  for (int j = 0; j < 4; ++j) {
    HV const &s0 = cs[j], &s1 = cs[(j + 1) % 4];
    double *sp   = q + 4 + 4 * j;
    if (!htolEqual(s1, s0)) {
      const HV sv       = hfix(hnormalized(hcross(n, hsub(s1, s0))));
      const double invl = 1. / hmag(sv);
      sp[0] = invl * sv[0], sp[1] = invl * sv[1], sp[2] = invl * sv[2];
      sp[3] = invl * -hdot(sv, s0);
    } else {
      sp[0] = sp[1] = sp[2] = sp[3] = 0.0;
    }
  }
  for (int j = 0; j < 4; ++j)
    for (int k = 0; k < 3; ++k) q[20 + 3 * j + k] = cs[j][k];
}
inline void FlipQuad(double *q)
{
  for (int k = 0; k < 4; ++k) q[k] = -q[k];
}
} // namespace

// source/UnplacedPolyhedron.cpp:273-303 (corners at rmin/rmax of every plane and phi step)
static void PolyhedronExtent(PolyhedronData const &P, std::vector<double> const &extra, double lo[3], double hi[3])
{
  for (int i = 0; i < 3; ++i) lo[i] = 1.7976931348623157e308, hi[i] = -1.7976931348623157e308;
  const double *b      = extra.data();
  const double sidePhi = P.phiDelta / P.sides;
  const double conv    = 1. / std::cos(0.5 * sidePhi);
  for (int iphi = 0; iphi <= P.sides; ++iphi) {
    const double phi  = P.phiStart + iphi * sidePhi;
    const double corx = conv * std::cos(phi), cory = conv * std::sin(phi);
    for (int k = 0; k < P.nz; ++k) {
      const double rn = b[P.rminOff + k], rx = b[P.rmaxOff + k], zk = b[P.zsOff + k];
      const double pa[3] = {rn * corx, rn * cory, zk}, pb[3] = {rx * corx, rx * cory, zk};
      for (int i = 0; i < 3; ++i) lo[i] = std::min(lo[i], pa[i]), hi[i] = std::max(hi[i], pa[i]);
      for (int i = 0; i < 3; ++i) lo[i] = std::min(lo[i], pb[i]), hi[i] = std::max(hi[i], pb[i]);
    }
  }
}

int MakePolyhedron(double const *a, int n, ShapeRecord &r, std::vector<double> &extra, std::string &err)
{
  if (n < 4) return err = "polyhedron needs sphi dphi nsides nz z[nz] rmin[nz] rmax[nz]", -1;
  const int sides = int(a[2]), nz = int(a[3]);
  if (sides < 1 || double(sides) != a[2] || nz < 2 || double(nz) != a[3] || n < 4 + 3 * nz)
    return err = "polyhedron needs nsides >= 1, nz >= 2 and 3*nz plane values", -1;
  const double phiStart0 = a[0], phiDelta0 = a[1];
  if (!(phiDelta0 > 0)) return err = "polyhedron dphi must be positive", -1;
  const double *z = a + 4, *rmn = a + 4 + nz, *rmx = a + 4 + 2 * nz;
  for (int i = 0; i < nz; ++i) {
    if (rmn[i] < 0 || rmx[i] < rmn[i]) return err = "polyhedron needs 0 <= rmin <= rmax at every plane", -1;
    if (i > 0 && z[i] < z[i - 1]) return err = "polyhedron z planes must not decrease", -1;
  }
  r                 = ShapeRecord();
  r.type            = kPolyhedron;
  PolyhedronData &P = r.u.phed;
  P.sides           = sides;
  P.nz              = nz;
  P.nseg            = nz - 1;
  P.hasPhiCutout      = phiDelta0 < 2. * hPi;
  P.hasLargePhiCutout = phiDelta0 < hPi;
  P.phiStart          = hnormAngle(phiStart0);
  P.phiDelta          = phiDelta0 > 2. * hPi ? 2. * hPi : phiDelta0;
  // wedge from (phiDelta, phiStart), source/Wedge_Evolution.cpp:100-108
  P.w1x       = std::cos(P.phiStart);
  P.w1y       = std::sin(P.phiStart);
  P.w2x       = std::cos(P.phiStart + P.phiDelta);
  P.w2y       = std::sin(P.phiStart + P.phiDelta);
  P.n1x       = -std::sin(P.phiStart);
  P.n1y       = std::cos(P.phiStart);
  P.n2x       = std::sin(P.phiStart + P.phiDelta);
  P.n2y       = -std::cos(P.phiStart + P.phiDelta);
  P.wedgeLtPi = P.phiDelta < hPi;

  extra.clear();
  auto put = [&](double v) { extra.push_back(v); };
  P.zsOff = 0;
  for (int i = 0; i < nz; ++i) put(z[i]);
  P.rminOff = int(extra.size());
  for (int i = 0; i < nz; ++i) put(rmn[i]);
  P.rmaxOff = int(extra.size());
  for (int i = 0; i < nz; ++i) put(rmx[i]);
  P.sameZOff = int(extra.size());
  for (int i = 0; i < nz; ++i) put((i > 0 && i < nz - 1 && z[i] == z[i + 1]) ? 1.0 : 0.0);

  const double ps = hnormAngle(phiStart0), pd = phiDelta0 > 2. * hPi ? 2. * hPi : phiDelta0;
  const double sidePhi = pd / sides;
  const bool cut       = P.hasPhiCutout != 0;
  auto getPhi = [&](int side) {
    if (!cut && side == sides) side = 0;
    return hnormAngle(ps + side * sidePhi);
  };
  P.phiSecOff = int(extra.size());
  std::vector<HV> phiSec(sides + 1);
  for (int i = 0; i < sides + 1; ++i) {
    const double ph = getPhi(i);
    const HV corner = hfix(hnormalized(hv(1. * std::cos(ph), 1. * std::sin(ph), 0)));
    phiSec[i]       = hcross(hnormalized(corner), hv(0, 0, -1));
    put(phiSec[i][0]), put(phiSec[i][1]), put(phiSec[i][2]);
  }
  const double cosHalf = std::cos(0.5 * sidePhi);
  double innerR = 1.7976931348623157e308, outerR = -1.7976931348623157e308;
  for (int i = 0; i < nz; ++i) {
    if (rmn[i] < innerR) innerR = rmn[i];
    if (rmx[i] > outerR) outerR = rmx[i];
  }
  outerR /= cosHalf;
  const double bz     = 0.5 * (z[nz - 1] - z[0]) + hTol;
  const double phiTol = 100 * hTol;
  const double bt[5]  = {innerR - hHTol, outerR + hHTol, bz, !cut ? 0 : ps - phiTol, !cut ? 2. * hPi : pd + 2 * phiTol};
  ShapeRecord tr;
  if (MakeTube(bt, 5, tr, err) != 0) return -1;
  P.bounding       = tr.u.tube;
  P.boundingOffset = 0.5 * (z[0] + z[nz - 1]);

  auto vertex = [&](const double *rr, int i, int j) {
    if (!cut && j == sides) j = 0;
    const double ph = getPhi(j), rad = rr[i] / cosHalf;
    return hfix(hv(rad * std::cos(ph), rad * std::sin(ph), z[i]));
  };
  // segment headers, then quadrilaterals
  P.segOff        = int(extra.size());
  const int nseg  = nz - 1;
  extra.resize(extra.size() + 6 * nseg, 0.0);
  for (int i = 0; i < nseg; ++i) {
    const bool hasInner = rmn[i] > 0 || rmn[i + 1] > 0;
    const int nOuter    = (z[i] == z[i + 1] && rmx[i] == rmx[i + 1]) ? 0 : sides;
    const int nPhi      = cut ? ((z[i] == z[i + 1]) ? 0 : 2) : 0;
    const int nInner    = (hasInner && !(z[i] == z[i + 1] && rmn[i] == rmn[i + 1])) ? sides : 0;
    const int oOff = int(extra.size()), pOff = oOff + kQuadDoubles * nOuter, iOff = pOff + kQuadDoubles * nPhi;
    extra.resize(iOff + kQuadDoubles * nInner, 0.0);
    double *hd = &extra[P.segOff + 6 * i];
    hd[0] = oOff, hd[1] = nOuter, hd[2] = pOff, hd[3] = nPhi, hd[4] = iOff, hd[5] = nInner;
    for (int is = 0; is < nOuter; ++is) {
      double *q = &extra[oOff + kQuadDoubles * is];
      PutQuad(q, vertex(rmx, i, is), vertex(rmx, i, is + 1), vertex(rmx, i + 1, is + 1), vertex(rmx, i + 1, is));
      const HV c = vertex(rmx, i, is);
      if (q[0] * c[0] + q[1] * c[1] < 0) FlipQuad(q);
    }
    for (int is = 0; is < nInner; ++is) {
      double *q = &extra[iOff + kQuadDoubles * is];
      PutQuad(q, vertex(rmn, i, is), vertex(rmn, i, is + 1), vertex(rmn, i + 1, is + 1), vertex(rmn, i + 1, is));
      const HV c = vertex(rmn, i, is);
      if (q[0] * c[0] + q[1] * c[1] < 0) FlipQuad(q);
    }
    if (cut && nPhi == 2) {
      double *q0 = &extra[pOff];
      PutQuad(q0, vertex(rmn, i, 0), vertex(rmn, i + 1, 0), vertex(rmx, i + 1, 0), vertex(rmx, i, 0));
      if (hdot(hv(q0[0], q0[1], q0[2]), phiSec[0]) > 0) FlipQuad(q0);
      double *q1 = &extra[pOff + kQuadDoubles];
      PutQuad(q1, vertex(rmx, i, sides), vertex(rmx, i + 1, sides), vertex(rmn, i + 1, sides),
              vertex(rmn, i, sides));
      if (hdot(hv(q1[0], q1[1], q1[2]), phiSec[sides]) < 0) FlipQuad(q1);
    }
  }
  P.base = extra.data();
  PolyhedronExtent(P, extra, P.ext, P.ext + 3);
  return 0;
}

// Torus2: TorusStruct2.h:36-43 and UnplacedTorus2.h:64-75 (wedge of the raw angles, bounding tube)
int MakeTorus2(double const *a, int n, ShapeRecord &r, std::string &err)
{
  if (n < 5) return err = "torus2 needs rmin rmax rtor sphi dphi", -1;
  if (!(a[1] > 0 && a[2] > 0 && a[4] > 0) || a[0] < 0 || !(a[0] < a[1]))
    return err = "torus2 needs 0 <= rmin < rmax, rtor > 0 and dphi > 0", -1;
  r            = ShapeRecord();
  r.type       = kTorus2;
  TorusData &t = r.u.tor;
  t.rmin = a[0], t.rmax = a[1], t.rtor = a[2], t.sphi = a[3], t.dphi = a[4];
  t.rmin2 = t.rmin * t.rmin, t.rmax2 = t.rmax * t.rmax, t.rtor2 = t.rtor * t.rtor;
  t.w1x       = std::cos(t.sphi);
  t.w1y       = std::sin(t.sphi);
  t.w2x       = std::cos(t.sphi + t.dphi);
  t.w2y       = std::sin(t.sphi + t.dphi);
  t.n1x       = -std::sin(t.sphi);
  t.n1y       = std::cos(t.sphi);
  t.n2x       = std::sin(t.sphi + t.dphi);
  t.n2y       = -std::cos(t.sphi + t.dphi);
  t.hasPhi    = t.dphi < 2. * hPi;
  t.wedgeLtPi = t.dphi < hPi;
  const double bt[5] = {t.rtor - t.rmax - hTol, t.rtor + t.rmax + hTol, t.rmax, t.sphi, t.dphi};
  ShapeRecord tr;
  if (MakeTube(bt, 5, tr, err) != 0) return -1;
  t.bounding = tr.u.tube;
  return 0;
}

int MakeShapeRecord(int type, double const *params, int nparams, ShapeRecord &r, std::vector<double> &extra,
                    std::string &err)
{
  extra.clear();
  switch (type) {
  case kBox:
    return MakeBox(params, nparams, r, err);
  case kTube:
    return MakeTube(params, nparams, r, err);
  case kTrapezoid:
    return MakeTrapezoid(params, nparams, r, err);
  case kOrb:
    return MakeOrb(params, nparams, r, err);
  case kCone:
    return MakeCone(params, nparams, r, err);
  case kPolycone:
    return MakePolycone(params, nparams, r, extra, err);
  case kSphere:
    return MakeSphere(params, nparams, r, err);
  case kPolyhedron:
    return MakePolyhedron(params, nparams, r, extra, err);
  case kTorus2:
    return MakeTorus2(params, nparams, r, err);
  case kTessellated:
    return MakeTessellated(params, nparams, r, extra, err);
  case kHype:
    return MakeHype(params, nparams, r, err);
  case kTrd:
    return MakeTrd(params, nparams, r, err);
  case kParaboloid:
    return MakeParaboloid(params, nparams, r, err);
  case kParallelepiped:
    return MakeParallelepiped(params, nparams, r, err);
  case kTet:
    return MakeTet(params, nparams, r, err);
  case kEllipticalTube:
    return MakeEllipticalTube(params, nparams, r, err);
  case kEllipsoid:
    return MakeEllipsoid(params, nparams, r, err);
  case kGenTrap:
    return MakeGenTrap(params, nparams, r, err);
  case kCutTube:
    return MakeCutTube(params, nparams, r, err);
  case kEllipticalCone:
    return MakeEllipticalCone(params, nparams, r, err);
  default:
    err = "shape type not supported";
    return -1;
  }
}

void SetShapeExtraBase(ShapeRecord &r, const double *device_base)
{
  if (r.type == kPolycone) r.u.pcon.base = device_base;
  if (r.type == kPolyhedron) r.u.phed.base = device_base;
  if (r.type == kTessellated) r.u.tess.base = device_base;
}

void ShapeExtent(ShapeRecord const &r, double lo[3], double hi[3])
{
  switch (r.type) {
  case kBox:
    for (int i = 0; i < 3; ++i) lo[i] = -r.u.box.h[i], hi[i] = r.u.box.h[i];
    return;
  case kTube: {
    // source/UnplacedTube.cpp:214-262
    TubeData const &t = r.u.tube;
    lo[0] = lo[1] = -t.rmax;
    hi[0] = hi[1] = t.rmax;
    lo[2]         = -t.z;
    hi[2]         = t.z;
    if (t.dphi == 2. * hPi) return;
    PhiExtent(t.a1x, t.a1y, t.a2x, t.a2y, t.wedgeLtPi != 0, t.rmax, t.rmin, t.sphi, t.dphi, lo, hi);
    return;
  }
  case kTrapezoid:
    TrapezoidExtent(r.u.trap, lo, hi);
    return;
  case kOrb:
    for (int i = 0; i < 3; ++i) lo[i] = -r.u.orb.r, hi[i] = r.u.orb.r;
    return;
  case kSphere:
    SphereExtent(r.u.sph, lo, hi);
    return;
  case kPolyhedron:
    for (int i = 0; i < 3; ++i) lo[i] = r.u.phed.ext[i], hi[i] = r.u.phed.ext[3 + i];
    return;
  case kTorus2: {
    // UnplacedTorus2::Extent is the bounding tube's extent (UnplacedTorus2.h:210)
    ShapeRecord tr;
    tr.type   = kTube;
    tr.u.tube = r.u.tor.bounding;
    ShapeExtent(tr, lo, hi);
    return;
  }
  case kPolycone:
    for (int i = 0; i < 3; ++i) lo[i] = r.u.pcon.ext[i], hi[i] = r.u.pcon.ext[3 + i];
    return;
  case kCone:
    ConeExtent(r.u.cone, lo, hi);
    return;
  case kTessellated:
    for (int i = 0; i < 3; ++i) lo[i] = r.u.tess.minExt[i], hi[i] = r.u.tess.maxExt[i];
    return;
  case kHype: {
    // source/UnplacedHype.cpp:94-100
    const double e = r.u.hype.endOuterRadius;
    lo[0] = lo[1] = -e, hi[0] = hi[1] = e;
    lo[2] = -r.u.hype.dz, hi[2] = r.u.hype.dz;
    return;
  }
  case kTrd: {
    // VecGeom/volumes/UnplacedTrd.h:218-222
    TrdData const &t = r.u.trd;
    const double mx = t.dx1 < t.dx2 ? t.dx2 : t.dx1, my = t.dy1 < t.dy2 ? t.dy2 : t.dy1;
    lo[0] = -mx, lo[1] = -my, lo[2] = -t.dz, hi[0] = mx, hi[1] = my, hi[2] = t.dz;
    return;
  }
  case kParallelepiped: {
    // source/UnplacedParallelepiped.cpp:79-89
    ParallelepipedData const &q = r.u.ppd;
    const double dx = q.dim[0] + q.dim[1] * std::fabs(q.tanAlpha) + q.dim[2] * std::fabs(q.tanThetaCosPhi);
    const double dy = q.dim[1] + q.dim[2] * std::fabs(q.tanThetaSinPhi);
    lo[0] = -dx, lo[1] = -dy, lo[2] = -q.dim[2], hi[0] = dx, hi[1] = dy, hi[2] = q.dim[2];
    return;
  }
  case kEllipticalTube: {
    // source/UnplacedEllipticalTube.cpp:64-71
    EllipticalTubeData const &e = r.u.eltu;
    lo[0] = -e.dx, lo[1] = -e.dy, lo[2] = -e.dz, hi[0] = e.dx, hi[1] = e.dy, hi[2] = e.dz;
    return;
  }
  case kGenTrap: {
    // GenTrapStruct::Extent (VecGeom/volumes/GenTrapStruct.h:199-214)
    GenTrapData const &g = r.u.gtra;
    lo[0] = hi[0] = g.vx[0], lo[1] = hi[1] = g.vy[0], lo[2] = -g.dz, hi[2] = g.dz;
    for (int i = 0; i < 8; ++i) {
      lo[0] = std::min(lo[0], g.vx[i]), hi[0] = std::max(hi[0], g.vx[i]);
      lo[1] = std::min(lo[1], g.vy[i]), hi[1] = std::max(hi[1], g.vy[i]);
    }
    return;
  }
  case kEllipticalCone: {
    // source/UnplacedEllipticalCone.cpp:67-75
    EllipticalConeData const &c = r.u.elco;
    const double xmax = c.dx * (c.dz + c.zCut), ymax = c.dy * (c.dz + c.zCut);
    lo[0] = -xmax, lo[1] = -ymax, lo[2] = -c.zCut, hi[0] = xmax, hi[1] = ymax, hi[2] = c.zCut;
    return;
  }
  case kCutTube: {
    // source/UnplacedCutTube.cpp:128-172 (ZlimitTop/ZlimitBottom: UnplacedCutTube.h:133-147)
    CutTubeData const &c = r.u.ctub;
    TubeData const &t    = c.tube;
    const double rmax = t.rmax, rmin = t.rmin, sphi = t.sphi, dphi = t.dphi;
    const double *T = c.n1, *B = c.n0;
    const double dztop    = rmax * std::sqrt(1. - T[2] * T[2]) / T[2];
    const double dzbottom = -rmax * std::sqrt(1. - B[2] * B[2]) / B[2];
    lo[0] = -rmax, lo[1] = -rmax, lo[2] = -c.dz - dzbottom;
    hi[0] = rmax, hi[1] = rmax, hi[2] = c.dz + dztop;
    if (dphi == 2. * hPi) return;
    auto zTop = [&](double rr, double ph) { return c.dz - (rr / T[2]) * (T[0] * std::cos(ph) + T[1] * std::sin(ph)); };
    auto zBot = [&](double rr, double ph) { return -c.dz - (rr / B[2]) * (B[0] * std::cos(ph) + B[1] * std::sin(ph)); };
    auto in = [&](double x, double y) { return WedgeContainsXY(t.a1x, t.a1y, t.a2x, t.a2y, t.wedgeLtPi, x, y); };
    if (!in(-T[0], -T[1])) {
      hi[2] = std::max(zTop(rmax, sphi), zTop(rmax, sphi + dphi));
      hi[2] = std::max(hi[2], std::max(zTop(rmin, sphi), zTop(rmin, sphi + dphi)));
    }
    if (!in(-B[0], -B[1])) {
      lo[2] = std::min(zBot(rmax, sphi), zBot(rmax, sphi + dphi));
      lo[2] = std::min(lo[2], std::min(zBot(rmin, sphi), zBot(rmin, sphi + dphi)));
    }
    if (!in(1., 0.)) hi[0] = std::max(std::max(rmax * c.cosPhi1, rmax * c.cosPhi2), std::max(rmin * c.cosPhi1, rmin * c.cosPhi2));
    if (!in(-1., 0.)) lo[0] = std::min(std::min(rmax * c.cosPhi1, rmax * c.cosPhi2), std::min(rmin * c.cosPhi1, rmin * c.cosPhi2));
    if (!in(0., 1.)) hi[1] = std::max(std::max(rmax * c.sinPhi1, rmax * c.sinPhi2), std::max(rmin * c.sinPhi1, rmin * c.sinPhi2));
    if (!in(0., -1.)) lo[1] = std::min(std::min(rmax * c.sinPhi1, rmax * c.sinPhi2), std::min(rmin * c.sinPhi1, rmin * c.sinPhi2));
    return;
  }
  case kEllipsoid: {
    // source/UnplacedEllipsoid.cpp:167-171
    EllipsoidData const &e = r.u.ells;
    lo[0] = -e.xmax, lo[1] = -e.ymax, lo[2] = e.zBottomCut, hi[0] = e.xmax, hi[1] = e.ymax, hi[2] = e.zTopCut;
    return;
  }
  case kTet: {
    // source/UnplacedTet.cpp:34-50
    TetData const &t = r.u.tet;
    for (int c = 0; c < 3; ++c) {
      lo[c] = hi[c] = t.v[0][c];
      for (int k = 1; k < 4; ++k) lo[c] = std::min(lo[c], t.v[k][c]), hi[c] = std::max(hi[c], t.v[k][c]);
    }
    return;
  }
  case kParaboloid: {
    // source/UnplacedParaboloid.cpp:68-76
    ParaboloidData const &q = r.u.para;
    lo[0] = lo[1] = -q.dx, hi[0] = hi[1] = q.dx, lo[2] = -q.dz, hi[2] = q.dz;
    return;
  }
  default:
    lo[0] = lo[1] = lo[2] = 0;
    hi[0] = hi[1] = hi[2] = 0;
  }
}

} // namespace vgmi
