// Oracle C ABI (ctypes-friendly) used by tests/ and by the CPU-baseline leg of bench.py.
// TEST INFRASTRUCTURE ONLY (see common.h).
//
// Shape parameter lists (same order as the reference constructors):
//   0 box         dx dy dz                                   (volumes/UnplacedBox.h)
//   1 tube        rmin rmax dz sphi dphi                     (volumes/UnplacedTube.h:29)
//   5 trapezoid   dz theta phi dy1 dx1 dx2 alpha1 dy2 dx3 dx4 alpha2 (volumes/UnplacedTrapezoid.h:56);
//                 also the reduced forms of 3, 4, 5, 6 parameters and 8 corners (24 doubles, the
//                 TrapCorners order of source/UnplacedTrapezoid.cpp:447-465)
//  10 hype        rmin rmax stIn stOut dz                    (volumes/UnplacedHype.h:56)
//  11 trd         x1 x2 y1 y2 z; also x1 x2 y z and x y z     (volumes/UnplacedTrd.h:51-80); a sixth
//                 value 1 selects the Trd1 kind (y1 == y2) of GeoManager::MakeInstance
//  12 paraboloid  rlo rhi dz                                 (volumes/UnplacedParaboloid.h:60)
//  13 parallelepiped dx dy dz alpha theta phi               (volumes/UnplacedParallelepiped.h:65-71)
//  14 tet         p0 p1 p2 p3 (12 doubles)                    (volumes/UnplacedTet.h:48-49)
//  15 elliptical tube dx dy dz                               (volumes/UnplacedEllipticalTube.h:51)
//  16 ellipsoid   dx dy dz [zBottomCut zTopCut]               (volumes/UnplacedEllipsoid.h:62; cuts default 0)
//  17 gentrap     x[8] y[8] dz (planar ones)                  (volumes/UnplacedGenTrap.h:46)
//  18 cuttube     rmin rmax dz sphi dphi bx by bz tx ty tz     (volumes/UnplacedCutTube.h:54-58)
//  19 ellipticalcone a b h zcut                                (volumes/UnplacedEllipticalCone.h)
//   9 tessellated ntri, then 9 doubles per triangle (v0 v1 v2, anticlockwise seen from outside;
//                 UnplacedTessellated::AddTriangularFacet, volumes/UnplacedTessellated.h:95-100)
#include "shapes.h"
#include <chrono>
#include <thread>
#include <vector>

using namespace vgo;

namespace vgo {
// dz theta phi dy1 dx1 dx2 alpha1 dy2 dx3 dx4 alpha2 of the reduced trapezoid constructors:
//  3: xbox ybox zbox (source/UnplacedTrapezoid.cpp:176-190)
//  4: dx1 dx2 dy dz (:148-154)
//  5: dx1 dx2 dy1 dy2 dz (:156-162)
//  6: dx dy dz alpha theta phi (:164-174)
// 11: the full list (volumes/UnplacedTrapezoid.h:56-63)
bool TrapezoidFullParams(const double *a, int n, double f[11])
{
  switch (n) {
  case 3: {
    double g[11] = {a[2], 0., 0., a[1], a[0], a[0], 0., a[1], a[0], a[0], 0.};
    for (int i = 0; i < 11; ++i) f[i] = g[i];
    return true;
  }
  case 4: {
    double g[11] = {a[3], 0., 0., a[2], a[0], a[0], 0., a[2], a[1], a[1], 0.};
    for (int i = 0; i < 11; ++i) f[i] = g[i];
    return true;
  }
  case 5: {
    double g[11] = {a[4], 0., 0., a[2], a[0], a[0], 0., a[3], a[1], a[1], 0.};
    for (int i = 0; i < 11; ++i) f[i] = g[i];
    return true;
  }
  case 6: {
    double g[11] = {a[2], a[4], a[5], a[1], a[0], a[0], a[3], a[1], a[0], a[0], a[3]};
    for (int i = 0; i < 11; ++i) f[i] = g[i];
    return true;
  }
  case 11:
    for (int i = 0; i < 11; ++i) f[i] = a[i];
    return true;
  default:
    return false;
  }
}

Shape *MakeShape(int type, const double *a, int n)
{
  switch (type) {
  case 0:
    return n >= 3 ? new Box(a[0], a[1], a[2]) : nullptr;
  case 1:
    return n >= 5 ? new Tube(a[0], a[1], a[2], a[3], a[4]) : nullptr;
  case 2:
    return n >= 7 ? new Cone(a[0], a[1], a[2], a[3], a[4], a[5], a[6]) : nullptr;
  case 8:
    return n >= 1 ? new Orb(a[0]) : nullptr;
  case 10:
    return n >= 5 ? new Hype(a[0], a[1], a[2], a[3], a[4]) : nullptr;
  case 11:
    if (n == 3) return new Trd(a[0], a[0], a[1], a[1], a[2]);
    if (n == 4) return new Trd(a[0], a[1], a[2], a[2], a[3]);
    if (n == 6 && a[5] == 1.0 && a[2] == a[3]) return new Trd(a[0], a[1], a[2], a[3], a[4], true);
    return n == 5 || (n == 6 && a[5] == 0.0) ? new Trd(a[0], a[1], a[2], a[3], a[4]) : nullptr;
  case 12:
    return n >= 3 ? new Paraboloid(a[0], a[1], a[2]) : nullptr;
  case 13:
    return n >= 6 ? new Parallelepiped(a[0], a[1], a[2], a[3], a[4], a[5]) : nullptr;
  case 15:
    return n >= 3 ? new EllipticalTube(a[0], a[1], a[2]) : nullptr;
  case 17: {
    if (n < 17) return nullptr;
    auto *g = new GenTrap(a, a + 8, a[16]);
    if (!g->ok) {
      delete g;
      return nullptr;
    }
    return g;
  }
  case 19:
    return n >= 4 ? new EllipticalCone(a[0], a[1], a[2], a[3]) : nullptr;
  case 18: { // rmin rmax dz sphi dphi, bottom normal (3), top normal (3)
    if (n < 11) return nullptr;
    auto *c = new CutTube(a[0], a[1], a[2], a[3], a[4], V3(a[5], a[6], a[7]), V3(a[8], a[9], a[10]));
    if (!c->ok) {
      delete c;
      return nullptr;
    }
    return c;
  }
  case 16:
    return n == 3 ? new Ellipsoid(a[0], a[1], a[2], 0., 0.)
                  : (n >= 5 ? new Ellipsoid(a[0], a[1], a[2], a[3], a[4]) : nullptr);
  case 14:
    return n >= 12 ? new Tet(V3(a[0], a[1], a[2]), V3(a[3], a[4], a[5]), V3(a[6], a[7], a[8]), V3(a[9], a[10], a[11]))
                   : nullptr;
  case 3: {
    if (n < 3) return nullptr;
    int nz = int(a[2]);
    if (nz < 2 || n < 3 + 3 * nz) return nullptr;
    return new Polycone(a[0], a[1], nz, a + 3, a + 3 + nz, a + 3 + 2 * nz);
  }
  case 4: {
    if (n < 4) return nullptr;
    int sides = int(a[2]), nz = int(a[3]);
    if (sides < 1 || nz < 2 || n < 4 + 3 * nz) return nullptr;
    return new Polyhedron(a[0], a[1], sides, nz, a + 4, a + 4 + nz, a + 4 + 2 * nz);
  }
  case 7:
    return n >= 5 ? new Torus2(a[0], a[1], a[2], a[3], a[4]) : nullptr;
  case 6:
    return n >= 6 ? new Sphere(a[0], a[1], a[2], a[3], a[4], a[5]) : nullptr;
  case 9: {
    if (n < 1) return nullptr;
    const int ntri = int(a[0]);
    if (ntri < 4 || n < 1 + 9 * ntri) return nullptr;
    auto *t = new Tessellated(a + 1, ntri);
    if (!t->ok) {
      delete t;
      return nullptr;
    }
    return t;
  }
  case 5: {
    // the UnplacedTrapezoid constructors by parameter count (volumes/UnplacedTrapezoid.h:56-102,
    // source/UnplacedTrapezoid.cpp:129-190); all but the corner form reduce to the full one
    double f[11];
    if (!TrapezoidFullParams(a, n, f)) {
      if (n != 24) return nullptr;
      V3 pt[8];
      for (int i = 0; i < 8; ++i) pt[i] = V3(a[3 * i], a[3 * i + 1], a[3 * i + 2]);
      return new Trapezoid(pt);
    }
    return new Trapezoid(f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], f[8], f[9], f[10]);
  }
  default:
    return nullptr;
  }
}
} // namespace vgo

enum { OP_CONTAINS = 0, OP_INSIDE = 1, OP_DIST_IN = 2, OP_DIST_OUT = 3, OP_SAFETY_IN = 4, OP_SAFETY_OUT = 5 };

static void run_range(const Shape *s, int op, const double *px, const double *py, const double *pz,
                      const double *dx, const double *dy, const double *dz, const double *step, void *out,
                      int64_t b, int64_t e)
{
  for (int64_t i = b; i < e; ++i) {
    V3 p(px[i], py[i], pz[i]);
    switch (op) {
    case OP_CONTAINS:
      ((uint8_t *)out)[i] = s->Contains(p) ? 1 : 0;
      break;
    case OP_INSIDE:
      ((int32_t *)out)[i] = s->Inside(p);
      break;
    case OP_DIST_IN:
      ((double *)out)[i] = s->DistanceToIn(p, V3(dx[i], dy[i], dz[i]), step ? step[i] : kInfLength);
      break;
    case OP_DIST_OUT:
      ((double *)out)[i] = s->DistanceToOut(p, V3(dx[i], dy[i], dz[i]), step ? step[i] : kInfLength);
      break;
    case OP_SAFETY_IN:
      ((double *)out)[i] = s->SafetyToIn(p);
      break;
    case OP_SAFETY_OUT:
      ((double *)out)[i] = s->SafetyToOut(p);
      break;
    }
  }
}

extern "C" {

void *vgo_shape_new(int type, const double *params, int nparams) { return MakeShape(type, params, nparams); }
void vgo_shape_free(void *s) { delete (Shape *)s; }

void vgo_shape_extent(void *s, double *lo3, double *hi3)
{
  V3 lo, hi;
  ((Shape *)s)->Extent(lo, hi);
  lo3[0] = lo.x; lo3[1] = lo.y; lo3[2] = lo.z;
  hi3[0] = hi.x; hi3[1] = hi.y; hi3[2] = hi.z;
}

// Runs op over n points with `threads` host threads (1 = the scalar reference loop).
int vgo_shape_query(void *s, int op, const double *px, const double *py, const double *pz, const double *dx,
                    const double *dy, const double *dz, const double *step, void *out, int64_t n, int threads)
{
  if (!s || op < 0 || op > 5) return -1;
  if ((op == OP_DIST_IN || op == OP_DIST_OUT) && (!dx || !dy || !dz)) return -2;
  if (threads <= 1) {
    run_range((Shape *)s, op, px, py, pz, dx, dy, dz, step, out, 0, n);
    return 0;
  }
  std::vector<std::thread> pool;
  int64_t chunk = (n + threads - 1) / threads;
  for (int t = 0; t < threads; ++t) {
    int64_t b = t * chunk, e = std::min<int64_t>(n, b + chunk);
    if (b >= e) break;
    pool.emplace_back(run_range, (Shape *)s, op, px, py, pz, dx, dy, dz, step, out, b, e);
  }
  for (auto &th : pool) th.join();
  return 0;
}

// Normal per point; n{x,y,z} hold the caller's vectors on entry (the reference leaves its normal
// argument unassigned in some branches) and the normals on return.
int vgo_shape_normal(void *s, const double *px, const double *py, const double *pz, double *nx, double *ny,
                     double *nz, uint8_t *valid, int64_t n)
{
  if (!s) return -1;
  for (int64_t i = 0; i < n; ++i) {
    V3 nv(nx[i], ny[i], nz[i]);
    valid[i] = ((Shape *)s)->Normal(V3(px[i], py[i], pz[i]), nv);
    nx[i] = nv.x, ny[i] = nv.y, nz[i] = nv.z;
  }
  return 0;
}

} // extern "C"
