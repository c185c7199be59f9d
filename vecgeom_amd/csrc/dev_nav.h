// Device-side BVH navigation (FP64), one lane per track.
//
//  - NavStateIndex arithmetic and cached global matrices (VecGeom/navigation/NavStateIndex.h:186-265)
//  - placed-volume transforms (VecGeom/base/Transformation3D.h:661-830, general case)
//  - BVH traversal CheckDaughterIntersections / LevelLocate / ComputeSafety (VecGeom/base/BVH.h:96-285)
//  - VNavigatorHelper step with PrepareOutState and Relocate (VecGeom/navigation/VNavigator.h:160-540,1066-1091)
//  - GlobalLocator (VecGeom/navigation/GlobalLocator.h:35-200)
//
// The traversal stack lives in LDS, one column per lane (stride = block size), so
// pushes/pops are bank-conflict free and no scratch memory is used.
#pragma once
#include "dev_shapes.h"
#include "geom_data.h"

namespace vgmi {
namespace nav {

constexpr int kNavBlock = 128;
constexpr int kStackDepth = 34; // BVH depth <= 32 -> at most 33 pending nodes

using V3 = Vec3<double>;

struct Stack {
  uint32_t *base; // LDS column of this lane
  VGMI_DEV uint32_t &at(int i) { return base[i * kNavBlock]; }
};

struct Mat {
  double t[3], r[9];
};

VGMI_DEV V3 rotate(const double *r, V3 const &m)
{
  double x = m.x * r[0], y = m.x * r[1], z = m.x * r[2];
  x += m.y * r[3];
  y += m.y * r[4];
  z += m.y * r[5];
  x += m.z * r[6];
  y += m.z * r[7];
  z += m.z * r[8];
  return V3(x, y, z);
}

VGMI_DEV V3 to_local(const double *t, const double *r, V3 const &m)
{
  return rotate(r, V3(m.x - t[0], m.y - t[1], m.z - t[2]));
}

VGMI_DEV V3 to_master(const double *t, const double *r, V3 const &l)
{
  double x = t[0], y = t[1], z = t[2];
  x += l.x * r[0];
  x += l.y * r[1];
  x += l.z * r[2];
  y += l.x * r[3];
  y += l.y * r[4];
  y += l.z * r[5];
  z += l.x * r[6];
  z += l.y * r[7];
  z += l.z * r[8];
  return V3(x, y, z);
}

VGMI_DEV int top(GeomView const &g, uint32_t nav) { return nav > 0 ? int(g.navind[nav + 1]) : -1; }
VGMI_DEV uint32_t pop(GeomView const &g, uint32_t nav) { return nav > 0 ? g.navind[nav] : 0u; }
VGMI_DEV uint32_t push(GeomView const &g, uint32_t nav, int pvol)
{
  return nav > 0 ? g.navind[nav + 3 + g.vols[pvol].childId] : 1u;
}

// Global matrix of the state's top volume, from the cached table entry.
VGMI_DEV void top_matrix(GeomView const &g, uint32_t nav, Mat &m)
{
  m.t[0] = m.t[1] = m.t[2] = 0.;
  for (int i = 0; i < 9; ++i) m.r[i] = (i % 4 == 0) ? 1. : 0.;
  if (nav == 0) return;
  const uint32_t w    = g.navind[nav + 2];
  const uint32_t hasm = (w >> 8) & 0xffu;
  if ((hasm & 0x03u) == 0) return;
  const uint32_t nd = w >> 16;
  uint32_t ti       = nav + 3 + nd + ((nd + 1) & 1);
  ti += ti & 1u; // 8-byte alignment of the matrix block
  // six 16-byte loads of the (8-byte aligned) 12 doubles, which every node stores
  typedef double D2 __attribute__((ext_vector_type(2), aligned(8)));
  const D2 *a  = reinterpret_cast<const D2 *>(g.navind + ti);
  const D2 v0 = a[0], v1 = a[1], v2 = a[2], v3 = a[3], v4 = a[4], v5 = a[5];
  if (hasm & 0x02u) m.t[0] = v0.x, m.t[1] = v0.y, m.t[2] = v1.x;
  if (hasm & 0x01u) {
    m.r[0] = v1.y, m.r[1] = v2.x, m.r[2] = v2.y, m.r[3] = v3.x, m.r[4] = v3.y;
    m.r[5] = v4.x, m.r[6] = v4.y, m.r[7] = v5.x, m.r[8] = v5.y;
  }
}

// ---- shape dispatch on a device shape record. Every function below is templated on TM,
// the bit mask of shape types the kernel compiles in (bit E for ShapeType E).
// The distances take the step limit the reference's navigator hands the shapes: the mother's
// DistanceToOut gets the step limit (navigation/VNavigator.h:246), a daughter's DistanceToIn the
// current step (base/BVH.h:115). Only the tessellated solid's, the GenTrap's and the CutTube's DistanceToIn depend on it.
template <unsigned TM>
VGMI_DEV double dist_in(ShapeRecord const &s, V3 const &p, V3 const &d, double stepMax)
{
  if constexpr ((TM >> kTessellated) & 1u)
    if (s.type == kTessellated) return tess::dist_in_step64(s.u.tess, tess::pool_walk(s.u.tess), p, d, stepMax);
  if constexpr ((TM >> kGenTrap) & 1u)
    if (s.type == kGenTrap) return gtra::DistanceToInStep(s.u.gtra, p, d, stepMax);
  if constexpr ((TM >> kCutTube) & 1u)
    if (s.type == kCutTube) return ctub::DistanceToInStep(s.u.ctub, p, d, stepMax);
  return rec_dist_in<TM>(s, p, d);
}
template <unsigned TM>
VGMI_DEV double dist_out(ShapeRecord const &s, V3 const &p, V3 const &d, double stepMax)
{
  if constexpr ((TM >> kTessellated) & 1u)
    if (s.type == kTessellated) return tess::dist_out_step64(s.u.tess, tess::pool_walk(s.u.tess), p, d, stepMax);
  return rec_dist_out<TM>(s, p, d);
}
template <unsigned TM>
VGMI_DEV bool contains(ShapeRecord const &s, V3 const &p) { return rec_contains<TM>(s, p); }
template <unsigned TM>
VGMI_DEV double safety_in(ShapeRecord const &s, V3 const &p) { return rec_safety_in<TM>(s, p); }
template <unsigned TM>
VGMI_DEV double safety_out(ShapeRecord const &s, V3 const &p) { return rec_safety_out<TM>(s, p); }

VGMI_DEV ShapeRecord const &shape_of_pvol(GeomView const &g, int pvol) { return g.shapes[g.vols[pvol].shape]; }

// AABB::ComputeIntersectionInvDir (VecGeom/base/AABB.h:136-154); B is any record with lo[3], hi[3]
template <class B>
VGMI_DEV void slab(B const &b, V3 const &p, V3 const &inv, double &tmin, double &tmax)
{
  double a0 = (b.lo[0] - p.x) * inv.x, a1 = (b.lo[1] - p.y) * inv.y, a2 = (b.lo[2] - p.z) * inv.z;
  double b0 = (b.hi[0] - p.x) * inv.x, b1 = (b.hi[1] - p.y) * inv.y, b2 = (b.hi[2] - p.z) * inv.z;
  if (a0 > b0) {
    double t = a0;
    a0       = b0;
    b0       = t;
  }
  if (a1 > b1) {
    double t = a1;
    a1       = b1;
    b1       = t;
  }
  if (a2 > b2) {
    double t = a2;
    a2       = b2;
    b2       = t;
  }
  tmin = vmax(a0, vmax(a1, a2));
  tmax = vmin(b0, vmin(b1, b2)) * (1. + 2 * DBL_EPSILON);
}

template <class B>
VGMI_DEV bool box_contains(B const &b, V3 const &p)
{
  return p.x >= b.lo[0] && p.x <= b.hi[0] && p.y >= b.lo[1] && p.y <= b.hi[1] && p.z >= b.lo[2] && p.z <= b.hi[2];
}

template <class B>
VGMI_DEV double box_safety(B const &b, V3 const &p)
{
  double c0 = 0.5 * (b.hi[0] + b.lo[0]), c1 = 0.5 * (b.hi[1] + b.lo[1]), c2 = 0.5 * (b.hi[2] + b.lo[2]);
  double s0 = vabs(p.x - c0) - 0.5 * (b.hi[0] - b.lo[0]);
  double s1 = vabs(p.y - c1) - 0.5 * (b.hi[1] - b.lo[1]);
  double s2 = vabs(p.z - c2) - 0.5 * (b.hi[2] - b.lo[2]);
  return vmax(s0, vmax(s1, s2));
}

// ---- mixed precision: FP32 box filters over the widened FP32 tables.
// TM bit 31 selects the mixed-precision traversal; the shape-type bits are unchanged.
constexpr unsigned kMixedFlag = 1u << 31;
template <unsigned TM>
constexpr bool kMixed = (TM & kMixedFlag) != 0;

using F3 = Vec3<float>;
// Relative and absolute slack of the FP32 slab comparisons (the FP32 arithmetic error is
// ~1e-7 relative; the spatial rounding is absorbed by the box widening).
constexpr float kSlabRel = 1e-5f, kSlabAbs = 1e-5f;

VGMI_DEV F3 to_f3(V3 const &v) { return F3(float(v.x), float(v.y), float(v.z)); }

// Wave-uniform access: when every active lane reads the same records, they are read through
// the scalar cache (constant address space, uniform address), which takes no vector-memory
// address cycles; the loop is then the same for all lanes.
VGMI_DEV bool wave_uniform(uint32_t x) { return __ballot(x != __builtin_amdgcn_readfirstlane(x)) == 0; }
template <class T>
using ConstPtr = const __attribute__((address_space(4))) T *;
// A slot read through a uniform constant-space pointer (field by field: the implicit copy
// constructor does not take the constant address space).
VGMI_DEV NavSlot32 load_slot(ConstPtr<NavSlot32> p, int i)
{
  NavSlot32 r;
  for (int k = 0; k < 3; ++k) r.lo[k] = p[i].lo[k], r.hi[k] = p[i].hi[k];
  r.pvol  = p[i].pvol;
  r.shape = p[i].shape;
  return r;
}
template <class T>
VGMI_DEV ConstPtr<T> uniform_ptr(const T *p)
{
  const uint64_t v  = reinterpret_cast<uint64_t>(p);
  const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(v)), hi = __builtin_amdgcn_readfirstlane(uint32_t(v >> 32));
  return reinterpret_cast<ConstPtr<T>>((uint64_t(hi) << 32) | lo);
}
VGMI_DEV float rcp32(float x) { return __builtin_amdgcn_rcpf(x); }

// FP32 slab test against a widened box: true unless the FP64 test of the unwidened box
// certainly fails (ray misses the box, box behind the point, or entry beyond `lim`).
template <class B>
VGMI_DEV bool slab32_pass(B const &b, F3 const &p, F3 const &inv, float lim, float &tmin_out)
{
  float a0 = (b.lo[0] - p.x) * inv.x, a1 = (b.lo[1] - p.y) * inv.y, a2 = (b.lo[2] - p.z) * inv.z;
  float b0 = (b.hi[0] - p.x) * inv.x, b1 = (b.hi[1] - p.y) * inv.y, b2 = (b.hi[2] - p.z) * inv.z;
  const float tmin = vmax(vmin(a0, b0), vmax(vmin(a1, b1), vmin(a2, b2)));
  const float tmax = vmin(vmax(a0, b0), vmin(vmax(a1, b1), vmax(a2, b2)));
  tmin_out = tmin;
  const float sl = kSlabRel * (vabs(tmin) + vabs(tmax)) + kSlabAbs;
  return tmin <= tmax + sl && tmax >= -sl && tmin < lim + (kSlabRel * vabs(lim) + kSlabAbs);
}

// Lower bound (in FP64 units) of the FP64 box_safety of the unwidened box.
template <class B>
VGMI_DEV double box_safety32(B const &b, F3 const &p)
{
  const float c0 = 0.5f * (b.hi[0] + b.lo[0]), c1 = 0.5f * (b.hi[1] + b.lo[1]), c2 = 0.5f * (b.hi[2] + b.lo[2]);
  const float s0 = vabs(p.x - c0) - 0.5f * (b.hi[0] - b.lo[0]);
  const float s1 = vabs(p.y - c1) - 0.5f * (b.hi[1] - b.lo[1]);
  const float s2 = vabs(p.z - c2) - 0.5f * (b.hi[2] - b.lo[2]);
  return double(vmax(s0, vmax(s1, s2)));
}

template <class B>
VGMI_DEV bool box_contains32(B const &b, F3 const &p)
{
  return p.x >= b.lo[0] && p.x <= b.hi[0] && p.y >= b.lo[1] && p.y <= b.hi[1] && p.z >= b.lo[2] && p.z <= b.hi[2];
}

// Lower bound of the distance from a local point to a shape (zero or negative inside): the
// largest signed distance to a half-space or cylinder containing the solid. -inf where no
// cheap bound is used.
VGMI_DEV double lower_bound_in(ShapeRecord const &s, V3 const &p)
{
  switch (s.type) {
  case kBox:
    return vmax(vabs(p.x) - s.u.box.h[0], vmax(vabs(p.y) - s.u.box.h[1], vabs(p.z) - s.u.box.h[2]));
  case kTube: {
    TubeData const &t = s.u.tube;
    const double rho  = vsqrt(p.x * p.x + p.y * p.y);
    return vmax(vmax(rho - t.rmax, t.rmin - rho), vabs(p.z) - t.z);
  }
  case kTrapezoid: {
    TrapezoidData const &t = s.u.trap;
    double b               = vabs(p.z) - t.dz;
    for (int i = 0; i < 4; ++i) b = vmax(b, t.a[i] * p.x + t.b[i] * p.y + t.c[i] * p.z + t.d[i]);
    return b;
  }
  case kOrb:
    return vsqrt(p.Mag2()) - s.u.orb.r;
  default:
    return -kInfLength;
  }
}

// BVH::CheckDaughterIntersections. `cap` (the step limit when the mother distance is finite,
// else kInfLength) prunes daughters that cannot change the outcome of the step: one whose
// distance exceeds both the current step and the limit ends the step at the limit either way.
// A daughter is skipped when its box entry distance or its distance lower bound exceeds
// cap + 1e-6, which leaves every distance below the limit, and the visiting order, unchanged.
// `bound` receives a lower bound of the Euclidean distance from lp to every daughter: the
// box distance of each subtree or daughter box the walk does not descend into, and the box
// distance (or shape lower bound, when computed) of each daughter it tests.
template <unsigned TM>
VGMI_DEV void check_daughters_mixed(GeomView const &g, NavVol const &M, V3 const &lp, V3 const &ld, double &step,
                                    int last, int &hit, double cap, double &bound, Stack stk);

template <unsigned TM>
VGMI_DEV void check_daughters(GeomView const &g, NavVol const &M, V3 const &lp, V3 const &ld, double &step, int last,
                              int &hit, double cap, double &bound, Stack stk)
{
  if constexpr (kMixed<TM>) {
    check_daughters_mixed<TM>(g, M, lp, ld, step, last, hit, cap, bound, stk);
    return;
  }
  const V3 inv(1.0 / NonZero(ld.x), 1.0 / NonZero(ld.y), 1.0 / NonZero(ld.z));
  const double capm = cap + 1e-6;
  const NavNode *nodes = g.bnodes + M.nodeBase;
  const NavSlot *slots = g.slots + M.slotBase;
  double bnd        = kInfLength;
  int sp            = 1;
  stk.at(0)         = 0;
  do {
    const uint32_t id = stk.at(--sp);
    NavNode const &N  = nodes[id];
    const int nc      = N.nchild;
    if (nc >= 0) {
      for (int i = 0; i < nc; ++i) {
        NavSlot const &box = slots[N.offset + i];
        double tmin, tmax;
        slab(box, lp, inv, tmin, tmax);
        double lb = box_safety(box, lp);
        if (tmin <= tmax && tmax >= 0.0 && tmin < step && !(tmin > capm)) {
          const int pv         = box.pvol;
          NavVol const &P      = g.vols[pv];
          ShapeRecord const &S = g.shapes[box.shape];
          const V3 loc         = to_local(P.t, P.r, lp);
          const double lbs     = lower_bound_in(S, loc);
          lb                   = vmax(lb, lbs);
          if (!(lbs > capm)) {
            const double dist = dist_in<TM>(S, loc, rotate(P.r, ld), step);
            if (dist < step && !(dist <= 0.0 && pv == last)) step = dist, hit = pv;
          }
        }
        bnd = vmin(bnd, lb);
      }
    } else {
      const uint32_t l = 2 * id + 1, r = 2 * id + 2;
      NavNode const &bl = nodes[l];
      NavNode const &br = nodes[r];
      double tminL, tmaxL, tminR, tmaxR;
      slab(bl, lp, inv, tminL, tmaxL);
      slab(br, lp, inv, tminR, tmaxR);
      const bool tl = tminL <= tmaxL && tmaxL >= 0.0 && tminL < step && !(tminL > capm);
      const bool tr = tminR <= tmaxR && tmaxR >= 0.0 && tminR < step && !(tminR > capm);
      if (!tl) bnd = vmin(bnd, box_safety(bl, lp));
      if (!tr) bnd = vmin(bnd, box_safety(br, lp));
      if (tminR < tminL) {
        if (tr) stk.at(sp++) = r;
        if (tl) stk.at(sp++) = l;
      } else {
        if (tl) stk.at(sp++) = l;
        if (tr) stk.at(sp++) = r;
      }
    }
  } while (sp > 0);
  bound = bnd;
}

// Mixed-precision CheckDaughterIntersections: the same walk with FP32 box tests as filters.
// A box the FP64 walk enters always passes its FP32 test (widened box, slack), so the walk
// visits a superset of the FP64 walk's candidates, in the FP64 order (sibling order is taken
// from FP64 entry distances whenever both siblings pass), and evaluates every candidate in
// FP64. An extra candidate cannot change the step: its FP64 box test fails, so its distance
// is not below the current step. Results are those of the FP64 walk, bit for bit.
template <unsigned TM>
VGMI_DEV void check_daughters_mixed(GeomView const &g, NavVol const &M, V3 const &lp, V3 const &ld, double &step,
                                    int last, int &hit, double cap, double &bound, Stack stk)
{
  const F3 pf = to_f3(lp);
  const F3 df = to_f3(ld);
  const F3 invf(rcp32(NonZero(df.x)), rcp32(NonZero(df.y)), rcp32(NonZero(df.z)));
  const double capm = cap + 1e-6;
  const NavNode32 *nodes = g.bnodes32 + M.nodeBase;
  const NavSlot32 *slots = g.slots32 + M.slotBase;
  double bnd = kInfLength;
  int sp     = 1;
  stk.at(0)  = 0;
  do {
    const uint32_t id  = stk.at(--sp);
    NavNode32 const &N = nodes[id];
    const int nc       = N.nchild;
    const float lim    = float(vmin(step, capm));
    if (nc >= 0) {
      // A leaf in two passes: the FP32 box tests of all its slots (each slot's load issued
      // while the previous one is tested), then the FP64 candidates in slot order. The first
      // pass uses the step at the leaf's entry, a superset of the FP64 walk's tests.
      for (int base = 0; base < nc; base += 32) {
        const int cnt = nc - base < 32 ? nc - base : 32;
        const NavSlot32 *ls = slots + N.offset + base;
        uint32_t pass = 0;
        if (wave_uniform(uint32_t(M.slotBase + N.offset + base)) && wave_uniform(uint32_t(cnt))) {
          ConstPtr<NavSlot32> us = uniform_ptr(ls);
          const int ucnt         = __builtin_amdgcn_readfirstlane(cnt);
          for (int i = 0; i < ucnt; ++i) {
            const NavSlot32 box = load_slot(us, i);
            float t;
            if (slab32_pass(box, pf, invf, lim, t))
              pass |= 1u << i;
            else
              bnd = vmin(bnd, box_safety32(box, pf));
          }
        } else {
          NavSlot32 cur = ls[0];
          for (int i = 0; i < cnt; ++i) {
            const NavSlot32 nxt = ls[i + 1 < cnt ? i + 1 : i];
            float t;
            if (slab32_pass(cur, pf, invf, lim, t))
              pass |= 1u << i;
            else
              bnd = vmin(bnd, box_safety32(cur, pf));
            cur = nxt;
          }
        }
        while (pass) {
          const int i = __builtin_ctz(pass);
          pass &= pass - 1u;
          NavSlot32 const &box = ls[i];
          double lb            = box_safety32(box, pf);
          const int pv         = box.pvol;
          NavVol const &P      = g.vols[pv];
          ShapeRecord const &S = g.shapes[box.shape];
          const V3 loc         = to_local(P.t, P.r, lp);
          const double lbs     = lower_bound_in(S, loc);
          lb                   = vmax(lb, lbs);
          if (!(lbs > capm)) {
            const double dist = dist_in<TM>(S, loc, rotate(P.r, ld), step);
            if (dist < step && !(dist <= 0.0 && pv == last)) step = dist, hit = pv;
          }
          bnd = vmin(bnd, lb);
        }
      }
    } else {
      const uint32_t l = 2 * id + 1, r = 2 * id + 2;
      float tl32, tr32;
      const bool tl = slab32_pass(nodes[l], pf, invf, lim, tl32);
      const bool tr = slab32_pass(nodes[r], pf, invf, lim, tr32);
      if (!tl) bnd = vmin(bnd, box_safety32(nodes[l], pf));
      if (!tr) bnd = vmin(bnd, box_safety32(nodes[r], pf));
      bool rfirst = false;
      if (tl && tr) {
        // sibling order from the FP64 entry distances, as in the FP64 walk
        const V3 inv(1.0 / NonZero(ld.x), 1.0 / NonZero(ld.y), 1.0 / NonZero(ld.z));
        double tminL, tmaxL, tminR, tmaxR;
        slab(g.bnodes[M.nodeBase + l], lp, inv, tminL, tmaxL);
        slab(g.bnodes[M.nodeBase + r], lp, inv, tminR, tmaxR);
        rfirst = tminR < tminL;
      }
      if (rfirst) {
        stk.at(sp++) = r;
        stk.at(sp++) = l;
      } else {
        if (tl) stk.at(sp++) = l;
        if (tr) stk.at(sp++) = r;
      }
    }
  } while (sp > 0);
  bound = bnd;
}

// BVH::LevelLocate(exclvol, localpoint, pvol, daughterlocalpoint)
// Mixed precision: FP32 containment filters over the widened boxes (a superset of the FP64
// box hits, visited in the same fixed order), FP64 shape containment.
template <unsigned TM>
VGMI_DEV bool level_locate_mixed(GeomView const &g, NavVol const &M, int excl, V3 const &lp, int &pvol, V3 &dlp,
                                 Stack stk)
{
  const F3 pf              = to_f3(lp);
  const NavNode32 *nodes   = g.bnodes32 + M.nodeBase;
  const NavSlot32 *slots   = g.slots32 + M.slotBase;
  int sp    = 1;
  stk.at(0) = 0;
  do {
    const uint32_t id  = stk.at(--sp);
    NavNode32 const &N = nodes[id];
    const int nc       = N.nchild;
    if (nc >= 0) {
      // FP32 containment of all slots first (loads issued ahead), then FP64 in slot order
      for (int base = 0; base < nc; base += 32) {
        const int cnt = nc - base < 32 ? nc - base : 32;
        const NavSlot32 *ls = slots + N.offset + base;
        uint32_t in = 0;
        if (wave_uniform(uint32_t(M.slotBase + N.offset + base)) && wave_uniform(uint32_t(cnt))) {
          ConstPtr<NavSlot32> us = uniform_ptr(ls);
          const int ucnt         = __builtin_amdgcn_readfirstlane(cnt);
          for (int i = 0; i < ucnt; ++i) {
            const NavSlot32 box = load_slot(us, i);
            if (box_contains32(box, pf)) in |= 1u << i;
          }
        } else {
          NavSlot32 cur = ls[0];
          for (int i = 0; i < cnt; ++i) {
            const NavSlot32 nxt = ls[i + 1 < cnt ? i + 1 : i];
            if (box_contains32(cur, pf)) in |= 1u << i;
            cur = nxt;
          }
        }
        while (in) {
          const int i = __builtin_ctz(in);
          in &= in - 1u;
          const int pv = ls[i].pvol;
          if (pv == excl) continue;
          NavVol const &P = g.vols[pv];
          const V3 loc    = to_local(P.t, P.r, lp);
          if (contains<TM>(g.shapes[ls[i].shape], loc)) {
            pvol = pv;
            dlp  = loc;
            return true;
          }
        }
      }
    } else {
      const uint32_t l = 2 * id + 1, r = 2 * id + 2;
      if (box_contains32(nodes[l], pf)) stk.at(sp++) = l;
      if (box_contains32(nodes[r], pf)) stk.at(sp++) = r;
    }
  } while (sp > 0);
  return false;
}

template <unsigned TM>
VGMI_DEV bool level_locate(GeomView const &g, NavVol const &M, int excl, V3 const &lp, int &pvol, V3 &dlp, Stack stk)
{
  if constexpr (kMixed<TM>) return level_locate_mixed<TM>(g, M, excl, lp, pvol, dlp, stk);
  const NavNode *nodes = g.bnodes + M.nodeBase;
  const NavSlot *slots = g.slots + M.slotBase;
  int sp    = 1;
  stk.at(0) = 0;
  do {
    const uint32_t id = stk.at(--sp);
    NavNode const &N  = nodes[id];
    const int nc      = N.nchild;
    if (nc >= 0) {
      for (int i = 0; i < nc; ++i) {
        NavSlot const &box = slots[N.offset + i];
        if (box_contains(box, lp)) {
          const int pv = box.pvol;
          if (pv == excl) continue;
          NavVol const &P = g.vols[pv];
          const V3 loc    = to_local(P.t, P.r, lp);
          if (contains<TM>(g.shapes[box.shape], loc)) {
            pvol = pv;
            dlp  = loc;
            return true;
          }
        }
      }
    } else {
      const uint32_t l = 2 * id + 1, r = 2 * id + 2;
      if (box_contains(nodes[l], lp)) stk.at(sp++) = l;
      if (box_contains(nodes[r], lp)) stk.at(sp++) = r;
    }
  } while (sp > 0);
  return false;
}

// BVH::ComputeSafety
// Mixed precision: a box whose FP32 safety lower bound is not below the current safety is
// skipped (the FP64 walk skips it too); every other box test, the sibling order and the
// shape safeties are FP64, so the result is the FP64 walk's.
template <unsigned TM>
VGMI_DEV double safety_to_daughters_mixed(GeomView const &g, NavVol const &M, V3 const &lp, double safety, Stack stk)
{
  const F3 pf              = to_f3(lp);
  const NavNode *nodes     = g.bnodes + M.nodeBase;
  const NavSlot *slots     = g.slots + M.slotBase;
  const NavNode32 *nodes32 = g.bnodes32 + M.nodeBase;
  const NavSlot32 *slots32 = g.slots32 + M.slotBase;
  int sp    = 1;
  stk.at(0) = 0;
  do {
    const uint32_t id = stk.at(--sp);
    const int nc      = nodes32[id].nchild;
    if (nc >= 0) {
      const int off = nodes32[id].offset;
      for (int i = 0; i < nc; ++i) {
        if (!(box_safety32(slots32[off + i], pf) < safety)) continue;
        NavSlot const &box = slots[off + i];
        if (box_safety(box, lp) < safety) {
          NavVol const &P = g.vols[box.pvol];
          const double s  = safety_in<TM>(g.shapes[box.shape], to_local(P.t, P.r, lp));
          if (s < safety) safety = s;
        }
      }
    } else {
      const uint32_t l = 2 * id + 1, r = 2 * id + 2;
      const bool cl = box_safety32(nodes32[l], pf) < safety, cr = box_safety32(nodes32[r], pf) < safety;
      const double sl = cl ? box_safety(nodes[l], lp) : kInfLength, sr = cr ? box_safety(nodes[r], lp) : kInfLength;
      const bool tl = sl < safety, tr = sr < safety;
      if (sr < sl) {
        if (tr) stk.at(sp++) = r;
        if (tl) stk.at(sp++) = l;
      } else {
        if (tl) stk.at(sp++) = l;
        if (tr) stk.at(sp++) = r;
      }
    }
  } while (sp > 0);
  return safety;
}

template <unsigned TM>
VGMI_DEV double safety_to_daughters(GeomView const &g, NavVol const &M, V3 const &lp, double safety, Stack stk)
{
  if constexpr (kMixed<TM>) return safety_to_daughters_mixed<TM>(g, M, lp, safety, stk);
  const NavNode *nodes = g.bnodes + M.nodeBase;
  const NavSlot *slots = g.slots + M.slotBase;
  int sp    = 1;
  stk.at(0) = 0;
  do {
    const uint32_t id = stk.at(--sp);
    NavNode const &N  = nodes[id];
    const int nc      = N.nchild;
    if (nc >= 0) {
      for (int i = 0; i < nc; ++i) {
        NavSlot const &box = slots[N.offset + i];
        if (box_safety(box, lp) < safety) {
          NavVol const &P = g.vols[box.pvol];
          const double s  = safety_in<TM>(g.shapes[box.shape], to_local(P.t, P.r, lp));
          if (s < safety) safety = s;
        }
      }
    } else {
      const uint32_t l = 2 * id + 1, r = 2 * id + 2;
      const double sl = box_safety(nodes[l], lp), sr = box_safety(nodes[r], lp);
      const bool tl = sl < safety, tr = sr < safety;
      if (sr < sl) {
        if (tr) stk.at(sp++) = r;
        if (tl) stk.at(sp++) = l;
      } else {
        if (tl) stk.at(sp++) = l;
        if (tr) stk.at(sp++) = r;
      }
    }
  } while (sp > 0);
  return safety;
}

// GlobalLocator::LocateGlobalPoint (top = false) / LocateGlobalPointExclVolume (excl >= 0)
template <unsigned TM>
VGMI_DEV uint32_t locate_down(GeomView const &g, int pvol, int excl, V3 p, uint32_t path, Stack stk)
{
  path              = push(g, path, pvol);
  const NavVol *M   = &g.vols[pvol];
  bool deeper       = true;
  while (M->ndaughters > 0 && deeper) {
    int next = -1;
    V3 dlp;
    deeper = level_locate<TM>(g, *M, excl, p, next, dlp, stk);
    if (deeper) {
      path = push(g, path, next);
      M    = &g.vols[next];
      p    = dlp;
    }
  }
  return path;
}

template <unsigned TM>
VGMI_DEV uint32_t locate_global(GeomView const &g, V3 const &p, Stack stk)
{
  if (!contains<TM>(shape_of_pvol(g, g.worldPvol), p)) return 0;
  return locate_down<TM>(g, g.worldPvol, -1, p, 0, stk);
}

// GlobalLocator::RelocatePointFromPathForceDifferent
template <unsigned TM>
VGMI_DEV uint32_t relocate_force_different(GeomView const &g, V3 const &lp, uint32_t path, Stack stk)
{
  int mother      = top(g, path);
  const int entry = mother;
  if (mother < 0) return path;
  V3 tmp = lp;
  while (mother >= 0) {
    NavVol const &P = g.vols[mother];
    if (mother == entry || !contains<TM>(g.shapes[P.shape], tmp)) {
      path   = pop(g, path);
      tmp    = to_master(P.t, P.r, tmp);
      mother = top(g, path);
    } else {
      break;
    }
  }
  if (mother >= 0) {
    path = pop(g, path);
    return locate_down<TM>(g, mother, entry, tmp, path, stk);
  }
  return path;
}

// BVHSafetyEstimator::ComputeSafetyForLocalPoint (VecGeom/navigation/BVHSafetyEstimator.h:38-45)
template <unsigned TM>
VGMI_DEV double compute_safety_local(GeomView const &g, V3 const &lp, NavVol const &M, Stack stk)
{
  double s = safety_out<TM>(g.shapes[M.shape], lp);
  if (s > 0.0 && M.ndaughters > 0) s = safety_to_daughters<TM>(g, M, lp, s, stk);
  return s;
}

// VSafetyEstimatorHelper<BVHSafetyEstimator>::ComputeSafety
template <unsigned TM>
VGMI_DEV double compute_safety(GeomView const &g, V3 const &gp, uint32_t nav, Stack stk)
{
  Mat m;
  top_matrix(g, nav, m);
  const V3 lp = to_local(m.t, m.r, gp);
  return compute_safety_local<TM>(g, lp, g.vols[top(g, nav)], stk);
}

struct State {
  uint32_t nav, last;
  bool boundary;
};

// Margin of the daughter-budget test: covers the rounding of the bounds and of the point
// updates over many steps, and the tolerance-level undershoot of DistanceToIn.
constexpr double kBudgetMargin = 1e-6;

// VNavigatorHelper<BVHNavigator>::ComputeStepAndPropagatedState / ComputeStep.
// `budget` carries, across the steps of one track, a lower bound of the distance from the
// current point to every daughter of the current volume (negative: unknown). While it exceeds
// the step limit and the mother distance is finite, no daughter can be hit within the limit,
// so the daughter walk cannot change the outcome of the step and is skipped; the walk itself
// refreshes the bound. Every change of state invalidates it.
// With `safety` non-null: VNavigatorHelper::ComputeStepAndSafetyAndPropagatedState
// (VecGeom/navigation/VNavigator.h:1014-1064), the isotropic safety of the local start point
// first (0 when !calc_safety), then the same step.
template <unsigned TM>
VGMI_DEV double compute_step(GeomView const &g, V3 const &gp, V3 const &gd, double limit, State const &in,
                             State &out, bool relocate, Stack stk, double &budget, double *safety = nullptr,
                             bool calc_safety = false)
{
  Mat m;
  top_matrix(g, in.nav, m);
  const V3 lp       = to_local(m.t, m.r, gp);
  const V3 ld       = rotate(m.r, gd);
  const int pvol    = top(g, in.nav);
  NavVol const &M   = g.vols[pvol];
  if (safety) *safety = calc_safety ? compute_safety_local<TM>(g, lp, M, stk) : 0.;
  double step       = dist_out<TM>(g.shapes[M.shape], lp, ld, limit);
  if (step < 0.) step = 0.;
  int hit = -1;
  if (M.ndaughters == 0) {
    budget = kInfLength;
  } else if (!(step < kInfLength && budget > limit + kBudgetMargin)) {
    check_daughters<TM>(g, M, lp, ld, step, top(g, in.last), hit, step < kInfLength ? limit : kInfLength, budget,
                        stk);
  }
  out = in;
  if (step == kInfLength && limit > 0.) {
    out.boundary = true;
    out.nav      = pop(g, out.nav);
    budget       = -1.;
    return kTolerance;
  }
  if (step > limit) {
    out.boundary = false;
    budget -= limit;
    return limit;
  }
  budget       = -1.;
  out.boundary = true;
  out.last     = out.nav;
  if (hit >= 0) out.nav = push(g, out.nav, hit);
  if (step < 0.) step = 0.;
  if (!relocate) return step;
  const V3 after = lp + (step + 1E-6) * ld;
  if (top(g, out.nav) == pvol) {
    out.nav = relocate_force_different<TM>(g, after, out.nav, stk);
  } else {
    NavVol const &N = g.vols[hit];
    out.nav         = pop(g, out.nav);
    out.nav         = locate_down<TM>(g, hit, -1, to_local(N.t, N.r, after), out.nav, stk);
  }
  return step;
}

} // namespace nav
} // namespace vgmi
