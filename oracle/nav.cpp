// Oracle navigation (see nav.h for the reference files followed). TEST INFRASTRUCTURE ONLY.
#include "nav.h"
#include <algorithm>
#include <cstring>
#include <numeric>
#include <stdexcept>

namespace vgo {

// ------------------------------------------------------------------ Transform
// source/Transformation3D.cpp:118-126, :174-185
void Transform::SetProperties()
{
  hasT = std::fabs(t[0]) > kTolerance || std::fabs(t[1]) > kTolerance || std::fabs(t[2]) > kTolerance;
  int code = 0;
  for (int i = 0; i < 9; ++i) code |= (1 << i) * (std::fabs(r[i]) > kTolerance);
  if (code == 0x111 && r[0] == 1. && r[4] == 1. && r[8] == 1.) code = 0x200;
  hasR  = code != 0x200;
  ident = !hasT && !hasR;
}

// Transformation3D.h:734-758 (general case: translate, then rotate with the row-major matrix)
V3 Transform::Transform3(V3 const &m) const
{
  V3 tmp(m.x - t[0], m.y - t[1], m.z - t[2]);
  return TransformDirection(tmp);
}

// Transformation3D.h:661-669 DoRotation general case
V3 Transform::TransformDirection(V3 const &m) const
{
  V3 l;
  l.x = m.x * r[0];
  l.y = m.x * r[1];
  l.z = m.x * r[2];
  l.x += m.y * r[3];
  l.y += m.y * r[4];
  l.z += m.y * r[5];
  l.x += m.z * r[6];
  l.y += m.z * r[7];
  l.z += m.z * r[8];
  return l;
}

// Transformation3D.h:808-823 InverseTransformKernel<true> (InverseTransformDirection, :858-864)
V3 Transform::InverseTransformDirection(V3 const &l) const
{
  V3 m;
  m.x = l.x * r[0];
  m.x += l.y * r[1];
  m.x += l.z * r[2];
  m.y = l.x * r[3];
  m.y += l.y * r[4];
  m.y += l.z * r[5];
  m.z = l.x * r[6];
  m.z += l.y * r[7];
  m.z += l.z * r[8];
  return m;
}

// Transformation3D.h:800-830 InverseTransformKernel<false>
V3 Transform::InverseTransform(V3 const &l) const
{
  V3 m;
  m.x = t[0];
  m.x += l.x * r[0];
  m.x += l.y * r[1];
  m.x += l.z * r[2];
  m.y = t[1];
  m.y += l.x * r[3];
  m.y += l.y * r[4];
  m.y += l.z * r[5];
  m.z = t[2];
  m.z += l.x * r[6];
  m.z += l.y * r[7];
  m.z += l.z * r[8];
  return m;
}

// Transformation3D.h:876-950
void Transform::MultiplyFromRight(Transform const &rhs)
{
  if (rhs.ident) return;
  ident = false;
  if (rhs.hasT) {
    hasT = true;
    for (int i = 0; i < 3; ++i) {
      t[i] += r[3 * i + 0] * rhs.t[0];
      t[i] += r[3 * i + 1] * rhs.t[1];
      t[i] += r[3 * i + 2] * rhs.t[2];
    }
  }
  if (rhs.hasR) {
    hasR = true;
    for (int row = 0; row < 3; ++row) {
      double x = r[3 * row], y = r[3 * row + 1], z = r[3 * row + 2];
      for (int c = 0; c < 3; ++c) r[3 * row + c] = x * rhs.r[c];
      for (int c = 0; c < 3; ++c) r[3 * row + c] += y * rhs.r[3 + c];
      for (int c = 0; c < 3; ++c) r[3 * row + c] += z * rhs.r[6 + c];
    }
  }
}

// ------------------------------------------------------------------ AABB
// VecGeom/base/AABB.h:136-154
void AABB::Intersect(V3 const &p, V3 const &invdir, double &tmin, double &tmax) const
{
  V3 t0 = (lo - p) * invdir;
  V3 t1 = (hi - p) * invdir;
  for (int i = 0; i < 3; ++i)
    if (t0[i] > t1[i]) std::swap(t0[i], t1[i]);
  tmin = t0.Max();
  tmax = t1.Min() * (1. + 2 * kEpsilon);
}

AABB AABB::Union(AABB const &a, AABB const &b)
{
  AABB u;
  u.lo = V3(std::min(a.lo.x, b.lo.x), std::min(a.lo.y, b.lo.y), std::min(a.lo.z, b.lo.z));
  u.hi = V3(std::max(a.hi.x, b.hi.x), std::max(a.hi.y, b.hi.y), std::max(a.hi.z, b.hi.z));
  return u;
}

// ------------------------------------------------------------------ BVH build (source/BVH.cpp)
namespace {
bool less3D(V3 const &l, V3 const &r, int ax)
{
  int a1 = (ax + 1) % 3, a2 = (ax + 2) % 3;
  return l[ax] < r[ax] || (l[ax] == r[ax] && (l[a1] < r[a1] || (l[a1] == r[a1] && l[a2] < r[a2])));
}

std::vector<std::pair<double, double>> sweep(AABB const *boxes, int const *begin, int const *end)
{
  std::vector<std::pair<double, double>> areas(end - begin, {0., 0.});
  AABB box = boxes[*begin];
  for (auto it = begin + 1; it < end; ++it) {
    areas[it - begin].first = box.SurfaceArea();
    box                     = AABB::Union(box, boxes[*it]);
  }
  AABB box2 = boxes[*(end - 1)];
  for (auto it = end - 1; it >= begin; --it) {
    box2                     = AABB::Union(box2, boxes[*it]);
    areas[it - begin].second = box2.SurfaceArea();
  }
  return areas;
}

// source/BVH.cpp:242-329 (surface-area heuristic with the integer balance term kept as written)
int *sah(AABB const *boxes, int *begin, int *end)
{
  int bestAxis       = -1;
  double bestMetric  = double(end - begin);
  int bestObj        = -1;
  const long nObj    = end - begin;
  int sortAxis       = 0;
  auto sorter        = [boxes, &sortAxis](int a, int b) {
    V3 ca = boxes[a].Center(), cb = boxes[b].Center();
    constexpr double shift = 0.01;
    return less3D(ca + shift * (ca - boxes[a].lo), cb + shift * (cb - boxes[b].lo), sortAxis);
  };
  for (int axis = 0; axis <= 2; ++axis) {
    sortAxis = axis;
    std::sort(begin, end, sorter);
    auto sw        = sweep(boxes, begin, end);
    double totArea = sw.front().second;
    for (int *s = begin; s < end; ++s) {
      double left   = sw[s - begin].first / totArea;
      double right  = sw[s - begin].second / totArea;
      double metric = left * double(s - begin) + right * double(end - s) +
                      0.1 * double(std::labs(nObj / 2 - (s - begin) / nObj));
      if (metric < bestMetric) {
        bestMetric = metric;
        bestAxis   = axis;
        bestObj    = *s;
      }
    }
  }
  if (bestAxis == -1) return end;
  sortAxis = bestAxis;
  return std::partition(begin, end, [sorter, bestObj](int i) { return sorter(i, bestObj); });
}
} // namespace

void Geometry::ComputeNodes(BVHo &b, unsigned id, int *first, int *last, unsigned nodes)
{
  if (id >= nodes) return;
  b.nChild[id] = int(last - first);
  b.offset[id] = int(first - b.primId.data());
  if (first == last) return;
  b.nodes[id] = b.aabbs[*first];
  for (int *it = first + 1; it != last; ++it) b.nodes[id] = AABB::Union(b.nodes[id], b.aabbs[*it]);
  if (first + 1 == last) return;
  int *pivot = sah(b.aabbs.data(), first, last);
  ComputeNodes(b, 2 * id + 1, first, pivot, nodes);
  ComputeNodes(b, 2 * id + 2, pivot, last, nodes);
}

// source/BVH.cpp:48-87 with ABBoxes from source/ABBoxManager.cpp:109-143
void Geometry::BuildBVH(LVol &lv)
{
  auto b    = std::make_unique<BVHo>();
  const int n = int(lv.daughters.size());
  b->aabbs.resize(n);
  for (int i = 0; i < n; ++i) {
    PVol const &pv = pvols[lv.daughters[i]];
    V3 lo, hi;
    shapes[lvs[pv.lv].shape]->Extent(lo, hi);
    V3 delta = hi - lo;
    double mn[3] = {kInfLength, kInfLength, kInfLength}, mx[3] = {-kInfLength, -kInfLength, -kInfLength};
    for (int x = 0; x <= 1; ++x)
      for (int y = 0; y <= 1; ++y)
        for (int z = 0; z <= 1; ++z) {
          V3 c(lo.x + x * delta.x, lo.y + y * delta.y, lo.z + z * delta.z);
          V3 tc = pv.tr.InverseTransform(c);
          for (int k = 0; k < 3; ++k) {
            mn[k] = std::min(mn[k], tc[k]);
            mx[k] = std::max(mx[k], tc[k]);
          }
        }
    b->aabbs[i].lo = V3(mn[0] - 1E-3, mn[1] - 1E-3, mn[2] - 1E-3);
    b->aabbs[i].hi = V3(mx[0] + 1E-3, mx[1] + 1E-3, mx[2] + 1E-3);
  }
  b->primId.resize(n);
  std::iota(b->primId.begin(), b->primId.end(), 0);
  b->depth       = std::min(std::max(0, n / 2 > 0 ? (int)std::log2(n / 2) : 0), 32);
  unsigned nodes = (2u << b->depth) - 1;
  b->nChild.assign(nodes, 0);
  b->offset.assign(nodes, -1);
  b->nodes.assign(nodes, AABB());
  ComputeNodes(*b, 0, b->primId.data(), b->primId.data() + n, nodes);
  for (unsigned id = 0; id < nodes / 2; ++id)
    if (b->nChild[id] > 8 && b->nChild[id] == b->nChild[2 * id + 1] + b->nChild[2 * id + 2]) b->nChild[id] = -1;
  lv.bvh = std::move(b);
}

// ------------------------------------------------------------------ NavIndex table
// source/NavIndexTable.cpp:9-99; visit order VecGeom/management/NavIndexTable.h:152-165
void Geometry::Close()
{
  for (auto &lv : lvs)
    if (!lv.daughters.empty()) BuildBVH(lv);
  navind.assign(1, 0u);
  uint32_t current = 1;
  std::vector<int> path;
  maxDepth = 0;
  struct Rec {
    static uint32_t visit(Geometry &g, std::vector<int> &path, uint32_t &current, int pvol, int level,
                          uint32_t mother, int dind)
    {
      path.push_back(pvol);
      g.maxDepth            = std::max(g.maxDepth, level + 1);
      LVol const &lv        = g.lvs[g.pvols[pvol].lv];
      const uint32_t nd     = uint32_t(lv.daughters.size());
      const uint32_t dIdx   = 3 + nd + ((nd + 1) & 1);
      const bool pad        = ((current + dIdx) * 4) % 8 != 0;
      const uint32_t newMom = current;
      size_t need           = current + dIdx + (pad ? 1 : 0) + 24;
      if (g.navind.size() < need) g.navind.resize(need, 0u);
      g.navind[current] = mother;
      if (mother > 0) g.navind[mother + 3 + dind] = current;
      g.navind[current + 1] = uint32_t(path[level]);
      Transform mat;
      for (int i = 1; i <= level; ++i) mat.MultiplyFromRight(g.pvols[path[i]].tr);
      uint32_t hasm          = 0x04 + 0x02 * uint32_t(mat.hasT) + uint32_t(mat.hasR);
      g.navind[current + 2] = uint32_t(level & 0xff) | (hasm << 8) | (nd << 16);
      for (uint32_t i = 0; i < nd; ++i) g.navind[current + 3 + i] = 0;
      current += dIdx;
      if (pad) current++;
      double m[12];
      for (int i = 0; i < 3; ++i) m[i] = mat.t[i];
      for (int i = 0; i < 9; ++i) m[3 + i] = mat.r[i];
      std::memcpy(&g.navind[current], m, sizeof(m));
      current += 24;
      for (uint32_t i = 0; i < nd; ++i) visit(g, path, current, lv.daughters[i], level + 1, newMom, int(i));
      path.pop_back();
      return newMom;
    }
  };
  Rec::visit(*this, path, current, worldPvol, 0, 0, 0);
  navind.resize(current);
}

// NavStateIndex::TopMatrixImpl (all levels carry a cached matrix)
void Geometry::TopMatrix(uint32_t nav, Transform &m) const
{
  m = Transform();
  if (nav == 0) return;
  uint32_t w    = navind[nav + 2];
  uint32_t hasm = (w >> 8) & 0xff;
  if ((hasm & 0x03) == 0) return;
  bool hasT   = (hasm & 0x02) > 0, hasR = (hasm & 0x01) > 0;
  uint32_t nd = w >> 16;
  uint32_t ti = nav + 3 + nd + ((nd + 1) & 1);
  if ((ti * 4) % 8 != 0) ti += 1;
  double a[12];
  std::memcpy(a, &navind[ti], sizeof(a));
  if (hasT)
    for (int i = 0; i < 3; ++i) m.t[i] = a[i];
  if (hasR)
    for (int i = 0; i < 9; ++i) m.r[i] = a[3 + i];
  m.hasT  = hasT;
  m.hasR  = hasR;
  m.ident = !hasT && !hasR;
}

// ------------------------------------------------------------------ traversal
// VecGeom/base/BVH.h:96-146 CheckDaughterIntersections; NewSimpleNavigator loop when useBVH is false
void Geometry::CheckDaughters(LVol const &lv, V3 const &lp, V3 const &ld, double &step, int last, int &hit) const
{
  if (lv.daughters.empty()) return;
  auto test = [&](int prim) {
    int pv      = lv.daughters[prim];
    PVol const &P = pvols[pv];
    double dist = ShapeOf(pv).DistanceToIn(P.tr.Transform3(lp), P.tr.TransformDirection(ld), step);
    return std::make_pair(pv, dist);
  };
  if (!useBVH) {
    for (size_t d = 0; d < lv.daughters.size(); ++d) {
      auto r     = test(int(d));
      bool valid = (r.second < step && !std::isinf(r.second)) && !(r.second <= 0. && last == r.first);
      if (valid) hit = r.first, step = r.second;
    }
    return;
  }
  BVHo const &b = *lv.bvh;
  unsigned stack[32], *ptr = &stack[1];
  stack[0] = 0;
  V3 invdir(1.0 / NonZero(ld.x), 1.0 / NonZero(ld.y), 1.0 / NonZero(ld.z));
  do {
    unsigned id = *--ptr;
    if (b.nChild[id] >= 0) {
      for (int i = 0; i < b.nChild[id]; ++i) {
        int prim = b.primId[b.offset[id] + i];
        double tmin, tmax;
        b.aabbs[prim].Intersect(lp, invdir, tmin, tmax);
        if (tmin <= tmax && tmax >= 0.0 && tmin < step) {
          auto r = test(prim);
          if (r.second < step && !(r.second <= 0.0 && r.first == last)) step = r.second, hit = r.first;
        }
      }
    } else {
      unsigned L = 2 * id + 1, R = 2 * id + 2;
      double tminL = kInfLength, tmaxL = -kInfLength, tminR = kInfLength, tmaxR = -kInfLength;
      b.nodes[L].Intersect(lp, invdir, tminL, tmaxL);
      b.nodes[R].Intersect(lp, invdir, tminR, tmaxR);
      bool tL = tminL <= tmaxL && tmaxL >= 0.0 && tminL < step;
      bool tR = tminR <= tmaxR && tmaxR >= 0.0 && tminR < step;
      if (tminR < tminL) {
        if (tR) *ptr++ = R;
        if (tL) *ptr++ = L;
      } else {
        if (tL) *ptr++ = L;
        if (tR) *ptr++ = R;
      }
    }
  } while (ptr > stack);
}

// VecGeom/base/BVH.h:255-285 LevelLocate (exclvol form); SimpleLevelLocator loop when useBVH is false
bool Geometry::LevelLocate(LVol const &lv, int excl, V3 const &lp, int &pvol, V3 &dlp) const
{
  auto contains = [&](int pv, V3 &loc) {
    loc = pvols[pv].tr.Transform3(lp);
    return ShapeOf(pv).Contains(loc);
  };
  if (!useBVH) {
    for (int pv : lv.daughters) {
      if (pv == excl) continue;
      V3 loc;
      if (contains(pv, loc)) return pvol = pv, dlp = loc, true;
    }
    return false;
  }
  BVHo const &b = *lv.bvh;
  unsigned stack[32], *ptr = &stack[1];
  stack[0] = 0;
  do {
    unsigned id = *--ptr;
    if (b.nChild[id] >= 0) {
      for (int i = 0; i < b.nChild[id]; ++i) {
        int prim = b.primId[b.offset[id] + i];
        if (b.aabbs[prim].Contains(lp)) {
          int pv = lv.daughters[prim];
          V3 loc;
          if (pv != excl && contains(pv, loc)) return pvol = pv, dlp = loc, true;
        }
      }
    } else {
      unsigned L = 2 * id + 1, R = 2 * id + 2;
      if (b.nodes[L].Contains(lp)) *ptr++ = L;
      if (b.nodes[R].Contains(lp)) *ptr++ = R;
    }
  } while (ptr > stack);
  return false;
}

// VecGeom/base/BVH.h:172-209 ComputeSafety
double Geometry::SafetyToDaughters(LVol const &lv, V3 const &lp, double safety) const
{
  auto sin = [&](int prim) {
    int pv = lv.daughters[prim];
    return ShapeOf(pv).SafetyToIn(pvols[pv].tr.Transform3(lp));
  };
  if (!useBVH) {
    for (size_t d = 0; d < lv.daughters.size(); ++d) safety = std::min(safety, sin(int(d)));
    return safety;
  }
  BVHo const &b = *lv.bvh;
  unsigned stack[32], *ptr = &stack[1];
  stack[0] = 0;
  do {
    unsigned id = *--ptr;
    if (b.nChild[id] >= 0) {
      for (int i = 0; i < b.nChild[id]; ++i) {
        int prim = b.primId[b.offset[id] + i];
        if (b.aabbs[prim].Safety(lp) < safety) {
          double d = sin(prim);
          if (d < safety) safety = d;
        }
      }
    } else {
      unsigned L = 2 * id + 1, R = 2 * id + 2;
      double sL = b.nodes[L].Safety(lp), sR = b.nodes[R].Safety(lp);
      bool tL = sL < safety, tR = sR < safety;
      if (sR < sL) {
        if (tR) *ptr++ = R;
        if (tL) *ptr++ = L;
      } else {
        if (tL) *ptr++ = L;
        if (tR) *ptr++ = R;
      }
    }
  } while (ptr > stack);
  return safety;
}

// ------------------------------------------------------------------ GlobalLocator
// GlobalLocator.h:35-82 LocateGlobalPoint (level locator attached to every volume with daughters)
uint32_t Geometry::LocateGlobalPoint(int pvol, V3 const &point, uint32_t path, bool top) const
{
  int cand = pvol;
  V3 cur   = point;
  if (top) cand = ShapeOf(pvol).Contains(point) ? pvol : -1;
  if (cand < 0) return path;
  path           = Push(path, cand);
  LVol const *lv = &lvs[pvols[cand].lv];
  bool deeper    = true;
  while (!lv->daughters.empty() && deeper) {
    int next = -1;
    V3 dlp;
    deeper = LevelLocate(*lv, -1, cur, next, dlp);
    if (deeper) {
      path = Push(path, next);
      lv   = &lvs[pvols[next].lv];
      cur  = dlp;
    }
  }
  return path;
}

// GlobalLocator.h:88-135 LocateGlobalPointExclVolume (top = false)
uint32_t Geometry::LocateExcl(int pvol, int excl, V3 const &point, uint32_t path) const
{
  int cand       = pvol;
  V3 cur         = point;
  path           = Push(path, cand);
  LVol const *lv = &lvs[pvols[cand].lv];
  bool deeper    = true;
  while (!lv->daughters.empty() && deeper) {
    deeper = false;
    int next = -1;
    V3 dlp;
    deeper = LevelLocate(*lv, excl, cur, next, dlp);
    if (deeper) {
      cand = next;
      lv   = &lvs[pvols[cand].lv];
      cur  = dlp;
      path = Push(path, cand);
    }
  }
  return path;
}

// GlobalLocator.h:171-200 RelocatePointFromPathForceDifferent
uint32_t Geometry::RelocateForceDifferent(V3 const &lp, uint32_t path) const
{
  int mother = Top(path);
  int entry  = mother;
  if (mother < 0) return path;
  V3 tmp = lp;
  while (mother >= 0) {
    if (mother == entry || !ShapeOf(mother).Contains(tmp)) {
      path   = Pop(path);
      tmp    = pvols[mother].tr.InverseTransform(tmp);
      mother = Top(path);
    } else
      break;
  }
  if (mother >= 0) {
    path = Pop(path);
    return LocateExcl(mother, entry, tmp, path);
  }
  return path;
}

// ------------------------------------------------------------------ navigator
// VNavigator.h:436-492 ComputeStepAndPropagatedState (relocate = true) / :494-540 ComputeStep,
// with PrepareOutState :192-235, TreatDistanceToMother :241-249, MovePointAfterBoundary :168-173,
// Relocate :1070-1091
double Geometry::ComputeStep(V3 const &gp, V3 const &gd, double limit, NavState const &in, NavState &out,
                             bool relocate) const
{
  Transform m;
  TopMatrix(in.nav, m);
  V3 lp      = m.Transform3(gp);
  V3 ld      = m.TransformDirection(gd);
  int pvol   = Top(in.nav);
  LVol const &lv = lvs[pvols[pvol].lv];
  double step = shapes[lv.shape]->DistanceToOut(lp, ld, limit);
  if (step < 0.) step = 0.;
  int hit  = -1;
  int last = Top(in.last);
  CheckDaughters(lv, lp, ld, step, last, hit);
  // PrepareOutState
  out = in;
  if (step == kInfLength && limit > 0.) {
    out.boundary = true;
    out.nav      = Pop(out.nav);
    return kTolerance;
  }
  if (step > limit) {
    out.boundary = false;
    return limit;
  }
  out.boundary = true;
  out.last     = out.nav;
  if (hit >= 0) out.nav = Push(out.nav, hit);
  if (step < 0.) step = 0.;
  if (!relocate) return step;
  V3 after = lp + (step + 1E-6) * ld;
  if (Top(out.nav) == Top(in.nav)) {
    out.nav = RelocateForceDifferent(after, out.nav);
  } else {
    int next = Top(out.nav);
    out.nav  = Pop(out.nav);
    out.nav  = LocateGlobalPoint(next, pvols[next].tr.Transform3(after), out.nav, false);
  }
  return step;
}

// VSafetyEstimator.h:98-107 + BVHSafetyEstimator::ComputeSafetyForLocalPoint (BVHSafetyEstimator.h:38-46)
double Geometry::ComputeSafety(V3 const &gp, NavState const &in) const
{
  Transform m;
  TopMatrix(in.nav, m);
  V3 lp          = m.Transform3(gp);
  int pvol       = Top(in.nav);
  LVol const &lv = lvs[pvols[pvol].lv];
  double safety  = shapes[lv.shape]->SafetyToOut(lp);
  if (safety > 0.0 && !lv.daughters.empty()) safety = SafetyToDaughters(lv, lp, safety);
  return safety;
}

} // namespace vgo
