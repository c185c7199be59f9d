// Oracle navigation C ABI for the tests and the CPU-baseline leg of the benchmark.
// TEST INFRASTRUCTURE ONLY (see common.h).
//
// Geometry description (shared with the product builder):
//   shapes     : type[i], parameters params[off[i] .. off[i+1])
//   logicals   : shape index per logical volume
//   placements : mother logical, daughter logical, 12 doubles (tx ty tz r0..r8) each;
//                placed-volume id k for placement k, the world placed volume gets id nplaced
#include "nav.h"
#include <thread>

using namespace vgo;

namespace {
template <class F>
void parallel_for(int64_t n, int threads, F f)
{
  if (threads <= 1) {
    f(int64_t(0), n);
    return;
  }
  std::vector<std::thread> pool;
  int64_t chunk = (n + threads - 1) / threads;
  for (int t = 0; t < threads; ++t) {
    int64_t b = t * chunk, e = std::min<int64_t>(n, b + chunk);
    if (b >= e) break;
    pool.emplace_back(f, b, e);
  }
  for (auto &th : pool) th.join();
}
} // namespace

extern "C" {

void *vgo_geom_new(int nshapes, const int *stype, const int *soff, const double *sparams, int nlv, const int *lvshape,
                   int npl, const int *plmother, const int *pllv, const double *pltr, int worldlv, int use_bvh)
{
  auto *g   = new Geometry();
  g->useBVH = use_bvh != 0;
  for (int i = 0; i < nshapes; ++i) {
    Shape *s = MakeShape(stype[i], sparams + soff[i], soff[i + 1] - soff[i]);
    if (!s) {
      delete g;
      return nullptr;
    }
    g->shapes.emplace_back(s);
  }
  g->lvs.resize(nlv);
  for (int i = 0; i < nlv; ++i) g->lvs[i].shape = lvshape[i];
  g->pvols.resize(npl + 1);
  for (int k = 0; k < npl; ++k) {
    PVol &p = g->pvols[k];
    p.lv    = pllv[k];
    for (int i = 0; i < 3; ++i) p.tr.t[i] = pltr[12 * k + i];
    for (int i = 0; i < 9; ++i) p.tr.r[i] = pltr[12 * k + 3 + i];
    p.tr.SetProperties();
    LVol &mom = g->lvs[plmother[k]];
    p.childId = int(mom.daughters.size());
    mom.daughters.push_back(k);
  }
  g->pvols[npl].lv = worldlv;
  g->worldPvol     = npl;
  g->Close();
  return g;
}

void vgo_geom_free(void *g) { delete (Geometry *)g; }

int64_t vgo_geom_navindex(void *g, uint32_t *out, int64_t cap)
{
  auto &t = ((Geometry *)g)->navind;
  if (out)
    for (int64_t i = 0; i < std::min<int64_t>(cap, int64_t(t.size())); ++i) out[i] = t[i];
  return int64_t(t.size());
}

int vgo_geom_maxdepth(void *g) { return ((Geometry *)g)->maxDepth; }

// BVH arrays of one logical volume: counts[0]=depth, counts[1]=nodes, counts[2]=n daughters.
int vgo_geom_bvh(void *gp, int lv, int *counts, int *nchild, int *offset, int *primid, double *nodes6, double *aabbs6)
{
  Geometry *g = (Geometry *)gp;
  if (lv < 0 || lv >= int(g->lvs.size()) || !g->lvs[lv].bvh) return -1;
  BVHo const &b = *g->lvs[lv].bvh;
  counts[0]     = b.depth;
  counts[1]     = int(b.nodes.size());
  counts[2]     = int(b.aabbs.size());
  if (nchild)
    for (size_t i = 0; i < b.nodes.size(); ++i) {
      nchild[i] = b.nChild[i];
      offset[i] = b.offset[i];
      for (int k = 0; k < 3; ++k) nodes6[6 * i + k] = b.nodes[i].lo[k], nodes6[6 * i + 3 + k] = b.nodes[i].hi[k];
    }
  if (primid)
    for (size_t i = 0; i < b.aabbs.size(); ++i) {
      primid[i] = b.primId[i];
      for (int k = 0; k < 3; ++k) aabbs6[6 * i + k] = b.aabbs[i].lo[k], aabbs6[6 * i + 3 + k] = b.aabbs[i].hi[k];
    }
  return 0;
}

int vgo_nav_locate(void *gp, const double *px, const double *py, const double *pz, uint32_t *nav_out, int64_t n,
                   int threads)
{
  Geometry *g = (Geometry *)gp;
  parallel_for(n, threads, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) nav_out[i] = g->LocateGlobalPoint(g->worldPvol, V3(px[i], py[i], pz[i]), 0, true);
  });
  return 0;
}

int vgo_nav_step(void *gp, const double *px, const double *py, const double *pz, const double *dx, const double *dy,
                 const double *dz, const double *limit, const uint32_t *nav_in, const uint32_t *last_in,
                 uint32_t *nav_out, uint32_t *last_out, uint8_t *boundary_out, double *step_out, int64_t n,
                 int relocate, int threads)
{
  Geometry *g = (Geometry *)gp;
  parallel_for(n, threads, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      NavState in, out;
      in.nav  = nav_in[i];
      in.last = last_in ? last_in[i] : 0;
      if (in.nav == 0) {
        nav_out[i] = 0, step_out[i] = 0;
        if (last_out) last_out[i] = in.last;
        if (boundary_out) boundary_out[i] = 0;
        continue;
      }
      step_out[i] = g->ComputeStep(V3(px[i], py[i], pz[i]), V3(dx[i], dy[i], dz[i]), limit[i], in, out, relocate);
      nav_out[i]  = out.nav;
      if (last_out) last_out[i] = out.last;
      if (boundary_out) boundary_out[i] = out.boundary;
    }
  });
  return 0;
}

int vgo_nav_safety(void *gp, const double *px, const double *py, const double *pz, const uint32_t *nav_in,
                   double *out, int64_t n, int threads)
{
  Geometry *g = (Geometry *)gp;
  parallel_for(n, threads, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      NavState in;
      in.nav = nav_in[i];
      out[i] = in.nav ? g->ComputeSafety(V3(px[i], py[i], pz[i]), in) : 0.;
    }
  });
  return 0;
}

// Transport loop: nsteps x (ComputeStepAndPropagatedState, p += step * d) per track; tracks whose
// state leaves the world stop. Updates p, nav, last in place; returns the number of steps taken.
int64_t vgo_nav_propagate(void *gp, double *px, double *py, double *pz, const double *dx, const double *dy,
                          const double *dz, double limit, uint32_t *nav, uint32_t *last, int64_t n, int nsteps,
                          int threads)
{
  Geometry *g = (Geometry *)gp;
  std::vector<int64_t> counts(std::max(1, threads), 0);
  int64_t chunk = (n + std::max(1, threads) - 1) / std::max(1, threads);
  parallel_for(n, threads, [&](int64_t b, int64_t e) {
    int64_t taken = 0;
    for (int64_t i = b; i < e; ++i) {
      NavState st;
      st.nav  = nav[i];
      st.last = last[i];
      V3 p(px[i], py[i], pz[i]), d(dx[i], dy[i], dz[i]);
      for (int s = 0; s < nsteps && st.nav != 0; ++s) {
        NavState out;
        double step = g->ComputeStep(p, d, limit, st, out, true);
        p           = p + step * d;
        st          = out;
        ++taken;
      }
      px[i] = p.x, py[i] = p.y, pz[i] = p.z;
      nav[i]  = st.nav;
      last[i] = st.last;
    }
    counts[chunk > 0 ? b / chunk : 0] = taken;
  });
  int64_t tot = 0;
  for (auto c : counts) tot += c;
  return tot;
}

// VLevelLocator::LevelLocate / LevelLocateExclVol (VecGeom/navigation/BVHLevelLocator.h:44-77) per
// point: lv[i] the logical volume, excl[i] the placed volume to skip (-1: none; excl may be null).
int vgo_nav_level_locate(void *gp, const int32_t *lv, const int32_t *excl, const double *px, const double *py,
                         const double *pz, int32_t *out, double *dx, double *dy, double *dz, int64_t n, int threads)
{
  Geometry *g = (Geometry *)gp;
  parallel_for(n, threads, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      const V3 lp(px[i], py[i], pz[i]);
      int pv = -1;
      V3 dlp = lp;
      const int l = lv[i];
      if (l >= 0 && l < int(g->lvs.size()) && !g->lvs[l].daughters.empty()) {
        if (!g->LevelLocate(g->lvs[l], excl ? excl[i] : -1, lp, pv, dlp)) pv = -1, dlp = lp;
      }
      out[i] = pv;
      if (dx) dx[i] = dlp.x, dy[i] = dlp.y, dz[i] = dlp.z;
    }
  });
  return 0;
}

// Placed-volume queries as CommonSpecializedVolImplHelper dispatches them
// (VecGeom/volumes/SpecializedPlacedVolImplHelper.h:98-179): Contains (0), Inside (1), DistanceToIn (2),
// SafetyToIn (4) and PlacedDistanceToOut (6) transform the point (Transformation3D::Transform) and the
// direction (TransformDirection); DistanceToOut (3) and SafetyToOut (5) take local coordinates
// (VecGeom/volumes/PlacedVolume.h:335-345, 408-416).
int vgo_placed_query(void *gp, int pvol, int op, const double *px, const double *py, const double *pz,
                     const double *dx, const double *dy, const double *dz, const double *step, void *out, int64_t n,
                     int threads)
{
  Geometry *g = (Geometry *)gp;
  if (pvol < 0 || pvol >= int(g->pvols.size()) || op < 0 || op > 6) return -1;
  Transform const &tr = g->pvols[pvol].tr;
  Shape const &s      = g->ShapeOf(pvol);
  const bool xf       = op != 3 && op != 5;
  parallel_for(n, threads, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      V3 p(px[i], py[i], pz[i]);
      if (xf) p = tr.Transform3(p);
      switch (op) {
      case 0: static_cast<uint8_t *>(out)[i] = s.Contains(p) ? 1 : 0; break;
      case 1: static_cast<int32_t *>(out)[i] = s.Inside(p); break;
      case 4: static_cast<double *>(out)[i] = s.SafetyToIn(p); break;
      case 5: static_cast<double *>(out)[i] = s.SafetyToOut(p); break;
      default: {
        V3 d(dx[i], dy[i], dz[i]);
        if (xf) d = tr.TransformDirection(d);
        const double sm = step ? step[i] : kInfLength;
        static_cast<double *>(out)[i] = op == 2 ? s.DistanceToIn(p, d, sm) : s.DistanceToOut(p, d, sm);
      }
      }
    }
  });
  return 0;
}

// VPlacedVolume::Normal (source/PlacedVolume.cpp:126-137): the point to the volume's frame, the
// unplaced Normal starting from a zero vector, the normal back to the mother's frame.
int vgo_placed_normal(void *gp, int pvol, const double *px, const double *py, const double *pz, double *nx,
                      double *ny, double *nz, uint8_t *valid, int64_t n)
{
  Geometry *g = (Geometry *)gp;
  if (pvol < 0 || pvol >= int(g->pvols.size())) return -1;
  Transform const &tr = g->pvols[pvol].tr;
  Shape const &s      = g->ShapeOf(pvol);
  for (int64_t i = 0; i < n; ++i) {
    const V3 lp = tr.Transform3(V3(px[i], py[i], pz[i]));
    V3 ln;
    valid[i]    = s.Normal(lp, ln);
    const V3 mn = tr.InverseTransformDirection(ln);
    nx[i] = mn.x, ny[i] = mn.y, nz[i] = mn.z;
  }
  return 0;
}

// ---- NavStateIndex <-> path interop. Level-major arrays: entry k of state i at [k * n + i].

// NavStateIndex::GetLevel / ValueAt / GetPathAsListOfIndices (VecGeom/navigation/NavStateIndex.h:194-211,
// 341-349; source/NavStateIndex.cpp:40-53): levels[i] = GetCurrentLevel (0 for the outside state),
// pvols[k] = the placed volume at level k, indices[k] = its position in its mother's daughter list
// (0 at level 0); -1 beyond the state's levels. Returns -1 when a state has more than maxl levels.
int vgo_nav_to_path(void *gp, const uint32_t *nav, int64_t n, int maxl, int32_t *levels, int32_t *pvols,
                    int32_t *indices)
{
  Geometry *g = (Geometry *)gp;
  for (int64_t i = 0; i < n; ++i) {
    uint32_t s   = nav[i];
    const int L  = g->Level(s);
    if (L + 1 > maxl) return -1;
    levels[i] = L + 1;
    for (int k = L + 1; k < maxl; ++k) {
      if (pvols) pvols[k * n + i] = -1;
      if (indices) indices[k * n + i] = -1;
    }
    for (int k = L; k >= 0; --k) {
      const int pv = g->Top(s);
      if (pvols) pvols[k * n + i] = pv;
      if (indices) indices[k * n + i] = k > 0 ? g->pvols[pv].childId : 0;
      s = g->Pop(s);
    }
  }
  return 0;
}

// kind 0: NavStateIndex::ResetPathFromListOfIndices (source/NavStateIndex.cpp:77-88): the world, then
// daughter `indices[k]` of the volume reached at each further level (the first entry is not read).
// kind 1: a NavStatePath's placed volumes, world first, pushed level by level (NavStateIndex::Push,
// NavStateIndex.h:282-287). A path the geometry does not hold (a daughter index beyond the volume's
// daughters, a placed volume that is not a daughter of the previous level's volume, a first entry
// other than the world) gives the outside state 0; so does levels[i] == 0.
int vgo_nav_from_path(void *gp, int kind, const int32_t *levels, const int32_t *path, int64_t n, int maxl,
                      uint32_t *nav)
{
  Geometry *g = (Geometry *)gp;
  const int np = int(g->pvols.size());
  for (int64_t i = 0; i < n; ++i) {
    const int L = levels[i];
    uint32_t s  = 0;
    bool ok     = L >= 1 && L <= maxl;
    for (int k = 0; ok && k < L; ++k) {
      const int v = path[k * n + i];
      if (k == 0) {
        ok = kind == 0 || v == g->worldPvol;
        s  = 1;
        continue;
      }
      int cid = v;
      if (kind == 1) {
        if (v < 0 || v >= np) {
          ok = false;
          break;
        }
        cid = g->pvols[v].childId;
      }
      if (cid < 0 || cid >= g->NDaughters(s)) {
        ok = false;
        break;
      }
      const uint32_t c = g->NavInd(s + 3 + uint32_t(cid));
      if (kind == 1 && g->Top(c) != v) {
        ok = false;
        break;
      }
      s = c;
    }
    nav[i] = ok ? s : 0u;
  }
  return 0;
}

// NavStateIndex::Distance (source/NavStateIndex.cpp:150-166): up moves from a to the deepest common
// level plus down moves to b. The outside state counts as having no level (-1).
int vgo_nav_distance(void *gp, const uint32_t *a, const uint32_t *b, int64_t n, int32_t *out)
{
  Geometry *g = (Geometry *)gp;
  auto at = [&](uint32_t s, int level) { // NavStateIndex::At (NavStateIndex.h:331-337)
    int up = g->Level(s) - level;
    while (s && up--) s = g->Pop(s);
    return s > 0 ? g->Top(s) : -1;
  };
  for (int64_t i = 0; i < n; ++i) {
    const int la = g->Level(a[i]), lb = g->Level(b[i]);
    const int maxlevel = std::min(la, lb);
    int common = -1;
    for (int k = 0; k < maxlevel + 1; ++k) {
      if (at(a[i], k) == at(b[i], k))
        common = k;
      else
        break;
    }
    out[i] = (la - common) + (lb - common);
  }
  return 0;
}

// NavStateIndex::GlobalToLocal(point, tolevel) (NavStateIndex.h:262-269, 368-377): the point in the
// frame of the state's ancestor at `tolevel` (tolevel < 0: the top volume), through the cached matrix
// of that ancestor. An ancestor level beyond the state's own level reaches the outside state, whose
// matrix is the identity (the reference's GetNavIndexImpl walks up past the world then).
int vgo_nav_global_to_local(void *gp, const uint32_t *nav, int tolevel, const double *px, const double *py,
                            const double *pz, double *lx, double *ly, double *lz, int64_t n)
{
  Geometry *g = (Geometry *)gp;
  for (int64_t i = 0; i < n; ++i) {
    uint32_t s = nav[i];
    if (tolevel >= 0) {
      int up = g->Level(s) - tolevel;
      if (up < 0) s = 0;
      while (s && up-- > 0) s = g->Pop(s);
    }
    Transform m;
    g->TopMatrix(s, m);
    const V3 l = m.Transform3(V3(px[i], py[i], pz[i]));
    lx[i] = l.x, ly[i] = l.y, lz[i] = l.z;
  }
  return 0;
}

} // extern "C"
