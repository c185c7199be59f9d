// Host geometry builder: geometry description -> flat tables (shape records,
// placed/logical volume tables, per-volume BVHs, NavIndex table), then upload.
//
// The table formats and the build algorithms follow the reference so that
// navigation states (NavIndex values) and BVH traversal order are identical:
//  - daughter AABBs: 8 extent corners mapped to the mother frame, +-1e-3 mm
//    (source/ABBoxManager.cpp:109-143)
//  - BVH: complete binary tree, depth min(log2(n/2), 32), surface-area-heuristic
//    split with the centroid ordering of source/BVH.cpp:217-329, nodes with more
//    than 8 children whose children partition them become internal (BVH.cpp:85-86)
//  - NavIndex: depth-first numbering, [mother, pvol, level|hasm|nd, daughters, pad]
//    followed by 12 cached doubles of the global matrix (source/NavIndexTable.cpp:9-99)
#include "geom_internal.h"
#include "shape_host.h"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

namespace vgmi {

namespace {

constexpr double kTol = 1e-9;

struct Mat {
  double t[3] = {0, 0, 0};
  double r[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  bool ident = true, hasT = false, hasR = false;
};

void set_properties(Mat &m)
{
  m.hasT   = std::fabs(m.t[0]) > kTol || std::fabs(m.t[1]) > kTol || std::fabs(m.t[2]) > kTol;
  int code = 0;
  for (int i = 0; i < 9; ++i) code |= (1 << i) * (std::fabs(m.r[i]) > kTol);
  bool identity = code == 0x111 && m.r[0] == 1. && m.r[4] == 1. && m.r[8] == 1.;
  m.hasR        = !identity;
  m.ident       = !m.hasT && !m.hasR;
}

// this := this * rhs (Transformation3D::MultiplyFromRight)
void compose(Mat &a, Mat const &b)
{
  if (b.ident) return;
  a.ident = false;
  if (b.hasT) {
    a.hasT = true;
    for (int i = 0; i < 3; ++i) {
      double acc = a.t[i];
      acc += a.r[3 * i] * b.t[0];
      acc += a.r[3 * i + 1] * b.t[1];
      acc += a.r[3 * i + 2] * b.t[2];
      a.t[i] = acc;
    }
  }
  if (b.hasR) {
    a.hasR = true;
    for (int row = 0; row < 3; ++row) {
      const double x = a.r[3 * row], y = a.r[3 * row + 1], z = a.r[3 * row + 2];
      double o[3];
      for (int c = 0; c < 3; ++c) o[c] = x * b.r[c];
      for (int c = 0; c < 3; ++c) o[c] += y * b.r[3 + c];
      for (int c = 0; c < 3; ++c) o[c] += z * b.r[6 + c];
      for (int c = 0; c < 3; ++c) a.r[3 * row + c] = o[c];
    }
  }
}

// local -> mother point (Transformation3D::InverseTransform)
void to_mother(PVolRec const &p, const double l[3], double m[3])
{
  for (int i = 0; i < 3; ++i) {
    double acc = p.t[i];
    acc += l[0] * p.r[3 * i];
    acc += l[1] * p.r[3 * i + 1];
    acc += l[2] * p.r[3 * i + 2];
    m[i] = acc;
  }
}

struct Box6 {
  double lo[3], hi[3];
  void center(double c[3]) const
  {
    for (int k = 0; k < 3; ++k) c[k] = 0.5 * (hi[k] + lo[k]);
  }
  double area() const
  {
    double e0 = hi[0] - lo[0], e1 = hi[1] - lo[1], e2 = hi[2] - lo[2];
    return 2. * (e0 * e1 + e1 * e2 + e2 * e0);
  }
};

Box6 merge(Box6 const &a, Box6 const &b)
{
  Box6 u;
  for (int k = 0; k < 3; ++k) {
    u.lo[k] = b.lo[k] < a.lo[k] ? b.lo[k] : a.lo[k];
    u.hi[k] = a.hi[k] < b.hi[k] ? b.hi[k] : a.hi[k];
  }
  return u;
}

// Strict ordering of shifted centroids along `axis`, ties broken by the next axes.
struct CentroidOrder {
  const Box6 *boxes;
  int axis;
  void key(int i, double k[3]) const
  {
    double c[3];
    boxes[i].center(c);
    for (int j = 0; j < 3; ++j) k[j] = c[j] + 0.01 * (c[j] - boxes[i].lo[j]);
  }
  bool operator()(int a, int b) const
  {
    double ka[3], kb[3];
    key(a, ka);
    key(b, kb);
    const int a0 = axis, a1 = (axis + 1) % 3, a2 = (axis + 2) % 3;
    if (ka[a0] != kb[a0]) return ka[a0] < kb[a0];
    if (ka[a1] != kb[a1]) return ka[a1] < kb[a1];
    return ka[a2] < kb[a2];
  }
};

struct BvhBuilder {
  const Box6 *boxes;
  int *prim; // permutation of 0..n-1
  std::vector<int> nchild, offset;
  std::vector<Box6> nodes;
  unsigned nnodes;

  int *split(int *begin, int *end)
  {
    const long n      = end - begin;
    double bestMetric = double(n);
    int bestAxis = -1, bestObj = -1;
    std::vector<double> left(n), right(n);
    for (int axis = 0; axis < 3; ++axis) {
      std::sort(begin, end, CentroidOrder{boxes, axis});
      // prefix/suffix surface areas of the sorted run
      left[0]  = 0.;
      Box6 acc = boxes[begin[0]];
      for (long i = 1; i < n; ++i) {
        left[i] = acc.area();
        acc     = merge(acc, boxes[begin[i]]);
      }
      Box6 acc2 = boxes[begin[n - 1]];
      for (long i = n - 1; i >= 0; --i) {
        acc2     = merge(acc2, boxes[begin[i]]);
        right[i] = acc2.area();
      }
      const double total = right[0];
      for (long i = 0; i < n; ++i) {
        // the balance term is integer arithmetic in the reference: 0.1 * |n/2 - i/n|
        double m = (left[i] / total) * double(i) + (right[i] / total) * double(n - i) +
                   0.1 * double(std::labs(n / 2 - i / n));
        if (m < bestMetric) bestMetric = m, bestAxis = axis, bestObj = begin[i];
      }
    }
    if (bestAxis < 0) return end;
    CentroidOrder ord{boxes, bestAxis};
    return std::partition(begin, end, [&](int i) { return ord(i, bestObj); });
  }

  void node(unsigned id, int *first, int *last)
  {
    if (id >= nnodes) return;
    nchild[id] = int(last - first);
    offset[id] = int(first - prim);
    if (first == last) return;
    Box6 b = boxes[*first];
    for (int *it = first + 1; it != last; ++it) b = merge(b, boxes[*it]);
    nodes[id] = b;
    if (first + 1 == last) return;
    int *pivot = split(first, last);
    node(2 * id + 1, first, pivot);
    node(2 * id + 2, pivot, last);
  }
};

} // namespace

int build_host_geometry(vgmi_geom_desc const &d, HostGeometry &h, std::string &err)
{
  if (d.nshapes <= 0 || d.nlogical <= 0) return err = "empty geometry description", -1;
  if (d.world_logical < 0 || d.world_logical >= d.nlogical) return err = "world logical volume out of range", -1;
  h = HostGeometry();
  // shapes
  h.shapes.resize(d.nshapes);
  h.extraOffset.assign(d.nshapes, -1);
  for (int i = 0; i < d.nshapes; ++i) {
    const int off = d.shape_param_offset[i], cnt = d.shape_param_offset[i + 1] - off;
    if (cnt < 0) return err = "shape parameter offsets must be non-decreasing", -1;
    std::string e;
    std::vector<double> extra;
    if (MakeShapeRecord(d.shape_type[i], d.shape_params + off, cnt, h.shapes[i], extra, e) != 0)
      return err = "shape " + std::to_string(i) + ": " + e, -1;
    if (!extra.empty()) {
      h.extraOffset[i] = int64_t(h.extraPool.size());
      h.extraPool.insert(h.extraPool.end(), extra.begin(), extra.end());
    }
  }
  // logical volumes and placements (placement k -> placed volume k, world -> nplaced)
  std::vector<std::vector<int>> dlist(d.nlogical);
  h.pvols.resize(d.nplaced + 1);
  for (int k = 0; k < d.nplaced; ++k) {
    const int m = d.placed_mother[k], l = d.placed_logical[k];
    if (m < 0 || m >= d.nlogical || l < 0 || l >= d.nlogical) return err = "placement logical index out of range", -1;
    PVolRec &p = h.pvols[k];
    p.lv       = l;
    p.childId  = int(dlist[m].size());
    Mat mt;
    for (int i = 0; i < 3; ++i) mt.t[i] = d.placed_transform[12 * k + i];
    for (int i = 0; i < 9; ++i) mt.r[i] = d.placed_transform[12 * k + 3 + i];
    set_properties(mt);
    std::memcpy(p.t, mt.t, sizeof(p.t));
    std::memcpy(p.r, mt.r, sizeof(p.r));
    p.hasT = mt.hasT;
    p.hasR = mt.hasR;
    dlist[m].push_back(k);
  }
  PVolRec &w = h.pvols[d.nplaced];
  w          = PVolRec();
  w.lv       = d.world_logical;
  w.r[0] = w.r[4] = w.r[8] = 1.;
  h.worldPvol = d.nplaced;

  // logical volumes, daughter AABBs and BVHs
  h.lvs.resize(d.nlogical);
  for (int lvi = 0; lvi < d.nlogical; ++lvi) {
    const int s = d.logical_shape[lvi];
    if (s < 0 || s >= d.nshapes) return err = "logical volume shape index out of range", -1;
    LVRec &L     = h.lvs[lvi];
    L            = LVRec();
    L.shape      = s;
    L.ndaughters = int(dlist[lvi].size());
    L.dOffset    = int(h.daughters.size());
    L.nodeBase   = int(h.nchild.size());
    L.primBase   = int(h.prim.size());
    const int n  = L.ndaughters;
    if (n == 0) continue;
    std::vector<Box6> boxes(n);
    for (int i = 0; i < n; ++i) {
      PVolRec const &pv = h.pvols[dlist[lvi][i]];
      double lo[3], hi[3];
      ShapeExtent(h.shapes[d.logical_shape[pv.lv]], lo, hi);
      const double delta[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
      double mn[3], mx[3];
      for (int k = 0; k < 3; ++k) mn[k] = DBL_MAX, mx[k] = -DBL_MAX;
      for (int c = 0; c < 8; ++c) {
        const int ix = (c >> 2) & 1, iy = (c >> 1) & 1, iz = c & 1;
        double corner[3] = {lo[0] + ix * delta[0], lo[1] + iy * delta[1], lo[2] + iz * delta[2]};
        double m[3];
        to_mother(pv, corner, m);
        for (int k = 0; k < 3; ++k) {
          mn[k] = std::min(mn[k], m[k]);
          mx[k] = std::max(mx[k], m[k]);
        }
      }
      for (int k = 0; k < 3; ++k) boxes[i].lo[k] = mn[k] - 1E-3, boxes[i].hi[k] = mx[k] + 1E-3;
    }
    for (int i = 0; i < n; ++i) {
      h.daughters.push_back(dlist[lvi][i]);
      AABBRec a;
      std::memcpy(a.lo, boxes[i].lo, sizeof(a.lo));
      std::memcpy(a.hi, boxes[i].hi, sizeof(a.hi));
      h.dAABB.push_back(a);
    }
    const int depth   = std::min(n / 2 > 0 ? std::max(0, int(std::log2(n / 2))) : 0, 32);
    const unsigned nn = (2u << depth) - 1;
    std::vector<int> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    BvhBuilder b{boxes.data(), perm.data(), std::vector<int>(nn, 0), std::vector<int>(nn, -1),
                 std::vector<Box6>(nn, Box6{{0, 0, 0}, {0, 0, 0}}), nn};
    b.node(0, perm.data(), perm.data() + n);
    for (unsigned id = 0; id < nn / 2; ++id)
      if (b.nchild[id] > 8 && b.nchild[id] == b.nchild[2 * id + 1] + b.nchild[2 * id + 2]) b.nchild[id] = -1;
    L.nNodes = int(nn);
    for (unsigned id = 0; id < nn; ++id) {
      h.nchild.push_back(b.nchild[id]);
      h.offset.push_back(b.offset[id]);
      AABBRec a;
      std::memcpy(a.lo, b.nodes[id].lo, sizeof(a.lo));
      std::memcpy(a.hi, b.nodes[id].hi, sizeof(a.hi));
      h.nodes.push_back(a);
    }
    for (int i = 0; i < n; ++i) h.prim.push_back(perm[i]);
    h.bvhDepth.push_back(depth);
  }

  // NavIndex table: depth-first over touchables, every level caches its global matrix
  h.navind.assign(1, 0u);
  h.maxDepth = 0;
  struct Frame {
    int pvol;
    int level;
    uint32_t mother;
    int dind;
    Mat mat;
  };
  std::vector<Frame> stack;
  stack.push_back(Frame{h.worldPvol, 0, 0, 0, Mat()});
  uint32_t cur = 1;
  while (!stack.empty()) {
    Frame f = stack.back();
    stack.pop_back();
    if (f.level > 255) return err = "geometry deeper than 255 levels", -1;
    const LVRec &L      = h.lvs[h.pvols[f.pvol].lv];
    const uint32_t nd   = uint32_t(L.ndaughters);
    if (nd >= 65535u) return err = "more than 65535 daughters", -1;
    const uint32_t dIdx = 3 + nd + ((nd + 1) & 1);
    const bool pad      = ((cur + dIdx) * 4) % 8 != 0;
    const uint32_t self = cur;
    h.navind.resize(cur + dIdx + (pad ? 1 : 0) + 24, 0u);
    h.navind[cur]     = f.mother;
    if (f.mother > 0) h.navind[f.mother + 3 + f.dind] = cur;
    h.navind[cur + 1] = uint32_t(f.pvol);
    const uint32_t hasm = 0x04u + 0x02u * uint32_t(f.mat.hasT) + uint32_t(f.mat.hasR);
    h.navind[cur + 2]   = uint32_t(f.level) | (hasm << 8) | (nd << 16);
    cur += dIdx + (pad ? 1 : 0);
    double m[12];
    std::memcpy(m, f.mat.t, 3 * sizeof(double));
    std::memcpy(m + 3, f.mat.r, 9 * sizeof(double));
    std::memcpy(&h.navind[cur], m, sizeof(m));
    cur += 24;
    h.maxDepth = std::max(h.maxDepth, f.level + 1);
    // push daughters in reverse so that they are numbered in daughter order
    for (int i = int(nd) - 1; i >= 0; --i) {
      const int child = h.daughters[L.dOffset + i];
      Frame c{child, f.level + 1, self, i, f.mat};
      Mat pl;
      PVolRec const &pv = h.pvols[child];
      std::memcpy(pl.t, pv.t, sizeof(pl.t));
      std::memcpy(pl.r, pv.r, sizeof(pl.r));
      pl.hasT  = pv.hasT;
      pl.hasR  = pv.hasR;
      pl.ident = !pl.hasT && !pl.hasR;
      compose(c.mat, pl);
      stack.push_back(c);
    }
  }
  h.navind.resize(cur);
  return 0;
}

} // namespace vgmi
