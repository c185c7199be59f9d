// Host construction of the tessellated solid: facets as Tile<3>::AddVertex computes them
// (VecGeom/volumes/Tile.h:73-140) with the side vectors of TessellatedCluster::AddFacet
// (VecGeom/volumes/TessellatedCluster.h:170-200), the extent of TessellatedStruct::AddFacet
// (VecGeom/volumes/TessellatedStruct.h:215-230), and a threaded BVH over the facets that the
// device kernels walk without a stack.
#include "shape_host.h"
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <numeric>

namespace vgmi {

namespace {

constexpr double hTol = 1e-9;

struct D3 {
  double x, y, z;
  D3 operator-(D3 const &o) const { return {x - o.x, y - o.y, z - o.z}; }
  D3 operator*(double s) const { return {x * s, y * s, z * s}; }
  double Dot(D3 const &o) const { return x * o.x + y * o.y + z * o.z; }
  double Mag2() const { return x * x + y * y + z * z; }
  D3 Cross(D3 const &o) const { return {y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x}; }
  double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
D3 Normalized(D3 const &v) { return v * (1. / std::sqrt(v.Mag2())); }

struct HostFacet {
  D3 v[3], side[3], n;
  double dist;
  double lo[3], hi[3];
  int index;
};

bool MakeFacet(D3 const &a, D3 const &b, D3 const &c, HostFacet &t)
{
  t.v[0] = a, t.v[1] = b, t.v[2] = c;
  int nvert = 3;
  for (int i = 0; i < 3; ++i)
    if ((t.v[(i + 1) % 3] - t.v[i]).Mag2() < hTol) nvert--;
  if (nvert < 3) return false;
  bool degenerated = true;
  for (int i = 0; i < 2 && degenerated; ++i) {
    D3 e1 = t.v[i + 1] - t.v[i];
    if (e1.Mag2() < hTol) continue;
    for (int j = i + 1; j < 3; ++j) {
      D3 e2 = t.v[(j + 1) % 3] - t.v[j];
      if (e2.Mag2() < hTol) continue;
      t.n = e1.Cross(e2);
      if (t.n.Mag2() < hTol) continue;
      t.n         = t.n * (1. / std::sqrt(t.n.Mag2()));
      degenerated = false;
      break;
    }
  }
  if (degenerated) return false;
  for (int i = 0; i < 3; ++i) {
    if ((t.v[(i + 1) % 3] - t.v[i]).Mag2() < hTol) continue;
    t.dist = -t.n.Dot(t.v[i]);
    break;
  }
  for (int i = 0; i < 3; ++i) t.side[i] = Normalized(t.n.Cross(t.v[(i + 1) % 3] - t.v[i]));
  for (int k = 0; k < 3; ++k) {
    t.lo[k] = std::min(t.v[0][k], std::min(t.v[1][k], t.v[2][k]));
    t.hi[k] = std::max(t.v[0][k], std::max(t.v[1][k], t.v[2][k]));
  }
  return true;
}

// First direction of a fixed list with |n.dir| >= 1e-8 for every facet.
D3 TestDir(std::vector<HostFacet> const &f)
{
  static const double cand[][3] = {{0.3917183, 0.5276331, 0.7537210},
                                   {-0.6417143, 0.2311529, 0.7312417},
                                   {0.1873307, -0.8215501, 0.5384472},
                                   {0.7741139, 0.4109378, -0.4813220}};
  for (auto const &c : cand) {
    const D3 d = Normalized(D3{c[0], c[1], c[2]});
    bool good  = true;
    for (auto const &t : f)
      if (std::fabs(t.n.Dot(d)) < 1.e-8) good = false;
    if (good) return d;
  }
  return Normalized(D3{cand[0][0], cand[0][1], cand[0][2]});
}

// Single-precision bounds of x from below and from above.
float FloatDown(double x)
{
  float f = float(x);
  return double(f) > x ? std::nextafter(f, -FLT_MAX) : f;
}
float FloatUp(double x)
{
  float f = float(x);
  return double(f) < x ? std::nextafter(f, FLT_MAX) : f;
}

struct BuildNode {
  double lo[3], hi[3];
  int first, count, escape;
};


// Depth-first build: the node is appended, its left then right subtree follow it, and its
// escape index is the first node after the subtree.
void Build(std::vector<HostFacet> &f, int b, int e, std::vector<BuildNode> &nodes)
{
  BuildNode nd;
  for (int k = 0; k < 3; ++k) nd.lo[k] = DBL_MAX, nd.hi[k] = -DBL_MAX;
  double clo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, chi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  for (int i = b; i < e; ++i)
    for (int k = 0; k < 3; ++k) {
      nd.lo[k]       = std::min(nd.lo[k], f[i].lo[k]);
      nd.hi[k]       = std::max(nd.hi[k], f[i].hi[k]);
      const double c = 0.5 * (f[i].lo[k] + f[i].hi[k]);
      clo[k]         = std::min(clo[k], c);
      chi[k]         = std::max(chi[k], c);
    }
  // enlarge by a margin that covers the single-precision facet boxes of the candidate test
  for (int k = 0; k < 3; ++k) {
    const double m = 1e-6 * (1. + std::max(std::fabs(nd.lo[k]), std::fabs(nd.hi[k])));
    nd.lo[k] -= m, nd.hi[k] += m;
  }
  const int self = int(nodes.size());
  nodes.push_back(nd);
  if (e - b <= kTessLeafFacets) {
    nodes[self].first  = b;
    nodes[self].count  = e - b;
    nodes[self].escape = self + 1;
    return;
  }
  int axis = 0;
  for (int k = 1; k < 3; ++k)
    if (chi[k] - clo[k] > chi[axis] - clo[axis]) axis = k;
  const int mid = (b + e) / 2;
  std::nth_element(f.begin() + b, f.begin() + mid, f.begin() + e, [axis](HostFacet const &x, HostFacet const &y) {
    const double cx = x.lo[axis] + x.hi[axis], cy = y.lo[axis] + y.hi[axis];
    return cx < cy || (cx == cy && x.index < y.index);
  });
  nodes[self].first = nodes[self].count = 0;
  Build(f, b, mid, nodes);
  Build(f, mid, e, nodes);
  nodes[self].escape = int(nodes.size());
}

} // namespace

int MakeTessellated(double const *a, int n, ShapeRecord &r, std::vector<double> &extra, std::string &err)
{
  if (n < 1) return err = "tessellated needs ntri followed by 9 doubles per triangle", -1;
  // at most 2^23 facets: every facet-array offset (facets x kFacetDoubles) stays an int
  if (!(a[0] >= 4.0 && a[0] <= double(1 << 23)))
    return err = "tessellated needs between 4 and 2^23 triangles", -1;
  const int ntri = int(a[0]);
  if (int64_t(n) < 1 + 9 * int64_t(ntri)) return err = "tessellated needs 9 doubles per triangle", -1;
  std::vector<HostFacet> f;
  for (int i = 0; i < ntri; ++i) {
    const double *q = a + 1 + 9 * i;
    HostFacet t;
    if (!MakeFacet(D3{q[0], q[1], q[2]}, D3{q[3], q[4], q[5]}, D3{q[6], q[7], q[8]}, t)) continue;
    t.index = int(f.size());
    f.push_back(t);
  }
  if (f.size() < 4) return err = "tessellated has fewer than 4 valid facets", -1;
  r.type              = kTessellated;
  TessellatedData &T  = r.u.tess;
  for (int k = 0; k < 3; ++k) T.minExt[k] = DBL_MAX, T.maxExt[k] = -DBL_MAX;
  for (auto const &t : f)
    for (int v = 0; v < 3; ++v)
      for (int k = 0; k < 3; ++k) {
        T.minExt[k] = std::min(T.minExt[k], t.v[v][k]);
        T.maxExt[k] = std::max(T.maxExt[k], t.v[v][k]);
      }
  const D3 td = TestDir(f);
  T.testDir[0] = td.x, T.testDir[1] = td.y, T.testDir[2] = td.z;
  std::vector<BuildNode> nodes;
  Build(f, 0, int(f.size()), nodes);
  T.nfacets  = int(f.size());
  T.nnodes   = int(nodes.size());
  T.facetOff = 0;
  T.nodeOff  = T.nfacets * kFacetDoubles;
  T.base     = nullptr;
  T.imageOff = T.nodeOff + T.nnodes * kNodeDoubles;
  T.imageDoubles = T.nfacets < 65536 ? T.nnodes * kImageNodeDoubles + T.nfacets * kImagePlaneDoubles +
                                          (T.nfacets * kImageBoxFloats + 1) / 2
                                    : 0;
  extra.assign(size_t(T.imageOff) + size_t(T.imageDoubles), 0.0);
  for (int k = 0; k < T.nfacets; ++k) {
    HostFacet const &t = f[k];
    double *F          = extra.data() + size_t(k) * kFacetDoubles;
    float box[6];
    for (int c = 0; c < 3; ++c) box[c] = float(t.lo[c]), box[3 + c] = float(t.hi[c]);
    std::memcpy(F, box, sizeof(box));
    F[3] = t.n.x, F[4] = t.n.y, F[5] = t.n.z, F[6] = t.dist;
    for (int v = 0; v < 3; ++v) {
      double *V = F + 7 + 6 * v;
      V[0] = t.v[v].x, V[1] = t.v[v].y, V[2] = t.v[v].z;
      V[3] = t.side[v].x, V[4] = t.side[v].y, V[5] = t.side[v].z;
    }
    F[25] = double(t.index);
  }
  for (int j = 0; j < T.nnodes; ++j) {
    double *N = extra.data() + T.nodeOff + size_t(j) * kNodeDoubles;
    for (int k = 0; k < 3; ++k) N[k] = nodes[j].lo[k], N[3 + k] = nodes[j].hi[k];
    int32_t ints[4] = {nodes[j].first, nodes[j].count, nodes[j].escape, 0};
    std::memcpy(N + 6, ints, sizeof(ints));
  }
  if (T.imageDoubles > 0) {
    double *img = extra.data() + T.imageOff;
    for (int j = 0; j < T.nnodes; ++j) {
      TessNode32 n;
      for (int k = 0; k < 3; ++k) n.lo[k] = FloatDown(nodes[j].lo[k]), n.hi[k] = FloatUp(nodes[j].hi[k]);
      n.first  = uint16_t(nodes[j].first);
      n.count  = uint16_t(nodes[j].count);
      n.escape = nodes[j].escape;
      std::memcpy(img + size_t(j) * kImageNodeDoubles, &n, sizeof(n));
    }
    double *pl = img + size_t(T.nnodes) * kImageNodeDoubles;
    for (int k = 0; k < T.nfacets; ++k)
      std::memcpy(pl + size_t(k) * kImagePlaneDoubles, extra.data() + size_t(k) * kFacetDoubles + 3,
                  kImagePlaneDoubles * sizeof(double));
    float *bx = reinterpret_cast<float *>(pl + size_t(T.nfacets) * kImagePlaneDoubles);
    for (int k = 0; k < T.nfacets; ++k)
      std::memcpy(bx + size_t(k) * kImageBoxFloats, extra.data() + size_t(k) * kFacetDoubles,
                  kImageBoxFloats * sizeof(float));
  }
  return 0;
}

} // namespace vgmi
