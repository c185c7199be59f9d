// Host program built against include/vgmi.hpp by tests/test_adapter.py: it reads a flat geometry
// description, closes it host-only through vgmi::Geometry, prints the NavIndex table and the counts,
// queries a shape's extent through vgmi::UnplacedVolume, and checks that errors surface as vgmi::Error.
// No GPU is touched.
//
// Input file (whitespace separated): nshapes, then per shape "type nparams p..."; nlogical, then the
// shape of each; nplaced, then per placement "mother logical t0..t11"; world logical.
#include "vgmi.hpp"

#include <cstdio>
#include <fstream>
#include <iostream>

int main(int argc, char **argv)
{
  if (argc != 2) {
    std::fprintf(stderr, "usage: %s description.txt\n", argv[0]);
    return 2;
  }
  std::ifstream in(argv[1]);
  int ns = 0;
  in >> ns;
  std::vector<int> stype(ns), soff(ns + 1, 0), lvshape, plm, pll;
  std::vector<double> sparams, plt;
  for (int i = 0; i < ns; ++i) {
    int np = 0;
    in >> stype[i] >> np;
    for (int k = 0; k < np; ++k) {
      double x;
      in >> x;
      sparams.push_back(x);
    }
    soff[i + 1] = soff[i] + np;
  }
  int nl = 0;
  in >> nl;
  lvshape.resize(nl);
  for (int i = 0; i < nl; ++i) in >> lvshape[i];
  int np = 0;
  in >> np;
  plm.resize(np);
  pll.resize(np);
  plt.resize(12 * size_t(np));
  for (int i = 0; i < np; ++i) {
    in >> plm[i] >> pll[i];
    for (int k = 0; k < 12; ++k) in >> plt[12 * size_t(i) + k];
  }
  int world = -1;
  in >> world;
  if (!in) {
    std::fprintf(stderr, "bad description\n");
    return 2;
  }

  vgmi_geom_desc d{};
  d.nshapes            = ns;
  d.shape_type         = stype.data();
  d.shape_param_offset = soff.data();
  d.shape_params       = sparams.data();
  d.nlogical           = nl;
  d.logical_shape      = lvshape.data();
  d.nplaced            = np;
  d.placed_mother      = plm.data();
  d.placed_logical     = pll.data();
  d.placed_transform   = plt.data();
  d.world_logical      = world;

  vgmi::Geometry g(d, /*host_only=*/true);
  std::vector<int64_t> info = g.Info();
  std::printf("info");
  for (int64_t c : info) std::printf(" %lld", (long long)c);
  std::printf("\nworld %d\nnavindex", g.WorldId());
  for (uint32_t w : g.NavIndexTable()) std::printf(" %u", w);
  std::printf("\n");

  // the first shape of the description, as an unplaced volume
  vgmi::UnplacedVolume s(stype[0], std::vector<double>(sparams.begin() + soff[0], sparams.begin() + soff[1]));
  double lo[3], hi[3];
  s.Extent(lo, hi);
  std::printf("extent %.17g %.17g %.17g %.17g %.17g %.17g\n", lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]);

  // errors: invalid shape parameters, a device query on a host-only geometry, a bad world index
  int caught = 0;
  try {
    vgmi::UnplacedVolume bad(VGMI_BOX, {1.0, -2.0, 3.0});
  } catch (vgmi::Error const &e) {
    caught += e.code() == VGMI_ERR_INVALID;
  }
  try {
    double x = 0, out = 0;
    vgmi::SOA3DView p{&x, &x, &x, 1};
    g.Placed(0).SafetyToIn(p, &out);
  } catch (vgmi::Error const &e) {
    caught += e.code() == VGMI_ERR_INVALID && std::string(e.what()).find("host-only") != std::string::npos;
  }
  try {
    vgmi_geom_desc bad = d;
    bad.world_logical  = nl + 5;
    vgmi::Geometry gb(bad, true);
  } catch (vgmi::Error const &e) {
    caught += e.code() == VGMI_ERR_INVALID;
  }
  std::printf("errors %d\n", caught);
  return 0;
}
