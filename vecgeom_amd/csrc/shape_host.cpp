// Host-side construction of the flat shape records (cached members as computed
// by the reference constructors). Runs once per shape at geometry build time.
#include "shape_host.h"
#include <algorithm>
#include <cmath>

namespace vgmi {

static constexpr double hTol  = 1e-9;
static constexpr double hHTol = 0.5e-9;
static constexpr double hPi   = 3.14159265358979323846;

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

int MakeShapeRecord(int type, double const *params, int nparams, ShapeRecord &r, std::string &err)
{
  switch (type) {
  case kBox:
    return MakeBox(params, nparams, r, err);
  case kTube:
    return MakeTube(params, nparams, r, err);
  default:
    err = "shape type not supported";
    return -1;
  }
}

void ShapeExtent(ShapeRecord const &r, double lo[3], double hi[3])
{
  switch (r.type) {
  case kBox:
    for (int i = 0; i < 3; ++i) lo[i] = -r.u.box.h[i], hi[i] = r.u.box.h[i];
    return;
  case kTube:
    lo[0] = lo[1] = -r.u.tube.rmax;
    hi[0] = hi[1] = r.u.tube.rmax;
    lo[2]         = -r.u.tube.z;
    hi[2]         = r.u.tube.z;
    return;
  default:
    lo[0] = lo[1] = lo[2] = 0;
    hi[0] = hi[1] = hi[2] = 0;
  }
}

} // namespace vgmi
