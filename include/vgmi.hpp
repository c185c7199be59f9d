/*
 * vgmi.hpp — header-only C++ adapter over the C ABI of libvgmi.so (vgmi.h).
 *
 * Classes that mirror the method names of the reference's query and navigation interfaces so a host
 * application can swap them in: UnplacedVolume (VUnplacedVolume SOA methods,
 * VecGeom/volumes/UnplacedVolume.h:88-223), PlacedVolume (VPlacedVolume SOA methods,
 * VecGeom/volumes/PlacedVolume.h:273-421) and Geometry (the closed geometry with its NavIndex table,
 * GlobalLocator::LocateGlobalPoint, VNavigator::ComputeStep / ComputeStepAndPropagatedState,
 * VSafetyEstimator::ComputeSafety, VLevelLocator::LevelLocate). They subclass nothing from the
 * reference and use no HIP or torch types: arrays are device pointers in SOA3DView, streams are the
 * opaque `void *` of the ABI. Failures throw vgmi::Error with vgmi_last_error()'s text.
 */
#ifndef VGMI_HPP
#define VGMI_HPP

#include "vgmi.h"

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace vgmi {

class Error : public std::runtime_error {
public:
  Error(std::string const &what, int code) : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

private:
  int code_;
};

inline void Check(vgmi_status rc, char const *call)
{
  if (rc != VGMI_OK) throw Error(std::string(call) + ": " + vgmi_last_error(), rc);
}

/* Non-owning view of three device arrays, as SOA3D<double> (VecGeom/base/SOA3D.h:177-181). */
struct SOA3DView {
  const double *x = nullptr, *y = nullptr, *z = nullptr;
  size_t n = 0;
  size_t size() const { return n; }
};

enum class Precision { kFP64 = VGMI_FP64, kMixed = VGMI_MIXED };

class UnplacedVolume {
public:
  /* type: vgmi_shape_type; params: the reference constructor's parameter list. */
  UnplacedVolume(int type, std::vector<double> const &params)
  {
    Check(vgmi_shape_create(type, params.data(), int(params.size()), &h_), "vgmi_shape_create");
  }
  ~UnplacedVolume() { vgmi_shape_destroy(h_); }
  UnplacedVolume(UnplacedVolume const &) = delete;
  UnplacedVolume &operator=(UnplacedVolume const &) = delete;
  UnplacedVolume(UnplacedVolume &&o) noexcept : h_(std::exchange(o.h_, nullptr)) {}

  void Extent(double lo[3], double hi[3]) const { Check(vgmi_shape_extent(h_, lo, hi), "vgmi_shape_extent"); }

  void Contains(SOA3DView const &p, uint8_t *out, void *stream = nullptr, Precision pr = Precision::kFP64) const
  {
    Query(VGMI_OP_CONTAINS, p, nullptr, nullptr, out, stream, pr);
  }
  void Inside(SOA3DView const &p, int32_t *out, void *stream = nullptr, Precision pr = Precision::kFP64) const
  {
    Query(VGMI_OP_INSIDE, p, nullptr, nullptr, out, stream, pr);
  }
  void DistanceToIn(SOA3DView const &p, SOA3DView const &d, const double *step_max, double *out,
                    void *stream = nullptr, Precision pr = Precision::kFP64) const
  {
    Query(VGMI_OP_DIST_IN, p, &d, step_max, out, stream, pr);
  }
  void DistanceToOut(SOA3DView const &p, SOA3DView const &d, const double *step_max, double *out,
                     void *stream = nullptr, Precision pr = Precision::kFP64) const
  {
    Query(VGMI_OP_DIST_OUT, p, &d, step_max, out, stream, pr);
  }
  void SafetyToIn(SOA3DView const &p, double *out, void *stream = nullptr, Precision pr = Precision::kFP64) const
  {
    Query(VGMI_OP_SAFETY_IN, p, nullptr, nullptr, out, stream, pr);
  }
  void SafetyToOut(SOA3DView const &p, double *out, void *stream = nullptr, Precision pr = Precision::kFP64) const
  {
    Query(VGMI_OP_SAFETY_OUT, p, nullptr, nullptr, out, stream, pr);
  }
  vgmi_shape const *handle() const { return h_; }

private:
  void Query(int op, SOA3DView const &p, SOA3DView const *d, const double *sm, void *out, void *stream,
             Precision pr) const
  {
    if (d && d->n != p.n) throw Error("vgmi adapter: points and directions differ in length", VGMI_ERR_INVALID);
    Check(vgmi_shape_query(h_, op, int(pr), p.x, p.y, p.z, d ? d->x : nullptr, d ? d->y : nullptr,
                           d ? d->z : nullptr, sm, out, p.n, stream),
          "vgmi_shape_query");
  }
  vgmi_shape *h_ = nullptr;
};

class Geometry;

/* One placed volume of a Geometry: VPlacedVolume's SOA queries. Contains, Inside, DistanceToIn and
 * SafetyToIn take points in the frame the volume is placed in; DistanceToOut and SafetyToOut in its
 * own frame; PlacedDistanceToOut in the placement frame. */
class PlacedVolume {
public:
  PlacedVolume(vgmi_geom const *g, int id) : g_(g), id_(id) {}
  int id() const { return id_; }
  void Contains(SOA3DView const &p, uint8_t *out, void *stream = nullptr) const
  {
    Q(VGMI_OP_CONTAINS, p, nullptr, out, stream);
  }
  void Inside(SOA3DView const &p, int32_t *out, void *stream = nullptr) const { Q(VGMI_OP_INSIDE, p, nullptr, out, stream); }
  void DistanceToIn(SOA3DView const &p, SOA3DView const &d, double *out, void *stream = nullptr) const
  {
    Q(VGMI_OP_DIST_IN, p, &d, out, stream);
  }
  void DistanceToOut(SOA3DView const &p, SOA3DView const &d, double *out, void *stream = nullptr) const
  {
    Q(VGMI_OP_DIST_OUT, p, &d, out, stream);
  }
  void PlacedDistanceToOut(SOA3DView const &p, SOA3DView const &d, double *out, void *stream = nullptr) const
  {
    Q(VGMI_OP_PLACED_DIST_OUT, p, &d, out, stream);
  }
  void SafetyToIn(SOA3DView const &p, double *out, void *stream = nullptr) const
  {
    Q(VGMI_OP_SAFETY_IN, p, nullptr, out, stream);
  }
  void SafetyToOut(SOA3DView const &p, double *out, void *stream = nullptr) const
  {
    Q(VGMI_OP_SAFETY_OUT, p, nullptr, out, stream);
  }

private:
  void Q(int op, SOA3DView const &p, SOA3DView const *d, void *out, void *stream) const
  {
    Check(vgmi_placed_query(g_, id_, op, VGMI_FP64, p.x, p.y, p.z, d ? d->x : nullptr, d ? d->y : nullptr,
                            d ? d->z : nullptr, nullptr, out, p.n, stream),
          "vgmi_placed_query");
  }
  vgmi_geom const *g_;
  int id_;
};

/* A closed geometry: vgmi_geom_build of a flat description (GeoManager::CloseGeometry). */
class Geometry {
public:
  explicit Geometry(vgmi_geom_desc const &d, bool host_only = false)
  {
    Check(vgmi_geom_build(&d, host_only ? VGMI_GEOM_HOST_ONLY : 0, &g_), "vgmi_geom_build");
  }
  ~Geometry() { vgmi_geom_destroy(g_); }
  Geometry(Geometry const &) = delete;
  Geometry &operator=(Geometry const &) = delete;

  /* counts[0..5] of vgmi_geom_info */
  std::vector<int64_t> Info() const
  {
    std::vector<int64_t> c(6);
    Check(vgmi_geom_info(g_, c.data()), "vgmi_geom_info");
    return c;
  }
  int WorldId() const { return int(Info()[4]); }
  /* NavIndexTable::CreateTable's words (VecGeom/management/NavIndexTable.h:128-143) */
  std::vector<uint32_t> NavIndexTable() const
  {
    int64_t n = 0;
    Check(vgmi_geom_navindex(g_, nullptr, 0, &n), "vgmi_geom_navindex");
    std::vector<uint32_t> w(static_cast<size_t>(n));
    Check(vgmi_geom_navindex(g_, w.data(), n, &n), "vgmi_geom_navindex");
    return w;
  }
  PlacedVolume Placed(int id) const { return PlacedVolume(g_, id); }

  void LocateGlobalPoint(SOA3DView const &p, uint32_t *nav_out, void *stream = nullptr,
                         Precision pr = Precision::kFP64) const
  {
    Check(vgmi_nav_locate(g_, int(pr), p.x, p.y, p.z, nav_out, p.n, stream), "vgmi_nav_locate");
  }
  void LevelLocate(int32_t const *logical, int32_t const *excl_placed, SOA3DView const &lp, int32_t *daughter,
                   SOA3DView const *daughter_lp = nullptr, void *stream = nullptr,
                   Precision pr = Precision::kFP64) const
  {
    double *dx = daughter_lp ? const_cast<double *>(daughter_lp->x) : nullptr;
    double *dy = daughter_lp ? const_cast<double *>(daughter_lp->y) : nullptr;
    double *dz = daughter_lp ? const_cast<double *>(daughter_lp->z) : nullptr;
    Check(vgmi_nav_level_locate(g_, int(pr), logical, excl_placed, lp.x, lp.y, lp.z, daughter, dx, dy, dz, lp.n,
                                stream),
          "vgmi_nav_level_locate");
  }
  /* ComputeStepAndPropagatedState (relocate) or ComputeStep per track */
  void ComputeStep(SOA3DView const &p, SOA3DView const &d, double const *limit, uint32_t const *nav_in,
                   uint32_t const *last_in, uint32_t *nav_out, uint32_t *last_out, uint8_t *boundary,
                   double *step, bool relocate = true, void *stream = nullptr, Precision pr = Precision::kFP64) const
  {
    Check(vgmi_nav_step(g_, int(pr), p.x, p.y, p.z, d.x, d.y, d.z, limit, nav_in, last_in, nav_out, last_out,
                        boundary, step, p.n, relocate ? 1 : 0, stream),
          "vgmi_nav_step");
  }
  void ComputeSafety(SOA3DView const &p, uint32_t const *nav, double *safety, void *stream = nullptr,
                     Precision pr = Precision::kFP64) const
  {
    Check(vgmi_nav_safety(g_, int(pr), p.x, p.y, p.z, nav, safety, p.n, stream), "vgmi_nav_safety");
  }
  /* the fused transport loop; p, nav and last are updated in place */
  void Propagate(double *px, double *py, double *pz, SOA3DView const &d, double limit, uint32_t *nav,
                 uint32_t *last, int32_t *steps_taken, int nsteps, void *stream = nullptr,
                 Precision pr = Precision::kFP64) const
  {
    Check(vgmi_nav_propagate(g_, int(pr), px, py, pz, d.x, d.y, d.z, limit, nav, last, steps_taken, d.n, nsteps,
                             stream),
          "vgmi_nav_propagate");
  }
  vgmi_geom const *handle() const { return g_; }

private:
  vgmi_geom *g_ = nullptr;
};

} // namespace vgmi

#endif /* VGMI_HPP */
