// Internal declarations shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>
#include "../../include/vgmi.h"
#include "shape_data.h"
#include <vector>

struct vgmi_shape {
  vgmi::ShapeRecord rec;
  std::vector<double> extra;   // variable-length shape data (host copy)
  void *extra_dev = nullptr;   // device copy, referenced by rec
};

namespace vgmi {
// Device arrays of one fused transport launch; tracks are updated in place.
struct TransportArgs {
  double *px, *py, *pz;
  const double *dx, *dy, *dz;
  double limit;
  uint32_t *nav, *last;
  int32_t *taken;              // steps done per track (required with track_steps)
  const int32_t *track_steps;  // per-track step counts, or null for a uniform count
  size_t n;
  int nsteps;                  // the uniform step count, or the largest of track_steps
  int precision;               // VGMI_FP64 or VGMI_MIXED
  const uint32_t *dn;          // device-side track count (<= n, then a capacity), or null: n tracks
};
// Words of device workspace launch_transport needs to regroup n tracks.
size_t transport_work_words(const vgmi_geom *geom, size_t n);
// The transport loop, regrouped by volume when `work` is given and the geometry asks for it.
vgmi_status launch_transport(const vgmi_geom *geom, TransportArgs const &a, uint32_t *work, hipStream_t st);
// Locates every point (p_in -> p_out copy, nav), clears last and taken.
vgmi_status launch_locate_init(const vgmi_geom *geom, int precision, const double *px, const double *py,
                               const double *pz, double *opx, double *opy, double *opz, uint32_t *nav,
                               uint32_t *last, int32_t *taken, size_t n, hipStream_t st,
                               const uint32_t *dn = nullptr);

// Records the message for vgmi_last_error() and returns `code`.
int set_error(int code, const char *fmt, ...);
vgmi_status mixed_shape_query(const vgmi_shape *shape, int op, const double *px, const double *py, const double *pz,
                              const double *dx, const double *dy, const double *dz, void *out, size_t n,
                              void *stream);
// Bytes of dynamic LDS a workgroup may give a tessellated solid's image on the current device: half
// of the device's LDS per CU (two workgroups per CU) less the kernels' static LDS.
int tess_lds_limit();
} // namespace vgmi
