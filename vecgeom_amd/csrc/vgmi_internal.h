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
// Device arrays of one serving launch (inputs, per-track step counts, outputs).
struct ServeArgs {
  const double *px, *py, *pz, *dx, *dy, *dz;
  const int32_t *nsteps;
  double limit;
  double *opx, *opy, *opz;
  uint32_t *onav;
  int32_t *otaken;
  size_t n;
};
vgmi_status launch_locate_propagate(const vgmi_geom *geom, ServeArgs const &a, hipStream_t st);

// Records the message for vgmi_last_error() and returns `code`.
int set_error(int code, const char *fmt, ...);
vgmi_status mixed_shape_query(const vgmi_shape *shape, int op, const double *px, const double *py, const double *pz,
                              const double *dx, const double *dy, const double *dz, void *out, size_t n,
                              void *stream);
} // namespace vgmi
