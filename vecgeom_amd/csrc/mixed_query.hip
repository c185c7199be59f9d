// Mixed-precision shape queries (FP32 bulk + FP64 re-evaluation): placeholder until implemented.
#include "vgmi_internal.h"

namespace vgmi {
vgmi_status mixed_shape_query(const vgmi_shape *, int, const double *, const double *, const double *,
                              const double *, const double *, const double *, void *, size_t, void *)
{
  return set_error(VGMI_ERR_UNSUPPORTED, "vgmi_shape_query: VGMI_MIXED is not implemented yet");
}
} // namespace vgmi
