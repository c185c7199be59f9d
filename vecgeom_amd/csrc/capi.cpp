// C ABI: library information, error reporting and shape lifetime.
#include "vgmi_internal.h"
#include "shape_host.h"
#include <cstdarg>
#include <cstdio>
#include <new>
#include <string>

namespace vgmi {
static thread_local std::string g_last_error;

int set_error(int code, const char *fmt, ...)
{
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}
} // namespace vgmi

using namespace vgmi;

extern "C" {

const char *vgmi_version(void) { return "vgmi 0.1 (gfx950)"; }

const char *vgmi_last_error(void) { return g_last_error.c_str(); }

vgmi_status vgmi_shape_create(int type, const double *params, int nparams, vgmi_shape **out)
{
  if (!out) return set_error(VGMI_ERR_INVALID, "vgmi_shape_create: null output handle");
  *out = nullptr;
  if (nparams > 0 && !params) return set_error(VGMI_ERR_INVALID, "vgmi_shape_create: null params");
  vgmi_shape *s = new (std::nothrow) vgmi_shape();
  if (!s) return set_error(VGMI_ERR_INVALID, "vgmi_shape_create: out of host memory");
  std::string err;
  if (MakeShapeRecord(type, params, nparams, s->rec, s->extra, err) != 0) {
    delete s;
    return set_error(VGMI_ERR_INVALID, "vgmi_shape_create: %s", err.c_str());
  }
  if (!s->extra.empty()) {
    const size_t bytes = s->extra.size() * sizeof(double);
    hipError_t e       = hipMalloc(&s->extra_dev, bytes);
    if (e == hipSuccess) e = hipMemcpy(s->extra_dev, s->extra.data(), bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      if (s->extra_dev) (void)hipFree(s->extra_dev);
      delete s;
      return set_error(VGMI_ERR_DEVICE, "vgmi_shape_create: device copy of the shape data: %s", hipGetErrorString(e));
    }
    SetShapeExtraBase(s->rec, static_cast<const double *>(s->extra_dev));
  }
  *out = s;
  return VGMI_OK;
}

vgmi_status vgmi_shape_destroy(vgmi_shape *shape)
{
  if (shape && shape->extra_dev) (void)hipFree(shape->extra_dev);
  delete shape;
  return VGMI_OK;
}

vgmi_status vgmi_shape_extent(const vgmi_shape *shape, double lo[3], double hi[3])
{
  if (!shape || !lo || !hi) return set_error(VGMI_ERR_INVALID, "vgmi_shape_extent: null argument");
  ShapeExtent(shape->rec, lo, hi);
  return VGMI_OK;
}

} // extern "C"
