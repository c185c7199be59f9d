/*
 * vgmi.h — C ABI of the MI355X-native batched shape-query and BVH-navigation
 * library (libvgmi.so). Plain pointers and sizes only; no C++ or torch types.
 *
 * Conventions (shared by every entry point):
 *  - All point/direction/output pointers are DEVICE pointers owned by the
 *    caller (hipMalloc'd or from any allocator of device memory), laid out as
 *    SOA: x[n], y[n], z[n] in separate arrays (VecGeom/base/SOA3D.h:35-41,
 *    non-owning view :177-181).
 *  - Work is enqueued on `stream` (a hipStream_t; NULL = the legacy default
 *    stream) and is asynchronous: results are visible after the stream syncs.
 *  - Return value is a vgmi_status; 0 is success. Failures never throw; the
 *    message of the last failure on the calling thread is vgmi_last_error().
 *  - Result sentinels follow the reference: a miss is kInfLength = DBL_MAX
 *    (VecGeom/base/Math.h:41); a query from the wrong side returns a negative
 *    value (DistanceToOut of an outside point is -1,
 *    VecGeom/volumes/kernel/BoxImplementation.h:150).
 *  - Inside results are int32 1/2/3 = kInside/kSurface/kOutside
 *    (VecGeom/base/Global.h:71-88); Contains results are uint8 0/1.
 */
#ifndef VGMI_H
#define VGMI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VGMI_API __attribute__((visibility("default")))

typedef int vgmi_status;
#define VGMI_OK 0
#define VGMI_ERR_INVALID (-1)     /* bad argument, unknown shape type or op */
#define VGMI_ERR_UNSUPPORTED (-2) /* valid request the library does not implement */
#define VGMI_ERR_DEVICE (-3)      /* HIP runtime error (launch, copy, allocation) */
#define VGMI_ERR_CAPACITY (-4)    /* request exceeds a documented maximum */

/* Shape types; parameter lists are those of the reference constructors. */
enum vgmi_shape_type {
  VGMI_BOX = 0,         /* dx dy dz                                volumes/UnplacedBox.h */
  VGMI_TUBE = 1,        /* rmin rmax dz sphi dphi                  volumes/UnplacedTube.h:29 */
  VGMI_CONE = 2,        /* rmin1 rmax1 rmin2 rmax2 dz sphi dphi    volumes/UnplacedCone.h */
  VGMI_POLYCONE = 3,    /* sphi dphi nz z[nz] rmin[nz] rmax[nz]    volumes/UnplacedPolycone.h */
  VGMI_POLYHEDRON = 4,  /* sphi dphi nsides nz z[] rmin[] rmax[]   volumes/UnplacedPolyhedron.h */
  VGMI_TRAPEZOID = 5,   /* dz theta phi dy1 dx1 dx2 alpha1 dy2 dx3 dx4 alpha2  volumes/UnplacedTrapezoid.h:56 */
  VGMI_SPHERE = 6,      /* rmin rmax sphi dphi stheta dtheta       volumes/UnplacedSphere.h:40 */
  VGMI_TORUS2 = 7,      /* rmin rmax rtor sphi dphi                volumes/UnplacedTorus2.h */
  VGMI_ORB = 8,         /* r                                       volumes/UnplacedOrb.h */
  VGMI_TESSELLATED = 9, /* ntri then 9 doubles per triangle        volumes/UnplacedTessellated.h */
  VGMI_HYPE = 10        /* rmin rmax stIn stOut dz                 volumes/UnplacedHype.h */
};

/* Query methods of VUnplacedVolume / VPlacedVolume (volumes/UnplacedVolume.h:88-223). */
enum vgmi_op {
  VGMI_OP_CONTAINS = 0,     /* uint8 out */
  VGMI_OP_INSIDE = 1,       /* int32 out */
  VGMI_OP_DIST_IN = 2,      /* double out, needs directions */
  VGMI_OP_DIST_OUT = 3,     /* double out, needs directions */
  VGMI_OP_SAFETY_IN = 4,    /* double out */
  VGMI_OP_SAFETY_OUT = 5    /* double out */
};

enum vgmi_precision {
  VGMI_FP64 = 0, /* plain FP64 kernels: the reference's scalar backend semantics */
  VGMI_MIXED = 1 /* FP32 bulk arithmetic, FP64 polishing and near-surface re-evaluation */
};

typedef struct vgmi_shape vgmi_shape; /* one unplaced shape: host record + device copy */

/* Library information and error reporting. */
VGMI_API const char *vgmi_version(void);
VGMI_API const char *vgmi_last_error(void);

/*
 * Create an unplaced shape from the reference constructor's parameters.
 * Replaces the Unplaced<Shape>(...) constructors (e.g. volumes/UnplacedBox.h,
 * volumes/UnplacedTube.h:29 → TubeStruct ctor volumes/TubeStruct.h:163-172).
 */
VGMI_API vgmi_status vgmi_shape_create(int type, const double *params, int nparams, vgmi_shape **out);
VGMI_API vgmi_status vgmi_shape_destroy(vgmi_shape *shape);
/* Bounding box, as VUnplacedVolume::Extent (volumes/UnplacedVolume.h:226). */
VGMI_API vgmi_status vgmi_shape_extent(const vgmi_shape *shape, double lo[3], double hi[3]);

/*
 * Batched SOA query of one unplaced shape: out[i] = op(p[i], d[i]).
 * Replaces the SOA overloads VUnplacedVolume::DistanceToOut(SOA3D, SOA3D,
 * step_max*, out*) and SafetyToOut(SOA3D, out*) (volumes/UnplacedVolume.h:144,179)
 * and the VPlacedVolume SOA overloads Contains/Inside/DistanceToIn/SafetyToIn
 * (volumes/PlacedVolume.h:273,299,332,372,405,421) for an identity placement.
 * dx/dy/dz may be NULL for Contains/Inside/Safety*; step_max may be NULL
 * (kInfLength). out is uint8*, int32* or double* by op.
 */
VGMI_API vgmi_status vgmi_shape_query(const vgmi_shape *shape, int op, int precision, const double *px, const double *py,
                             const double *pz, const double *dx, const double *dy, const double *dz,
                             const double *step_max, void *out, size_t n, void *stream);

/*
 * All six methods in one pass over the stream (points and directions are read
 * once). Any output pointer may be NULL to skip that method.
 * No single reference counterpart: equivalent to the six SOA calls above.
 */
VGMI_API vgmi_status vgmi_shape_query_all(const vgmi_shape *shape, int precision, const double *px, const double *py,
                                 const double *pz, const double *dx, const double *dy, const double *dz,
                                 uint8_t *contains, int32_t *inside, double *dist_in, double *dist_out,
                                 double *safety_in, double *safety_out, size_t n, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* VGMI_H */
