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
  VGMI_HYPE = 10,       /* rmin rmax stIn stOut dz                 volumes/UnplacedHype.h */
  VGMI_TRD = 11,        /* x1 x2 y1 y2 z (or x1 x2 y z, or x y z; a sixth value 1: the Trd1 kind)  volumes/UnplacedTrd.h:51-80 */
  VGMI_PARABOLOID = 12, /* rlo rhi dz (rhi > rlo, as G4Paraboloid)  volumes/UnplacedParaboloid.h:60 */
  VGMI_PARALLELEPIPED = 13, /* dx dy dz alpha theta phi            volumes/UnplacedParallelepiped.h:65-71 */
  VGMI_TET = 14,        /* the four corners p0 p1 p2 p3 (12 values) volumes/UnplacedTet.h:48-49 */
  VGMI_ELLIPTICALTUBE = 15, /* dx dy dz                             volumes/UnplacedEllipticalTube.h:51 */
  VGMI_ELLIPSOID = 16,  /* dx dy dz [zBottomCut zTopCut]           volumes/UnplacedEllipsoid.h:62 */
  VGMI_GENTRAP = 17,    /* x[8] y[8] dz, planar faces only         volumes/UnplacedGenTrap.h:46 */
  VGMI_CUTTUBE = 18,    /* rmin rmax dz sphi dphi bx by bz tx ty tz volumes/UnplacedCutTube.h:54 */
  VGMI_ELLIPTICALCONE = 19 /* a b h zcut                          volumes/UnplacedEllipticalCone.h */
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
 * dx/dy/dz may be NULL for Contains/Inside/Safety*. step_max (per point) may be NULL
 * (kInfLength); as in the reference only the tessellated solid uses it
 * (VecGeom/volumes/kernel/TessellatedImplementation.h:191-199, 213-216, 250: the extent distance is
 * returned when it is not below the limit, facet clusters beyond it are not searched); the other
 * shapes' kernels ignore it. out is uint8*, int32* or double* by op.
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

/*
 * Surface normal per point: VUnplacedVolume::Normal (VecGeom/volumes/UnplacedVolume.h:252), the
 * shape's own Normal (e.g. BoxImplementation::NormalKernel, BoxImplementation.h:172-191;
 * UnplacedTube::Normal, source/UnplacedTube.cpp:181-186). Writes the normal to n{x,y,z}[i] and the
 * validity flag (1: the point is on the surface as the shape's Normal decides it) to valid[i].
 * Where the reference leaves its normal argument unassigned (tube and hype points near none of the
 * surfaces their kernels test, torus points on no surface), n{x,y,z}[i] keep the caller's values.
 * FP64 only.
 */
VGMI_API vgmi_status vgmi_shape_normal(const vgmi_shape *shape, const double *px, const double *py, const double *pz,
                                       double *nx, double *ny, double *nz, uint8_t *valid, size_t n, void *stream);

/* ------------------------------------------------------------------------
 * One-launch batches over many shapes with uneven point counts.
 * No single reference counterpart: it batches the per-shape SOA calls above
 * (the reference issues one kernel per volume and method,
 * test/VecGeomBenchmark/Benchmarker.cu:140-216). Points of shape s occupy
 * [sum(counts[0..s)), +counts[s]) of the SOA arrays. The shapes must outlive the batch.
 * ------------------------------------------------------------------------ */
typedef struct vgmi_batch vgmi_batch;

VGMI_API vgmi_status vgmi_batch_create(const vgmi_shape *const *shapes, const int64_t *counts, int nshapes,
                                       vgmi_batch **out);
VGMI_API vgmi_status vgmi_batch_destroy(vgmi_batch *batch);
/* Launch shape: points per tile (64 .. 2^24), threads per block (64/128/256/512), resident blocks per
 * CU. A batch holding a tessellated solid whose LDS image (nodes, planes, candidate boxes; about
 * 72 B per facet) fits in 76 KiB reserves that much dynamic LDS per workgroup, which holds the
 * launch to two resident workgroups per CU. */
VGMI_API vgmi_status vgmi_batch_set_tuning(vgmi_batch *batch, int tile_points, int block, int blocks_per_cu);
/* info = nshapes, total points, tiles, tile points, block, blocks per CU */
VGMI_API vgmi_status vgmi_batch_info(const vgmi_batch *batch, int64_t info[6]);
/* All six methods for every point of the batch in one persistent launch (NULL outputs are skipped).
 * precision VGMI_FP64, or VGMI_MIXED: each point's answers equal vgmi_shape_query's mixed answers
 * for its shape bit for bit. */
VGMI_API vgmi_status vgmi_batch_query_all(vgmi_batch *batch, int precision, const double *px, const double *py,
                                          const double *pz, const double *dx, const double *dy, const double *dz,
                                          uint8_t *contains, int32_t *inside, double *dist_in, double *dist_out,
                                          double *safety_in, double *safety_out, void *stream);
/* Capture the same launch into a HIP graph bound to these pointers; replay with vgmi_batch_graph_launch. */
VGMI_API vgmi_status vgmi_batch_graph_capture(vgmi_batch *batch, int precision, const double *px, const double *py,
                                              const double *pz, const double *dx, const double *dy, const double *dz,
                                              uint8_t *contains, int32_t *inside, double *dist_in, double *dist_out,
                                              double *safety_in, double *safety_out, void *stream);
VGMI_API vgmi_status vgmi_batch_graph_launch(vgmi_batch *batch, void *stream);

/* ------------------------------------------------------------------------
 * Geometry and BVH navigation.
 *
 * A geometry is described by plain arrays (the same data a VecGeom user feeds
 * to the Unplaced* constructors, LogicalVolume and LogicalVolume::PlaceDaughter):
 * placement k becomes placed volume id k, the world placed volume gets id
 * nplaced. vgmi_geom_build closes it (ABBoxes, per-volume BVHs, NavIndex table;
 * GeoManager::CloseGeometry source/GeoManager.cpp:192-241, BVHManager::Init
 * source/BVHManager.cpp:24-30, NavIndexTable::CreateTable
 * management/NavIndexTable.h:128-143) and uploads the flat tables to the
 * current HIP device (CudaManager::Synchronize source/CudaManager.cpp:101-332).
 *
 * Navigation states are NavStateIndex values (navigation/NavStateIndex.h:34-41):
 * a uint32 index into the NavIndex table, 0 = outside the world; the "last
 * exited" state is a second uint32 and the boundary flag a uint8.
 * ------------------------------------------------------------------------ */
typedef struct vgmi_geom_desc {
  int nshapes;
  const int *shape_type;         /* vgmi_shape_type per shape */
  const int *shape_param_offset; /* nshapes + 1 offsets into shape_params */
  const double *shape_params;
  int nlogical;
  const int *logical_shape;      /* shape index per logical volume */
  int nplaced;
  const int *placed_mother;      /* mother logical volume per placement */
  const int *placed_logical;     /* placed logical volume per placement */
  const double *placed_transform;/* 12 doubles per placement: tx ty tz r0..r8 (row major) */
  int world_logical;
} vgmi_geom_desc;

typedef struct vgmi_geom vgmi_geom;

#define VGMI_GEOM_HOST_ONLY 1 /* build the host tables only (no device upload) */

VGMI_API vgmi_status vgmi_geom_build(const vgmi_geom_desc *desc, int flags, vgmi_geom **out);
VGMI_API vgmi_status vgmi_geom_destroy(vgmi_geom *geom);
/* counts[0..5] = placed volumes, logical volumes, NavIndex words, max depth, world placed id, device bytes */
VGMI_API vgmi_status vgmi_geom_info(const vgmi_geom *geom, int64_t counts[6]);
/* Host copy of the NavIndex table (uint32 words); returns the word count in *nwords. */
VGMI_API vgmi_status vgmi_geom_navindex(const vgmi_geom *geom, uint32_t *out, int64_t cap, int64_t *nwords);
/* BVH arrays of one logical volume (counts = depth, nodes, daughters); any output may be NULL. */
VGMI_API vgmi_status vgmi_geom_bvh(const vgmi_geom *geom, int logical, int counts[3], int *nchild, int *offset,
                                   int *primid, double *nodes6, double *aabbs6);

/*
 * GDML ingestion: vgdml::Frontend::Load / Parser::Load (persistency/gdml/source/include/Frontend.h:12-21,
 * persistency/gdml/source/src/Middleware.cpp) without Xerces. The file's solids, logical volumes and
 * placements become a vgmi_geom_desc for vgmi_geom_build: logical volumes in document order,
 * placement k from the k-th <physvol>, solids as GeoManager::MakeInstance returns them. In the
 * reference's default build (VECGEOM_NO_SPECIALIZATION=ON, CMakeLists.txt:155) that is every solid as
 * written; flags VGMI_GDML_SPECIALIZED applies the substitutions of a build with it OFF (a full sphere
 * as an orb, a cone or hype with straight sides as a tube, a Trd or trapezoid with equal ends as a
 * box, a trapezoid of equal widths and tilts as a parallelepiped, ...). Lengths are scaled by mm_unit
 * per millimetre (the reference's default 0.1: centimetres; 1.0 keeps millimetres). Supported solids:
 * box, orb, tube, cone, polycone, polyhedra, torus, sphere, trd, trap, para, tet, paraboloid, hype,
 * tessellated; any other solid, a reflected placement or a malformed document fails with
 * VGMI_ERR_INVALID and a message. Materials and auxiliary data are skipped.
 */
#define VGMI_GDML_SPECIALIZED 1
typedef struct vgmi_gdml vgmi_gdml;
VGMI_API vgmi_status vgmi_gdml_read(const char *path, double mm_unit, int flags, vgmi_gdml **out);
/* Fills *desc with arrays owned by the reader handle (valid until vgmi_gdml_destroy). */
VGMI_API vgmi_status vgmi_gdml_desc(const vgmi_gdml *gdml, vgmi_geom_desc *desc);
/* Name of logical volume (kind 0) or placed volume (kind 1) `index`; the string is owned by the handle. */
VGMI_API vgmi_status vgmi_gdml_name(const vgmi_gdml *gdml, int kind, int index, const char **name);
VGMI_API vgmi_status vgmi_gdml_destroy(vgmi_gdml *gdml);

/*
 * Locate global points from the world down (GlobalLocator::LocateGlobalPoint(world, p,
 * state, true), navigation/GlobalLocator.h:35-82, with BVHLevelLocator at every level,
 * navigation/BVHLevelLocator.h:44-95). nav_out[i] = 0 for points outside the world.
 */
VGMI_API vgmi_status vgmi_nav_locate(const vgmi_geom *geom, int precision, const double *px, const double *py,
                                     const double *pz, uint32_t *nav_out, size_t n, void *stream);

/*
 * Level locate per point: VLevelLocator::LevelLocate(lvol, localpoint, pvol, daughterlocalpoint) and
 * LevelLocateExclVol(lvol, exclvol, ...) of the BVH level locator (navigation/BVHLevelLocator.h:44-77,
 * base/BVH.h:219-344). logical[i] is the logical volume the point (in its frame) is in; excl_placed
 * (may be NULL) the placed volume to skip, -1 for none. daughter_out[i] is the placed volume of the first
 * daughter (BVH order) containing the point, -1 when none does or logical[i] is out of range; dpx/dpy/dpz
 * (all or none NULL) receive the point in that daughter's frame (the input point when none).
 */
VGMI_API vgmi_status vgmi_nav_level_locate(const vgmi_geom *geom, int precision, const int32_t *logical,
                                           const int32_t *excl_placed, const double *px, const double *py,
                                           const double *pz, int32_t *daughter_out, double *dpx, double *dpy,
                                           double *dpz, size_t n, void *stream);

/*
 * Placed-volume SOA queries: VPlacedVolume::Contains / Inside / DistanceToIn / DistanceToOut /
 * SafetyToIn / SafetyToOut over SOA3D (volumes/PlacedVolume.h:273,299,332,372,405,421) and
 * PlacedDistanceToOut (op VGMI_OP_PLACED_DIST_OUT), dispatched as CommonSpecializedVolImplHelper
 * (volumes/SpecializedPlacedVolImplHelper.h:98-179): Contains, Inside, DistanceToIn, SafetyToIn and
 * PlacedDistanceToOut take points (and directions) in the frame the volume is placed in and transform
 * them with its placement; DistanceToOut and SafetyToOut take points in the volume's own frame
 * (PlacedVolume.h:335-345, 408-416). placed_id is a placement index of the description, or the
 * world's id (vgmi_geom_info counts[4]). Other arguments as vgmi_shape_query.
 */
#define VGMI_OP_PLACED_DIST_OUT 6
VGMI_API vgmi_status vgmi_placed_query(const vgmi_geom *geom, int placed_id, int op, int precision,
                                       const double *px, const double *py, const double *pz, const double *dx,
                                       const double *dy, const double *dz, const double *step_max, void *out,
                                       size_t n, void *stream);

/*
 * Surface normal of a placed volume: VPlacedVolume::Normal (source/PlacedVolume.cpp:126-137). Points
 * are in the frame the volume is placed in; the normal is returned in that frame (the unplaced
 * Normal of the local point, started from a zero vector, rotated back). placed_id as in
 * vgmi_placed_query. valid[i] as in vgmi_shape_normal. FP64 only.
 */
VGMI_API vgmi_status vgmi_placed_normal(const vgmi_geom *geom, int placed_id, const double *px, const double *py,
                                        const double *pz, double *nx, double *ny, double *nz, uint8_t *valid,
                                        size_t n, void *stream);

/*
 * One navigation step per track: VNavigatorHelper<BVHNavigator>::ComputeStepAndPropagatedState
 * (relocate = 1) or ComputeStep (relocate = 0) (navigation/VNavigator.h:436-540).
 * Inputs are global points/directions, per-track step limits and in-states;
 * outputs are the step, the out-state (nav, last exited, boundary flag).
 * last_in may be NULL (no last exited volume).
 */
VGMI_API vgmi_status vgmi_nav_step(const vgmi_geom *geom, int precision, const double *px, const double *py,
                                   const double *pz, const double *dx, const double *dy, const double *dz,
                                   const double *step_limit, const uint32_t *nav_in, const uint32_t *last_in,
                                   uint32_t *nav_out, uint32_t *last_out, uint8_t *boundary_out, double *step_out,
                                   size_t n, int relocate, void *stream);

/*
 * Step and isotropic safety per track: VNavigatorHelper<BVHNavigator>::
 * ComputeStepAndSafetyAndPropagatedState (navigation/VNavigator.h:1014-1064), the batched form
 * ComputeStepsAndSafetiesAndPropagatedStates (:136-139, :947-...). safety_out[i] is the safety of the
 * start point (BVHSafetyEstimator::ComputeSafetyForLocalPoint) when calc_safety[i] is non-zero,
 * else 0; calc_safety may be NULL (every track). The step and out-state are those of vgmi_nav_step
 * with relocate = 1.
 */
VGMI_API vgmi_status vgmi_nav_step_safety(const vgmi_geom *geom, int precision, const double *px, const double *py,
                                          const double *pz, const double *dx, const double *dy, const double *dz,
                                          const double *step_limit, const uint32_t *nav_in, const uint32_t *last_in,
                                          const uint8_t *calc_safety, uint32_t *nav_out, uint32_t *last_out,
                                          uint8_t *boundary_out, double *step_out, double *safety_out, size_t n,
                                          void *stream);

/* precision (all vgmi_nav_* calls): VGMI_FP64, or VGMI_MIXED, which tests BVH boxes in FP32 against
 * widened FP32 copies as conservative filters in front of the same FP64 decisions (results are the
 * FP64 results bit for bit). */

/* ---- NavStateIndex <-> path interop (callers that keep NavStatePath-style states: one placed
 * volume per level, VecGeom/navigation/NavStatePath.h:104-400). Per-level arrays are level-major:
 * entry k of state i is at [k * n + i]. No precision argument: these walk the NavIndex table and
 * (global_to_local) apply its FP64 cached matrices. */
#define VGMI_PATH_INDICES 0 /* daughter indices below the world (NavStateIndex::GetPathAsListOfIndices) */
#define VGMI_PATH_PLACED 1  /* placed-volume ids, world first (NavStatePath::ValueAt per level) */
/* levels[i] = GetCurrentLevel() (NavStateIndex.h:315-317; 0 for the outside state); pvols[k] = ValueAt(k)
 * (NavStateIndex.h:341-349); indices[k] = GetPathAsListOfIndices (source/NavStateIndex.cpp:40-53: 0 at
 * level 0, then each level's position in its mother's daughter list); -1 beyond the state's levels.
 * pvols and indices may be NULL. max_levels must be at least the geometry's depth (VGMI_ERR_CAPACITY). */
VGMI_API vgmi_status vgmi_nav_to_path(const vgmi_geom *geom, const uint32_t *nav, size_t n, int max_levels,
                                      int32_t *levels, int32_t *pvols, int32_t *indices, void *stream);
/* The NavIndex of a path: kind VGMI_PATH_INDICES is NavStateIndex::ResetPathFromListOfIndices
 * (source/NavStateIndex.cpp:77-88: the world, then daughter path[k] of each level; the first entry is
 * not read); kind VGMI_PATH_PLACED pushes placed volumes level by level (NavStateIndex::Push,
 * NavStateIndex.h:282-287), the first being the world. levels[i] entries of path are read (at most
 * max_levels). A path the geometry does not hold (an index beyond a volume's daughters, a placed volume
 * that is not a daughter of the previous level's volume, a first placed volume other than the world)
 * and levels[i] == 0 give the outside state 0; the reference does not check these. */
VGMI_API vgmi_status vgmi_nav_from_path(const vgmi_geom *geom, int kind, const int32_t *levels, const int32_t *path,
                                        size_t n, int max_levels, uint32_t *nav_out, void *stream);
/* NavStateIndex::Distance (source/NavStateIndex.cpp:150-166): moves up from nav_a[i] to the deepest
 * level both paths share, plus moves down to nav_b[i]. The outside state counts as having no level
 * (the reference reads an undefined level for it). */
VGMI_API vgmi_status vgmi_nav_distance(const vgmi_geom *geom, const uint32_t *nav_a, const uint32_t *nav_b,
                                       size_t n, int32_t *out, void *stream);
/* NavStateIndex::GlobalToLocal(point, tolevel) (NavStateIndex.h:262-269, 368-377): the global point in
 * the frame of the state's ancestor at level tolevel (tolevel < 0: of the state's top volume), through
 * that ancestor's cached matrix. A level deeper than the state's own reaches the outside state (identity), as
 * the reference's GetNavIndexImpl does. */
VGMI_API vgmi_status vgmi_nav_global_to_local(const vgmi_geom *geom, const uint32_t *nav, int tolevel,
                                              const double *px, const double *py, const double *pz, double *lx,
                                              double *ly, double *lz, size_t n, void *stream);

/* Isotropic safety: VSafetyEstimatorHelper<BVHSafetyEstimator>::ComputeSafety
 * (navigation/VSafetyEstimator.h:98-107, navigation/BVHSafetyEstimator.h:38-46). */
VGMI_API vgmi_status vgmi_nav_safety(const vgmi_geom *geom, int precision, const double *px, const double *py,
                                     const double *pz, const uint32_t *nav_in, double *safety_out, size_t n,
                                     void *stream);

/*
 * Transport loop fused on the device: for each track, up to nsteps times
 * { step = ComputeStepAndPropagatedState(p, d, limit, state); p += step * d; state = out },
 * stopping when the track leaves the world. p, nav and last are updated in place;
 * steps_taken (may be NULL) receives the per-track number of steps.
 * No single reference counterpart: the loop a transport engine runs around
 * VNavigator::ComputeStepAndPropagatedState.
 */
VGMI_API vgmi_status vgmi_nav_propagate(const vgmi_geom *geom, int precision, double *px, double *py, double *pz,
                                        const double *dx, const double *dy, const double *dz, double step_limit,
                                        uint32_t *nav, uint32_t *last, int32_t *steps_taken, size_t n, int nsteps,
                                        void *stream);

/*
 * vgmi_nav_propagate that allocates nothing. No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). The regrouping workspace is a fixed device array of the library (8.6 MB per GPU, held
 * until the process ends) instead of a per-call allocation, so at most 2^21 tracks and 2^16 logical
 * volumes (VGMI_ERR_CAPACITY beyond). Every call on one GPU uses that GPU's one array: calls that
 * overlap on one GPU (other streams, other threads) race and must be serialised by the caller.
 * Results are those of vgmi_nav_propagate.
 */
VGMI_API vgmi_status vgmi_nav_propagate_noalloc(const vgmi_geom *geom, int precision, double *px, double *py,
                                                double *pz, const double *dx, const double *dy, const double *dz,
                                                double step_limit, uint32_t *nav, uint32_t *last,
                                                int32_t *steps_taken, size_t n, int nsteps, void *stream);

/*
 * Regrouping of the fused transport loops (vgmi_nav_propagate, the serving loop) of at least
 * 65536 tracks: every chunk_steps steps the tracks are counting-sorted on the device by the
 * logical volume they are in, so the lanes of a wavefront run the same volume's code; 0
 * disables it (default 10).
 * Results per track do not depend on it. No reference counterpart (a GPU scheduling knob).
 */
VGMI_API vgmi_status vgmi_geom_set_regroup(vgmi_geom *geom, int chunk_steps);
/*
 * Tracks per wavefront in the same loops (1, 2, 4, ..., 64; 0 = automatic, the default:
 * as few as keep every wavefront slot of the GPU busy). A wavefront executes the union of
 * its tracks' code paths, so fewer tracks per wavefront cut the latency of small batches.
 * Results per track do not depend on it. No reference counterpart (a GPU scheduling knob).
 */
VGMI_API vgmi_status vgmi_geom_set_lanes_per_wave(vgmi_geom *geom, int lanes);

/* ------------------------------------------------------------------------
 * Serving loop (continuous batching). No reference counterpart: the reference
 * is called synchronously by a transport engine; this is the request queue a
 * tracking service puts in front of vgmi_nav_propagate-style work.
 *
 * Each request is a set of tracks given as host arrays: global points and
 * directions plus a per-track step count. A track is located from the world
 * (vgmi_nav_locate), then stepped up to nsteps[i] times
 * (ComputeStepAndPropagatedState, navigation/VNavigator.h:436-540) or until it
 * leaves the world. Results go to the caller's host arrays, which must stay
 * valid (inputs unmodified) until vgmi_server_wait returns for the ticket.
 *
 * Requests from many threads are packed into launches of at most lane_tracks
 * tracks; `lanes` launches (each with its own stream, pinned staging and device
 * buffers) are in flight at once. Admission: a request with more than
 * lane_tracks tracks, or one that would put more than queue_tracks tracks in
 * the queue, is refused with VGMI_ERR_CAPACITY; creation fails with
 * VGMI_ERR_CAPACITY when the lanes would take more than half of the free HBM;
 * vgmi_server_set_hbm_reserve adds a live free-HBM check to every submission and batch.
 * precision is VGMI_FP64 or VGMI_MIXED (the navigation kernels' mixed mode: FP32 box filters,
 * results identical to FP64).
 * ------------------------------------------------------------------------ */
typedef struct vgmi_server vgmi_server;

VGMI_API vgmi_status vgmi_server_create(const vgmi_geom *geom, int precision, double step_limit,
                                        int64_t lane_tracks, int lanes, int64_t queue_tracks, vgmi_server **out);
/* nav_out and steps_out may be NULL. *ticket identifies the request for vgmi_server_wait. */
VGMI_API vgmi_status vgmi_server_submit(vgmi_server *server, const double *px, const double *py, const double *pz,
                                        const double *dx, const double *dy, const double *dz, const int32_t *nsteps,
                                        int64_t n, double *px_out, double *py_out, double *pz_out,
                                        uint32_t *nav_out, int32_t *steps_out, uint64_t *ticket);
/* Blocks until the request is done; latency_us (may be NULL) = submit-to-completion time.
 * Each ticket is waited for exactly once. Returns VGMI_ERR_CAPACITY for a request the server was
 * destroyed before launching. */
VGMI_API vgmi_status vgmi_server_wait(vgmi_server *server, uint64_t ticket, double *latency_us);
/* Non-blocking completion test: *done = 1 once the request has finished (vgmi_server_wait then
 * returns at once), else 0. The ticket stays valid until it is waited for. */
VGMI_API vgmi_status vgmi_server_poll(vgmi_server *server, uint64_t ticket, int *done);
/* Cancels a request that is still queued: it finishes at once and vgmi_server_wait returns
 * VGMI_ERR_CAPACITY for it (the ticket must still be waited for). A request already on the GPU
 * or done is not cancelled (VGMI_ERR_UNSUPPORTED). */
VGMI_API vgmi_status vgmi_server_cancel(vgmi_server *server, uint64_t ticket);
/* Admission by free HBM: while the runtime reports less than `bytes` of free device memory
 * (hipMemGetInfo, queried at each submission and before each batch is launched), submissions are
 * refused with VGMI_ERR_CAPACITY, and queued requests whose batch comes up are finished with
 * VGMI_ERR_CAPACITY instead of being launched. 0 (the default) turns the check off. */
VGMI_API vgmi_status vgmi_server_set_hbm_reserve(vgmi_server *server, int64_t bytes);
/* stats[0..3] = launches, tracks done, steps done, requests refused */
VGMI_API vgmi_status vgmi_server_stats(vgmi_server *server, int64_t stats[4]);
/* Launches whose tracks need at most 100 steps replay HIP graphs captured at creation (one per
 * power-of-two track-count bucket from 4096 up to lane_tracks; the kernels read the actual count
 * from device memory). stats[0] = graph launches so far, stats[1] = graphs per lane. */
VGMI_API vgmi_status vgmi_server_graph_stats(vgmi_server *server, int64_t stats[2]);
/* Stops the threads after the launches in flight complete. Requests still queued are never launched:
 * their waiters (threads blocked in vgmi_server_wait) return VGMI_ERR_CAPACITY, and destroy returns
 * once every waiter has left. No other call on the server may start once destroy has begun. */
VGMI_API vgmi_status vgmi_server_destroy(vgmi_server *server);

#ifdef __cplusplus
}
#endif
#endif /* VGMI_H */
