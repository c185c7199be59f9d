// All device shape kernels, a per-record-type trait for statically dispatched
// launches, and run-time dispatch over a ShapeRecord for navigation and mixed batches.
#pragma once
#include "dev_box.h"
#include "dev_cone.h"
#include "dev_cuttube.h"
#include "dev_ellipsoids.h"
#include "dev_gentrap.h"
#include "dev_hype.h"
#include "dev_orb.h"
#include "dev_para_tet.h"
#include "dev_paraboloid.h"
#include "dev_polycone.h"
#include "dev_polyhedron.h"
#include "dev_sphere.h"
#include "dev_tessellated.h"
#include "dev_torus.h"
#include "dev_trapezoid.h"
#include "dev_trd.h"
#include "dev_tube.h"

// (enum value, record type, union member, kernel namespace)
#define VGMI_FOR_EACH_SHAPE(X)                                                                                       \
  VGMI_FOR_EACH_SHAPE_BUT_TESSELLATED(X)                                                                             \
  X(kTessellated, TessellatedData, tess, tess)

#define VGMI_FOR_EACH_SHAPE_BUT_TESSELLATED(X)                                                                       \
  X(kBox, BoxData, box, box)                                                                                         \
  X(kTube, TubeData, tube, tube)                                                                                     \
  X(kCone, ConeData, cone, cone)                                                                                     \
  X(kTrapezoid, TrapezoidData, trap, trap)                                                                           \
  X(kOrb, OrbData, orb, orb)                                                                                         \
  X(kPolycone, PolyconeData, pcon, pcon)                                                                             \
  X(kSphere, SphereData, sph, sph)                                                                                   \
  X(kPolyhedron, PolyhedronData, phed, phed)                                                                         \
  X(kTorus2, TorusData, tor, tor)                                                                                    \
  X(kHype, HypeData, hype, hype)                                                                                     \
  X(kTrd, TrdData, trd, trd)                                                                                         \
  X(kParaboloid, ParaboloidData, para, para)                                                                         \
  X(kParallelepiped, ParallelepipedData, ppd, ppd)                                                                   \
  X(kTet, TetData, tet, tet)                                                                                         \
  X(kEllipticalTube, EllipticalTubeData, eltu, eltu)                                                                 \
  X(kEllipsoid, EllipsoidData, ells, ells)                                                                           \
  X(kGenTrap, GenTrapData, gtra, gtra)                                                                               \
  X(kCutTube, CutTubeData, ctub, ctub)                                                                               \
  X(kEllipticalCone, EllipticalConeData, elco, elco)

namespace vgmi {

template <class Data>
struct ShapeOps;

#define VGMI_SHAPE_OPS(E, T, M, NS)                                                                                   \
  template <>                                                                                                        \
  struct ShapeOps<T> {                                                                                               \
    static constexpr int kType = E;                                                                                  \
    static T const &get(ShapeRecord const &r) { return r.u.M; }                                                     \
    template <class R>                                                                                               \
    static VGMI_DEV bool Contains(T const &s, Vec3<R> const &p)                                                      \
    {                                                                                                                \
      return NS::Contains(s, p);                                                                                     \
    }                                                                                                                \
    template <class R>                                                                                               \
    static VGMI_DEV int Inside(T const &s, Vec3<R> const &p)                                                         \
    {                                                                                                                \
      return NS::Inside(s, p);                                                                                       \
    }                                                                                                                \
    template <class R>                                                                                               \
    static VGMI_DEV R DistIn(T const &s, Vec3<R> const &p, Vec3<R> const &d)                                         \
    {                                                                                                                \
      return NS::DistanceToIn(s, p, d);                                                                              \
    }                                                                                                                \
    template <class R>                                                                                               \
    static VGMI_DEV R DistOut(T const &s, Vec3<R> const &p, Vec3<R> const &d)                                        \
    {                                                                                                                \
      return NS::DistanceToOut(s, p, d);                                                                             \
    }                                                                                                                \
    template <class R>                                                                                               \
    static VGMI_DEV R SafeIn(T const &s, Vec3<R> const &p)                                                           \
    {                                                                                                                \
      return NS::SafetyToIn(s, p);                                                                                   \
    }                                                                                                                \
    template <class R>                                                                                               \
    static VGMI_DEV R SafeOut(T const &s, Vec3<R> const &p)                                                          \
    {                                                                                                                \
      return NS::SafetyToOut(s, p);                                                                                  \
    }                                                                                                                \
  };
VGMI_FOR_EACH_SHAPE(VGMI_SHAPE_OPS)
#undef VGMI_SHAPE_OPS

// ---- run-time dispatch on a record (FP64). TypeMask is a bit mask of the shape types compiled
// in (bit E for enum value E); a kernel that serves a geometry made of a few simple types
// leaves the large solvers out of its register allocation.
template <unsigned TypeMask = ~0u>
VGMI_DEV double rec_dist_in(ShapeRecord const &s, Vec3<double> const &p, Vec3<double> const &d)
{
  switch (s.type) {
#define X(E, T, M, NS)                                                                                               \
  case E:                                                                                                            \
    if constexpr ((TypeMask >> E) & 1u) return NS::DistanceToIn(s.u.M, p, d);                                        \
    break;
    VGMI_FOR_EACH_SHAPE(X)
#undef X
  }
  return kInfLength;
}
template <unsigned TypeMask = ~0u>
VGMI_DEV double rec_dist_out(ShapeRecord const &s, Vec3<double> const &p, Vec3<double> const &d)
{
  switch (s.type) {
#define X(E, T, M, NS)                                                                                               \
  case E:                                                                                                            \
    if constexpr ((TypeMask >> E) & 1u) return NS::DistanceToOut(s.u.M, p, d);                                       \
    break;
    VGMI_FOR_EACH_SHAPE(X)
#undef X
  }
  return -1.;
}
template <unsigned TypeMask = ~0u>
VGMI_DEV bool rec_contains(ShapeRecord const &s, Vec3<double> const &p)
{
  switch (s.type) {
#define X(E, T, M, NS)                                                                                               \
  case E:                                                                                                            \
    if constexpr ((TypeMask >> E) & 1u) return NS::Contains(s.u.M, p);                                               \
    break;
    VGMI_FOR_EACH_SHAPE(X)
#undef X
  }
  return false;
}
template <unsigned TypeMask = ~0u>
VGMI_DEV int rec_inside(ShapeRecord const &s, Vec3<double> const &p)
{
  switch (s.type) {
#define X(E, T, M, NS)                                                                                               \
  case E:                                                                                                            \
    if constexpr ((TypeMask >> E) & 1u) return NS::Inside(s.u.M, p);                                                 \
    break;
    VGMI_FOR_EACH_SHAPE(X)
#undef X
  }
  return kOutside;
}
template <unsigned TypeMask = ~0u>
VGMI_DEV double rec_safety_in(ShapeRecord const &s, Vec3<double> const &p)
{
  switch (s.type) {
#define X(E, T, M, NS)                                                                                               \
  case E:                                                                                                            \
    if constexpr ((TypeMask >> E) & 1u) return NS::SafetyToIn(s.u.M, p);                                             \
    break;
    VGMI_FOR_EACH_SHAPE(X)
#undef X
  }
  return kInfLength;
}
template <unsigned TypeMask = ~0u>
VGMI_DEV double rec_safety_out(ShapeRecord const &s, Vec3<double> const &p)
{
  switch (s.type) {
#define X(E, T, M, NS)                                                                                               \
  case E:                                                                                                            \
    if constexpr ((TypeMask >> E) & 1u) return NS::SafetyToOut(s.u.M, p);                                            \
    break;
    VGMI_FOR_EACH_SHAPE(X)
#undef X
  }
  return 0.;
}

} // namespace vgmi
