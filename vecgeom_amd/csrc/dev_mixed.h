// Mixed-precision evaluation of one point: FP32 bulk arithmetic with FP64 re-evaluation
// of every point within the near-surface band and of every FP32 answer that cannot be
// decided from FP32 inputs.
//
// The lane converts its point (and direction) to FP32 once and evaluates the method and
// the two safeties in FP32. A point whose FP32 safeties put it within kMixedBand of a
// surface is evaluated again by the FP64 kernel, so every decision that depends on the
// tolerance bands (Inside/Contains, on-surface conventions, wrong-side -1) is the FP64
// one; elsewhere the FP32 rounding error (about 1e-7 of the local coordinates) is far
// below the distance to any surface, so Inside/Contains cannot flip. Directed queries can
// still meet a surface edge or a tangent point at grazing incidence far from the start
// point: the certified FP32 kernels (CertifiedF32Dist below) carry a rounding bound with every
// quantity that feeds a decision and answer NaN (the helper Undecided) when a test's margin
// does not exceed its bound or a distance's bound exceeds kMixDistErr, and that point is
// evaluated in FP64 as well, so a hit or miss, and the surface that is hit, are the FP64
// ones.
#pragma once
#include "dev_shapes.h"

namespace vgmi {

// Shapes whose FP32 DistanceToIn / DistanceToOut are certified: every hit/miss decision, every
// choice of the surface that is hit and every returned distance is the FP64 kernel's (the distance
// within kMixDistErr) or the FP32 pass answers NaN. Directed queries of the other shapes (cone,
// polycone, sphere, polyhedron, torus, hype, tessellated) run the FP64 kernel in the mixed path.
template <class Data>
struct CertifiedF32Dist {
  static constexpr bool value = false;
};
template <>
struct CertifiedF32Dist<BoxData> {
  static constexpr bool value = true;
};
template <>
struct CertifiedF32Dist<TubeData> {
  static constexpr bool value = true;
};
template <>
struct CertifiedF32Dist<TrapezoidData> {
  static constexpr bool value = true;
};
template <>
struct CertifiedF32Dist<OrbData> {
  static constexpr bool value = true;
};

// Points an FP32 pass cannot classify although they are far from the surface (the reference's
// answer changes across an internal plane): evaluated in FP64 like near-surface points.
template <class Data>
VGMI_DEV bool F32PointUndecided(Data const &, Vec3<float> const &)
{
  return false;
}
template <>
VGMI_DEV bool F32PointUndecided<PolyhedronData>(PolyhedronData const &s, Vec3<float> const &p)
{
  return phed::F32PointUndecided(s, p);
}

constexpr float kMixBandF = float(kMixedBand);
constexpr float kMixFltBig = 1e30f;

// Contains (uint8), Inside (int32) or a distance/safety (double) of one point.
template <class Data, int OP>
VGMI_DEV double mixed_point(Data const &s, Vec3<double> const &p, Vec3<double> const &d)
{
  using S = ShapeOps<Data>;
  if constexpr ((OP == kOpDistIn || OP == kOpDistOut) && !CertifiedF32Dist<Data>::value)
    return OP == kOpDistIn ? S::template DistIn<double>(s, p, d) : S::template DistOut<double>(s, p, d);
  const Vec3<float> pf(float(p.x), float(p.y), float(p.z));
  const float sin32  = S::template SafeIn<float>(s, pf);
  const float sout32 = S::template SafeOut<float>(s, pf);
  // A consistent FP32 state has safeties of opposite sign (inside: SafetyToIn < 0 < SafetyToOut,
  // outside: the reverse). Equal signs mean the FP32 tolerance bands collapsed onto the surface
  // (a point within FP32 rounding of it is then both inside and outside): treat as near.
  // An FP32 safety that cannot be decided (NaN) sends the point to FP64 as well.
  const bool near = vabs(sin32) < kMixBandF || vabs(sout32) < kMixBandF || ((sin32 < 0.f) == (sout32 < 0.f)) ||
                    sin32 != sin32 || sout32 != sout32 || F32PointUndecided(s, pf);
  if constexpr (OP == kOpContains) {
    return (near ? S::template Contains<double>(s, p) : S::template Contains<float>(s, pf)) ? 1.0 : 0.0;
  } else if constexpr (OP == kOpInside) {
    return double(near ? S::template Inside<double>(s, p) : S::template Inside<float>(s, pf));
  } else if constexpr (OP == kOpSafetyIn) {
    return near ? S::template SafeIn<double>(s, p) : double(sin32);
  } else if constexpr (OP == kOpSafetyOut) {
    return near ? S::template SafeOut<double>(s, p) : double(sout32);
  } else {
    double r  = 0.;
    bool fp64 = near;
    if (!near) {
      const Vec3<float> df(float(d.x), float(d.y), float(d.z));
      const float f = OP == kOpDistIn ? S::template DistIn<float>(s, pf, df) : S::template DistOut<float>(s, pf, df);
      r             = f >= kMixFltBig ? kInfLength : double(f);
      fp64          = f != f;
    }
    if (fp64) r = OP == kOpDistIn ? S::template DistIn<double>(s, p, d) : S::template DistOut<double>(s, p, d);
    return r;
  }
}

} // namespace vgmi
