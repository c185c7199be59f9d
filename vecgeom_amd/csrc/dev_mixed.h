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

// The FP32 classification of one point, shared by every method asked of it.
struct MixedState {
  Vec3<float> pf;
  float sin32, sout32;
  bool near;
};

template <class Data>
VGMI_DEV MixedState mixed_state(Data const &s, Vec3<double> const &p)
{
  using S = ShapeOps<Data>;
  MixedState m;
  m.pf     = Vec3<float>(float(p.x), float(p.y), float(p.z));
  m.sin32  = S::template SafeIn<float>(s, m.pf);
  m.sout32 = S::template SafeOut<float>(s, m.pf);
  // A consistent FP32 state has safeties of opposite sign (inside: SafetyToIn < 0 < SafetyToOut,
  // outside: the reverse). Equal signs mean the FP32 tolerance bands collapsed onto the surface
  // (a point within FP32 rounding of it is then both inside and outside): treat as near.
  // An FP32 safety that cannot be decided (NaN) sends the point to FP64 as well.
  m.near = vabs(m.sin32) < kMixBandF || vabs(m.sout32) < kMixBandF || ((m.sin32 < 0.f) == (m.sout32 < 0.f)) ||
           m.sin32 != m.sin32 || m.sout32 != m.sout32 || F32PointUndecided(s, m.pf);
  return m;
}

// Contains (uint8), Inside (int32) or a distance/safety (double) of one point whose FP32 state is m.
// Directed queries of shapes without certified FP32 distances do not use m.
template <class Data, int OP>
VGMI_DEV double mixed_answer(Data const &s, MixedState const &m, Vec3<double> const &p, Vec3<double> const &d)
{
  using S = ShapeOps<Data>;
  if constexpr ((OP == kOpDistIn || OP == kOpDistOut) && !CertifiedF32Dist<Data>::value) {
    return OP == kOpDistIn ? S::template DistIn<double>(s, p, d) : S::template DistOut<double>(s, p, d);
  } else if constexpr (OP == kOpContains) {
    return (m.near ? S::template Contains<double>(s, p) : S::template Contains<float>(s, m.pf)) ? 1.0 : 0.0;
  } else if constexpr (OP == kOpInside) {
    return double(m.near ? S::template Inside<double>(s, p) : S::template Inside<float>(s, m.pf));
  } else if constexpr (OP == kOpSafetyIn) {
    return m.near ? S::template SafeIn<double>(s, p) : double(m.sin32);
  } else if constexpr (OP == kOpSafetyOut) {
    return m.near ? S::template SafeOut<double>(s, p) : double(m.sout32);
  } else {
    double r  = 0.;
    bool fp64 = m.near;
    if (!m.near) {
      const Vec3<float> df(float(d.x), float(d.y), float(d.z));
      const float f =
          OP == kOpDistIn ? S::template DistIn<float>(s, m.pf, df) : S::template DistOut<float>(s, m.pf, df);
      r    = f >= kMixFltBig ? kInfLength : double(f);
      fp64 = f != f;
    }
    if (fp64) r = OP == kOpDistIn ? S::template DistIn<double>(s, p, d) : S::template DistOut<double>(s, p, d);
    return r;
  }
}

// Contains (uint8), Inside (int32) or a distance/safety (double) of one point.
template <class Data, int OP>
VGMI_DEV double mixed_point(Data const &s, Vec3<double> const &p, Vec3<double> const &d)
{
  if constexpr ((OP == kOpDistIn || OP == kOpDistOut) && !CertifiedF32Dist<Data>::value)
    return mixed_answer<Data, OP>(s, MixedState{}, p, d);
  else
    return mixed_answer<Data, OP>(s, mixed_state(s, p), p, d);
}

} // namespace vgmi
