// Oracle: CPU scalar FP64 restatement of VecGeom's scalar backend.
//
// TEST INFRASTRUCTURE ONLY. This directory is the checker: the parity tests,
// the smoke check and the CPU-baseline leg of the benchmark call it. The
// product library (vecgeom_amd/) never links, loads or calls anything here.
//
// Constants and helpers follow VecGeom/base/Math.h:26-53, VecGeom/base/Global.h
// :71-88,152-177 and VecGeom/base/Vector3D.h:144-290 (operation order kept so
// FP64 results can be compared bit for bit with the GPU kernels, which are built
// with FP contraction disabled).
#pragma once
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <algorithm>

namespace vgo {

constexpr double kTolerance     = 1e-9;          // base/Math.h:27
constexpr double kHalfTolerance = 0.5 * kTolerance;
constexpr double kPushTolerance = 1e-6;          // base/Math.h:28
constexpr double kSqrtTolerance = 3.1622777e-5;  // base/Math.h:29
constexpr double kAngTolerance  = 1e-9;          // base/Math.h:30
constexpr double kConeTolerance = 1e-7;          // base/Math.h:31
constexpr double kFarAway       = 1e10;          // base/Math.h:32
constexpr double kInfLength     = DBL_MAX;       // base/Math.h:41
constexpr double kMinimum       = DBL_MIN;       // base/Math.h:43
constexpr double kEpsilon       = DBL_EPSILON;   // base/Math.h:38
constexpr double kPi            = 3.14159265358979323846;
constexpr double kHalfPi        = 0.5 * kPi;
constexpr double kTwoPi         = 2. * kPi;
constexpr double kDegToRad      = kPi / 180.;
constexpr double kTiny          = 1e-30;

// EInside values, base/Global.h:71-88
constexpr int kInside  = 1;
constexpr int kSurface = 2;
constexpr int kOutside = 3;

struct V3 {
  double x, y, z;
  V3() : x(0), y(0), z(0) {}
  V3(double a, double b, double c) : x(a), y(b), z(c) {}
  double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
  double &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
  V3 operator+(V3 const &o) const { return V3(x + o.x, y + o.y, z + o.z); }
  V3 operator-(V3 const &o) const { return V3(x - o.x, y - o.y, z - o.z); }
  V3 operator*(V3 const &o) const { return V3(x * o.x, y * o.y, z * o.z); }
  V3 operator*(double s) const { return V3(x * s, y * s, z * s); }
  V3 operator-() const { return V3(-x, -y, -z); }
  double Dot(V3 const &o) const { return x * o.x + y * o.y + z * o.z; }
  double Mag2() const { return x * x + y * y + z * z; }
  double Mag() const { return std::sqrt(Mag2()); }
  double Perp2() const { return x * x + y * y; }
  double Perp() const { return std::sqrt(Perp2()); }
  V3 Abs() const { return V3(std::fabs(x), std::fabs(y), std::fabs(z)); }
  double Max() const { return std::max(x, std::max(y, z)); }
  double Min() const { return std::min(x, std::min(y, z)); }
  V3 Cross(V3 const &o) const { return V3(y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x); }
};
inline V3 operator*(double s, V3 const &v) { return V3(s * v.x, s * v.y, s * v.z); }

// base/Global.h:163-177
inline double NonZero(double x) { return x + std::copysign(1.0e-30, x); }
inline double NonZeroAbs(double x) { return std::fabs(x) + 1.0e-30; }
// VecCore math::Sign is CopySign(1, x): -0.0 maps to -1.
inline double Sign(double x) { return std::copysign(1.0, x); }
// MakePlusTolerantSquare<true> / MakeMinusTolerantSquare<true> (volumes/kernel/GenericKernels.h:65-82)
inline double PlusTolSq(double x) { return x * (x + 2.0 * kTolerance); }
inline double MinusTolSq(double x) { return std::fabs(x * (x - 2.0 * kTolerance)); }

} // namespace vgo
