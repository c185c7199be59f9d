// Device-side constants and small vector helpers for the CDNA4 kernels.
//
// The kernels are templated on the arithmetic type R (double for the FP64 path,
// float for the bulk of the mixed path). Tolerances follow
// VecGeom/base/Math.h:26-53; min/max use the (a < b) ? b : a form so FP64
// results match a scalar host evaluation bit for bit (the library is compiled
// with -ffp-contract=off; FP32 code asks for fused multiply-add explicitly).
#pragma once
#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

namespace vgmi {

#ifndef VGMI_DEV
#define VGMI_DEV __device__ __forceinline__
#endif

constexpr double kTolerance     = 1e-9;
constexpr double kHalfTolerance = 0.5 * kTolerance;
constexpr double kPushTolerance = 1e-6;
constexpr double kConeTolerance = 1e-7;
// Mixed precision: points whose FP32 safeties put them within this distance (mm) of a
// surface, and FP32 section hand-offs this close to an internal plane, are evaluated in FP64.
constexpr double kMixedBand      = 1e-3;
constexpr double kMixedPlaneBand = 1e-3;
constexpr double kInfLength     = DBL_MAX;
constexpr double kTiny          = 1e-30; // base/Global.h kTiny: keeps a denominator off zero
constexpr double kPi            = 3.14159265358979323846;
constexpr double kTwoPi         = 2. * kPi;

constexpr int kInside  = 1;
constexpr int kSurface = 2;
constexpr int kOutside = 3;

template <class R>
struct Vec3 {
  R x, y, z;
  VGMI_DEV Vec3() {}
  VGMI_DEV Vec3(R a, R b, R c) : x(a), y(b), z(c) {}
  VGMI_DEV Vec3 operator+(Vec3 const &o) const { return Vec3(x + o.x, y + o.y, z + o.z); }
  VGMI_DEV Vec3 operator-(Vec3 const &o) const { return Vec3(x - o.x, y - o.y, z - o.z); }
  VGMI_DEV Vec3 operator*(Vec3 const &o) const { return Vec3(x * o.x, y * o.y, z * o.z); }
  VGMI_DEV Vec3 operator*(R s) const { return Vec3(x * s, y * s, z * s); }
  VGMI_DEV Vec3 operator-() const { return Vec3(-x, -y, -z); }
  VGMI_DEV R Dot(Vec3 const &o) const { return x * o.x + y * o.y + z * o.z; }
  VGMI_DEV R Mag2() const { return x * x + y * y + z * z; }
  VGMI_DEV R Perp2() const { return x * x + y * y; }
};

template <class R>
VGMI_DEV Vec3<R> operator*(R s, Vec3<R> const &v)
{
  return Vec3<R>(s * v.x, s * v.y, s * v.z);
}

template <class R>
VGMI_DEV R vmax(R a, R b)
{
  return a < b ? b : a;
}
template <class R>
VGMI_DEV R vmin(R a, R b)
{
  return b < a ? b : a;
}
template <class R>
VGMI_DEV R vabs(R a)
{
  return __builtin_fabs(a);
}
VGMI_DEV double vsqrt(double a) { return __builtin_sqrt(a); }
VGMI_DEV float vsqrt(float a) { return __builtin_sqrtf(a); }
VGMI_DEV double vcopysign(double a, double b) { return __builtin_copysign(a, b); }
// library transcendental functions (Torus2 cubic solver)
VGMI_DEV double vcbrt(double a) { return ::cbrt(a); }
VGMI_DEV float vcbrt(float a) { return ::cbrtf(a); }
VGMI_DEV double vpow(double a, double b) { return ::pow(a, b); }
VGMI_DEV float vpow(float a, float b) { return ::powf(a, b); }
VGMI_DEV double vcos(double a) { return ::cos(a); }
VGMI_DEV float vcos(float a) { return ::cosf(a); }
VGMI_DEV double vatan2(double a, double b) { return ::atan2(a, b); }
VGMI_DEV float vatan2(float a, float b) { return ::atan2f(a, b); }
VGMI_DEV float vcopysign(float a, float b) { return __builtin_copysignf(a, b); }

template <class R>
VGMI_DEV R Max3(Vec3<R> const &v)
{
  return vmax(v.x, vmax(v.y, v.z));
}
template <class R>
VGMI_DEV R Min3(Vec3<R> const &v)
{
  return vmin(v.x, vmin(v.y, v.z));
}
template <class R>
VGMI_DEV Vec3<R> Abs3(Vec3<R> const &v)
{
  return Vec3<R>(vabs(v.x), vabs(v.y), vabs(v.z));
}

// Guarded reciprocal argument and sign (VecGeom/base/Global.h:163-177; Sign is CopySign(1, x))
template <class R>
VGMI_DEV R NonZero(R x)
{
  return x + vcopysign(R(1.0e-30), x);
}
template <class R>
VGMI_DEV R NonZeroAbs(R x)
{
  return vabs(x) + R(1.0e-30);
}
template <class R>
VGMI_DEV R Sign(R x)
{
  return vcopysign(R(1), x);
}

// Band for re-checking an intersection point a kernel has just computed (is the hit
// on this plane / half-plane?). FP64 uses the reference's band; in FP32 the point is
// only known to ~1e-7 of its coordinates, so the FP32 pass uses kMixedBand, which
// accepts every genuine hit of that surface.
template <class R>
VGMI_DEV R HitTol(double fp64_band)
{
  return sizeof(R) == sizeof(float) ? R(kMixedBand) : R(fp64_band);
}

// Root (-b + sgn*sqrt|D|) / a of a*t^2 + 2*b*t + c = 0. FP64 keeps the reference's form
// (multiplying by the given 1/a); the FP32 pass switches to the equal c / (-b - sgn*sqrt|D|)
// where -b and sgn*sqrt|D| would cancel.
template <class R>
VGMI_DEV R QuadRoot(R b, R D, R ainv, R c, int sgn)
{
  const R sq = vsqrt(vabs(D));
  if (sizeof(R) == sizeof(float) && ((sgn > 0) == (b > R(0)))) return c / NonZero(-b - R(sgn) * sq);
  return (-b + R(sgn) * sq) * ainv;
}

// ---------------------------------------------------------------- certified FP32 (mixed precision)
// The FP32 pass of a directed query returns either NaN (undecided: evaluate in FP64) or a distance
// with the FP64 kernel's hit/miss decision and within kMixDistErr of its value. Every FP32 quantity
// that feeds a decision carries a rounding bound: with u = 2^-24, a length computed from FP32
// copies of the inputs by a few roundings is within c*u*(sum of the magnitudes involved) of its
// FP64 value (c counted per expression, rounded up). A test is decided only when the margin exceeds
// its bound; parameters along the ray are carried as intervals [lo, hi] holding the FP64 value.
constexpr float kU32        = 5.9604645e-8f;
constexpr float kMixDistErr = float(kMixedBand) * 0.25f;

struct FInt {
  float lo, hi;
};
VGMI_DEV FInt fint(float v, float e) { return FInt{v - e, v + e}; }
VGMI_DEV FInt fmaxi(FInt a, FInt b) { return FInt{a.lo < b.lo ? b.lo : a.lo, a.hi < b.hi ? b.hi : a.hi}; }
VGMI_DEV FInt fmini(FInt a, FInt b) { return FInt{b.lo < a.lo ? b.lo : a.lo, b.hi < a.hi ? b.hi : a.hi}; }
VGMI_DEV float Undecided() { return __builtin_nanf(""); }
// three-valued tests of certified FP32 code: a > b is decided (kYes / kNo) when |a - b| exceeds e,
// the bound of the rounding error of a - b
enum { kNo = 0, kYes = 1, kUnsure = 2 };
VGMI_DEV int cmpGT(float a, float b, float e)
{
  const float m = a - b;
  return m > e ? kYes : (m < -e ? kNo : kUnsure);
}
VGMI_DEV int and3(int a, int b) { return (a == kNo || b == kNo) ? kNo : ((a == kYes && b == kYes) ? kYes : kUnsure); }
VGMI_DEV int or3(int a, int b) { return (a == kYes || b == kYes) ? kYes : ((a == kNo && b == kNo) ? kNo : kUnsure); }
VGMI_DEV int not3(int a) { return a == kUnsure ? kUnsure : (a == kYes ? kNo : kYes); }

template <class R>
VGMI_DEV R InfLen();
template <>
VGMI_DEV double InfLen<double>()
{
  return DBL_MAX;
}
template <>
VGMI_DEV float InfLen<float>()
{
  return FLT_MAX;
}

} // namespace vgmi
