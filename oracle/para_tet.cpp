// Oracle: Parallelepiped and Tet (see shapes.h for the reference files followed).
// TEST INFRASTRUCTURE ONLY (see common.h).
#include "shapes.h"

namespace vgo {
namespace {
// vecCore::math::Max / Min and Vector3D::Max / Min (VecGeom/base/Vector3D.h)
inline double Mx(double a, double b) { return a > b ? a : b; }
inline double Mn(double a, double b) { return a < b ? a : b; }
inline double VMax(V3 const &v) { return Mx(Mx(v.x, v.y), v.z); }
inline double VMin(V3 const &v) { return Mn(Mn(v.x, v.y), v.z); }
inline V3 Unit(V3 const &v)
{
  const double m = std::sqrt(v.Mag2() + kMinimum); // Vector3D::Unit (Vector3D.h:284-290)
  return V3(v.x / m, v.y / m, v.z / m);
}
} // namespace

// ---------------------------------------------------------------- Parallelepiped
// ParallelepipedStruct(dx, dy, dz, alpha, theta, phi): SetAlpha, SetThetaAndPhi, ComputeNormals
// (ParallelepipedStruct.h:62-124)
Parallelepiped::Parallelepiped(double dx, double dy, double dz, double alpha, double theta, double phi)
    : dim(dx, dy, dz)
{
  tanAlpha       = std::tan(alpha);
  tanThetaCosPhi = std::tan(theta) * std::cos(phi);
  tanThetaSinPhi = std::tan(theta) * std::sin(phi);
  cosTheta       = std::cos(theta);
  const V3 vx(1., 0., 0.), vy(tanAlpha, 1., 0.), vz(tanThetaCosPhi, tanThetaSinPhi, 1.);
  normals[0] = vy.Cross(vz);
  normals[1] = vz.Cross(vx);
  normals[2] = V3(0., 0., 1.);
  normals[0] = normals[0] * (1. / normals[0].Mag()); // Vector3D::Normalize (Vector3D.h:198)
  normals[1] = normals[1] * (1. / normals[1].Mag());
  ctx = std::fabs(normals[0].x);
  cty = std::fabs(normals[1].y);
}

// ParallelepipedImplementation::Transform (ParallelepipedImplementation.h:60-66): to the oblique frame
V3 Parallelepiped::Oblique(V3 p) const
{
  p.y -= tanThetaSinPhi * p.z;
  p.x -= tanThetaCosPhi * p.z + tanAlpha * p.y;
  return p;
}

// ::SafetyVector (:68-76)
V3 Parallelepiped::SafetyVector(V3 const &lp) const
{
  V3 s = lp.Abs() - dim;
  s.x *= ctx;
  s.y *= cty;
  return s;
}

bool Parallelepiped::Contains(V3 const &p) const { return VMax(SafetyVector(Oblique(p))) < 0.0; }

int Parallelepiped::Inside(V3 const &p) const
{
  const double safety = VMax(SafetyVector(Oblique(p)));
  int in              = safety < 0.0 ? kInside : kOutside;
  if (std::fabs(safety) < kHalfTolerance) in = kSurface;
  return in;
}

double Parallelepiped::SafetyToIn(V3 const &p) const
{
  double safety = VMax(SafetyVector(Oblique(p)));
  if (std::fabs(safety) < kHalfTolerance) safety = 0.0;
  return safety;
}

double Parallelepiped::SafetyToOut(V3 const &p) const
{
  double safety = -VMax(SafetyVector(Oblique(p)));
  if (std::fabs(safety) < kHalfTolerance) safety = 0.0;
  return safety;
}

// :130-162
double Parallelepiped::DistanceToIn(V3 const &point, V3 const &direction, double) const
{
  const V3 p = Oblique(point), v = Oblique(direction), sv = SafetyVector(p);
  bool leaving = false;
  leaving      = leaving || (sv.x >= -kHalfTolerance && p.x * v.x >= 0.);
  leaving      = leaving || (sv.y >= -kHalfTolerance && p.y * v.y >= 0.);
  leaving      = leaving || (sv.z >= -kHalfTolerance && p.z * v.z >= 0.);
  const V3 inv(1. / NonZero(v.x), 1. / NonZero(v.y), 1. / NonZero(v.z));
  const V3 sgn(Sign(inv.x), Sign(inv.y), Sign(inv.z));
  const V3 temp        = sgn * dim;
  const double distIn  = VMax((-temp - p) * inv);
  const double distOut = VMin((temp - p) * inv);
  double distance      = kInfLength;
  if (!leaving && distOut > distIn + kHalfTolerance) distance = distIn;
  return distance;
}

// :164-183
double Parallelepiped::DistanceToOut(V3 const &point, V3 const &direction, double) const
{
  const V3 p = Oblique(point), v = Oblique(direction), sv = SafetyVector(p);
  const V3 dir(NonZero(v.x), NonZero(v.y), NonZero(v.z));
  const V3 sgn(Sign(dir.x), Sign(dir.y), Sign(dir.z));
  const V3 num    = sgn * dim - p;
  double distance = VMin(V3(num.x / dir.x, num.y / dir.y, num.z / dir.z));
  if (VMax(sv) > kHalfTolerance) distance = -1.0;
  return distance;
}

// NormalKernel (:185-232)
bool Parallelepiped::Normal(V3 const &point, V3 &n) const
{
  const V3 p = Oblique(point), sv = SafetyVector(p);
  const V3 signs(Sign(p.x), Sign(p.y), Sign(p.z));
  V3 normal(0., 0., 0.);
  if (std::fabs(sv.z) <= kHalfTolerance) normal = V3(0., 0., signs.z);
  if (std::fabs(sv.y) <= kHalfTolerance) normal = normal + signs.y * normals[1];
  if (std::fabs(sv.x) <= kHalfTolerance) normal = normal + signs.x * normals[0];
  const double mag2 = normal.Mag2();
  if (mag2 > 1.) normal = Unit(normal);
  if (mag2 > 0.) {
    n = normal;
    return true;
  }
  const double safety = VMax(sv);
  normal              = signs.x * normals[0];
  if (sv.y == safety) normal = signs.y * normals[1];
  if (sv.z == safety) normal = signs.z * normals[2];
  n = normal;
  return false;
}

// UnplacedParallelepiped::Extent (source/UnplacedParallelepiped.cpp:79-89)
void Parallelepiped::Extent(V3 &lo, V3 &hi) const
{
  const double dx = dim.x + dim.y * std::fabs(tanAlpha) + dim.z * std::fabs(tanThetaCosPhi);
  const double dy = dim.y + dim.z * std::fabs(tanThetaSinPhi);
  const double dz = dim.z;
  lo = V3(-dx, -dy, -dz), hi = V3(dx, dy, dz);
}

// ---------------------------------------------------------------- Tet
// TetStruct::CalculateCached (TetStruct.h:57-90)
Tet::Tet(V3 const &p0, V3 const &p1, V3 const &p2, V3 const &p3)
{
  vertex[0] = p0, vertex[1] = p1, vertex[2] = p2, vertex[3] = p3;
  V3 n0 = Unit((vertex[1] - vertex[0]).Cross(vertex[2] - vertex[0]));
  V3 n1 = Unit((vertex[2] - vertex[1]).Cross(vertex[3] - vertex[1]));
  V3 n2 = Unit((vertex[3] - vertex[2]).Cross(vertex[0] - vertex[2]));
  V3 n3 = Unit((vertex[0] - vertex[3]).Cross(vertex[1] - vertex[3]));
  if (n0.Dot(vertex[3] - vertex[0]) > 0) n0 = -n0;
  if (n1.Dot(vertex[0] - vertex[1]) > 0) n1 = -n1;
  if (n2.Dot(vertex[1] - vertex[2]) > 0) n2 = -n2;
  if (n3.Dot(vertex[2] - vertex[3]) > 0) n3 = -n3;
  n[0] = n0, n[1] = n1, n[2] = n2, n[3] = n3;
  for (int i = 0; i < 4; ++i) d[i] = -n[i].Dot(vertex[i]);
}

// GenericKernelForContainsAndInside (TetImplementation.h:85-107): the largest plane distance
double Tet::MaxPlaneDistance(V3 const &p) const
{
  double dist[4];
  for (int i = 0; i < 4; ++i) dist[i] = n[i].Dot(p) + d[i];
  return Mx(Mx(Mx(dist[0], dist[1]), dist[2]), dist[3]);
}

bool Tet::Contains(V3 const &p) const { return !(MaxPlaneDistance(p) > kHalfTolerance); }

int Tet::Inside(V3 const &p) const
{
  const double safety = MaxPlaneDistance(p);
  int in              = kSurface;
  if (safety > kHalfTolerance) in = kOutside;
  if (safety <= -kHalfTolerance) in = kInside;
  return in;
}

// :109-135 (absSafe is a minimum of absolute values, so its "<= -kHalfTolerance" test never holds)
double Tet::DistanceToIn(V3 const &p, V3 const &dir, double) const
{
  double distance = -kInfLength, distanceOut = kInfLength, absSafe = kInfLength;
  double cosa[4], safe[4], dist[4];
  for (int i = 0; i < 4; ++i) {
    cosa[i] = NonZero(n[i].Dot(dir));
    safe[i] = n[i].Dot(p) + d[i];
    dist[i] = -safe[i] / cosa[i];
  }
  for (int i = 0; i < 4; ++i) {
    if (cosa[i] < 0.) distance = Mx(distance, dist[i]);
    if (cosa[i] > 0.) distanceOut = Mn(distanceOut, dist[i]);
    if (cosa[i] > 0.) absSafe = Mn(absSafe, std::fabs(safe[i]));
  }
  if (distance >= distanceOut || distanceOut <= kHalfTolerance || absSafe <= -kHalfTolerance) distance = kInfLength;
  return distance;
}

// :137-157
double Tet::DistanceToOut(V3 const &p, V3 const &dir, double) const
{
  double distance = kInfLength, safety = -kInfLength;
  double cosa[4], safe[4];
  for (int i = 0; i < 4; ++i) {
    cosa[i] = NonZero(n[i].Dot(dir));
    safe[i] = n[i].Dot(p) + d[i];
    safety  = Mx(safety, safe[i]);
  }
  for (int i = 0; i < 4; ++i)
    if (cosa[i] > 0.) distance = Mn(distance, -safe[i] / cosa[i]);
  if (safety > kHalfTolerance) distance = -1.;
  return distance;
}

// :159-185 (Dot(point, n): the point's components first)
double Tet::SafetyToIn(V3 const &p) const
{
  double dist[4];
  for (int i = 0; i < 4; ++i) dist[i] = p.Dot(n[i]) + d[i];
  double safety = Mx(Mx(Mx(dist[0], dist[1]), dist[2]), dist[3]);
  if (std::fabs(safety) <= kHalfTolerance) safety = 0.;
  return safety;
}

double Tet::SafetyToOut(V3 const &p) const
{
  double dist[4];
  for (int i = 0; i < 4; ++i) dist[i] = p.Dot(n[i]) + d[i];
  double safety = -Mx(Mx(Mx(dist[0], dist[1]), dist[2]), dist[3]);
  if (std::fabs(safety) <= kHalfTolerance) safety = 0.;
  return safety;
}

// NormalKernel (:187-218)
bool Tet::Normal(V3 const &p, V3 &nrm) const
{
  V3 normal(0., 0., 0.);
  double dist[4];
  for (int i = 0; i < 4; ++i) {
    dist[i] = n[i].Dot(p) + d[i];
    if (std::fabs(dist[i]) <= kHalfTolerance) normal = normal + n[i];
  }
  if (normal.Mag2() > 1) normal = Unit(normal);
  if (normal.Mag2() > 0.) {
    nrm = normal;
    return true;
  }
  double safety = -kInfLength;
  for (int i = 0; i < 4; ++i)
    if (dist[i] > safety) normal = n[i], safety = dist[i];
  nrm = normal;
  return false;
}

// UnplacedTet::Extent (source/UnplacedTet.cpp:34-50)
void Tet::Extent(V3 &lo, V3 &hi) const
{
  lo = hi = vertex[0];
  for (int i = 1; i < 4; ++i) {
    lo.x = std::min(lo.x, vertex[i].x), lo.y = std::min(lo.y, vertex[i].y), lo.z = std::min(lo.z, vertex[i].z);
    hi.x = std::max(hi.x, vertex[i].x), hi.y = std::max(hi.y, vertex[i].y), hi.z = std::max(hi.z, vertex[i].z);
  }
}

} // namespace vgo
