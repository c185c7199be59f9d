// Polyhedron kernels (VecGeom/volumes/kernel/PolyhedronImplementation.h:298-1200, generic
// EInnerRadii/EPhiCutout specialisation) over quadrilateral plane sets
// (VecGeom/volumes/Quadrilaterals.h:105-597, VecGeom/volumes/Planes.h:189-275; scalar loops).
//
// The shape's z planes, radii, phi-section vectors, segment headers and quadrilaterals are
// read from the variable-length block; in a single-shape launch every lane reads the same
// addresses, so these loads are uniform across the wavefront.
#pragma once
#include "dev_common.h"
#include "dev_tube.h"
#include "shape_data.h"

namespace vgmi {
namespace phed {

struct QSet {
  const double *q; // kQuadDoubles per quadrilateral
  int n;
};

VGMI_DEV double zp(PolyhedronData const &P, int i) { return P.base[P.zsOff + i]; }
VGMI_DEV double rmn(PolyhedronData const &P, int i) { return P.base[P.rminOff + i]; }
VGMI_DEV double rmx(PolyhedronData const &P, int i) { return P.base[P.rmaxOff + i]; }
VGMI_DEV bool sameZ(PolyhedronData const &P, int i) { return P.base[P.sameZOff + i] != 0.0; }
// which = 0 outer, 1 phi, 2 inner
VGMI_DEV QSet qset(PolyhedronData const &P, int seg, int which)
{
  const double *h = P.base + P.segOff + 6 * seg + 2 * which;
  return QSet{P.base + int(h[0]), int(h[1])};
}

template <class R>
VGMI_DEV Vec3<R> qn(const double *q)
{
  return Vec3<R>(R(q[0]), R(q[1]), R(q[2]));
}
template <class R>
VGMI_DEV R planeDist(const double *q, Vec3<R> const &p)
{
  return p.Dot(qn<R>(q)) + R(q[3]);
}
template <class R>
VGMI_DEV R sideDist(const double *q, int j, Vec3<R> const &x)
{
  const double *s = q + 4 + 4 * j;
  return x.Dot(Vec3<R>(R(s[0]), R(s[1]), R(s[2]))) + R(s[3]);
}
template <class R>
VGMI_DEV Vec3<R> corner(const double *q, int j)
{
  const double *c = q + 20 + 3 * j;
  return Vec3<R>(R(c[0]), R(c[1]), R(c[2]));
}

// Planes::Contains
template <class R>
VGMI_DEV bool QContains(QSet const &s, Vec3<R> const &p)
{
  bool r = true;
  for (int i = 0; i < s.n; ++i) r = r && (planeDist(s.q + kQuadDoubles * i, p) < R(0));
  return r;
}

// Planes::Inside
template <class R>
VGMI_DEV int QInside(QSet const &s, Vec3<R> const &p)
{
  int r = kInside;
  for (int i = 0; i < s.n; ++i) {
    const double *q = s.q + kQuadDoubles * i;
    const R dist    = R(q[0]) * p.x + R(q[1]) * p.y + R(q[2]) * p.z + R(q[3]);
    if (dist > R(kHalfTolerance)) r = kOutside;
    if (r == kInside && dist > R(-kHalfTolerance)) r = kSurface;
    if (r == kOutside) break;
  }
  return r;
}

template <class R>
VGMI_DEV int QInsideAt(QSet const &s, Vec3<R> const &p, int i)
{
  const double *q = s.q + kQuadDoubles * i;
  int r           = kInside;
  const R dist    = R(q[0]) * p.x + R(q[1]) * p.y + R(q[2]) * p.z + R(q[3]);
  if (dist > R(kHalfTolerance)) r = kOutside;
  if (r == kInside && dist > R(-kHalfTolerance)) r = kSurface;
  return r;
}

// Quadrilaterals::DistanceToIn<behind>
template <class R>
VGMI_DEV R QDistanceToIn(QSet const &s, bool behind, Vec3<R> const &p, Vec3<R> const &d)
{
  R best = InfLen<R>();
  for (int i = 0; i < s.n; ++i) {
    const double *q = s.q + kQuadDoubles * i;
    const Vec3<R> n = qn<R>(q);
    R distance      = p.Dot(n) + R(q[3]);
    bool valid      = (behind ? -distance : distance) > R(-kTolerance);
    if (!valid) continue;
    const R proj = d.Dot(n);
    valid        = valid && ((behind ? proj : -proj) > R(0));
    if (!valid) continue;
    distance /= -(proj + vcopysign(R(1E-20), proj));
    const Vec3<R> x = p + d * distance;
    for (int j = 0; j < 4; ++j) {
      valid = valid && (sideDist(q, j, x) >= -HitTol<R>(kTolerance));
      if (!valid) break;
    }
    if (valid) best = distance;
    if (best < InfLen<R>()) break;
  }
  return vmax(R(0), best);
}

// Quadrilaterals::DistanceToOut(point, dir, zMin, zMax)
template <class R>
VGMI_DEV R QDistanceToOut(QSet const &s, Vec3<R> const &p, Vec3<R> const &d, double zMin, double zMax)
{
  R best = InfLen<R>();
  for (int i = 0; i < s.n; ++i) {
    const double *q = s.q + kQuadDoubles * i;
    const Vec3<R> n = qn<R>(q);
    R dt            = p.Dot(n) + R(q[3]);
    bool valid      = dt < R(kTolerance);
    if (!valid) continue;
    const R proj = d.Dot(n);
    valid        = valid && proj > R(0);
    if (!valid) continue;
    dt /= -proj;
    valid = valid && dt < best;
    if (!valid) continue;
    if (zMin == zMax) {
      const Vec3<R> x = p + d * dt;
      valid           = true;
      for (int j = 0; j < 4; ++j) {
        valid = valid && (sideDist(q, j, x) >= -HitTol<R>(kTolerance));
        if (!valid) break;
      }
    } else {
      const R zq = p.z + dt * d.z;
      valid      = valid && (zq >= R(zMin) - HitTol<R>(kTolerance)) && (zq < R(zMax) + HitTol<R>(kTolerance));
    }
    if (!valid) continue;
    best = dt;
  }
  if (best > R(-kTolerance)) best = vmax(best, R(0));
  return best;
}

// GenericKernels.h:188-215
template <class R>
VGMI_DEV R SegmentDistSq(Vec3<R> const &c0, Vec3<R> const &c1, Vec3<R> const &p)
{
  R result         = InfLen<R>();
  const Vec3<R> ln = c1 - c0;
  const R dot0     = (p - c0).Dot(ln);
  if (dot0 <= R(0)) return (p - c0).Mag2();
  const R dot1 = ln.Mag2();
  if (dot1 <= dot0) result = (p - c1).Mag2();
  if (result < InfLen<R>()) return result;
  return ((c0 + ln * (dot0 / dot1)) - p).Mag2();
}

// Quadrilaterals::ScalarDistanceSquared
template <class R>
VGMI_DEV R QDistanceSquared(QSet const &s, int i, Vec3<R> const &p)
{
  const double *q = s.q + kQuadDoubles * i;
  const Vec3<R> n = qn<R>(q);
  R distance      = p.Dot(n) + R(q[3]);
  const Vec3<R> x = p - n * distance;
  bool within[4];
  for (int j = 0; j < 4; ++j) {
    const double *sv = q + 4 + 4 * j;
    within[j]        = x.x * R(sv[0]) + x.y * R(sv[1]) + x.z * R(sv[2]) + R(sv[3]) >= R(0);
  }
  if (within[0] && within[1] && within[2] && within[3]) return distance * distance;
  R dsq = InfLen<R>();
  for (int j = 0; j < 4; ++j) {
    if (!within[j]) {
      distance = SegmentDistSq(corner<R>(q, j), corner<R>(q, (j + 1) % 4), p);
      if (distance < dsq) dsq = distance;
    }
  }
  return dsq;
}

template <class R>
VGMI_DEV int FindZSegment(PolyhedronData const &P, R z)
{
  int index = -1;
  int k     = 0;
  while (k < P.nz - 1 && z - R(kTolerance) > R(zp(P, k))) {
    ++index;
    ++k;
  }
  if (z + R(kTolerance) > R(zp(P, k))) return index + 1;
  return index;
}

template <class R>
VGMI_DEV int FindPhiSegment(PolyhedronData const &P, Vec3<R> const &p)
{
  const double *ps = P.base + P.phiSecOff;
  int index        = -1;
  R first          = p.x * R(ps[0]) + p.y * R(ps[1]) + p.z * R(ps[2]);
  for (int i = 1; i < P.sides + 1; ++i) {
    const R second = p.x * R(ps[3 * i]) + p.y * R(ps[3 * i + 1]) + p.z * R(ps[3 * i + 2]);
    if (first > R(-kTolerance) && second < R(kTolerance)) index = i - 1;
    if (index >= 0) break;
    first = second;
  }
  return index;
}

template <class R>
VGMI_DEV int WedgeInside(PolyhedronData const &P, Vec3<R> const &p)
{
  const R startCheck = (-p.x * R(P.w1y) + p.y * R(P.w1x));
  const R endCheck   = (-R(P.w2x) * p.y + R(P.w2y) * p.x);
  bool co            = startCheck < R(0);
  if (P.wedgeLtPi)
    co = co || (endCheck < R(0));
  else
    co = co && (endCheck < R(0));
  const bool on = tube::OnWedgePlane(R(P.w1x), R(P.w1y), R(P.n1x), R(P.n1y), p) ||
                  tube::OnWedgePlane(R(P.w2x), R(P.w2y), R(P.n2x), R(P.n2y), p);
  co            = co && !on;
  const bool ci = !on && !co;
  int r         = kSurface;
  if (co) r = kOutside;
  if (ci) r = kInside;
  return r;
}

template <class R>
VGMI_DEV bool InPhiCutoutWedge(PolyhedronData const &P, int seg, Vec3<R> const &p)
{
  const QSet ph = qset(P, seg, 1);
  const bool s0 = planeDist(ph.q, p) >= R(0);
  const bool s1 = planeDist(ph.q + kQuadDoubles, p) >= R(0);
  if (P.hasLargePhiCutout) return s0 || s1;
  return s0 && s1;
}

template <class R>
VGMI_DEV bool InBoundingTube(PolyhedronData const &P, Vec3<R> const &p)
{
  return tube::Contains(P.bounding, Vec3<R>(p.x, p.y, p.z - R(P.boundingOffset)));
}

template <class R>
VGMI_DEV R DistanceToInZSegment(PolyhedronData const &P, int i, Vec3<R> const &p, Vec3<R> const &d)
{
  R distance = QDistanceToIn(qset(P, i, 0), false, p, d);
  bool done  = distance < InfLen<R>();
  if (done) return distance;
  if (P.hasPhiCutout) distance = QDistanceToIn(qset(P, i, 1), false, p, d);
  done = done || distance < InfLen<R>();
  if (done) return distance;
  const QSet in = qset(P, i, 2);
  if (in.n > 0) distance = QDistanceToIn(in, true, p, d);
  return distance;
}

template <class R>
VGMI_DEV R DistanceToOutZSegment(PolyhedronData const &P, int i, double zMin, double zMax, Vec3<R> const &p,
                                 Vec3<R> const &d)
{
  bool done     = false;
  R distance    = InfLen<R>();
  const QSet in = qset(P, i, 2);
  if (in.n > 0) {
    distance = QDistanceToIn(in, false, p, d);
    done     = distance < InfLen<R>();
    if (done) return distance;
  }
  if (P.hasPhiCutout) {
    const R dphi = QDistanceToIn(qset(P, i, 1), true, p, d);
    if (!done && distance > R(-kTolerance)) distance = dphi;
    done = distance > R(-kTolerance) && distance < InfLen<R>();
    if (done) return distance;
  }
  const R dout = QDistanceToOut(qset(P, i, 0), p, d, zMin, zMax);
  if (!done) distance = dout;
  return distance;
}

template <class R>
VGMI_DEV R SafetyToZSegmentSquared(PolyhedronData const &P, int i, int &phiIndex, Vec3<R> const &p, bool ptInside,
                                   int *iSurf = nullptr)
{
  // iSurf (when given) receives the surface kind of the minimum, as ScalarNormalKernel uses it
  // (PolyhedronImplementation.h:448-557): 0 / 1 the phi planes, 2 the inner, 3 the outer quadrilaterals
  auto mark = [iSurf](int k) {
    if (iSurf) *iSurf = k;
  };
  const QSet outer = qset(P, i, 0), phi = qset(P, i, 1), inner = qset(P, i, 2);
  const bool inCutout = phiIndex < 0;
  R sq                = InfLen<R>();
  if (P.hasPhiCutout && phi.n == 2) {
    if (ptInside || inCutout) {
      mark(0);
      sq          = QDistanceSquared(phi, 0, p);
      const R saf = QDistanceSquared(phi, 1, p);
      if (saf < sq) {
        sq = saf;
        mark(1);
      }
      if (inCutout) {
        if (P.hasLargePhiCutout && inner.n > 0) {
          R si = QDistanceSquared(inner, 0, p);
          if (si < sq) {
            mark(2);
            phiIndex = 0;
            sq       = si;
          }
          if (inner.n > 1) {
            si = QDistanceSquared(inner, inner.n - 1, p);
            if (si < sq) {
              mark(2);
              phiIndex = inner.n - 1;
              sq       = si;
            }
          }
        }
        return sq;
      }
    }
  }
  if (inCutout && i > 0 && i < P.nseg - 1 && zp(P, i) == zp(P, i + 1)) {
    if (outer.n > 0) {
      R so = QDistanceSquared(outer, 0, p);
      if (so < sq) {
        mark(3);
        phiIndex = 0;
        sq       = so;
      }
      if (outer.n > 1) {
        so = QDistanceSquared(outer, outer.n - 1, p);
        if (so < sq) {
          mark(3);
          phiIndex = outer.n - 1;
          sq       = so;
        }
      }
    }
    if (inner.n > 0) {
      R si = QDistanceSquared(inner, 0, p);
      if (si < sq) {
        mark(2);
        phiIndex = 0;
        sq       = si;
      }
      if (inner.n > 1) {
        si = QDistanceSquared(inner, inner.n - 1, p);
        if (si < sq) {
          mark(2);
          phiIndex = inner.n - 1;
          sq       = si;
        }
      }
    }
    return sq;
  }
  R so = InfLen<R>();
  if (outer.n > 0) so = QDistanceSquared(outer, phiIndex, p);
  R si = InfLen<R>();
  if (inner.n > 0) si = QDistanceSquared(inner, phiIndex, p);
  if (si < sq) {
    mark(2);
    sq = si;
  }
  if (so < sq) {
    mark(3);
    sq = so;
  }
  return sq;
}

template <class R>
VGMI_DEV void DistanceToEndcaps(PolyhedronData const &P, bool inside, Vec3<R> const &p, Vec3<R> const &d, R &distance)
{
  const bool first = inside ? (d.z < R(0)) : !(d.z < R(0));
  const int seg    = first ? 0 : P.nseg - 1;
  const R zPlane   = R(first ? zp(P, 0) : zp(P, P.nseg));
  const R dt       = (zPlane - p.z) / NonZero(d.z);
  if (dt < R(-kTolerance) || dt >= distance) return;
  const Vec3<R> x = p + d * dt;
  if (!QContains(qset(P, seg, 0), x)) return;
  const QSet in = qset(P, seg, 2);
  if (in.n > 0 && QContains(in, x)) return;
  if (P.hasPhiCutout && InPhiCutoutWedge(P, seg, x)) return;
  distance = dt;
}

template <class R>
VGMI_DEV void SafetyToEndcapsSquared(PolyhedronData const &P, Vec3<R> const &p, R &dsq, int *iz = nullptr)
{
  if (iz) *iz = 0;
  const R firstD     = R(zp(P, 0)) - p.z;
  const R lastD      = R(zp(P, P.nseg)) - p.z;
  const bool isFirst = vabs(firstD) < vabs(lastD);
  const int seg      = isFirst ? 0 : P.nseg - 1;
  const R dt         = isFirst ? firstD : lastD;
  const R dt2        = dt * dt;
  if (dt2 >= dsq) return;
  const Vec3<R> x(p.x, p.y, p.z + dt);
  if (!QContains(qset(P, seg, 0), x)) return;
  const QSet in = qset(P, seg, 2);
  if (in.n > 0 && QContains(in, x)) return;
  if (P.hasPhiCutout && InPhiCutoutWedge(P, seg, x)) return;
  if (iz) *iz = isFirst ? -1 : 1;
  dsq = dt2;
}

template <class R>
VGMI_DEV Vec3<R> SideNormal0(PolyhedronData const &P, int seg, int phiIndex)
{
  const QSet outer = qset(P, seg, 0);
  const double *q  = (outer.n ? outer.q : qset(P, seg, 2).q) + kQuadDoubles * phiIndex;
  return Vec3<R>(R(q[4]), R(q[5]), R(q[6]));
}

template <class R>
VGMI_DEV int InsideSegBorder(PolyhedronData const &P, Vec3<R> const &p, int z)
{
  const int phiIndex = FindPhiSegment(P, p);
  if (phiIndex < 0) return kOutside;
  const R rdv    = vabs(p.Dot(SideNormal0<R>(P, z, phiIndex)));
  const R tol    = R(kTolerance);
  const bool in1 = (rdv > R(rmn(P, z)) - tol) && (rdv < R(rmx(P, z)) + tol);
  const bool in2 = (rdv > R(rmn(P, z + 1)) - tol) && (rdv < R(rmx(P, z + 1)) + tol);
  if (in1 && in2) {
    if ((rdv < R(rmn(P, z)) + tol) || (rdv > R(rmx(P, z)) - tol) || (rdv < R(rmn(P, z + 1)) + tol) ||
        (rdv > R(rmx(P, z + 1)) - tol))
      return kSurface;
    if (P.hasPhiCutout) return WedgeInside(P, p);
    return kInside;
  }
  if (!in1 && !in2) return kOutside;
  return kSurface;
}

template <class R>
VGMI_DEV int InsideSegPhi(PolyhedronData const &P, Vec3<R> const &p, int z, int phiIndex)
{
  if (phiIndex < 0) return kOutside;
  const R dz = vabs(p.z - R(P.boundingOffset)) - R(0.5) * (R(zp(P, P.nseg)) - R(zp(P, 0)));
  if (dz > R(kHalfTolerance)) return kOutside;
  if (sameZ(P, z)) return InsideSegBorder(P, p, z);
  const int io = QInsideAt(qset(P, z, 0), p, phiIndex);
  if (io != kInside) return io;
  const QSet in = qset(P, z, 2);
  if (in.n > 0) {
    const int ii = QInsideAt(in, p, phiIndex);
    if (ii == kInside) return kOutside;
    if (ii == kSurface) return kSurface;
  }
  if (P.hasPhiCutout) {
    const int ip = WedgeInside(P, p);
    if (ip != kInside) return ip;
  }
  if (vabs(dz) < R(kHalfTolerance)) return kSurface;
  return kInside;
}

template <class R>
VGMI_DEV bool Contains(PolyhedronData const &P, Vec3<R> const &p)
{
  if (!InBoundingTube(P, p)) return false;
  const int z = FindZSegment(P, p.z);
  if (!((z >= 0) && (z < P.nseg))) return false;
  if (sameZ(P, z)) {
    const int phiIndex = FindPhiSegment(P, p);
    if (phiIndex < 0) return false;
    const R rdv    = vabs(p.Dot(SideNormal0<R>(P, z, phiIndex)));
    const bool in1 = (rdv >= R(rmn(P, z))) && (rdv <= R(rmx(P, z)));
    const bool in2 = (rdv >= R(rmn(P, z + 1))) && (rdv <= R(rmx(P, z + 1)));
    return in1 | in2;
  }
  if (!QContains(qset(P, z, 0), p)) return false;
  const QSet in = qset(P, z, 2);
  if (in.n > 0 && QContains(in, p)) return false;
  if (P.hasPhiCutout && !QContains(qset(P, z, 1), p)) return false;
  return true;
}

template <class R>
VGMI_DEV int Inside(PolyhedronData const &P, Vec3<R> const &p)
{
  if (!InBoundingTube(P, p)) return kOutside;
  int z = FindZSegment(P, p.z);
  if (z > (P.nseg - 1)) z = P.nseg - 1;
  if (z < 0) z = 0;
  if (sameZ(P, z)) return InsideSegBorder(P, p, z);
  const int io = QInside(qset(P, z, 0), p);
  if (io != kInside) return io;
  const QSet in = qset(P, z, 2);
  if (in.n > 0) {
    const int ii = QInside(in, p);
    if (ii == kInside) return kOutside;
    if (ii == kSurface) return kSurface;
  }
  if (P.hasPhiCutout) {
    const int ip = QInside(qset(P, z, 1), p);
    if (ip != kInside) return ip;
  }
  const R dz = vabs(vabs(p.z - R(P.boundingOffset)) - R(0.5) * (R(zp(P, P.nseg)) - R(zp(P, 0))));
  if (dz < R(kTolerance)) return kSurface;
  return kInside;
}

template <class R>
VGMI_DEV R DistanceToIn(PolyhedronData const &P, Vec3<R> const &p, Vec3<R> const &d)
{
  if ((p.z < R(zp(P, 0)) + R(kTolerance)) && d.z <= R(0)) return InfLen<R>();
  if ((p.z > R(zp(P, P.nseg)) - R(kTolerance)) && d.z >= R(0)) return InfLen<R>();
  if (Inside(P, p) == kInside) return R(-1);
  R tubeDistance = R(0);
  const Vec3<R> bp(p.x, p.y, p.z - R(P.boundingOffset));
  if (!tube::Contains(P.bounding, bp)) {
    tubeDistance = tube::DistanceToIn(P.bounding, bp, d);
    if (tubeDistance == InfLen<R>()) return InfLen<R>();
  }
  int z                 = FindZSegment(P, p.z);
  const int zMax        = P.nseg;
  z                     = z < 0 ? 0 : (z >= zMax ? zMax - 1 : z);
  const bool goingRight = d.z >= R(0);
  R distance            = InfLen<R>();
  if (goingRight) {
    for (; z < zMax; ++z) {
      distance = DistanceToInZSegment(P, z, p, d);
      if (distance >= R(0) && distance < InfLen<R>()) break;
    }
  } else {
    for (; z >= 0; --z) {
      distance = DistanceToInZSegment(P, z, p, d);
      if (distance >= R(0) && distance < InfLen<R>()) break;
    }
  }
  DistanceToEndcaps(P, false, p, d, distance);
  return (distance >= tubeDistance - HitTol<R>(1E-6)) ? distance : InfLen<R>();
}

template <class R>
VGMI_DEV R DistanceToOut(PolyhedronData const &P, Vec3<R> const &p, Vec3<R> const &d)
{
  const int zMax = P.nseg;
  if ((p.z < R(zp(P, 0)) - R(kTolerance)) || (p.z > R(zp(P, zMax)) + R(kTolerance))) return R(-1);
  if (Inside(P, p) == kOutside) return R(-1);
  int z                 = FindZSegment(P, p.z);
  z                     = z < 0 ? 0 : (z >= zMax ? zMax - 1 : z);
  const bool goingRight = d.z >= R(0);
  R distance            = InfLen<R>();
  if (goingRight) {
    for (; z < zMax; ++z) {
      distance = DistanceToOutZSegment(P, z, zp(P, z), zp(P, z + 1), p, d);
      if (distance >= R(0) && distance < InfLen<R>()) break;
      if (R(zp(P, z)) - p.z > distance) break;
    }
  } else {
    for (; z >= 0; --z) {
      distance = DistanceToOutZSegment(P, z, zp(P, z), zp(P, z + 1), p, d);
      if (distance >= R(0) && distance < InfLen<R>()) break;
      if (p.z - R(zp(P, z)) > distance) break;
    }
  }
  DistanceToEndcaps(P, true, p, d, distance);
  if (distance >= InfLen<R>()) distance = R(0);
  return distance;
}

template <class R>
VGMI_DEV R Safety(PolyhedronData const &P, Vec3<R> const &p, bool ptInside)
{
  R safety       = InfLen<R>();
  const int zMax = P.nseg;
  int z          = FindZSegment(P, p.z);
  z              = z < 0 ? 0 : (z >= zMax ? zMax - 1 : z);
  int phiIndex   = FindPhiSegment(P, p);
  const int in   = InsideSegPhi(P, p, z, phiIndex);
  if (in == kSurface) return R(0);
  const bool contains = (in == kInside);
  if (contains ^ ptInside) return R(-1);
  for (int iz = z; iz < zMax;) {
    safety = vmin(safety, SafetyToZSegmentSquared(P, iz, phiIndex, p, ptInside));
    ++iz;
    const R dz = R(zp(P, iz)) - p.z;
    if (dz * dz > safety) break;
  }
  for (int iz = z - 1; iz >= 0; --iz) {
    safety     = vmin(safety, SafetyToZSegmentSquared(P, iz, phiIndex, p, ptInside));
    const R dz = p.z - R(zp(P, iz));
    if (dz * dz > safety) break;
  }
  SafetyToEndcapsSquared(P, p, safety);
  return vsqrt(safety);
}

// The safeties take the z segment of the point's z and change abruptly across a z plane: an FP32
// z within kMixedPlaneBand of a plane answers NaN (evaluated in FP64).
template <class R>
VGMI_DEV bool NearZPlaneF32(PolyhedronData const &P, R z)
{
  if constexpr (sizeof(R) == sizeof(float)) {
    for (int k = 0; k < P.nz; ++k)
      if (vabs(z - R(zp(P, k))) < R(kMixedPlaneBand)) return true;
  }
  return false;
}

// The reference's Inside and Contains test the phi quadrilaterals as two half-spaces (Planes::Inside
// and Contains on segment.phi, PolyhedronImplementation.h:728,826), so for a phi range above pi the
// extended phi planes also pass through the solid's interior and the answer changes there. An FP32
// point within kMixedPlaneBand of a phi plane or of a z plane is left to FP64.
VGMI_DEV bool F32PointUndecided(PolyhedronData const &P, Vec3<float> const &p)
{
  if (NearZPlaneF32(P, p.z)) return true;
  if (P.hasPhiCutout) {
    const QSet ph = qset(P, 0, 1);
    for (int i = 0; i < ph.n; ++i)
      if (vabs(planeDist(ph.q + kQuadDoubles * i, p)) < float(kMixedPlaneBand)) return true;
  }
  return false;
}

template <class R>
VGMI_DEV R SafetyToIn(PolyhedronData const &P, Vec3<R> const &p)
{
  if (NearZPlaneF32(P, p.z)) return R(__builtin_nanf(""));
  return Safety(P, p, false);
}
template <class R>
VGMI_DEV R SafetyToOut(PolyhedronData const &P, Vec3<R> const &p)
{
  if (NearZPlaneF32(P, p.z)) return R(__builtin_nanf(""));
  return Safety(P, p, true);
}

} // namespace phed
} // namespace vgmi
