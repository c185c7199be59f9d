// Oracle: Polyhedron (regular n-sided z segments with optional inner radii and phi cutout)
// restated from VecGeom's scalar backend. TEST INFRASTRUCTURE ONLY (see common.h).
//
// Sources: source/Quadrilaterals.cpp:36-118 (Set), source/Planes.cpp:30-52, VecGeom/volumes/Quadrilaterals.h
// :105-630 and VecGeom/volumes/Planes.h:189-300 (plane-set kernels, scalar loops: no Vc acceleration),
// VecGeom/volumes/PolyhedronStruct.h:91-430 (construction), VecGeom/volumes/kernel/PolyhedronImplementation.h
// :298-1200 with the generic EInnerRadii/EPhiCutout traits (TreatInner/TreatPhi/LargePhiCutout pass the
// run-time flags through), VecGeom/volumes/kernel/GenericKernels.h:139-215 (NormalizeAngle,
// DistanceToLineSegmentSquared1), source/UnplacedPolyhedron.cpp:273-303 (extent).
#include "shapes.h"

namespace vgo {

namespace {
// GenericKernels.h:139-145 (int truncation of a / 2pi)
inline double NormalizeAngle(double a) { return a + kTwoPi * (int(a < 0) - int(a / kTwoPi)); }
inline V3 FromCylindrical(double r, double phi, double z) { return V3(r * std::cos(phi), r * std::sin(phi), z); }
inline V3 FixZeroes(V3 v)
{
  for (int i = 0; i < 3; ++i)
    if (std::fabs(v[i]) < kTolerance) v[i] = 0.0;
  return v;
}
inline V3 Normalized(V3 const &v) { return v * (1. / v.Mag()); }
// Vector3D operator== (base/Vector3D.h:387-390): equal within kTolerance per component
inline bool TolEqual(V3 const &a, V3 const &b)
{
  return std::fabs(a.x - b.x) < kTolerance && std::fabs(a.y - b.y) < kTolerance && std::fabs(a.z - b.z) < kTolerance;
}
// GenericKernels.h:188-215
double DistanceToLineSegmentSquared1(V3 const &c0, V3 const &c1, V3 const &p)
{
  double result      = kInfLength;
  const V3 line      = c1 - c0;
  const V3 proj      = p - c0;
  const double dot0  = proj.Dot(line);
  if (dot0 <= 0) return (p - c0).Mag2();
  const double dot1 = line.Mag2();
  if (dot1 <= dot0) result = (p - c1).Mag2();
  if (result < kInfLength) return result;
  return ((c0 + (dot0 / dot1) * line) - p).Mag2();
}
} // namespace

// ================================================================ Quads
// source/Quadrilaterals.cpp:36-110
void Quads::Set(int i, V3 const &c0, V3 const &c1, V3 const &c2, V3 const &c3)
{
  Quad &Q = q[i];
  Q.c[0] = c0, Q.c[1] = c1, Q.c[2] = c2, Q.c[3] = c3;
  V3 ch[4];
  ch[0]        = Q.c[0];
  int assigned = 1;
  for (int t = 1; t < 4; ++t) {
    bool chosen = true;
    for (int j = 0; j < assigned; ++j)
      if (TolEqual(ch[j], Q.c[t])) {
        chosen = false;
        break;
      }
    if (chosen) ch[assigned++] = Q.c[t];
  }
  double a = ch[0][1] * (ch[1][2] - ch[2][2]) + ch[1][1] * (ch[2][2] - ch[0][2]) + ch[2][1] * (ch[0][2] - ch[1][2]);
  double b = ch[0][2] * (ch[1][0] - ch[2][0]) + ch[1][2] * (ch[2][0] - ch[0][0]) + ch[2][2] * (ch[0][0] - ch[1][0]);
  double c = ch[0][0] * (ch[1][1] - ch[2][1]) + ch[1][0] * (ch[2][1] - ch[0][1]) + ch[2][0] * (ch[0][1] - ch[1][1]);
  double d = -ch[0][0] * (ch[1][1] * ch[2][2] - ch[2][1] * ch[1][2]) -
             ch[1][0] * (ch[2][1] * ch[0][2] - ch[0][1] * ch[2][2]) -
             ch[2][0] * (ch[0][1] * ch[1][2] - ch[1][1] * ch[0][2]);
  V3 normal(a, b, c);
  const double inv = 1. / normal.Mag();
  normal           = normal * inv;
  d *= inv;
  normal = FixZeroes(normal);
  Q.n    = normal;
  Q.d    = d;
  auto side = [&](int j, V3 const &s0, V3 const &s1) {
    if (!TolEqual(s1, s0)) {
      // Planes::Set(index, normal, origin), source/Planes.cpp:30-37
      const V3 sv       = FixZeroes(Normalized(normal.Cross(s1 - s0)));
      const double invl = 1. / sv.Mag();
      Q.sn[j]           = invl * sv;
      Q.sd[j]           = invl * -sv.Dot(s0);
    } else {
      Q.sn[j] = V3(0, 0, 0);
      Q.sd[j] = 0;
    }
  };
  side(0, c0, c1);
  side(1, c1, c2);
  side(2, c2, c3);
  side(3, c3, c0);
}

// Planes::Contains (Planes.h:189-205)
bool Quads::Contains(V3 const &p) const
{
  bool r = true;
  for (auto const &Q : q) r = r && (p.Dot(Q.n) + Q.d < 0);
  return r;
}

// Planes::Inside (Planes.h:241-261)
int Quads::Inside(V3 const &p) const
{
  int r = kInside;
  for (auto const &Q : q) {
    const double dist = Q.n.x * p.x + Q.n.y * p.y + Q.n.z * p.z + Q.d;
    if (dist > kHalfTolerance) r = kOutside;
    if (r == kInside && dist > -kHalfTolerance) r = kSurface;
    if (r == kOutside) break;
  }
  return r;
}

// Planes::Inside(point, i) (Planes.h:263-275)
int Quads::InsideAt(V3 const &p, int i) const
{
  Quad const &Q     = q[i];
  int r             = kInside;
  const double dist = Q.n.x * p.x + Q.n.y * p.y + Q.n.z * p.z + Q.d;
  if (dist > kHalfTolerance) r = kOutside;
  if (r == kInside && dist > -kHalfTolerance) r = kSurface;
  return r;
}

// Quadrilaterals.h:105-114
bool Quads::RayHits(int i, V3 const &x) const
{
  bool valid = true;
  for (int j = 0; j < 4; ++j) {
    valid = valid && (x.Dot(q[i].sn[j]) + q[i].sd[j] >= -kTolerance);
    if (!valid) break;
  }
  return valid;
}

// Quadrilaterals.h:398-431 (scalar loop)
double Quads::DistanceToIn(bool behind, V3 const &p, V3 const &d) const
{
  double best = kInfLength;
  for (int i = 0; i < size(); ++i) {
    Quad const &Q   = q[i];
    double distance = p.Dot(Q.n) + Q.d;
    bool valid      = (behind ? -distance : distance) > -kTolerance;
    if (!valid) continue;
    const double proj = d.Dot(Q.n);
    valid             = valid && ((behind ? proj : -proj) > 0);
    if (!valid) continue;
    distance /= -(proj + std::copysign(1E-20, proj));
    const V3 x = p + d * distance;
    for (int j = 0; j < 4; ++j) {
      valid = valid && (x.Dot(Q.sn[j]) + Q.sd[j] >= -kTolerance);
      if (!valid) break;
    }
    if (valid) best = distance;
    if (best < kInfLength) break;
  }
  return std::max(0., best);
}

// Quadrilaterals.h:517-558 (scalar loop)
double Quads::DistanceToOut(V3 const &p, V3 const &d, double zMin, double zMax) const
{
  double best = kInfLength;
  for (int i = 0; i < size(); ++i) {
    Quad const &Q = q[i];
    double dt     = p.Dot(Q.n) + Q.d;
    bool valid    = dt < kTolerance;
    if (!valid) continue;
    const double proj = d.Dot(Q.n);
    valid             = valid && proj > 0;
    if (!valid) continue;
    dt /= -proj;
    valid = valid && dt < best;
    if (!valid) continue;
    if (zMin == zMax) {
      valid = RayHits(i, p + dt * d);
    } else {
      const double zp = p[2] + dt * d[2];
      valid           = valid && (zp >= zMin - kTolerance) && (zp < zMax + kTolerance);
    }
    if (!valid) continue;
    best = dt;
  }
  if (best > -kTolerance) best = std::max(best, 0.);
  return best;
}

// Quadrilaterals.h:571-597
double Quads::DistanceSquared(int i, V3 const &p) const
{
  Quad const &Q   = q[i];
  double distance = p.Dot(Q.n) + Q.d;
  const V3 x      = p - distance * Q.n;
  bool within[4];
  for (int j = 0; j < 4; ++j) within[j] = x[0] * Q.sn[j].x + x[1] * Q.sn[j].y + x[2] * Q.sn[j].z + Q.sd[j] >= 0;
  if (within[0] && within[1] && within[2] && within[3]) return distance * distance;
  double dsq = kInfLength;
  for (int j = 0; j < 4; ++j) {
    if (!within[j]) {
      distance = DistanceToLineSegmentSquared1(Q.c[j], Q.c[(j + 1) % 4], p);
      if (distance < dsq) dsq = distance;
    }
  }
  return dsq;
}

// ================================================================ Polyhedron
// PolyhedronStruct.h:91-103 (member initialisers) and :248-429 (Initialize)
Polyhedron::Polyhedron(double phiStart_, double phiDelta_, int sides_, int nz_, const double *z, const double *rmn,
                       const double *rmx)
    : sides(sides_), nz(nz_)
{
  hasPhiCutout      = phiDelta_ < kTwoPi;
  hasLargePhiCutout = phiDelta_ < kPi;
  phiStart          = NormalizeAngle(phiStart_);
  phiDelta          = (phiDelta_ > kTwoPi) ? kTwoPi : phiDelta_;
  wedge.Set(phiDelta, phiStart);
  zs.assign(z, z + nz);
  rmin.assign(rmn, rmn + nz);
  rmax.assign(rmx, rmx + nz);
  sameZ.assign(nz, 0);
  for (int i = 0; i < nz; ++i)
    if (i > 0 && i < nz - 1 && zs[i] == zs[i + 1]) sameZ[i] = 1;
  segs.resize(nz - 1);
  for (int i = 0; i < nz - 1; ++i) {
    const bool hasInnerRadius = rmin[i] > 0 || rmin[i + 1] > 0;
    int mult                  = (zs[i] == zs[i + 1] && rmax[i] == rmax[i + 1]) ? 0 : 1;
    segs[i].outer.q.resize(sides * mult);
    if (hasPhiCutout) {
      mult = (zs[i] == zs[i + 1]) ? 0 : 1;
      segs[i].phi.q.resize(2 * mult);
    }
    mult = (zs[i] == zs[i + 1] && rmin[i] == rmin[i + 1]) ? 0 : 1;
    if (hasInnerRadius && mult > 0) {
      segs[i].inner.q.resize(sides * mult);
      hasInnerRadii = true;
    }
  }
  double ps = NormalizeAngle(phiStart_);
  double pd = phiDelta_ > kTwoPi ? kTwoPi : phiDelta_;
  const double sidePhi = pd / sides;
  auto getPhi = [&](int side) {
    if (!hasPhiCutout && side == sides) side = 0;
    return NormalizeAngle(ps + side * sidePhi);
  };
  phiSections.resize(sides + 1);
  for (int i = 0; i < sides + 1; ++i) {
    const V3 corner = FixZeroes(Normalized(FromCylindrical(1., getPhi(i), 0)));
    phiSections[i]  = Normalized(corner).Cross(V3(0, 0, -1));
  }
  const double cosHalf = std::cos(0.5 * sidePhi);
  double innerR = kInfLength, outerR = -kInfLength;
  for (int i = 0; i < nz; ++i) {
    if (rmin[i] < innerR) innerR = rmin[i];
    if (rmax[i] > outerR) outerR = rmax[i];
  }
  outerR /= cosHalf;
  const double bz       = 0.5 * (zs[nz - 1] - zs[0]) + kTolerance;
  const double phiTol   = 100 * kTolerance;
  const double bPhiS    = !hasPhiCutout ? 0 : ps - phiTol;
  const double bPhiD    = !hasPhiCutout ? kTwoPi : pd + 2 * phiTol;
  bounding              = Tube(innerR - kHalfTolerance, outerR + kHalfTolerance, bz, bPhiS, bPhiD);
  boundingOffset        = 0.5 * (zs[0] + zs[nz - 1]);
  auto vertex = [&](std::vector<double> const &r, int i, int j) {
    if (!hasPhiCutout && j == sides) j = 0;
    return FixZeroes(FromCylindrical(r[i] / cosHalf, getPhi(j), zs[i]));
  };
  auto wrongNormal = [](V3 const &n, V3 const &c) { return n[0] * c[0] + n[1] * c[1] < 0; };
  for (int ip = 0; ip < nz - 1; ++ip) {
    ZSeg &S = segs[ip];
    for (int is = 0; is < S.outer.size(); ++is) {
      S.outer.Set(is, vertex(rmax, ip, is), vertex(rmax, ip, is + 1), vertex(rmax, ip + 1, is + 1),
                  vertex(rmax, ip + 1, is));
      if (wrongNormal(S.outer.q[is].n, vertex(rmax, ip, is))) S.outer.FlipSign(is);
    }
    for (int is = 0; is < S.inner.size(); ++is) {
      S.inner.Set(is, vertex(rmin, ip, is), vertex(rmin, ip, is + 1), vertex(rmin, ip + 1, is + 1),
                  vertex(rmin, ip + 1, is));
      if (wrongNormal(S.inner.q[is].n, vertex(rmin, ip, is))) S.inner.FlipSign(is);
    }
    if (hasPhiCutout && S.phi.size() == 2) {
      S.phi.Set(0, vertex(rmin, ip, 0), vertex(rmin, ip + 1, 0), vertex(rmax, ip + 1, 0), vertex(rmax, ip, 0));
      if (S.phi.q[0].n.Dot(phiSections[0]) > 0) S.phi.FlipSign(0);
      S.phi.Set(1, vertex(rmax, ip, sides), vertex(rmax, ip + 1, sides), vertex(rmin, ip + 1, sides),
                vertex(rmin, ip, sides));
      if (S.phi.q[1].n.Dot(phiSections[sides]) < 0) S.phi.FlipSign(1);
    }
  }
}

// PolyhedronImplementation.h:298-324 (FindZSegmentKernel)
int Polyhedron::FindZSegment(double z) const
{
  int index         = -1;
  const double *b   = zs.data();
  const double *end = zs.data() + zs.size();
  while (b < end - 1 && z - kTolerance > *b) {
    ++index;
    ++b;
  }
  if (z + kTolerance > *b) return index + 1;
  return index;
}

// PolyhedronImplementation.h:326-346
int Polyhedron::FindPhiSegment(V3 const &p) const
{
  int index     = -1;
  double first  = p[0] * phiSections[0].x + p[1] * phiSections[0].y + p[2] * phiSections[0].z;
  for (int i = 1; i < sides + 1; ++i) {
    const double second = p[0] * phiSections[i].x + p[1] * phiSections[i].y + p[2] * phiSections[i].z;
    if (first > -kTolerance && second < kTolerance) index = i - 1;
    if (index >= 0) break;
    first = second;
  }
  return index;
}

// PolyhedronImplementation.h:348-375
double Polyhedron::DistanceToInZSegment(int i, V3 const &p, V3 const &d) const
{
  ZSeg const &S   = segs[i];
  double distance = S.outer.DistanceToIn(false, p, d);
  bool done       = distance < kInfLength;
  if (done) return distance;
  if (hasPhiCutout) distance = S.phi.DistanceToIn(false, p, d);
  done = done || distance < kInfLength;
  if (done) return distance;
  if (S.hasInner()) distance = S.inner.DistanceToIn(true, p, d);
  return distance;
}

// PolyhedronImplementation.h:377-409
double Polyhedron::DistanceToOutZSegment(int i, double zMin, double zMax, V3 const &p, V3 const &d) const
{
  bool done       = false;
  double distance = kInfLength;
  ZSeg const &S   = segs[i];
  if (S.hasInner()) {
    distance = S.inner.DistanceToIn(false, p, d);
    done     = distance < kInfLength;
    if (done) return distance;
  }
  if (hasPhiCutout) {
    const double dphi = S.phi.DistanceToIn(true, p, d);
    if (!done && distance > -kTolerance) distance = dphi;
    done = distance > -kTolerance && distance < kInfLength;
    if (done) return distance;
  }
  const double dout = S.outer.DistanceToOut(p, d, zMin, zMax);
  if (!done) distance = dout;
  return distance;
}

// PolyhedronImplementation.h:411-557
double Polyhedron::SafetyToZSegmentSquared(int i, int &phiIndex, V3 const &p, bool ptInside, int *iSurf) const
{
  // iSurf (when given) receives the surface kind of the minimum as the reference's Normal uses it:
  // 0 / 1 the phi planes, 2 the inner, 3 the outer quadrilaterals (left unchanged if none is below kInfLength)
  auto mark = [iSurf](int k) {
    if (iSurf) *iSurf = k;
  };
  ZSeg const &S       = segs[i];
  const bool inCutout = phiIndex < 0;
  double sq           = kInfLength;
  if (hasPhiCutout && S.phi.size() == 2) {
    if (ptInside || inCutout) {
      mark(0);
      sq               = S.phi.DistanceSquared(0, p);
      const double saf = S.phi.DistanceSquared(1, p);
      if (saf < sq) {
        sq = saf;
        mark(1);
      }
      if (inCutout) {
        if (hasLargePhiCutout && S.hasInner()) {
          if (S.inner.size() > 0) {
            double si = S.inner.DistanceSquared(0, p);
            if (si < sq) {
              mark(2);
              phiIndex = 0;
              sq       = si;
            }
            if (S.inner.size() > 1) {
              si = S.inner.DistanceSquared(S.inner.size() - 1, p);
              if (si < sq) {
                mark(2);
                phiIndex = S.inner.size() - 1;
                sq       = si;
              }
            }
          }
        }
        return sq;
      }
    }
  }
  if (inCutout && i > 0 && i < NSeg() - 1 && zs[i] == zs[i + 1]) {
    double so = kInfLength;
    if (S.outer.size() > 0) {
      so = S.outer.DistanceSquared(0, p);
      if (so < sq) {
        mark(3);
        phiIndex = 0;
        sq       = so;
      }
      if (S.outer.size() > 1) {
        so = S.outer.DistanceSquared(S.outer.size() - 1, p);
        if (so < sq) {
          mark(3);
          phiIndex = S.outer.size() - 1;
          sq       = so;
        }
      }
    }
    double si = kInfLength;
    if (S.hasInner()) {
      if (S.inner.size() > 0) {
        si = S.inner.DistanceSquared(0, p);
        if (si < sq) {
          mark(2);
          phiIndex = 0;
          sq       = si;
        }
        if (S.inner.size() > 1) {
          si = S.inner.DistanceSquared(S.inner.size() - 1, p);
          if (si < sq) {
            mark(2);
            phiIndex = S.inner.size() - 1;
            sq       = si;
          }
        }
      }
    }
    return sq;
  }
  double so = kInfLength;
  if (S.outer.size() > 0) so = S.outer.DistanceSquared(phiIndex, p);
  double si = kInfLength;
  if (S.hasInner()) {
    if (S.inner.size() > 0) si = S.inner.DistanceSquared(phiIndex, p);
  }
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

// PolyhedronImplementation.h:559-614
void Polyhedron::DistanceToEndcaps(bool inside, V3 const &p, V3 const &d, double &distance) const
{
  ZSeg const *S;
  double zPlane;
  const bool first = inside ? (d[2] < 0) : !(d[2] < 0);
  if (first) {
    S      = &segs[0];
    zPlane = zs[0];
  } else {
    S      = &segs[NSeg() - 1];
    zPlane = zs[NSeg()];
  }
  const double dt = (zPlane - p[2]) / NonZero(d[2]);
  if (dt < -kTolerance || dt >= distance) return;
  const V3 x = p + dt * d;
  if (!S->outer.Contains(x)) return;
  if (S->hasInner() && S->inner.Contains(x)) return;
  if (hasPhiCutout && InPhiCutoutWedge(*S, x)) return;
  distance = dt;
}

// PolyhedronImplementation.h:616-653
void Polyhedron::SafetyToEndcapsSquared(V3 const &p, double &dsq, int *iz) const
{
  if (iz) *iz = 0;
  const double firstD = zs[0] - p[2];
  const double lastD  = zs[NSeg()] - p[2];
  const bool isFirst  = std::fabs(firstD) < std::fabs(lastD);
  ZSeg const &S       = isFirst ? segs[0] : segs[NSeg() - 1];
  const double dt     = isFirst ? firstD : lastD;
  const double dt2    = dt * dt;
  if (dt2 >= dsq) return;
  const V3 x(p[0], p[1], p[2] + dt);
  if (!S.outer.Contains(x)) return;
  if (S.hasInner() && S.inner.Contains(x)) return;
  if (hasPhiCutout && InPhiCutoutWedge(S, x)) return;
  if (iz) *iz = isFirst ? -1 : 1;
  dsq = dt2;
}

// PolyhedronImplementation.h:655-672
bool Polyhedron::InPhiCutoutWedge(ZSeg const &S, V3 const &p) const
{
  const bool s0 = p.Dot(S.phi.q[0].n) + S.phi.q[0].d >= 0;
  const bool s1 = p.Dot(S.phi.q[1].n) + S.phi.q[1].d >= 0;
  if (hasLargePhiCutout) return s0 || s1;
  return s0 && s1;
}

int Polyhedron::WedgeInside(V3 const &p) const
{
  bool ci = false, co = false;
  wedge.ContainsInside(p, true, ci, co);
  int r = kSurface;
  if (co) r = kOutside;
  if (ci) r = kInside;
  return r;
}

// PolyhedronImplementation.h:674-733
bool Polyhedron::Contains(V3 const &p) const
{
  if (!bounding.Contains(V3(p[0], p[1], p[2] - boundingOffset))) return false;
  const int z = FindZSegment(p[2]);
  if (!((z >= 0) && (z < NSeg()))) return false;
  ZSeg const &S = segs[z];
  if (sameZ[z]) {
    const int phiIndex = FindPhiSegment(p);
    if (phiIndex < 0) return false;
    V3 const &vout      = S.outer.size() ? S.outer.q[phiIndex].sn[0] : S.inner.q[phiIndex].sn[0];
    const double rdv    = std::fabs(p.Dot(vout));
    const bool in1      = (rdv >= rmin[z]) && (rdv <= rmax[z]);
    const bool in2      = (rdv >= rmin[z + 1]) && (rdv <= rmax[z + 1]);
    return in1 | in2;
  }
  if (!S.outer.Contains(p)) return false;
  if (S.hasInner() && S.inner.Contains(p)) return false;
  if (hasPhiCutout && !S.phi.Contains(p)) return false;
  return true;
}

// PolyhedronImplementation.h:735-793
int Polyhedron::Inside(V3 const &p) const
{
  if (!bounding.Contains(V3(p[0], p[1], p[2] - boundingOffset))) return kOutside;
  int z = FindZSegment(p[2]);
  if (z > (NSeg() - 1)) z = NSeg() - 1;
  if (z < 0) z = 0;
  ZSeg const &S = segs[z];
  if (sameZ[z]) return InsideSegBorder(p, z);
  const int io = S.outer.Inside(p);
  if (io != kInside) return io;
  if (S.hasInner()) {
    const int ii = S.inner.Inside(p);
    if (ii == kInside) return kOutside;
    if (ii == kSurface) return kSurface;
  }
  if (hasPhiCutout) {
    const int ip = S.phi.Inside(p);
    if (ip != kInside) return ip;
  }
  const double dz = std::fabs(std::fabs(p[2] - boundingOffset) - 0.5 * (zs[NSeg()] - zs[0]));
  if (dz < kTolerance) return kSurface;
  return kInside;
}

// PolyhedronImplementation.h:795-833
int Polyhedron::InsideSegBorder(V3 const &p, int z) const
{
  ZSeg const &S      = segs[z];
  const int phiIndex = FindPhiSegment(p);
  if (phiIndex < 0) return kOutside;
  V3 const &vout   = S.outer.size() ? S.outer.q[phiIndex].sn[0] : S.inner.q[phiIndex].sn[0];
  const double rdv = std::fabs(p.Dot(vout));
  const bool in1   = (rdv > rmin[z] - kTolerance) && (rdv < rmax[z] + kTolerance);
  const bool in2   = (rdv > rmin[z + 1] - kTolerance) && (rdv < rmax[z + 1] + kTolerance);
  if (in1 && in2) {
    if ((rdv < rmin[z] + kTolerance) || (rdv > rmax[z] - kTolerance) || (rdv < rmin[z + 1] + kTolerance) ||
        (rdv > rmax[z + 1] - kTolerance))
      return kSurface;
    if (hasPhiCutout) return WedgeInside(p);
    return kInside;
  }
  if (!in1 && !in2) return kOutside;
  return kSurface;
}

// PolyhedronImplementation.h:835-875
int Polyhedron::InsideSegPhi(V3 const &p, int z, int phiIndex) const
{
  if (phiIndex < 0) return kOutside;
  const double dz = std::fabs(p[2] - boundingOffset) - 0.5 * (zs[NSeg()] - zs[0]);
  if (dz > kHalfTolerance) return kOutside;
  if (sameZ[z]) return InsideSegBorder(p, z);
  ZSeg const &S = segs[z];
  const int io  = S.outer.InsideAt(p, phiIndex);
  if (io != kInside) return io;
  if (S.hasInner()) {
    const int ii = S.inner.InsideAt(p, phiIndex);
    if (ii == kInside) return kOutside;
    if (ii == kSurface) return kSurface;
  }
  if (hasPhiCutout) {
    const int ip = WedgeInside(p);
    if (ip != kInside) return ip;
  }
  if (std::fabs(dz) < kHalfTolerance) return kSurface;
  return kInside;
}

// PolyhedronImplementation.h:877-948
double Polyhedron::DistanceToIn(V3 const &p, V3 const &d, double step_max) const
{
  if ((p[2] < zs[0] + kTolerance) && d[2] <= 0) return kInfLength;
  if ((p[2] > zs[NSeg()] - kTolerance) && d[2] >= 0) return kInfLength;
  if (Inside(p) == kInside) return -1.;
  double tubeDistance = 0.;
  const V3 bp(p[0], p[1], p[2] - boundingOffset);
  if (!bounding.Contains(bp)) {
    tubeDistance = bounding.DistanceToIn(bp, d, step_max);
    if (tubeDistance == kInfLength) return kInfLength;
  }
  int z          = FindZSegment(p[2]);
  const int zMax = NSeg();
  z              = z < 0 ? 0 : (z >= zMax ? zMax - 1 : z);
  const bool goingRight = d[2] >= 0;
  double distance       = kInfLength;
  if (goingRight) {
    for (; z < zMax; ++z) {
      distance = DistanceToInZSegment(z, p, d);
      if (distance >= 0 && distance < kInfLength) break;
    }
  } else {
    for (; z >= 0; --z) {
      distance = DistanceToInZSegment(z, p, d);
      if (distance >= 0 && distance < kInfLength) break;
    }
  }
  DistanceToEndcaps(false, p, d, distance);
  return (distance >= tubeDistance - 1E-6) ? distance : kInfLength;
}

// PolyhedronImplementation.h:950-985
double Polyhedron::Safety(V3 const &p, bool ptInside) const
{
  double safety  = kInfLength;
  const int zMax = NSeg();
  int z          = FindZSegment(p[2]);
  z              = z < 0 ? 0 : (z >= zMax ? zMax - 1 : z);
  int phiIndex   = FindPhiSegment(p);
  const int in   = InsideSegPhi(p, z, phiIndex);
  if (in == kSurface) return 0.;
  const bool contains = (in == kInside);
  if (contains ^ ptInside) return -1.;
  for (int iz = z; iz < zMax;) {
    safety = std::min(safety, SafetyToZSegmentSquared(iz, phiIndex, p, ptInside));
    ++iz;
    const double dz = zs[iz] - p[2];
    if (dz * dz > safety) break;
  }
  for (int iz = z - 1; iz >= 0; --iz) {
    safety          = std::min(safety, SafetyToZSegmentSquared(iz, phiIndex, p, ptInside));
    const double dz = p[2] - zs[iz];
    if (dz * dz > safety) break;
  }
  SafetyToEndcapsSquared(p, safety);
  return std::sqrt(safety);
}

// PolyhedronImplementation.h:1051-1085
double Polyhedron::DistanceToOut(V3 const &p, V3 const &d, double) const
{
  const int zMax = NSeg();
  if ((p[2] < zs[0] - kTolerance) || (p[2] > zs[zMax] + kTolerance)) return -1.;
  if (Inside(p) == kOutside) return -1.;
  int z = FindZSegment(p[2]);
  z     = z < 0 ? 0 : (z >= zMax ? zMax - 1 : z);
  const bool goingRight = d[2] >= 0;
  double distance       = kInfLength;
  if (goingRight) {
    for (; z < zMax; ++z) {
      distance = DistanceToOutZSegment(z, zs[z], zs[z + 1], p, d);
      if (distance >= 0 && distance < kInfLength) break;
      if (zs[z] - p[2] > distance) break;
    }
  } else {
    for (; z >= 0; --z) {
      distance = DistanceToOutZSegment(z, zs[z], zs[z + 1], p, d);
      if (distance >= 0 && distance < kInfLength) break;
      if (p[2] - zs[z] > distance) break;
    }
  }
  DistanceToEndcaps(true, p, d, distance);
  if (distance >= kInfLength) distance = 0.;
  return distance;
}

// source/UnplacedPolyhedron.cpp:273-303
void Polyhedron::Extent(V3 &lo, V3 &hi) const
{
  lo                   = V3(kInfLength, kInfLength, kInfLength);
  hi                   = V3(-kInfLength, -kInfLength, -kInfLength);
  const double sidePhi = phiDelta / sides;
  const double conv    = 1. / std::cos(0.5 * sidePhi);
  for (int iphi = 0; iphi <= sides; ++iphi) {
    const double phi  = phiStart + iphi * sidePhi;
    const double corx = conv * std::cos(phi), cory = conv * std::sin(phi);
    for (int k = 0; k < nz; ++k) {
      const V3 a(rmin[k] * corx, rmin[k] * cory, zs[k]);
      const V3 b(rmax[k] * corx, rmax[k] * cory, zs[k]);
      for (int i = 0; i < 3; ++i) {
        lo[i] = std::min(lo[i], a[i]);
        hi[i] = std::max(hi[i], a[i]);
      }
      for (int i = 0; i < 3; ++i) {
        lo[i] = std::min(lo[i], b[i]);
        hi[i] = std::max(hi[i], b[i]);
      }
    }
  }
}

} // namespace vgo
