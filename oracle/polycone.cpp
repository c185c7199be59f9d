// Oracle: Polycone. TEST INFRASTRUCTURE ONLY (see common.h).
#include "shapes.h"

namespace vgo {

// PolyconeStruct::Init (VecGeom/volumes/PolyconeStruct.h:144-290)
Polycone::Polycone(double sphi_, double dphi_, int nz, const double *z, const double *rmin, const double *rmax)
    : sphi(sphi_), dphi(dphi_)
{
  std::vector<double> Z(z, z + nz), Rmin(rmin, rmin + nz), Rmax(rmax, rmax + nz);
  if (Z[0] > Z[nz - 1]) {
    for (int i = 0; i < nz; ++i) Z[i] = z[nz - 1 - i], Rmin[i] = rmin[nz - 1 - i], Rmax[i] = rmax[nz - 1 - i];
  }
  double prevZ = Z[0], prevRmax = 0, prevRmin = 0;
  int dirZ = Z[1] < Z[0] ? -1 : 1;
  for (int i = 0; i < nz; ++i) {
    rmaxExtent = std::max(rmaxExtent, Rmax[i]);
    if (i > 0) {
      if (((Z[i] > prevZ + kTolerance) && (dirZ > 0)) || ((Z[i] < prevZ - kTolerance) && (dirZ < 0))) {
        double dz = (Z[i] - prevZ) / 2;
        secs.emplace_back(prevRmin, prevRmax, Rmin[i], Rmax[i], dz, sphi_, dphi_);
        zs.push_back(Z[i]);
        size_t zi = zs.size() - 1;
        shift.push_back(zs[zi - 1] + 0.5 * (zs[zi] - zs[zi - 1]));
      }
    } else {
      zs.push_back(Z[i]);
    }
    prevZ    = Z[i];
    prevRmin = Rmin[i];
    prevRmax = Rmax[i];
  }
  zlo = zs.front();
  zhi = zs.back();
  // fStartPhi / fDeltaPhi: SetAndCheckDPhiAngle, SetAndCheckSPhiAngle, then the open/closed
  // conversion (PolyconeStruct.h:152-184, :355-390)
  deltaPhi = dphi_;
  startPhi = sphi_;
  if (dphi_ >= kTwoPi - 0.5 * kAngTolerance) {
    deltaPhi = kTwoPi;
    startPhi = 0;
  } else if (dphi_ > 0) {
    deltaPhi = dphi_;
  }
  startPhi = sphi_ < 0 ? kTwoPi - std::fmod(std::fabs(sphi_), kTwoPi) : std::fmod(sphi_, kTwoPi);
  if (startPhi + deltaPhi > kTwoPi) startPhi -= kTwoPi;
  if (dphi_ <= 0. || dphi_ > kTwoPi - kTolerance) {
    startPhi = 0;
    deltaPhi = kTwoPi;
  } else {
    startPhi = sphi_;
    while (startPhi < 0) startPhi += kTwoPi;
  }
}

// source/UnplacedPolycone.cpp:865-898: z from the planes, x/y from the extent of the cone
// (minR, maxR, minR, maxR, 1, fStartPhi, fDeltaPhi) over the sections' constructor radii
void Polycone::Extent(V3 &lo, V3 &hi) const
{
  double maxR = 0, minR = kInfLength;
  for (auto const &c : secs) {
    maxR = std::max(maxR, std::max(c.origRmax1, c.origRmax2));
    minR = std::min(minR, std::min(c.origRmin1, c.origRmin2));
  }
  const Cone tmp(minR, maxR, minR, maxR, 1, startPhi, deltaPhi);
  tmp.Extent(lo, hi);
  lo.z = zlo;
  hi.z = zhi;
}

// PolyconeStruct.h:299-308 GetSectionIndex
int Polycone::SectionIndex(double z) const
{
  if (z < zs[0]) return -1;
  for (int i = 0; i < NSec(); ++i)
    if (z >= zs[i] && z <= zs[i + 1]) return i;
  return -2;
}

// PolyconeImplementation.h:72-96 GenericKernelForASection
void Polycone::SectionInOut(int isec, V3 const &p, bool forInside, bool &cin, bool &cout) const
{
  if (isec < 0) {
    cin  = false;
    cout = true;
    return;
  }
  V3 lp = p - V3(0, 0, shift[isec]);
  secs[isec].ContainsInside(lp, forInside, cin, cout);
}

// PolyconeImplementation.h:118-160
void Polycone::ContainsInside(V3 const &p, bool forInside, bool &cin, bool &cout) const
{
  cin = cout = false;
  int lo = SectionIndex(p.z - kTolerance), hi = SectionIndex(p.z + kTolerance);
  if (lo < 0 && hi < 0) {
    cout = true;
    return;
  }
  if (lo < 0 && hi == 0) return SectionInOut(0, p, forInside, cin, cout);
  if (hi < 0 && lo == NSec() - 1) return SectionInOut(NSec() - 1, p, forInside, cin, cout);
  if (lo >= 0 && hi >= 0) {
    if (lo == hi) return SectionInOut(lo, p, forInside, cin, cout);
    bool inL = false, outL = false, inH = false, outH = false;
    SectionInOut(lo, p, forInside, inL, outL);
    SectionInOut(hi, p, forInside, inH, outH);
    bool surfL = !inL && !outL, surfH = !inH && !outH;
    if (surfL && surfH) {
      cin = true;
      return;
    }
    if (outL && outH) cout = true;
  }
}

bool Polycone::Contains(V3 const &p) const
{
  bool ci = false, co = false;
  ContainsInside(p, false, ci, co);
  return !co;
}

int Polycone::Inside(V3 const &p) const
{
  bool ci = false, co = false;
  ContainsInside(p, true, ci, co);
  int r = kSurface;
  if (co) r = kOutside;
  if (ci) r = kInside;
  return r;
}

// PolyconeImplementation.h:162-200
double Polycone::DistanceToIn(V3 const &p, V3 const &v, double step) const
{
  double distance = kInfLength;
  int inc         = v.z > 0 ? 1 : -1;
  if (std::fabs(v.z) < kTolerance) inc = 0;
  int index = SectionIndex(p.z);
  if (index == -1) index = 0;
  if (index == -2) index = NSec() - 1;
  do {
    distance = secs[index].DistanceToIn(p - V3(0, 0, shift[index]), v, step);
    if (distance < kInfLength || !inc) break;
    index += inc;
  } while (index >= 0 && index < NSec());
  return distance;
}

// PolyconeImplementation.h:202-330
double Polycone::DistanceToOut(V3 const &point, V3 const &dir, double step) const
{
  if (NSec() == 1) return secs[0].DistanceToOut(point - V3(0, 0, shift[0]), dir, step);
  int lo = SectionIndex(point.z - kTolerance), hi = SectionIndex(point.z + kTolerance);
  int index = 0;
  if (lo < 0 && hi < 0) return -1;
  if (lo < 0 && hi >= 0) {
    index = hi;
    if (secs[index].Inside(point - V3(0, 0, shift[index])) == kOutside) return -1;
  } else if (lo != hi && lo >= 0) {
    index = secs[lo].Inside(point - V3(0, 0, shift[lo])) == kOutside ? hi : lo;
  } else {
    index = lo;
    if (index < 0) index = SectionIndex(point.z);
  }
  if (index < 0) return 0.;
  if (secs[index].Inside(point - V3(0, 0, shift[index])) == kOutside) return -1;
  double total = 0.;
  int inc      = dir.z > 0 ? 1 : -1;
  if (std::fabs(dir.z) < kTolerance) inc = 0;
  int istep = 0;
  V3 pn     = point;
  do {
    Cone const &sec = secs[index];
    if (total != 0 || istep < 2) {
      pn = point + total * dir;
      pn.z -= shift[index];
      if (sec.Inside(pn) == kOutside) break;
    } else {
      pn.z -= shift[index];
    }
    istep++;
    double dist = sec.DistanceToOut(pn, dir, step);
    if (dist == -1) return kInfLength; // the reference returns with its output still at kInfLength
    if (std::fabs(dist) < 0.5 * kTolerance) {
      int index1 = index;
      if (index > 0 && index < NSec() - 1) {
        index1 += inc;
      } else {
        if (index == 0 && inc > 0) index1 += inc;
        if (index == NSec() - 1 && inc < 0) index1 += inc;
      }
      V3 pte = point + (total + dist) * dir;
      pte.z -= shift[index1];
      if (secs[index1].Inside(pte) == kOutside || inc == 0) break;
    }
    total += dist;
    index += inc;
  } while (inc != 0 && index >= 0 && index < NSec());
  return total;
}

// PolyconeImplementation.h:332-380
double Polycone::SafetyToIn(V3 const &p) const
{
  int index  = SectionIndex(p.z);
  bool needZ = false;
  if (index < 0) {
    needZ = true;
    if (index == -1) index = 0;
    if (index == -2) index = NSec() - 1;
  }
  double safety = secs[index].SafetyToIn(p - V3(0, 0, shift[index]));
  if (needZ) return safety;
  if (safety < kTolerance) return safety;
  double minSafety = safety;
  double zbase     = zs[index + 1];
  for (int i = index + 1; i < NSec(); ++i) {
    double dz = zs[i] - zbase;
    if (dz >= minSafety) break;
    safety = secs[i].SafetyToIn(p - V3(0, 0, shift[i]));
    if (safety < minSafety) minSafety = safety;
  }
  if (index > 0) {
    zbase = zs[index - 1];
    for (int i = index - 1; i >= 0; --i) {
      double dz = zbase - zs[i];
      if (dz >= minSafety) break;
      safety = secs[i].SafetyToIn(p - V3(0, 0, shift[i]));
      if (safety < minSafety) minSafety = safety;
    }
  }
  return minSafety;
}

// PolyconeImplementation.h:382-460
double Polycone::SafetyToOut(V3 const &point) const
{
  bool cin = false, cout = false;
  ContainsInside(point, true, cin, cout);
  if (cout) return -1;
  int index = SectionIndex(point.z);
  if (index < 0) return -1;
  double safety    = secs[index].SafetyToOut(point - V3(0, 0, shift[index]));
  double minSafety = safety;
  if (minSafety == kInfLength) return 0.;
  if (minSafety < kTolerance) return 0.;
  double zbase = zs[index + 1];
  for (int i = index + 1; i < NSec(); ++i) {
    double dz = zs[i] - zbase;
    if (dz >= minSafety) break;
    safety = secs[i].SafetyToIn(point - V3(0, 0, shift[i]));
    if (safety < minSafety) minSafety = safety;
  }
  if (index > 0) {
    zbase = zs[index - 1];
    for (int i = index - 1; i >= 0; --i) {
      double dz = zbase - zs[i];
      if (dz >= minSafety) break;
      safety = secs[i].SafetyToIn(point - V3(0, 0, shift[i]));
      if (safety < minSafety) minSafety = safety;
    }
  }
  return minSafety;
}

} // namespace vgo
