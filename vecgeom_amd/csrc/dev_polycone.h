// Polycone kernels (VecGeom/volumes/kernel/PolyconeImplementation.h:70-520): a stack
// of cone sections, each evaluated in its own frame shifted by the section's mid z.
#pragma once
#include "dev_cone.h"

namespace vgmi {
namespace pcon {

VGMI_DEV ConeData const &sec(PolyconeData const &P, int i)
{
  return reinterpret_cast<ConeData const *>(P.base + P.secOff)[i];
}
VGMI_DEV double shift(PolyconeData const &P, int i) { return P.base[P.shiftOff + i]; }
VGMI_DEV double zb(PolyconeData const &P, int i) { return P.base[P.zsOff + i]; }

// PolyconeStruct::GetSectionIndex (VecGeom/volumes/PolyconeStruct.h:299-308)
template <class R>
VGMI_DEV int SectionIndex(PolyconeData const &P, R z)
{
  if (z < R(zb(P, 0))) return -1;
  for (int i = 0; i < P.nsec; ++i)
    if (z >= R(zb(P, i)) && z <= R(zb(P, i + 1))) return i;
  return -2;
}

template <class R>
VGMI_DEV Vec3<R> local(PolyconeData const &P, int i, Vec3<R> const &p)
{
  return Vec3<R>(p.x - R(0), p.y - R(0), p.z - R(shift(P, i)));
}

template <class R>
VGMI_DEV void SectionInOut(PolyconeData const &P, int i, Vec3<R> const &p, bool forInside, bool &cin, bool &cout)
{
  if (i < 0) {
    cin  = false;
    cout = true;
    return;
  }
  cone::ContainsInside(sec(P, i), local(P, i, p), forInside, cin, cout);
}

template <class R>
VGMI_DEV void ContainsInside(PolyconeData const &P, Vec3<R> const &p, bool forInside, bool &cin, bool &cout)
{
  cin = cout = false;
  const int lo = SectionIndex(P, p.z - R(kTolerance)), hi = SectionIndex(P, p.z + R(kTolerance));
  if (lo < 0 && hi < 0) {
    cout = true;
    return;
  }
  if (lo < 0 && hi == 0) return SectionInOut(P, 0, p, forInside, cin, cout);
  if (hi < 0 && lo == P.nsec - 1) return SectionInOut(P, P.nsec - 1, p, forInside, cin, cout);
  if (lo >= 0 && hi >= 0) {
    if (lo == hi) return SectionInOut(P, lo, p, forInside, cin, cout);
    bool inL = false, outL = false, inH = false, outH = false;
    SectionInOut(P, lo, p, forInside, inL, outL);
    SectionInOut(P, hi, p, forInside, inH, outH);
    if (!inL && !outL && !inH && !outH) {
      cin = true;
      return;
    }
    if (outL && outH) cout = true;
  }
}

template <class R>
VGMI_DEV bool Contains(PolyconeData const &P, Vec3<R> const &p)
{
  bool ci, co;
  ContainsInside(P, p, false, ci, co);
  return !co;
}

template <class R>
VGMI_DEV int Inside(PolyconeData const &P, Vec3<R> const &p)
{
  bool ci, co;
  ContainsInside(P, p, true, ci, co);
  return ci ? kInside : (co ? kOutside : kSurface);
}

template <class R>
VGMI_DEV R DistanceToIn(PolyconeData const &P, Vec3<R> const &p, Vec3<R> const &v)
{
  R distance = InfLen<R>();
  int inc    = v.z > R(0) ? 1 : -1;
  if (vabs(v.z) < R(kTolerance)) inc = 0;
  int index = SectionIndex(P, p.z);
  if (index == -1) index = 0;
  if (index == -2) index = P.nsec - 1;
  do {
    distance = cone::DistanceToIn(sec(P, index), local(P, index, p), v);
    if (distance < InfLen<R>() || !inc) break;
    index += inc;
  } while (index >= 0 && index < P.nsec);
  return distance;
}

template <class R>
VGMI_DEV R DistanceToOut(PolyconeData const &P, Vec3<R> const &point, Vec3<R> const &dir)
{
  if (P.nsec == 1) return cone::DistanceToOut(sec(P, 0), local(P, 0, point), dir);
  const int lo = SectionIndex(P, point.z - R(kTolerance)), hi = SectionIndex(P, point.z + R(kTolerance));
  int index    = 0;
  if (lo < 0 && hi < 0) return R(-1);
  if (lo < 0 && hi >= 0) {
    index = hi;
    if (cone::Inside(sec(P, index), local(P, index, point)) == kOutside) return R(-1);
  } else if (lo != hi && lo >= 0) {
    index = cone::Inside(sec(P, lo), local(P, lo, point)) == kOutside ? hi : lo;
  } else {
    index = lo;
    if (index < 0) index = SectionIndex(P, point.z);
  }
  if (index < 0) return R(0);
  if (cone::Inside(sec(P, index), local(P, index, point)) == kOutside) return R(-1);
  R total  = R(0);
  int inc  = dir.z > R(0) ? 1 : -1;
  if (vabs(dir.z) < R(kTolerance)) inc = 0;
  int istep  = 0;
  Vec3<R> pn = point;
  do {
    ConeData const &S = sec(P, index);
    if constexpr (sizeof(R) == sizeof(float)) {
      // FP32 pass of the mixed path: the hand-off below tests the exit point against
      // the next section with the 1e-9 mm band, but in FP32 that point lies ~1e-6 of
      // its magnitude off the shared z plane. Exits within kMixedPlaneBand of an internal
      // plane are answered NaN, which sends the point to the FP64 re-evaluation.
      if (istep > 0) {
        const R plane = R(zb(P, inc > 0 ? index : index + 1));
        if (vabs(point.z + total * dir.z - plane) < R(kMixedPlaneBand)) return R(__builtin_nanf(""));
      }
    }
    if (total != R(0) || istep < 2) {
      pn = point + total * dir;
      pn.z -= R(shift(P, index));
      if (cone::Inside(S, pn) == kOutside) break;
    } else {
      pn.z -= R(shift(P, index));
    }
    istep++;
    R dist = cone::DistanceToOut(S, pn, dir);
    if (dist == R(-1)) return InfLen<R>();
    if (vabs(dist) < R(0.5 * kTolerance)) {
      int index1 = index;
      if (index > 0 && index < P.nsec - 1) {
        index1 += inc;
      } else {
        if (index == 0 && inc > 0) index1 += inc;
        if (index == P.nsec - 1 && inc < 0) index1 += inc;
      }
      Vec3<R> pte = point + (total + dist) * dir;
      pte.z -= R(shift(P, index1));
      if (cone::Inside(sec(P, index1), pte) == kOutside || inc == 0) break;
    }
    total += dist;
    index += inc;
  } while (inc != 0 && index >= 0 && index < P.nsec);
  return total;
}

// The safeties take the section of the point's z and change abruptly across a z plane, even far
// from the surface: an FP32 z within kMixedPlaneBand of a plane answers NaN (evaluated in FP64).
template <class R>
VGMI_DEV bool NearZPlaneF32(PolyconeData const &P, R z)
{
  if constexpr (sizeof(R) == sizeof(float)) {
    for (int i = 0; i <= P.nsec; ++i)
      if (vabs(z - R(zb(P, i))) < R(kMixedPlaneBand)) return true;
  }
  return false;
}

template <class R>
VGMI_DEV R SafetyToIn(PolyconeData const &P, Vec3<R> const &p)
{
  if (NearZPlaneF32(P, p.z)) return R(__builtin_nanf(""));
  int index  = SectionIndex(P, p.z);
  bool needZ = false;
  if (index < 0) {
    needZ = true;
    if (index == -1) index = 0;
    if (index == -2) index = P.nsec - 1;
  }
  R safety = cone::SafetyToIn(sec(P, index), local(P, index, p));
  if (needZ || safety < R(kTolerance)) return safety;
  R minSafety = safety;
  R zbase     = R(zb(P, index + 1));
  for (int i = index + 1; i < P.nsec; ++i) {
    if (R(zb(P, i)) - zbase >= minSafety) break;
    safety = cone::SafetyToIn(sec(P, i), local(P, i, p));
    if (safety < minSafety) minSafety = safety;
  }
  if (index > 0) {
    zbase = R(zb(P, index - 1));
    for (int i = index - 1; i >= 0; --i) {
      if (zbase - R(zb(P, i)) >= minSafety) break;
      safety = cone::SafetyToIn(sec(P, i), local(P, i, p));
      if (safety < minSafety) minSafety = safety;
    }
  }
  return minSafety;
}

template <class R>
VGMI_DEV R SafetyToOut(PolyconeData const &P, Vec3<R> const &p)
{
  if (NearZPlaneF32(P, p.z)) return R(__builtin_nanf(""));
  bool cin, cout;
  ContainsInside(P, p, true, cin, cout);
  if (cout) return R(-1);
  const int index = SectionIndex(P, p.z);
  if (index < 0) return R(-1);
  R safety    = cone::SafetyToOut(sec(P, index), local(P, index, p));
  R minSafety = safety;
  if (minSafety == InfLen<R>() || minSafety < R(kTolerance)) return R(0);
  R zbase = R(zb(P, index + 1));
  for (int i = index + 1; i < P.nsec; ++i) {
    if (R(zb(P, i)) - zbase >= minSafety) break;
    safety = cone::SafetyToIn(sec(P, i), local(P, i, p));
    if (safety < minSafety) minSafety = safety;
  }
  if (index > 0) {
    zbase = R(zb(P, index - 1));
    for (int i = index - 1; i >= 0; --i) {
      if (zbase - R(zb(P, i)) >= minSafety) break;
      safety = cone::SafetyToIn(sec(P, i), local(P, i, p));
      if (safety < minSafety) minSafety = safety;
    }
  }
  return minSafety;
}

} // namespace pcon
} // namespace vgmi
