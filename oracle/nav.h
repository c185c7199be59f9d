// Oracle navigation: scalar FP64 restatement of VecGeom's BVH navigation path.
// TEST INFRASTRUCTURE ONLY (see common.h).
//
//  - Transformation3D (VecGeom/base/Transformation3D.h:600-960, source/Transformation3D.cpp:118-185)
//  - ABBoxManager::ComputeABBox (source/ABBoxManager.cpp:109-143)
//  - AABB (VecGeom/base/AABB.h) and BVH build/traversal (source/BVH.cpp:48-385, VecGeom/base/BVH.h:96-344)
//  - NavIndex table (source/NavIndexTable.cpp:9-99, VecGeom/management/NavIndexTable.h:114-165)
//  - NavStateIndex (VecGeom/navigation/NavStateIndex.h:186-265)
//  - VNavigatorHelper<BVHNavigator> step, PrepareOutState, Relocate (VecGeom/navigation/VNavigator.h:160-592,1066-1091)
//  - GlobalLocator (VecGeom/navigation/GlobalLocator.h:35-200), BVHLevelLocator, BVHSafetyEstimator
//  - NewSimpleNavigator brute-force daughter loop (VecGeom/navigation/NewSimpleNavigator.h), the
//    reference's own cross-check for BVH navigation (persistency/gdml/test/TestVGDMLNavigation/
//    TestVGDMLNavigation.cpp:72-90)
#pragma once
#include "shapes.h"
#include <memory>
#include <vector>

namespace vgo {

struct Transform {
  double t[3] = {0, 0, 0};
  double r[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  bool ident = true, hasT = false, hasR = false;
  void SetProperties();
  V3 Transform3(V3 const &m) const;        // master -> local point
  V3 TransformDirection(V3 const &m) const; // master -> local direction
  V3 InverseTransform(V3 const &l) const;  // local -> master point
  V3 InverseTransformDirection(V3 const &l) const; // local -> master direction
  void MultiplyFromRight(Transform const &rhs);
};

struct AABB {
  V3 lo, hi;
  V3 Center() const { return 0.5 * (hi + lo); }
  V3 Size() const { return hi - lo; }
  double SurfaceArea() const
  {
    V3 e = Size();
    return 2. * (e.x * e.y + e.y * e.z + e.z * e.x);
  }
  bool Contains(V3 const &p) const
  {
    return p.x >= lo.x && p.x <= hi.x && p.y >= lo.y && p.y <= hi.y && p.z >= lo.z && p.z <= hi.z;
  }
  double Safety(V3 const &p) const { return ((p - Center()).Abs() - 0.5 * Size()).Max(); }
  void Intersect(V3 const &p, V3 const &invdir, double &tmin, double &tmax) const;
  static AABB Union(AABB const &a, AABB const &b);
};

struct PVol {
  int lv      = -1;
  int childId = 0;
  Transform tr;
};

struct BVHo {
  int depth = 0;
  std::vector<int> primId, offset, nChild;
  std::vector<AABB> nodes, aabbs;
};

struct LVol {
  int shape = -1;
  std::vector<int> daughters; // placed-volume ids in placement order
  std::unique_ptr<BVHo> bvh;  // only for volumes with daughters
};

struct NavState {
  uint32_t nav = 0, last = 0;
  bool boundary = false;
};

class Geometry {
public:
  std::vector<std::unique_ptr<Shape>> shapes;
  std::vector<LVol> lvs;
  std::vector<PVol> pvols;
  int worldPvol = -1;
  int maxDepth  = 0;
  std::vector<uint32_t> navind;
  bool useBVH = true; // false: brute-force daughter loops (NewSimpleNavigator / SimpleLevelLocator)

  // Geometry close: ABBoxes, BVHs, NavIndex table.
  void Close();

  // NavStateIndex accessors
  uint32_t NavInd(uint32_t i) const { return navind[i]; }
  int Top(uint32_t nav) const { return nav > 0 ? int(navind[nav + 1]) : -1; }
  uint32_t Pop(uint32_t nav) const { return nav > 0 ? navind[nav] : 0; }
  uint32_t Push(uint32_t nav, int pvol) const { return nav > 0 ? navind[nav + 3 + pvols[pvol].childId] : 1; }
  int Level(uint32_t nav) const { return nav > 0 ? int(navind[nav + 2] & 0xff) : -1; }
  int NDaughters(uint32_t nav) const { return int(navind[nav + 2] >> 16); }
  void TopMatrix(uint32_t nav, Transform &m) const;

  // placed-volume helpers (VecGeom/volumes/SpecializedPlacedVolImplHelper.h:98-179)
  Shape const &ShapeOf(int pvol) const { return *shapes[lvs[pvols[pvol].lv].shape]; }

  // navigation entry points
  double ComputeStep(V3 const &gp, V3 const &gd, double limit, NavState const &in, NavState &out,
                     bool relocate) const;
  double ComputeSafety(V3 const &gp, NavState const &in) const;
  uint32_t LocateGlobalPoint(int pvol, V3 const &p, uint32_t path, bool top) const;
  // BVH::LevelLocate with an excluded placed volume (VecGeom/base/BVH.h:219-344)
  bool LevelLocate(LVol const &lv, int excl, V3 const &lp, int &pvol, V3 &dlp) const;

private:
  void BuildBVH(LVol &lv);
  void ComputeNodes(BVHo &b, unsigned id, int *first, int *last, unsigned nodes);
  void CheckDaughters(LVol const &lv, V3 const &lp, V3 const &ld, double &step, int last, int &hit) const;
  uint32_t LocateExcl(int pvol, int excl, V3 const &p, uint32_t path) const;
  uint32_t RelocateForceDifferent(V3 const &lp, uint32_t path) const;
  double SafetyToDaughters(LVol const &lv, V3 const &lp, double safety) const;
};

} // namespace vgo
