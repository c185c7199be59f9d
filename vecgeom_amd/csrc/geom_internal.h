// Host-side geometry tables and the device copy behind a vgmi_geom handle.
#pragma once
#include "../../include/vgmi.h"
#include "geom_data.h"
#include <cfloat>
#include <string>
#include <vector>

namespace vgmi {

struct HostGeometry {
  std::vector<ShapeRecord> shapes;
  std::vector<PVolRec> pvols;
  std::vector<LVRec> lvs;
  std::vector<int32_t> daughters;
  std::vector<AABBRec> dAABB;
  std::vector<int32_t> nchild, offset, prim;
  std::vector<AABBRec> nodes;
  std::vector<int32_t> bvhDepth;
  std::vector<uint32_t> navind;
  std::vector<double> extraPool;      // variable-length shape data of all shapes
  std::vector<int64_t> extraOffset;   // per shape offset into extraPool (-1: none)
  int worldPvol = -1;
  int maxDepth  = 0;
};

int build_host_geometry(vgmi_geom_desc const &d, HostGeometry &h, std::string &err);

} // namespace vgmi

struct vgmi_geom {
  vgmi::HostGeometry host;
  vgmi::GeomView view{};   // device pointers (all null for a host-only build)
  void *device_block = nullptr;
  size_t device_bytes = 0;
  int device = -1;
  std::vector<vgmi::ShapeRecord> devShapes; // shape records as uploaded (variable data at device addresses)
  unsigned typeMask = 0; // bit t set when a shape of ShapeType t is present
  int nlogical = 0;
  int regroupSteps = 10; // transport loops regroup tracks by volume every this many steps (0: never)
  int lanesPerWave = 0;  // tracks per wavefront of the transport loops (0: chosen from the track count)
};
