// GDML reader (gdml.cpp): the flat geometry description a GDML file defines.
#pragma once
#include <string>
#include <vector>

namespace vgmi {
namespace gdml {

struct Model {
  std::vector<int> shape_type, shape_offset; // shape_offset: nshapes + 1 offsets into shape_params
  std::vector<double> shape_params;
  std::vector<int> logical_shape;
  std::vector<std::string> logical_names;
  std::vector<int> placed_mother, placed_logical;
  std::vector<double> placed_transform; // 12 per placement: tx ty tz r0..r8
  std::vector<std::string> placed_names;
  int world = -1;
};

// Reads `path`; lengths are scaled by `mm` per millimetre; `specialized` applies GeoManager::MakeInstance's
// substitutions of a build with VECGEOM_NO_SPECIALIZATION=OFF. False with a message on failure.
bool Read(std::string const &path, double mm, bool specialized, Model &out, std::string &err);

} // namespace gdml
} // namespace vgmi
