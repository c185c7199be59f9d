// Host-side shape record construction (product code, no device dependency).
#pragma once
#include "shape_data.h"
#include <string>

namespace vgmi {
int MakeShapeRecord(int type, double const *params, int nparams, ShapeRecord &r, std::string &err);
void ShapeExtent(ShapeRecord const &r, double lo[3], double hi[3]);
} // namespace vgmi
