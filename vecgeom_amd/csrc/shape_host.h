// Host-side shape record construction (product code, no device dependency).
#pragma once
#include "shape_data.h"
#include <string>
#include <vector>

namespace vgmi {
// Builds the record and, for shapes with variable-length data (polycone sections,
// ...), fills `extra` with that data; offsets inside the record index `extra`.
int MakeShapeRecord(int type, double const *params, int nparams, ShapeRecord &r, std::vector<double> &extra,
                    std::string &err);
// Points the record at the device copy of its variable-length block.
void SetShapeExtraBase(ShapeRecord &r, const double *device_base);
void ShapeExtent(ShapeRecord const &r, double lo[3], double hi[3]);
// Tessellated solid: ntri, then 9 doubles per triangle (tessellated_host.cpp).
int MakeTessellated(double const *a, int n, ShapeRecord &r, std::vector<double> &extra, std::string &err);
} // namespace vgmi
