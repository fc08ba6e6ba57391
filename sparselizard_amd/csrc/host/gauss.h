// Gauss quadrature rules of the reference (src/gausspoint/).
#pragma once
#include <vector>
#include "common.h"

namespace slx {

struct GaussRule {
    std::vector<double> coords;   // 3 reference coordinates per point
    std::vector<double> weights;
    int count() const { return (int)weights.size(); }
};

// gausspoints(elementtypenumber, integrationorder) (src/gausspoint/gausspoints.cpp:4-54).
GaussRule gauss_rule(int type, int order);
// Number of line points for an integration order (gpline.cpp:3-9).
int line_point_count(int order);

}  // namespace slx
