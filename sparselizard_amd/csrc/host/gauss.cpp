// Gauss quadrature rules (src/gausspoint/gausspoints.cpp:4-54).
//
// Line, triangle and tetrahedron points/weights are the reference's tabulated
// data (see gauss_tables.inc); quadrangle and hexahedron rules are tensor
// products of the line rule with the first index outermost
// (gpquadrangle.cpp:12-35, gphexahedron.cpp:12-39).
#include "gauss.h"
#include <cmath>

namespace slx {

#include "gauss_tables.inc"

namespace {
void from_table(const int* off, const double* xyz, const double* w, int k, GaussRule& r) {
    int b = off[k], e = off[k + 1];
    r.coords.assign(xyz + 3 * b, xyz + 3 * e);
    r.weights.assign(w + b, w + e);
}
}  // namespace

int line_point_count(int order) { return (int)std::ceil(0.5 * (order + 1)); }

GaussRule gauss_rule(int type, int order) {
    SLX_CHECK(order >= 0, "cannot get the Gauss points for negative integration order " << order);
    GaussRule r;
    switch (type) {
        case POINT:
            r.coords = {0.0, 0.0, 0.0};
            r.weights = {1.0};
            return r;
        case LINE: {
            int n = line_point_count(order);
            SLX_CHECK(n <= SLX_GP_MAX_LINE_POINTS, "line Gauss rule for order " << order << " is not tabulated");
            from_table(gp_line_off, gp_line_xyz, gp_line_w, n, r);
            return r;
        }
        case TRI:
            SLX_CHECK(order <= SLX_GP_MAX_TRI_ORDER, "triangle Gauss rule for order " << order << " is not tabulated");
            from_table(gp_tri_off, gp_tri_xyz, gp_tri_w, order, r);
            return r;
        case TET:
            SLX_CHECK(order <= SLX_GP_MAX_TET_ORDER, "tetrahedron Gauss rule for order " << order << " is not tabulated");
            from_table(gp_tet_off, gp_tet_xyz, gp_tet_w, order, r);
            return r;
        case QUAD:
        case HEX: {
            GaussRule l = gauss_rule(LINE, order);
            int n = (int)l.weights.size();
            if (type == QUAD) {
                for (int i = 0; i < n; i++)
                    for (int j = 0; j < n; j++) {
                        r.coords.insert(r.coords.end(), {l.coords[3 * i], l.coords[3 * j], 0.0});
                        r.weights.push_back(l.weights[i] * l.weights[j]);
                    }
            } else {
                for (int i = 0; i < n; i++)
                    for (int j = 0; j < n; j++)
                        for (int k = 0; k < n; k++) {
                            r.coords.insert(r.coords.end(), {l.coords[3 * i], l.coords[3 * j], l.coords[3 * k]});
                            r.weights.push_back(l.weights[i] * l.weights[j] * l.weights[k]);
                        }
            }
            return r;
        }
    }
    if (type == PRISM) {
        // line points outer, triangle points inner; (ki, eta) from the
        // triangle, phi from the line (gpprism.cpp:13-49)
        GaussRule l = gauss_rule(LINE, order), t = gauss_rule(TRI, order);
        for (size_t i = 0; i < l.weights.size(); i++)
            for (size_t j = 0; j < t.weights.size(); j++) {
                r.coords.insert(r.coords.end(), {t.coords[3 * j], t.coords[3 * j + 1], l.coords[3 * i]});
                r.weights.push_back(l.weights[i] * t.weights[j]);
            }
        return r;
    }
    fail("Gauss rules are provided for point, line, triangle, quadrangle, tetrahedron, hexahedron and prism elements");
}

}  // namespace slx
