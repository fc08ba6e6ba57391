// Reference-element topology tables (src/element.cpp:170-445, 823-861, 999-1096).
#include "common.h"

namespace slx {
namespace topo {

int dimension(int type) {
    switch (type) {
        case POINT: return 0;
        case LINE: return 1;
        case TRI: case QUAD: return 2;
        case TET: case HEX: case PRISM: case PYRAMID: return 3;
    }
    fail("unknown element type " + std::to_string(type));
}

// counttype(typenum) for the straight element (src/element.cpp:235-398).
int count_type(int type, int sub) {
    static const int table[8][8] = {
        // sub:  pt ln tri quad tet hex pri pyr
        {1, 0, 0, 0, 0, 0, 0, 0},   // point
        {2, 1, 0, 0, 0, 0, 0, 0},   // line
        {3, 3, 1, 0, 0, 0, 0, 0},   // triangle
        {4, 4, 0, 1, 0, 0, 0, 0},   // quadrangle
        {4, 6, 4, 0, 1, 0, 0, 0},   // tetrahedron
        {8, 12, 0, 6, 0, 1, 0, 0},  // hexahedron
        {6, 9, 2, 3, 0, 0, 1, 0},   // prism
        {5, 8, 4, 1, 0, 0, 0, 1},   // pyramid
    };
    SLX_CHECK(type >= 0 && type < 8 && sub >= 0 && sub < 8, "bad element type " << type << "/" << sub);
    return table[type][sub];
}

int count_nodes(int type) { return count_type(type, 0); }
// As in the reference, a line counts itself as its one edge and a triangle or
// quadrangle itself as its one face (orientation and basis entities).
int count_edges(int type) { return count_type(type, 1); }
int count_tri_faces(int type) { return count_type(type, 2); }
int count_quad_faces(int type) { return count_type(type, 3); }

int count_dim(int type, int dim) {
    switch (dim) {
        case 0: return count_type(type, 0);
        case 1: return count_type(type, 1);
        case 2: return count_type(type, 2) + count_type(type, 3);
        case 3: return count_type(type, 4) + count_type(type, 5) + count_type(type, 6) + count_type(type, 7);
    }
    fail("bad dimension");
}

const std::vector<int>& edge_nodes(int type) {
    static const std::vector<int> t[8] = {
        {},
        {0, 1},
        {0, 1, 1, 2, 2, 0},
        {0, 1, 1, 2, 2, 3, 3, 0},
        {0, 1, 1, 2, 2, 0, 3, 0, 3, 2, 3, 1},
        {0, 1, 0, 3, 0, 4, 1, 2, 1, 5, 2, 3, 2, 6, 3, 7, 4, 5, 4, 7, 5, 6, 6, 7},
        {0, 1, 0, 2, 0, 3, 1, 2, 1, 4, 2, 5, 3, 4, 3, 5, 4, 5},
        {0, 1, 0, 3, 0, 4, 1, 2, 1, 4, 2, 3, 2, 4, 3, 4},
    };
    return t[type];
}

const std::vector<int>& face_nodes(int type) {
    static const std::vector<int> t[8] = {
        {},
        {},
        {0, 1, 2},
        {0, 1, 2, 3},
        {0, 2, 1, 0, 1, 3, 0, 3, 2, 3, 1, 2},
        {0, 3, 2, 1, 0, 1, 5, 4, 0, 4, 7, 3, 1, 2, 6, 5, 2, 3, 7, 6, 4, 5, 6, 7},
        {0, 2, 1, 3, 4, 5, 0, 1, 4, 3, 0, 3, 5, 2, 1, 2, 5, 4},
        {0, 1, 4, 3, 0, 4, 1, 2, 4, 2, 3, 4, 0, 3, 2, 1},
    };
    return t[type];
}

const std::vector<int>& face_edges(int type) {
    static const std::vector<int> t[8] = {
        {},
        {},
        {1, 2, 3},
        {1, 2, 3, 4},
        {-3, -2, -1, 1, -6, 4, -4, 5, 3, 6, 2, -5},
        {2, -6, -4, -1, 1, 5, -9, -3, 3, 10, -8, -2, 4, 7, -11, -5, 6, 8, -12, -7, 9, 11, 12, -10},
        {2, -4, -1, 7, 9, -8, 1, 5, -7, -3, 3, 8, -6, -2, 4, 6, -9, -5},
        {1, 5, -3, -2, 3, -8, 4, 7, -5, 6, 8, -7, 2, -6, -4, -1},
    };
    return t[type];
}

bool is_triangular_face(int type, int facenum) {
    switch (type) {
        case TRI: case TET: return true;
        case PRISM: return facenum < 2;
        case PYRAMID: return facenum < 4;
        default: return false;
    }
}

int from_gmsh(int g) {
    switch (g) {
        case 15: return POINT;
        case 1: case 8: return LINE;
        case 2: case 9: return TRI;
        case 3: case 10: return QUAD;
        case 4: case 11: return TET;
        case 5: case 12: return HEX;
        case 6: case 13: return PRISM;
        case 7: return PYRAMID;
    }
    fail("unsupported gmsh element type " + std::to_string(g) + " (order-1 geometry, or order 2 without pyramids)");
}

int gmsh_order(int g) { return (g >= 8 && g <= 13) ? 2 : 1; }

int to_gmsh(int type) {
    static const int g[8] = {15, 1, 2, 3, 4, 5, 6, 7};
    return g[type];
}

}  // namespace topo
}  // namespace slx
