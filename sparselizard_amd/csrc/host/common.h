// Shared host-side definitions: error reporting and reference-element topology.
//
// Element type numbers follow the reference (src/element.cpp:8, 999-1096):
//   0 point, 1 line, 2 triangle, 3 quadrangle, 4 tetrahedron, 5 hexahedron,
//   6 prism, 7 pyramid.  Node/edge/face definitions are the reference's
//   tables because DOF numbering, orientation and basis ordering all derive
//   from them.
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>
#include <sstream>

namespace slx {

// Errors are thrown as std::runtime_error like logs::error() (src/logs.cpp:9-12);
// the C ABI converts them into status codes.
[[noreturn]] inline void fail(const std::string& msg) { throw std::runtime_error(msg); }

#define SLX_CHECK(cond, msg)                                                   \
    do {                                                                       \
        if (!(cond)) {                                                         \
            std::ostringstream _slx_os;                                        \
            _slx_os << msg;                                                    \
            ::slx::fail(_slx_os.str());                                        \
        }                                                                      \
    } while (0)

enum ElemType { POINT = 0, LINE = 1, TRI = 2, QUAD = 3, TET = 4, HEX = 5, PRISM = 6, PYRAMID = 7 };

namespace topo {

int dimension(int type);
int count_nodes(int type);                 // corner nodes
int count_edges(int type);
int count_tri_faces(int type);
int count_quad_faces(int type);
inline int count_faces(int type) { return count_tri_faces(type) + count_quad_faces(type); }
// Number of sub-entities of dimension `dim` (element::countdim, src/element.cpp:400).
int count_dim(int type, int dim);
// Number of sub-elements of type `subtype` (element::counttype, src/element.cpp:235).
int count_type(int type, int subtype);
// Edge definitions as corner-node pairs (src/element.cpp:999-1030).
const std::vector<int>& edge_nodes(int type);
// Face definitions as corner-node lists, triangles then quadrangles (src/element.cpp:1032-1063).
const std::vector<int>& face_nodes(int type);
// Face definitions as signed 1-based edge lists (src/element.cpp:1065-1096).
const std::vector<int>& face_edges(int type);
bool is_triangular_face(int type, int facenum);
// Gmsh element type → (type, curvature order) for the straight elements handled here.
int from_gmsh(int gmshtype);
int gmsh_order(int gmshtype);  // geometry (curvature) order of a gmsh element type: 1 or 2
int to_gmsh(int type);

}  // namespace topo

}  // namespace slx
