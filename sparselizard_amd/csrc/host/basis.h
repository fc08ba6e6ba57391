// Hierarchical H1 shape functions (Zaglmayr) for line, triangle, quadrangle,
// tetrahedron and hexahedron elements, evaluated with their reference
// derivatives at arbitrary points.
//
// Follows src/shapefunction/hierarchical/h1/h1{line,triangle,quadrangle,
// tetrahedron,hexahedron}.cpp, legendre.cpp, orientation.cpp and the form
// function iterator order of hierarchicalformfunctioniterator.cpp:20-56
// (vertex, edge, face, volume; inside an entity by increasing order).
#pragma once
#include <vector>
#include "common.h"

namespace slx {

// Number of H1 functions of order <= `order` on the whole element
// (h1*::count(int), e.g. src/shapefunction/hierarchical/h1/h1triangle.cpp:6-12).
int h1_count(int type, int order);
// Number of H1 functions of order <= `order` on ONE sub-entity of dimension `dim`
// (h1*::count(order, dim, num), e.g. h1triangle.cpp:14-48).
int h1_count(int type, int order, int dim);
// h1*::isorientationdependent (h1triangle.h:44): orders above 2.
inline bool h1_orientation_dependent(int order) { return order > 2; }

struct FormFunction {
    int dim;            // 0 vertex, 1 edge, 2 face, 3 volume
    int num;            // node/edge/face index in the element
    int index;          // index of the function inside its entity
    int order;          // polynomial order of the function
    int l;              // index among the functions of that order in the entity
    int assoc_type;     // element type of the sub-entity carrying the function
    int sub;            // index of that sub-entity among the element's sub-entities of
                        // type assoc_type (a prism's quadrangular faces 2-4 are quadrangles 0-2)
};

// Form functions in the reference iterator order.
std::vector<FormFunction> h1_iterate(int type, int order);

// Total orientation helpers (src/shapefunction/hierarchical/orientation.cpp:66-291).
int orientation_of_edge(long long a, long long b);
int orientation_of_triangle(const long long* n);
int orientation_of_quadrangle(const long long* n);
int total_orientation(int type, const long long* corner_nodes);
void orientations_from_total(int type, int total, std::vector<int>& edges, std::vector<int>& faces);

// tomatrix (hierarchicalformfunctioncontainer.cpp:78-112): values of every form
// function (rows, iterator order) at every point (columns), for the given total
// orientation and derivative (0 value, 1 d/dki, 2 d/deta, 3 d/dphi).
// `pts` holds 3 reference coordinates per point.  out[deriv][ff][pt].
void h1_tables(int type, int order, int totalorientation, const std::vector<double>& pts,
               std::vector<double>& out);

// Orientation folding.  Where every edge/face function under every orientation
// of its entity equals, at the given points and up to sign, one function of the
// same entity under orientation 0 (hexahedra and quadrangles: Legendre parity
// and exchanged face axes), an element of any total orientation can be
// integrated with the orientation-0 tables: its functions are renumbered
// (absorbed in the address tables) and signed.  The relation is found
// numerically on the values and reference derivatives at `pts`.
struct H1Fold {
    bool ok = false;
    // map[f][v] = (orientation-0 function, sign) that function f equals when
    // its entity has orientation v (vertex/volume functions: v = 0 only)
    std::vector<std::vector<std::pair<int, int>>> map;
    std::vector<int> dim, num;  // entity of every function
};
H1Fold h1_fold(int type, int order, const std::vector<double>& pts);
// perm[f], sign[f] of every function for the given total orientation.
void h1_fold_apply(const H1Fold& F, int type, int total, std::vector<int>& perm, std::vector<int8_t>& sign);
// The same folding for hcurl functions (components and their reference
// gradients compared): found on quadrangles and hexahedra; h1_fold_apply
// applies it.
H1Fold hcurl_fold(int type, int order, const std::vector<double>& pts);

// Hierarchical H(curl) functions (hcurl.cpp; src/shapefunction/hierarchical/
// hcurl/hcurl{line,triangle,quadrangle,tetrahedron,hexahedron}.cpp) for lines,
// triangles, quadrangles, tetrahedra and hexahedra, orders >= 0.
// Count on the whole element (hcurl*::count(order)) and on one sub-entity
// (hcurl*::count(order, dim, num)).
int hcurl_count(int type, int order);
int hcurl_count(int type, int order, int dim, int num);
// Form functions in the reference iterator order (edge, face, volume; inside
// an entity by increasing order).
std::vector<FormFunction> hcurl_iterate(int type, int order);
// hcurl functions are orientation dependent at every order
// (hierarchicalformfunction::isorientationdependent default).
enum { HCURL_VALUE = 1, HCURL_CURL = 2 };
// out[slot][ff][pt], slot 0 unused, slots 1..3 the reference components of
// the function value (HCURL_VALUE) or of its reference curl (HCURL_CURL).
void hcurl_tables(int type, int order, int totalorientation, int family, const std::vector<double>& pts,
                  std::vector<double>& out);

// Order-1 Lagrange geometry functions (corner nodes) and their reference
// derivatives: out[deriv][node][pt] (src/shapefunction/lagrange/lagrange*.cpp, order 1).
void lagrange1_tables(int type, const std::vector<double>& pts, std::vector<double>& out);
// Order-2 (curved) geometry: node reference coordinates in gmsh order
// (lagrange*.cpp getnodecoordinates(2)), node count (3 line, 6 triangle,
// 9 quadrangle, 10 tetrahedron, 27 hexahedron, 18 prism) and the Lagrange
// functions with their reference derivatives, out[deriv][node][pt].
const std::vector<double>& lagrange2_nodes(int type);
int lagrange2_count(int type);
void lagrange2_tables(int type, const std::vector<double>& pts, std::vector<double>& out);
// Geometry functions of an element type with n_geo nodes per element: order 1
// (corner nodes) or order 2 (every node of a curved element).
void geometry_tables(int type, int n_geo, const std::vector<double>& pts, std::vector<double>& out);

}  // namespace slx
