// Mesh container and the loading pipeline that fixes every numbering the
// assembly depends on: node and sub-element (edge, face) numbers, disjoint
// regions and their element ranges, and element total orientations.
//
// Pipeline (src/mesh/rawmesh.cpp:254-325, without splitting, DDM or curvature):
//   read (file order) -> elements::explode (elements.cpp:1115-1208)
//   -> rawmesh::removeduplicates (rawmesh.cpp:160-182; gentools.cpp:89-246)
//   -> elements::definedisjointregions (elements.cpp:1257-1342)
//   -> elements::reorderbydisjointregions (elements.cpp:1344-1372)
//   -> elements::definedisjointregionsranges (elements.cpp:1374-1391)
//   -> physicalregions::definewithdisjointregions (physicalregions.cpp:101-105)
//   -> elements::orient (elements.cpp:1412-1441)
#pragma once
#include <array>
#include <string>
#include <vector>
#include "common.h"

namespace slx {

class Mesh;
// The pipeline's steps after the duplicate nodes on the device
// (kernels/meshpipe.hip): the mesh's arrays end exactly as the host steps
// leave them.
void finalize_on_device(Mesh& m, void* hip_stream);
// Stream of the device pipeline for meshes finalized on this thread (null:
// the host pipeline); set around a problem's construction.
void set_mesh_pipeline_stream(void* hip_stream);
void* mesh_pipeline_stream();

class Mesh {
   public:
    // ---- loading ---------------------------------------------------------
    // Nodes are numbered in the order they are given (gmsh node tags 1..N).
    void set_nodes(const std::vector<double>& xyz);
    // elements::add + physicalregion::addelement for one element read from a
    // file (gmshinterface.cpp:239-288): returns the element number.
    int add_element(int type, const std::vector<int>& nodes, int physreg);
    // An element of a curved (order-2) mesh: every node in gmsh order, the
    // corner nodes first.  The corners define the topology; every node
    // defines the geometry (rawmesh curvature, lagrange*.cpp order 2).
    int add_curved_element(int type, const std::vector<int>& nodes, int physreg);
    // A repeated element line of a file: the previous element also joins
    // another physical region (gmshinterface.cpp:274-281).
    void add_to_physreg(int type, int elem, int physreg);
    // Before finalize: node i becomes node newindex[i] (coordinates and every
    // element node list follow).
    void permute_nodes(const std::vector<int>& newindex);
    // Run the pipeline once every element has been added.
    void finalize();
    // Gmsh 2.2 ASCII reader / writer (gmshinterface.cpp:157-298).
    static Mesh read_msh(const std::string& filename);
    void write_msh(const std::string& filename) const;

    // ---- queries ---------------------------------------------------------
    int dimension() const { return meshdim_; }
    int count(int type) const;  // elements of a type (nodes for type 0)
    const std::vector<double>& coords() const { return coords_; }
    // subelementsinelements: sub-element `idx` of type `subtype` in element `elem`.
    int subelement(int subtype, int type, int elem, int idx) const;
    const std::vector<int>& subelements(int type, int subtype) const { return sub_[type][subtype]; }
    int disjreg_of(int type, int elem) const { return indisjreg_[type][elem]; }
    int total_orientation(int type, int elem) const { return type == 0 ? 0 : totalorient_[type][elem]; }
    // Geometry nodes of the elements of a type: every node of a curved element
    // (elements of the mesh dimension read with order-2 geometry), else the corners.
    int geometry_nodes(int type) const;
    int geometry_node(int type, int elem, int k) const;
    bool curved() const;

    int count_disjregs() const { return (int)dr_type_.size(); }
    int disjreg_type(int dr) const { return dr_type_[dr]; }
    int disjreg_begin(int dr) const { return dr_begin_[dr]; }
    int disjreg_end(int dr) const { return dr_end_[dr]; }
    int disjreg_count(int dr) const { return dr_end_[dr] - dr_begin_[dr] + 1; }

    // physicalregion::getdisjointregions(dim) (physicalregion.cpp:85-111):
    // dim = -1 all, dim = -2 only those of the region's own dimension.
    std::vector<int> physreg_disjregs(int physnum, int dim = -1) const;
    int physreg_dim(int physnum) const;
    bool has_physreg(int physnum) const;
    std::vector<int> physreg_numbers() const { return phys_number_; }
    // Number of the elements of `type` of a physical region (after finalize).
    std::vector<int> physreg_elements(int physnum, int type) const;

   private:
    friend void finalize_on_device(Mesh& m, void* hip_stream);
    int physreg_index(int physnum) const;
    std::vector<int> remove_duplicates(int type);   // returns the renumbering
    void renumber(int type, const std::vector<int>& renum);
    void reorder(int type, const std::vector<int>& reordering);
    void explode();
    void define_disjoint_regions();
    void reorder_by_disjoint_regions();
    void define_disjoint_region_ranges();
    void orient();

    int meshdim_ = 0;
    bool finalized_ = false;
    std::vector<double> coords_;
    std::vector<int> sub_[8][4];
    std::vector<int> curv_[8];   // every node of the curved elements of a type (gmsh order)
    int curv_n_[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    std::vector<int> indisjreg_[8];
    std::vector<int> totalorient_[8];
    // physical regions in creation order
    std::vector<int> phys_number_;
    std::vector<int> phys_dim_;
    std::vector<std::array<std::vector<int>, 8>> phys_elems_;
    std::vector<std::vector<char>> phys_includes_;   // [phys index][disjreg]
    // disjoint regions
    std::vector<int> dr_type_;
    std::vector<std::vector<char>> dr_def_;          // [disjreg][phys index]
    std::vector<int> dr_begin_, dr_end_;
};

}  // namespace slx
