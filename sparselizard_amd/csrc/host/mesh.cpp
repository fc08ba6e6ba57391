// Mesh loading pipeline; see mesh.h for the reference steps it follows.
#include "mesh.h"
#include <algorithm>
#include <cmath>
#include <fstream>
#include <map>
#include <numeric>
#include "basis.h"

namespace slx {

void Mesh::set_nodes(const std::vector<double>& xyz) {
    SLX_CHECK(!finalized_, "mesh already finalized");
    SLX_CHECK(xyz.size() % 3 == 0, "node coordinates must come in triplets");
    coords_ = xyz;
}

void Mesh::permute_nodes(const std::vector<int>& newindex) {
    SLX_CHECK(!finalized_, "mesh already finalized");
    const int n = count(0);
    SLX_CHECK((int)newindex.size() == n, "node permutation of size " << newindex.size() << " for " << n << " nodes");
    std::vector<char> seen(n, 0);
    for (int v : newindex) {
        SLX_CHECK(v >= 0 && v < n && !seen[v], "not a node permutation");
        seen[v] = 1;
    }
    std::vector<double> c(coords_.size());
    for (int i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) c[3 * (size_t)newindex[i] + d] = coords_[3 * (size_t)i + d];
    coords_.swap(c);
    for (int t = 1; t < 8; t++) {
        for (int& v : sub_[t][0]) v = newindex[v];
        for (int& v : curv_[t]) v = newindex[v];
    }
}

int Mesh::count(int type) const {
    if (type == 0) return (int)coords_.size() / 3;
    int nn = topo::count_nodes(type);
    return (int)sub_[type][0].size() / nn;
}

int Mesh::subelement(int subtype, int type, int elem, int idx) const {
    if (subtype == type) return elem;
    int n = subtype == 0 ? topo::count_nodes(type) : topo::count_type(type, subtype);
    SLX_CHECK(idx >= 0 && idx < n, "sub-entity " << idx << " of type " << subtype << " out of range for element type " << type);
    return sub_[type][subtype][(size_t)elem * n + idx];
}

int Mesh::physreg_index(int physnum) const {
    for (size_t i = 0; i < phys_number_.size(); i++)
        if (phys_number_[i] == physnum) return (int)i;
    return -1;
}

bool Mesh::has_physreg(int physnum) const { return physreg_index(physnum) >= 0; }

int Mesh::physreg_dim(int physnum) const {
    int i = physreg_index(physnum);
    SLX_CHECK(i >= 0, "physical region " << physnum << " does not exist");
    return phys_dim_[i];
}

int Mesh::add_element(int type, const std::vector<int>& nodes, int physreg) {
    SLX_CHECK(!finalized_, "mesh already finalized");
    SLX_CHECK(type >= 0 && type <= PRISM, "unsupported element type " << type);
    SLX_CHECK((int)nodes.size() == topo::count_nodes(type), "element of type " << type << " needs " << topo::count_nodes(type) << " nodes");
    for (int n : nodes) SLX_CHECK(n >= 0 && n < count(0), "element node " << n << " out of range");
    // physicalregions::get: regions are created on first use (physicalregions.cpp:112-127).
    int pi = physreg_index(physreg);
    int dim = topo::dimension(type);
    if (pi < 0) {
        pi = (int)phys_number_.size();
        phys_number_.push_back(physreg);
        phys_dim_.push_back(dim);
        phys_elems_.emplace_back();
    }
    SLX_CHECK(phys_dim_[pi] == dim, "trying to add a " << dim << "D element to physical region " << physreg
                                                        << " which has " << phys_dim_[pi] << "D elements");
    int num;
    if (type == POINT)
        num = nodes[0];  // point elements are the nodes themselves (elements.cpp:18-21)
    else {
        sub_[type][0].insert(sub_[type][0].end(), nodes.begin(), nodes.end());
        num = count(type) - 1;
    }
    phys_elems_[pi][type].push_back(num);
    return num;
}

int Mesh::add_curved_element(int type, const std::vector<int>& nodes, int physreg) {
    const int nn = topo::count_nodes(type);
    SLX_CHECK(type != POINT && (int)nodes.size() > nn, "not a curved element");
    SLX_CHECK(curv_n_[type] == 0 || curv_n_[type] == (int)nodes.size(), "elements of one type with different curvature");
    SLX_CHECK(curv_[type].size() == (size_t)count(type) * curv_n_[type], "a type mixes straight and curved elements");
    for (int n : nodes) SLX_CHECK(n >= 0 && n < count(0), "element node " << n << " out of range");
    const int num = add_element(type, std::vector<int>(nodes.begin(), nodes.begin() + nn), physreg);
    curv_n_[type] = (int)nodes.size();
    curv_[type].insert(curv_[type].end(), nodes.begin(), nodes.end());
    return num;
}

bool Mesh::curved() const {
    for (int t = 1; t < 8; t++)
        if (!curv_[t].empty()) return true;
    return false;
}

int Mesh::geometry_nodes(int type) const { return curv_[type].empty() ? topo::count_nodes(type) : curv_n_[type]; }

int Mesh::geometry_node(int type, int elem, int k) const {
    if (curv_[type].empty()) return subelement(0, type, elem, k);
    return curv_[type][(size_t)elem * curv_n_[type] + k];
}

void Mesh::add_to_physreg(int type, int elem, int physreg) {
    SLX_CHECK(!finalized_, "mesh already finalized");
    SLX_CHECK(elem >= 0 && elem < count(type), "element out of range");
    int pi = physreg_index(physreg);
    int dim = topo::dimension(type);
    if (pi < 0) {
        pi = (int)phys_number_.size();
        phys_number_.push_back(physreg);
        phys_dim_.push_back(dim);
        phys_elems_.emplace_back();
    }
    SLX_CHECK(phys_dim_[pi] == dim, "physical region " << physreg << " mixes element dimensions");
    phys_elems_[pi][type].push_back(elem);
}

void Mesh::renumber(int type, const std::vector<int>& renum) {
    // elements::renumber (elements.cpp:1005-1033): only node, line, triangle and
    // quadrangle references exist in the sub-element tables.
    if (type > 3) return;
    for (int t = 0; t <= 7; t++)
        for (int& v : sub_[t][type]) v = renum[v];
    if (type == 0)
        for (int t = 0; t <= 7; t++)
            for (int& v : curv_[t]) v = renum[v];
}

std::vector<int> Mesh::remove_duplicates(int type) {
    int n = count(type);
    std::vector<int> renum(n, -1);
    if (n == 0) return renum;
    int nunique = 0;
    if (type == 0) {
        // gentools::removeduplicates(coordinates) (gentools.cpp:89-134): merge
        // nodes closer than the noise threshold, numbering by first occurrence.
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int i = 0; i < n; i++)
            for (int d = 0; d < 3; d++) {
                lo[d] = std::min(lo[d], coords_[3 * i + d]);
                hi[d] = std::max(hi[d], coords_[3 * i + d]);
            }
        double nt[3];
        for (int d = 0; d < 3; d++) nt[d] = 1e-10 * std::abs(hi[d] - lo[d]);  // nodes::getnoisethreshold
        std::vector<int> byx(n);
        std::iota(byx.begin(), byx.end(), 0);
        std::stable_sort(byx.begin(), byx.end(), [&](int a, int b) { return coords_[3 * a] < coords_[3 * b]; });
        std::vector<double> xs(n);
        for (int i = 0; i < n; i++) xs[i] = coords_[3 * byx[i]];
        for (int i = 0; i < n; i++) {
            if (renum[i] != -1) continue;
            renum[i] = nunique;
            double x = coords_[3 * i], y = coords_[3 * i + 1], z = coords_[3 * i + 2];
            int b = (int)(std::lower_bound(xs.begin(), xs.end(), x - nt[0]) - xs.begin());
            for (int k = b; k < n && xs[k] <= x + nt[0]; k++) {
                int j = byx[k];
                if (renum[j] == -1 && std::abs(y - coords_[3 * j + 1]) <= nt[1] && std::abs(z - coords_[3 * j + 2]) <= nt[2])
                    renum[j] = nunique;
            }
            nunique++;
        }
        std::vector<double> nc(3 * (size_t)nunique);
        for (int i = 0; i < n; i++)
            for (int d = 0; d < 3; d++) nc[3 * (size_t)renum[i] + d] = coords_[3 * (size_t)i + d];
        coords_ = nc;
        renumber(0, renum);
        return renum;
    }
    // gentools::removeduplicates(int blocks) (gentools.cpp:136-246): elements
    // with the same node set are merged; unique elements are numbered in the
    // order of their first occurrence.
    int nn = topo::count_nodes(type);
    std::map<std::vector<int>, int> seen;
    std::vector<int> key(nn);
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < nn; k++) key[k] = sub_[type][0][(size_t)i * nn + k];
        std::sort(key.begin(), key.end());
        auto it = seen.find(key);
        if (it == seen.end()) {
            seen.emplace(key, nunique);
            renum[i] = nunique++;
        } else
            renum[i] = it->second;
    }
    // elements::removeduplicates (elements.cpp:965-998): the data of element i is
    // copied to slot renum[i] in increasing i (a later duplicate overwrites).
    for (int st = 0; st < 4; st++) {
        std::vector<int>& v = sub_[type][st];
        if (v.empty()) continue;
        int per = st == 0 ? nn : topo::count_type(type, st);
        for (int i = 0; i < n; i++)
            if (renum[i] != i)
                for (int j = 0; j < per; j++) v[(size_t)renum[i] * per + j] = v[(size_t)i * per + j];
        v.resize((size_t)nunique * per);
    }
    if (!curv_[type].empty()) {
        // the geometry nodes follow the corner nodes (same rows)
        std::vector<int>& v = curv_[type];
        const int per = curv_n_[type];
        for (int i = 0; i < n; i++)
            if (renum[i] != i)
                for (int j = 0; j < per; j++) v[(size_t)renum[i] * per + j] = v[(size_t)i * per + j];
        v.resize((size_t)nunique * per);
    }
    renumber(type, renum);
    return renum;
}

void Mesh::explode() {
    // elements::explode (elements.cpp:1115-1208): for 2D and 3D elements, in
    // increasing type order, add every edge (and for 3D every face) as a new
    // line/triangle/quadrangle element; duplicates are removed afterwards.
    for (int type = 2; type <= 7; type++) {
        int ne = count(type);
        if (ne == 0) continue;
        int nn = topo::count_nodes(type), nedges = topo::count_edges(type);
        int ntri = topo::count_tri_faces(type), nquad = topo::count_quad_faces(type);
        bool is3d = topo::dimension(type) == 3;
        const std::vector<int>& en = topo::edge_nodes(type);
        const std::vector<int>& fn = topo::face_nodes(type);
        const std::vector<int>& fe = topo::face_edges(type);
        for (int e = 0; e < ne; e++) {
            const int* nodes = &sub_[type][0][(size_t)e * nn];
            int firstline = count(LINE);
            for (int l = 0; l < nedges; l++) {
                sub_[LINE][0].push_back(nodes[en[2 * l]]);
                sub_[LINE][0].push_back(nodes[en[2 * l + 1]]);
                sub_[type][LINE].push_back(count(LINE) - 1);
            }
            if (!is3d) continue;
            for (int f = 0; f < ntri; f++) {
                for (int k = 0; k < 3; k++) sub_[TRI][0].push_back(nodes[fn[3 * f + k]]);
                sub_[type][TRI].push_back(count(TRI) - 1);
                for (int k = 0; k < 3; k++) sub_[TRI][LINE].push_back(firstline + std::abs(fe[3 * f + k]) - 1);
            }
            for (int f = 0; f < nquad; f++) {
                for (int k = 0; k < 4; k++) sub_[QUAD][0].push_back(nodes[fn[3 * ntri + 4 * f + k]]);
                sub_[type][QUAD].push_back(count(QUAD) - 1);
                for (int k = 0; k < 4; k++) sub_[QUAD][LINE].push_back(firstline + std::abs(fe[3 * ntri + 4 * f + k]) - 1);
            }
        }
    }
}

void Mesh::define_disjoint_regions() {
    // elements::definedisjointregions (elements.cpp:1257-1342).
    int np = (int)phys_number_.size();
    std::vector<std::vector<char>> isin(8);
    for (int st = 0; st <= 7; st++) {
        isin[st].assign((size_t)count(st) * np, 0);
        for (int pi = 0; pi < np; pi++)
            for (int t = 0; t <= 7; t++) {
                int nsub = st == 0 ? topo::count_nodes(t) : topo::count_type(t, st);
                for (int elem : phys_elems_[pi][t])
                    for (int s = 0; s < nsub; s++) {
                        int se = (st == t) ? elem : sub_[t][st][(size_t)elem * nsub + s];
                        isin[st][(size_t)se * np + pi] = 1;
                    }
            }
    }
    dr_type_.clear();
    dr_def_.clear();
    for (int t = 0; t <= 7; t++) {
        int n = count(t);
        indisjreg_[t].assign(n, -1);
        std::map<std::vector<char>, int> drmap;
        std::vector<char> key(np);
        for (int e = 0; e < n; e++) {
            for (int i = 0; i < np; i++) key[i] = isin[t][(size_t)e * np + i];
            auto it = drmap.find(key);
            if (it == drmap.end()) {
                int dr = (int)dr_type_.size();
                dr_type_.push_back(t);
                dr_def_.push_back(key);
                drmap.emplace(key, dr);
                indisjreg_[t][e] = dr;
            } else
                indisjreg_[t][e] = it->second;
        }
    }
}

void Mesh::reorder(int type, const std::vector<int>& ro) {
    // elements::reorder (elements.cpp:1035-1113); nodes::reorder for type 0.
    int n = count(type);
    if (type == 0) {
        std::vector<double> c = coords_;
        for (int i = 0; i < n; i++)
            for (int d = 0; d < 3; d++) coords_[3 * (size_t)i + d] = c[3 * (size_t)ro[i] + d];
    } else {
        for (int st = 0; st < 4; st++) {
            std::vector<int>& v = sub_[type][st];
            if (v.empty()) continue;
            int per = st == 0 ? topo::count_nodes(type) : topo::count_type(type, st);
            std::vector<int> c = v;
            for (int i = 0; i < n; i++)
                for (int j = 0; j < per; j++) v[(size_t)i * per + j] = c[(size_t)ro[i] * per + j];
        }
        if (!curv_[type].empty()) {
            std::vector<int>& v = curv_[type];
            const int per = curv_n_[type];
            std::vector<int> c = v;
            for (int i = 0; i < n; i++)
                for (int j = 0; j < per; j++) v[(size_t)i * per + j] = c[(size_t)ro[i] * per + j];
        }
    }
    std::vector<int> d = indisjreg_[type];
    for (int i = 0; i < n; i++) indisjreg_[type][i] = d[ro[i]];
}

void Mesh::reorder_by_disjoint_regions() {
    // elements::reorderbydisjointregions (elements.cpp:1350-1372): stable sort of
    // every type by disjoint region number, then renumber every reference.
    for (int t = 0; t <= 7; t++) {
        int n = count(t);
        std::vector<int> ro(n);
        std::iota(ro.begin(), ro.end(), 0);
        const std::vector<int>& dr = indisjreg_[t];
        std::stable_sort(ro.begin(), ro.end(), [&](int a, int b) { return dr[a] < dr[b]; });
        std::vector<int> renum(n);
        for (int i = 0; i < n; i++) renum[ro[i]] = i;
        reorder(t, ro);
        renumber(t, renum);
        for (auto& pe : phys_elems_)
            for (int& e : pe[t]) e = renum[e];
    }
}

void Mesh::define_disjoint_region_ranges() {
    // elements::definedisjointregionsranges (elements.cpp:1374-1391).
    int nd = (int)dr_type_.size();
    dr_begin_.assign(nd, 0);
    dr_end_.assign(nd, -1);
    for (int t = 0; t <= 7; t++) {
        const std::vector<int>& dr = indisjreg_[t];
        int n = (int)dr.size();
        for (int i = 0; i < n; i++) {
            if (i == 0 && dr[i] >= 0) dr_begin_[dr[i]] = 0;
            if (i != 0 && dr[i] != dr[i - 1]) {
                if (dr[i - 1] >= 0) dr_end_[dr[i - 1]] = i - 1;
                if (dr[i] >= 0) dr_begin_[dr[i]] = i;
            }
            if (i == n - 1 && dr[i] >= 0) dr_end_[dr[i]] = n - 1;
        }
    }
}

void Mesh::orient() {
    // elements::orient (elements.cpp:1412-1441) with orientation::gettotalorientation.
    for (int t = 1; t <= 7; t++) {
        int n = count(t);
        totalorient_[t].assign(n, 0);
        int nn = topo::count_nodes(t);
        std::vector<long long> cn(nn);
        for (int e = 0; e < n; e++) {
            for (int j = 0; j < nn; j++) cn[j] = sub_[t][0][(size_t)e * nn + j];
            totalorient_[t][e] = slx::total_orientation(t, cn.data());
        }
    }
}

namespace {
thread_local void* g_mesh_stream = nullptr;
}

void set_mesh_pipeline_stream(void* hip_stream) { g_mesh_stream = hip_stream; }
void* mesh_pipeline_stream() { return g_mesh_stream; }

void Mesh::finalize() {
    SLX_CHECK(!finalized_, "mesh already finalized");
    SLX_CHECK(count(0) > 0, "mesh has no nodes");
    int maxt = 0;
    for (int t = 1; t <= 7; t++)
        if (count(t) > 0) maxt = t;
    meshdim_ = topo::dimension(maxt);
    // curvature is kept for the elements of the mesh dimension only: integrals
    // run over them, lower-dimensional elements only carry topology here
    for (int t = 1; t <= 7; t++)
        if (!curv_[t].empty() && topo::dimension(t) < meshdim_) {
            curv_[t].clear();
            curv_n_[t] = 0;
        }
    if (g_mesh_stream && phys_number_.size() <= 16 && !curved()) {
        // duplicate nodes first (they only rename node numbers, which commutes
        // with the explosion), then every other step on the device
        std::vector<int> renum = remove_duplicates(0);
        for (auto& pe : phys_elems_)
            for (int& e : pe[0]) e = renum[e];
        finalize_on_device(*this, g_mesh_stream);
        define_disjoint_region_ranges();
        int np = (int)phys_number_.size(), nd = (int)dr_type_.size();
        phys_includes_.assign(np, std::vector<char>(nd, 0));
        for (int pi = 0; pi < np; pi++)
            for (int d = 0; d < nd; d++) phys_includes_[pi][d] = dr_def_[d][pi];
        finalized_ = true;
        return;
    }
    explode();
    // rawmesh::removeduplicates (rawmesh.cpp:160-182)
    for (int t = 0; t <= 7; t++) {
        if (count(t) == 0) continue;
        std::vector<int> renum = remove_duplicates(t);
        for (auto& pe : phys_elems_)
            for (int& e : pe[t]) e = renum[e];
    }
    for (auto& pe : phys_elems_)
        for (int t = 0; t <= 7; t++) {
            std::sort(pe[t].begin(), pe[t].end());
            pe[t].erase(std::unique(pe[t].begin(), pe[t].end()), pe[t].end());
        }
    define_disjoint_regions();
    reorder_by_disjoint_regions();
    define_disjoint_region_ranges();
    // physicalregions::definewithdisjointregions (physicalregion.cpp:58-78)
    int np = (int)phys_number_.size(), nd = (int)dr_type_.size();
    phys_includes_.assign(np, std::vector<char>(nd, 0));
    for (int pi = 0; pi < np; pi++)
        for (int d = 0; d < nd; d++) phys_includes_[pi][d] = dr_def_[d][pi];
    orient();
    finalized_ = true;
}

std::vector<int> Mesh::physreg_disjregs(int physnum, int dim) const {
    int pi = physreg_index(physnum);
    SLX_CHECK(pi >= 0, "physical region " << physnum << " does not exist");
    std::vector<int> out;
    for (int d = 0; d < (int)dr_type_.size(); d++) {
        if (!phys_includes_[pi][d]) continue;
        int ddim = topo::dimension(dr_type_[d]);
        if (dim == -1 || (dim == -2 && ddim == phys_dim_[pi]) || ddim == dim) out.push_back(d);
    }
    return out;
}

std::vector<int> Mesh::physreg_elements(int physnum, int type) const {
    std::vector<int> out;
    for (int d : physreg_disjregs(physnum, -1))
        if (dr_type_[d] == type)
            for (int e = dr_begin_[d]; e <= dr_end_[d]; e++) out.push_back(e);
    return out;
}

Mesh Mesh::read_msh(const std::string& filename) {
    std::ifstream f(filename);
    SLX_CHECK(f.is_open(), "Unable to open file " << filename << " or file not found");
    std::string line;
    double version = 0;
    while (std::getline(f, line)) {
        if (line.rfind("$MeshFormat", 0) == 0) {
            std::getline(f, line);
            version = std::stod(line);
            break;
        }
    }
    SLX_CHECK(version > 0 && version < 4, "GMSH format " << version << " is not supported (export as GMSH 2 format)");
    Mesh m;
    while (std::getline(f, line))
        if (line.rfind("$Nodes", 0) == 0) break;
    std::getline(f, line);
    int nnodes = std::stoi(line);
    std::vector<double> xyz(3 * (size_t)nnodes);
    for (int i = 0; i < nnodes; i++) {
        std::getline(f, line);
        std::istringstream is(line);
        long tag;
        is >> tag >> xyz[3 * i] >> xyz[3 * i + 1] >> xyz[3 * i + 2];
    }
    m.set_nodes(xyz);
    while (std::getline(f, line))
        if (line.rfind("$Elements", 0) == 0) break;
    std::getline(f, line);
    int nelems = std::stoi(line);
    std::vector<int> prev;
    int previdx = -1, prevtype = -1;
    for (int i = 0; i < nelems; i++) {
        std::getline(f, line);
        std::istringstream is(line);
        int tag, gtype, ntags, phys = 0, dummy;
        is >> tag >> gtype >> ntags;
        for (int j = 0; j < ntags; j++) {
            if (j == 0) is >> phys;
            else is >> dummy;
        }
        int type = topo::from_gmsh(gtype);
        const int order = topo::gmsh_order(gtype);
        std::vector<int> nodes(order == 2 ? lagrange2_count(type) : topo::count_nodes(type));
        for (int& n : nodes) {
            is >> n;
            n -= 1;
        }
        // A repeated element only joins the new physical region (gmshinterface.cpp:274-281).
        if (nodes == prev && type == prevtype) {
            m.add_to_physreg(type, previdx, phys);
            continue;
        }
        previdx = order == 2 ? m.add_curved_element(type, nodes, phys) : m.add_element(type, nodes, phys);
        prev = nodes;
        prevtype = type;
    }
    m.finalize();
    return m;
}

void Mesh::write_msh(const std::string& filename) const {
    // Writes nodes and the elements of every physical region as given
    // (no sub-elements), enough to reload the same mesh.
    std::ofstream f(filename);
    SLX_CHECK(f.is_open(), "cannot write " << filename);
    f.precision(17);
    f << "$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n" << count(0) << "\n";
    for (int i = 0; i < count(0); i++) f << i + 1 << " " << coords_[3 * i] << " " << coords_[3 * i + 1] << " " << coords_[3 * i + 2] << "\n";
    f << "$EndNodes\n$Elements\n";
    size_t total = 0;
    for (size_t pi = 0; pi < phys_number_.size(); pi++)
        for (int t = 0; t <= 7; t++) total += phys_elems_[pi][t].size();
    f << total << "\n";
    size_t tag = 1;
    for (size_t pi = 0; pi < phys_number_.size(); pi++)
        for (int t = 0; t <= 7; t++)
            for (int e : phys_elems_[pi][t]) {
                f << tag++ << " " << topo::to_gmsh(t) << " 2 " << phys_number_[pi] << " " << phys_number_[pi];
                for (int k = 0; k < topo::count_nodes(t); k++) f << " " << subelement(0, t, e, k) + 1;
                f << "\n";
            }
    f << "$EndElements\n";
}

}  // namespace slx
