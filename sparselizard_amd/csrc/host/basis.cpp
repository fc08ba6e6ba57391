// Hierarchical H1 functions evaluated with forward-mode derivatives.
//
// The reference builds each function as an explicit 3-variate polynomial and
// evaluates the polynomial and its derivative polynomials
// (hierarchicalformfunctioncontainer.cpp:49-76).  Here each function is
// evaluated directly from its defining formula on dual numbers carrying the
// three reference partial derivatives; values agree to rounding.
#include "basis.h"
#include <cmath>
#include <functional>

namespace slx {

namespace {

struct D {
    double v, d[3];
    D(double x = 0.0) : v(x), d{0, 0, 0} {}
};
inline D var(double x, int k) { D r(x); r.d[k] = 1.0; return r; }
inline D operator+(const D& a, const D& b) { D r(a.v + b.v); for (int k = 0; k < 3; k++) r.d[k] = a.d[k] + b.d[k]; return r; }
inline D operator-(const D& a, const D& b) { D r(a.v - b.v); for (int k = 0; k < 3; k++) r.d[k] = a.d[k] - b.d[k]; return r; }
inline D operator*(const D& a, const D& b) { D r(a.v * b.v); for (int k = 0; k < 3; k++) r.d[k] = a.d[k] * b.v + a.v * b.d[k]; return r; }
inline D operator*(double s, const D& a) { D r(s * a.v); for (int k = 0; k < 3; k++) r.d[k] = s * a.d[k]; return r; }
[[maybe_unused]] inline D operator+(double s, const D& a) { D r(a); r.v += s; return r; }
inline D operator-(double s, const D& a) { D r(s - a.v); for (int k = 0; k < 3; k++) r.d[k] = -a.d[k]; return r; }

// Legendre families of src/shapefunction/hierarchical/legendre.cpp.
// l: Legendre polynomials; L: integrated Legendre; ls/Ls: their scaled forms.
std::vector<D> leg_l(int maxn, const D& x) {
    std::vector<D> l(maxn + 1);
    l[0] = D(1.0);
    if (maxn > 0) l[1] = x;
    for (int n = 1; n < maxn; n++) l[n + 1] = (1.0 / (n + 1)) * ((double)(2 * n + 1) * l[n] * x - (double)n * l[n - 1]);
    return l;
}
std::vector<D> leg_L(int maxn, const D& x) {
    std::vector<D> L(std::max(maxn + 1, 3));
    L[1] = x;
    L[2] = 0.5 * (x * x - D(1.0));
    for (int n = 2; n < maxn; n++) L[n + 1] = (1.0 / (n + 1)) * ((double)(2 * n - 1) * x * L[n] - (double)(n - 2) * L[n - 1]);
    return L;
}
std::vector<D> leg_ls(int maxn, const D& x, const D& t) {
    std::vector<D> ls(maxn + 1);
    ls[0] = D(1.0);
    if (maxn > 0) ls[1] = x;
    for (int n = 1; n < maxn; n++) ls[n + 1] = (1.0 / (n + 1)) * ((double)(2 * n + 1) * x * ls[n] - (double)n * t * t * ls[n - 1]);
    return ls;
}
std::vector<D> leg_Ls(int maxn, const D& x, const D& t) {
    std::vector<D> Ls(std::max(maxn + 1, 3));
    Ls[1] = x;
    Ls[2] = 0.5 * (x * x - t * t);
    for (int n = 2; n < maxn; n++) Ls[n + 1] = (1.0 / (n + 1)) * ((double)(2 * n - 1) * x * Ls[n] - (double)(n - 2) * t * t * Ls[n - 1]);
    return Ls;
}

// Node reorderings bringing an edge / face to its reference orientation 0
// (orientation.cpp:164-180).
const int edge_reorder[2][2] = {{0, 1}, {1, 0}};
const int tri_reorder[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 2, 0}, {1, 0, 2}, {2, 0, 1}, {2, 1, 0}};
const int quad_reorder[8][4] = {{0, 1, 2, 3}, {0, 3, 2, 1}, {1, 2, 3, 0}, {1, 0, 3, 2},
                                {2, 3, 0, 1}, {2, 1, 0, 3}, {3, 0, 1, 2}, {3, 2, 1, 0}};

// Barycentric / bilinear 'lambda' and 'sigma' functions of the element (1-based
// like the reference; index 0 unused).
void lambdas(int type, const D& ki, const D& eta, const D& phi, std::vector<D>& lam, std::vector<D>& sig) {
    lam.assign(9, D(0.0));
    sig.assign(9, D(0.0));
    if (type == LINE) {
        D x = 0.5 * (ki + 1.0);
        lam[1] = 1.0 - x; lam[2] = x;
        sig[1] = 1.0 - x; sig[2] = x;
    } else if (type == TRI) {
        lam[1] = 1.0 - ki - eta; lam[2] = ki; lam[3] = eta;
    } else if (type == TET) {
        lam[1] = 1.0 - ki - eta - phi; lam[2] = ki; lam[3] = eta; lam[4] = phi;
    } else if (type == QUAD) {
        D x = 0.5 * (ki + 1.0), y = 0.5 * (eta + 1.0);
        lam[1] = (1.0 - x) * (1.0 - y); lam[2] = x * (1.0 - y); lam[3] = x * y; lam[4] = (1.0 - x) * y;
        sig[1] = (1.0 - x) + (1.0 - y); sig[2] = x + (1.0 - y); sig[3] = x + y; sig[4] = (1.0 - x) + y;
    } else if (type == HEX) {
        D x = 0.5 * (ki + 1.0), y = 0.5 * (eta + 1.0), z = 0.5 * (phi + 1.0);
        lam[1] = (1.0 - x) * (1.0 - y) * (1.0 - z); lam[2] = x * (1.0 - y) * (1.0 - z);
        lam[3] = x * y * (1.0 - z); lam[4] = (1.0 - x) * y * (1.0 - z);
        lam[5] = (1.0 - x) * (1.0 - y) * z; lam[6] = x * (1.0 - y) * z;
        lam[7] = x * y * z; lam[8] = (1.0 - x) * y * z;
        sig[1] = (1.0 - x) + (1.0 - y) + (1.0 - z); sig[2] = x + (1.0 - y) + (1.0 - z);
        sig[3] = x + y + (1.0 - z); sig[4] = (1.0 - x) + y + (1.0 - z);
        sig[5] = (1.0 - x) + (1.0 - y) + z; sig[6] = x + (1.0 - y) + z;
        sig[7] = x + y + z; sig[8] = (1.0 - x) + y + z;
    } else if (type == PRISM) {
        // triangle barycentrics times the linear functions of the prism axis,
        // phi mapped to [0, 1] (h1prism.cpp: lambda and mu, vertex functions)
        D z = 0.5 * (phi + 1.0);
        D t[3] = {1.0 - ki - eta, ki, eta};
        for (int n = 0; n < 3; n++) {
            lam[1 + n] = t[n] * (1.0 - z);
            lam[4 + n] = t[n] * z;
        }
    } else {
        fail("H1 functions are provided for line, triangle, quadrangle, tetrahedron, hexahedron and prism elements");
    }
}

bool simplex(int type) { return type == TRI || type == TET; }

// Evaluate form function `f` (for entity orientation `orient`) at one point.
// Prism functions (h1prism.cpp): triangle barycentrics lambda_n and axis
// functions mu_n per node (1-based, nodes 1-3 at phi = -1, 4-6 at phi = +1).
D eval_prism(int maxorder, const FormFunction& f, int orient, const D& ki, const D& eta, const D& phi) {
    const D z = 0.5 * (phi + 1.0);
    D lam[7], mu[7];
    const D t[3] = {1.0 - ki - eta, ki, eta};
    for (int n = 0; n < 3; n++) {
        lam[1 + n] = t[n];
        lam[4 + n] = t[n];
        mu[1 + n] = 1.0 - z;
        mu[4 + n] = z;
    }
    auto same_layer = [](int a, int b) { return (a <= 3) == (b <= 3); };
    const int h = f.order;
    if (f.dim == 0) return lam[f.num + 1] * mu[f.num + 1];
    if (f.dim == 1) {
        const std::vector<int>& en = topo::edge_nodes(PRISM);
        int e1 = en[2 * f.num + edge_reorder[orient][0]] + 1;
        int e2 = en[2 * f.num + edge_reorder[orient][1]] + 1;
        if (same_layer(e1, e2)) return leg_Ls(maxorder, lam[e1] - lam[e2], lam[e1] + lam[e2])[h] * mu[e1];
        return leg_L(maxorder, 2.0 * mu[e1] - 1.0)[h] * lam[e1];
    }
    const std::vector<int>& fn = topo::face_nodes(PRISM);
    if (f.dim == 2 && f.num < 2) {
        int f1 = fn[3 * f.num + tri_reorder[orient][0]] + 1;
        int f2 = fn[3 * f.num + tri_reorder[orient][1]] + 1;
        int f3 = fn[3 * f.num + tri_reorder[orient][2]] + 1;
        std::vector<D> Ls = leg_Ls(maxorder, lam[f1] - lam[f2], lam[f1] + lam[f2]);
        std::vector<D> l = leg_l(maxorder, 2.0 * lam[f3] - 1.0);
        // i + j = h - 3, i outer
        int i = f.l, j = h - 3 - f.l;
        return Ls[i + 2] * lam[f3] * l[j] * mu[f1];
    }
    if (f.dim == 2) {
        const int off = 6 + 4 * (f.num - 2);
        int f1 = fn[off + quad_reorder[orient][0]] + 1;
        int f2 = fn[off + quad_reorder[orient][1]] + 1;
        int f4 = fn[off + quad_reorder[orient][3]] + 1;
        // the face direction along the triangle side starts at f1 and goes to
        // f2 or f4, whichever lies in f1's layer
        int f2s = same_layer(f1, f2) ? f2 : f4;
        std::vector<D> Ls = leg_Ls(maxorder, lam[f1] - lam[f2s], lam[f1] + lam[f2s]);
        std::vector<D> L = leg_L(maxorder, 2.0 * mu[f1] - 1.0);
        int cnt = 0;
        for (int i = 0; i <= h - 2; i++)
            for (int j = 0; j <= h - 2; j++) {
                if (i != h - 2 && j != h - 2) continue;
                if (cnt == f.l) return f2 == f2s ? Ls[i + 2] * L[j + 2] : L[i + 2] * Ls[j + 2];
                cnt++;
            }
        fail("prism quad face function index out of range");
    }
    std::vector<D> Ls = leg_Ls(maxorder, lam[1] - lam[2], lam[1] + lam[2]);
    std::vector<D> l = leg_l(maxorder, 2.0 * lam[3] - 1.0);
    std::vector<D> L = leg_L(maxorder, 2.0 * mu[1] - 1.0);
    int cnt = 0;
    for (int i = 0; i <= h - 3; i++)
        for (int j = 0; j <= h - 3 - i; j++)
            for (int k = 0; k <= h - 2; k++) {
                if (i + j != h - 3 && k != h - 2) continue;
                if (cnt == f.l) return Ls[i + 2] * lam[3] * l[j] * L[k + 2];
                cnt++;
            }
    fail("prism volume function index out of range");
}

D eval_ff(int type, int maxorder, const FormFunction& f, int orient, const D& ki, const D& eta, const D& phi) {
    if (type == PRISM) return eval_prism(maxorder, f, orient, ki, eta, phi);
    std::vector<D> lam, sig;
    lambdas(type, ki, eta, phi, lam, sig);
    const int h = f.order;
    if (f.dim == 0) return lam[f.num + 1];
    if (f.dim == 1 && type != HEX && type != QUAD && type != LINE && type != TRI && type != TET) fail("bad type");
    if (f.dim == 1) {
        const std::vector<int>& en = topo::edge_nodes(type);
        int e1 = en[2 * f.num + edge_reorder[orient][0]] + 1;
        int e2 = en[2 * f.num + edge_reorder[orient][1]] + 1;
        if (simplex(type)) {
            std::vector<D> Ls = leg_Ls(maxorder, lam[e1] - lam[e2], lam[e1] + lam[e2]);
            return Ls[h];
        }
        std::vector<D> L = leg_L(maxorder, sig[e1] - sig[e2]);
        return L[h] * (lam[e1] + lam[e2]);
    }
    if (f.dim == 2 && (type == TRI || type == TET)) {
        const std::vector<int>& fn = topo::face_nodes(type);
        int f1 = fn[3 * f.num + tri_reorder[orient][0]] + 1;
        int f2 = fn[3 * f.num + tri_reorder[orient][1]] + 1;
        int f3 = fn[3 * f.num + tri_reorder[orient][2]] + 1;
        D lamF = lam[f1] + lam[f2] + lam[f3];
        std::vector<D> Ls = leg_Ls(maxorder, lam[f1] - lam[f2], lam[f1] + lam[f2]);
        std::vector<D> ls = leg_ls(maxorder, 2.0 * lam[f3] - lamF, lamF);
        // (i, j) with i + j = h - 3, i outer (h1triangle.cpp:159-185)
        int i = f.l, j = h - 3 - f.l;
        return Ls[i + 2] * lam[f3] * ls[j];
    }
    if (f.dim == 2) {  // quadrangle face (QUAD itself or a HEX face)
        const std::vector<int>& fn = topo::face_nodes(type);
        int f1 = fn[4 * f.num + quad_reorder[orient][0]] + 1;
        int f2 = fn[4 * f.num + quad_reorder[orient][1]] + 1;
        int f3 = fn[4 * f.num + quad_reorder[orient][2]] + 1;
        int f4 = fn[4 * f.num + quad_reorder[orient][3]] + 1;
        D lamF = lam[f1] + lam[f2] + lam[f3] + lam[f4];
        std::vector<D> Lx = leg_L(maxorder, sig[f1] - sig[f2]);
        std::vector<D> Ly = leg_L(maxorder, sig[f1] - sig[f4]);
        // pairs (i, j) in [0, h-2]^2 with i == h-2 or j == h-2, i outer (h1quadrangle.cpp:176-201)
        int cnt = 0;
        for (int i = 0; i <= h - 2; i++)
            for (int j = 0; j <= h - 2; j++) {
                if (i != h - 2 && j != h - 2) continue;
                if (cnt == f.l) return Lx[i + 2] * Ly[j + 2] * lamF;
                cnt++;
            }
        fail("quad face function index out of range");
    }
    // volume functions
    if (type == TET) {
        std::vector<D> Ls = leg_Ls(maxorder, lam[1] - lam[2], lam[1] + lam[2]);
        std::vector<D> ls = leg_ls(maxorder, 2.0 * lam[3] - (1.0 - lam[4]), 1.0 - lam[4]);
        std::vector<D> l = leg_l(maxorder, 2.0 * lam[4] - 1.0);
        int cnt = 0;
        for (int i = 0; i <= h - 4; i++)
            for (int j = 0; j <= h - 4; j++)
                for (int k = 0; k <= h - 4; k++) {
                    if (i + j + k != h - 4) continue;
                    if (cnt == f.l) return Ls[i + 2] * lam[3] * ls[j] * lam[4] * l[k];
                    cnt++;
                }
        fail("tet volume function index out of range");
    }
    if (type == HEX) {
        D x = 0.5 * (ki + 1.0), y = 0.5 * (eta + 1.0), z = 0.5 * (phi + 1.0);
        std::vector<D> Lx = leg_L(maxorder, 2.0 * x - 1.0);
        std::vector<D> Ly = leg_L(maxorder, 2.0 * y - 1.0);
        std::vector<D> Lz = leg_L(maxorder, 2.0 * z - 1.0);
        int cnt = 0;
        for (int i = 0; i <= h - 2; i++)
            for (int j = 0; j <= h - 2; j++)
                for (int k = 0; k <= h - 2; k++) {
                    if (i != h - 2 && j != h - 2 && k != h - 2) continue;
                    if (cnt == f.l) return Lx[i + 2] * Ly[j + 2] * Lz[k + 2];
                    cnt++;
                }
        fail("hex volume function index out of range");
    }
    fail("no volume functions for this element type");
}

}  // namespace

int h1_count(int type, int order, int dim) {
    if (order <= 0) return 0;
    switch (dim) {
        case 0: return 1;
        case 1: return order - 1;
        case 2:
            if (type == TRI || type == TET) return (order - 2) * (order - 1) / 2;
            if (type == QUAD || type == HEX) return (order - 1) * (order - 1);
            return 0;
        case 3:
            if (type == TET) return (order - 3) * (order - 2) * (order - 1) / 6;
            if (type == HEX) return (order - 1) * (order - 1) * (order - 1);
            if (type == PRISM) return (order - 2) * (order - 1) * (order - 1) / 2;
            return 0;
    }
    return 0;
}

int h1_count(int type, int order) {
    if (order <= 0) return 0;
    switch (type) {
        case POINT: return 1;
        case LINE: return order + 1;
        case TRI: return (order + 1) * (order + 2) / 2;
        case QUAD: return (order + 1) * (order + 1);
        case TET: return (order + 1) * (order + 2) * (order + 3) / 6;
        case HEX: return (order + 1) * (order + 1) * (order + 1);
        case PRISM: return (order + 1) * (order + 1) * (order + 2) / 2;  // line x triangle
    }
    fail("H1 count: unsupported element type " + std::to_string(type));
}

std::vector<FormFunction> h1_iterate(int type, int order) {
    SLX_CHECK(type != PYRAMID, "pyramid H1 functions are not provided");
    std::vector<FormFunction> out;
    for (int dim = 0; dim <= 3; dim++) {
        int nent = topo::count_dim(type, dim);
        for (int num = 0; num < nent; num++) {
            // counts follow the entity's own type (a prism has triangular and
            // quadrangular faces, h1prism.cpp count(order, dim, num))
            int assoc = dim == 0 ? POINT : dim == 1 ? LINE : dim == 2 ? (topo::is_triangular_face(type, num) ? TRI : QUAD) : type;
            int nff = h1_count(assoc, order, dim);
            // entity counts are only defined up to the element dimension
            if (dim > topo::dimension(type)) nff = 0;
            for (int idx = 0; idx < nff; idx++) {
                FormFunction f;
                f.dim = dim;
                f.num = num;
                f.index = idx;
                int h = 0;
                while (idx >= h1_count(assoc, h, dim)) h++;
                f.order = h;
                f.l = idx - h1_count(assoc, h - 1, dim);
                f.assoc_type = assoc;
                f.sub = (dim == 2 && assoc == QUAD) ? num - topo::count_tri_faces(type) : num;
                out.push_back(f);
            }
        }
    }
    SLX_CHECK((int)out.size() == h1_count(type, order), "H1 iterator count mismatch");
    return out;
}

int orientation_of_edge(long long a, long long b) { return a > b ? 0 : 1; }

int orientation_of_triangle(const long long* n) {
    int maxpos = 0;
    for (int i = 1; i < 3; i++)
        if (n[i] > n[maxpos]) maxpos = i;
    return n[(maxpos + 1) % 3] > n[(maxpos + 2) % 3] ? 2 * maxpos : 2 * maxpos + 1;
}

int orientation_of_quadrangle(const long long* n) {
    int maxpos = 0;
    for (int i = 1; i < 4; i++)
        if (n[i] > n[maxpos]) maxpos = i;
    return n[(maxpos + 1) % 4] > n[(maxpos + 3) % 4] ? 2 * maxpos : 2 * maxpos + 1;
}

int total_orientation(int type, const long long* nodes) {
    const std::vector<int>& en = topo::edge_nodes(type);
    const std::vector<int>& fn = topo::face_nodes(type);
    int ne = topo::count_edges(type), ntf = topo::count_tri_faces(type), nqf = topo::count_quad_faces(type);
    int total = 0, factor = 1;
    for (int i = 0; i < ne; i++) {
        total += orientation_of_edge(nodes[en[2 * i]], nodes[en[2 * i + 1]]) * factor;
        factor *= 2;
    }
    for (int i = 0; i < ntf + nqf; i++) {
        int o;
        if (i < ntf) {
            long long f[3] = {nodes[fn[3 * i]], nodes[fn[3 * i + 1]], nodes[fn[3 * i + 2]]};
            o = orientation_of_triangle(f);
        } else {
            int off = 3 * ntf + 4 * (i - ntf);
            long long f[4] = {nodes[fn[off]], nodes[fn[off + 1]], nodes[fn[off + 2]], nodes[fn[off + 3]]};
            o = orientation_of_quadrangle(f);
        }
        total += o * factor;
        factor *= 8;
    }
    return total;
}

void orientations_from_total(int type, int total, std::vector<int>& edges, std::vector<int>& faces) {
    int ne = topo::count_edges(type), nf = topo::count_faces(type);
    edges.resize(ne);
    faces.resize(nf);
    for (int e = 0; e < ne; e++) { edges[e] = total % 2; total /= 2; }
    for (int f = 0; f < nf; f++) { faces[f] = total % 8; total /= 8; }
}

void h1_tables(int type, int order, int totalorientation, const std::vector<double>& pts, std::vector<double>& out) {
    std::vector<FormFunction> ffs = h1_iterate(type, order);
    int nff = (int)ffs.size(), np = (int)pts.size() / 3;
    std::vector<int> eo, fo;
    orientations_from_total(type, totalorientation, eo, fo);
    for (size_t f = 0; f < fo.size(); f++)
        SLX_CHECK(fo[f] < (topo::is_triangular_face(type, (int)f) ? 6 : 8),
                  "total orientation " << totalorientation << " gives face " << f << " orientation " << fo[f]);
    out.assign((size_t)4 * nff * np, 0.0);
    for (int p = 0; p < np; p++) {
        D ki = var(pts[3 * p], 0), eta = var(pts[3 * p + 1], 1), phi = var(pts[3 * p + 2], 2);
        for (int f = 0; f < nff; f++) {
            const FormFunction& ff = ffs[f];
            int orient = 0;
            // Only edges and faces carry an orientation (container.cpp:96-105);
            // the element's own edge (line) and face (tri/quad) included.
            if (ff.dim == 1) orient = eo[ff.num];
            if (ff.dim == 2) orient = fo[ff.num];
            D v = eval_ff(type, std::max(order, 2), ff, orient, ki, eta, phi);
            out[((size_t)0 * nff + f) * np + p] = v.v;
            out[((size_t)1 * nff + f) * np + p] = v.d[0];
            out[((size_t)2 * nff + f) * np + p] = v.d[1];
            out[((size_t)3 * nff + f) * np + p] = v.d[2];
        }
    }
}

H1Fold h1_fold(int type, int order, const std::vector<double>& pts) {
    H1Fold F;
    std::vector<FormFunction> ffs = h1_iterate(type, order);
    const int nff = (int)ffs.size(), np = (int)pts.size() / 3;
    auto values = [&](const FormFunction& ff, int orient, std::vector<double>& v) {
        v.resize((size_t)4 * np);
        for (int p = 0; p < np; p++) {
            D ki = var(pts[3 * p], 0), eta = var(pts[3 * p + 1], 1), phi = var(pts[3 * p + 2], 2);
            D x = eval_ff(type, std::max(order, 2), ff, orient, ki, eta, phi);
            v[4 * p] = x.v;
            v[4 * p + 1] = x.d[0];
            v[4 * p + 2] = x.d[1];
            v[4 * p + 3] = x.d[2];
        }
    };
    auto oriented = [](const FormFunction& ff) { return ff.dim == 1 || ff.dim == 2; };
    std::vector<std::vector<double>> v0(nff);
    for (int f = 0; f < nff; f++)
        if (oriented(ffs[f])) values(ffs[f], 0, v0[f]);
    F.map.assign(nff, {});
    F.dim.resize(nff);
    F.num.resize(nff);
    for (int f = 0; f < nff; f++) {
        F.dim[f] = ffs[f].dim;
        F.num[f] = ffs[f].num;
    }
    std::vector<double> a;
    for (int f = 0; f < nff; f++) {
        const FormFunction& ff = ffs[f];
        const int nv = ff.dim == 1 ? 2 : (ff.dim == 2 ? (ff.assoc_type == TRI ? 6 : 8) : 1);
        F.map[f].assign(nv, {f, 1});
        if (!oriented(ff)) continue;
        for (int v = 1; v < nv; v++) {
            values(ff, v, a);
            double amax = 0.0;
            for (double x : a) amax = std::max(amax, std::fabs(x));
            const double tol = 1e-10 * std::max(1.0, amax);
            int best = -1, bsign = 0, nmatch = 0;
            for (int g = 0; g < nff; g++) {
                if (ffs[g].dim != ff.dim || ffs[g].num != ff.num) continue;
                for (int sg = 1; sg >= -1; sg -= 2) {
                    double err = 0.0;
                    for (size_t k = 0; k < a.size() && err <= tol; k++) err = std::max(err, std::fabs(a[k] - sg * v0[g][k]));
                    if (err <= tol) {
                        best = g;
                        bsign = sg;
                        nmatch++;
                    }
                }
            }
            if (nmatch != 1) return F;  // no (or no unique) signed match: not foldable
            F.map[f][v] = {best, bsign};
        }
    }
    // every entity orientation must renumber the entity's functions one to one
    for (int f = 0; f < nff; f++)
        for (size_t v = 0; v < F.map[f].size(); v++)
            for (int g = f + 1; g < nff; g++)
                if (ffs[g].dim == ffs[f].dim && ffs[g].num == ffs[f].num && F.map[g].size() > v &&
                    F.map[g][v].first == F.map[f][v].first)
                    return F;
    F.ok = true;
    return F;
}

void h1_fold_apply(const H1Fold& F, int type, int total, std::vector<int>& perm, std::vector<int8_t>& sign) {
    SLX_CHECK(F.ok, "orientation folding is not available for this basis");
    std::vector<int> eo, fo;
    orientations_from_total(type, total, eo, fo);
    const int nff = (int)F.map.size();
    perm.resize(nff);
    sign.resize(nff);
    for (int f = 0; f < nff; f++) {
        const int v = F.dim[f] == 1 ? eo[F.num[f]] : (F.dim[f] == 2 ? fo[F.num[f]] : 0);
        SLX_CHECK(v >= 0 && v < (int)F.map[f].size(), "orientation " << v << " out of range");
        perm[f] = F.map[f][v].first;
        sign[f] = (int8_t)F.map[f][v].second;
    }
}

// Reference coordinates of the nodes of order-2 elements, gmsh node order
// (lagrange*.cpp getnodecoordinates(2); the mesh files list nodes in this order).
const std::vector<double>& lagrange2_nodes(int type) {
    static const std::vector<double> line = {-1, 0, 0, 1, 0, 0, 0, 0, 0};
    static const std::vector<double> tri = {0, 0, 0, 1, 0, 0, 0, 1, 0, 0.5, 0, 0, 0.5, 0.5, 0, 0, 0.5, 0};
    static const std::vector<double> quad = {-1, -1, 0, 1, -1, 0, 1, 1, 0, -1, 1, 0, 0, -1, 0, 1, 0, 0, 0, 1, 0, -1, 0, 0, 0, 0, 0};
    static const std::vector<double> tet = {0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1, 0.5, 0, 0, 0.5, 0.5, 0, 0, 0.5, 0, 0, 0, 0.5, 0, 0.5, 0.5, 0.5, 0, 0.5};
    static const std::vector<double> hex = {-1, -1, -1, 1, -1, -1, 1, 1, -1, -1, 1, -1, -1, -1, 1, 1, -1, 1, 1, 1, 1, -1, 1, 1,
                                            0,  -1, -1, -1, 0, -1, -1, -1, 0, 1, 0, -1, 1, -1, 0, 0, 1, -1, 1, 1, 0, -1, 1, 0,
                                            0,  -1, 1, -1, 0, 1, 1, 0, 1, 0, 1, 1, 0, 0, -1, 0, -1, 0, -1, 0, 0, 1, 0, 0,
                                            0,  1, 0, 0, 0, 1, 0, 0, 0};
    static const std::vector<double> prism = {0, 0, -1, 1, 0, -1, 0, 1, -1, 0, 0, 1, 1, 0, 1, 0, 1, 1, 0.5, 0, -1, 0, 0.5, -1,
                                              0, 0, 0, 0.5, 0.5, -1, 1, 0, 0, 0, 1, 0, 0.5, 0, 1, 0, 0.5, 1, 0.5, 0.5, 1,
                                              0.5, 0, 0, 0, 0.5, 0, 0.5, 0.5, 0};
    switch (type) {
        case LINE: return line;
        case TRI: return tri;
        case QUAD: return quad;
        case TET: return tet;
        case HEX: return hex;
        case PRISM: return prism;
    }
    fail("order-2 geometry is not provided for element type " + std::to_string(type));
}

int lagrange2_count(int type) { return (int)lagrange2_nodes(type).size() / 3; }

namespace {
// quadratic 1D Lagrange function of the node at c in {-1, 0, 1} on [-1, 1]
D quad1(const D& x, double c) {
    if (c < -0.5) return 0.5 * (x * (x - 1.0));
    if (c > 0.5) return 0.5 * (x * (x + 1.0));
    return 1.0 - x * x;
}
// quadratic simplex function of the node with barycentric coordinates b
// (corner: lambda (2 lambda - 1); edge midpoint: 4 lambda_i lambda_j)
D quad_simplex(const std::vector<D>& lam, const double* b, int n) {
    int one = -1, h1 = -1, h2 = -1;
    for (int i = 0; i < n; i++) {
        if (std::fabs(b[i] - 1.0) < 1e-12) one = i;
        else if (std::fabs(b[i] - 0.5) < 1e-12) (h1 < 0 ? h1 : h2) = i;
    }
    if (one >= 0) return lam[one] * (2.0 * lam[one] - 1.0);
    SLX_CHECK(h1 >= 0 && h2 >= 0, "internal: order-2 simplex node is neither a corner nor an edge midpoint");
    return 4.0 * (lam[h1] * lam[h2]);
}
}  // namespace

void lagrange2_tables(int type, const std::vector<double>& pts, std::vector<double>& out) {
    const std::vector<double>& c = lagrange2_nodes(type);
    const int nn = (int)c.size() / 3, np = (int)pts.size() / 3;
    out.assign((size_t)4 * nn * np, 0.0);
    for (int p = 0; p < np; p++) {
        const D ki = var(pts[3 * p], 0), eta = var(pts[3 * p + 1], 1), phi = var(pts[3 * p + 2], 2);
        for (int n = 0; n < nn; n++) {
            const double* cn = &c[3 * n];
            D v(0.0);
            if (type == LINE) {
                v = quad1(ki, cn[0]);
            } else if (type == QUAD) {
                v = quad1(ki, cn[0]) * quad1(eta, cn[1]);
            } else if (type == HEX) {
                v = quad1(ki, cn[0]) * quad1(eta, cn[1]) * quad1(phi, cn[2]);
            } else if (type == TRI || type == PRISM) {
                const std::vector<D> lam = {1.0 - ki - eta, ki, eta};
                const double b[3] = {1.0 - cn[0] - cn[1], cn[0], cn[1]};
                v = quad_simplex(lam, b, 3);
                if (type == PRISM) v = v * quad1(phi, cn[2]);
            } else {  // TET
                const std::vector<D> lam = {1.0 - ki - eta - phi, ki, eta, phi};
                const double b[4] = {1.0 - cn[0] - cn[1] - cn[2], cn[0], cn[1], cn[2]};
                v = quad_simplex(lam, b, 4);
            }
            out[((size_t)0 * nn + n) * np + p] = v.v;
            out[((size_t)1 * nn + n) * np + p] = v.d[0];
            out[((size_t)2 * nn + n) * np + p] = v.d[1];
            out[((size_t)3 * nn + n) * np + p] = v.d[2];
        }
    }
}

void lagrange1_tables(int type, const std::vector<double>& pts, std::vector<double>& out) {
    int nn = topo::count_nodes(type), np = (int)pts.size() / 3;
    out.assign((size_t)4 * nn * np, 0.0);
    for (int p = 0; p < np; p++) {
        D ki = var(pts[3 * p], 0), eta = var(pts[3 * p + 1], 1), phi = var(pts[3 * p + 2], 2);
        std::vector<D> lam, sig;
        if (type == POINT) {
            out[p] = 1.0;
            continue;
        }
        lambdas(type, ki, eta, phi, lam, sig);
        for (int n = 0; n < nn; n++) {
            const D& v = lam[n + 1];
            out[((size_t)0 * nn + n) * np + p] = v.v;
            out[((size_t)1 * nn + n) * np + p] = v.d[0];
            out[((size_t)2 * nn + n) * np + p] = v.d[1];
            out[((size_t)3 * nn + n) * np + p] = v.d[2];
        }
    }
}

void geometry_tables(int type, int n_geo, const std::vector<double>& pts, std::vector<double>& out) {
    if (n_geo == topo::count_nodes(type)) return lagrange1_tables(type, pts, out);
    SLX_CHECK(type != POINT && type != PYRAMID && n_geo == lagrange2_count(type),
              "element type " << type << " with " << n_geo << " geometry nodes is not supported (order 1 or 2 geometry)");
    lagrange2_tables(type, pts, out);
}

}  // namespace slx
