// Hierarchical H(curl) form functions (Zaglmayr's Nedelec-type basis) on
// lines, triangles, quadrangles, tetrahedra and hexahedra, as
// src/shapefunction/hierarchical/hcurl/hcurl{line,triangle,quadrangle,
// tetrahedron,hexahedron}.cpp define them, for the edge and face orientations
// of orientation.cpp:
//   edges     order 0 the lowest-order (Whitney) function, order h >= 1 a
//             gradient of an integrated Legendre polynomial of the edge;
//   faces     triangles from order 2, quadrangles from order 1: "type 1"
//             gradients, "type 2" and "type 3" functions;
//   volumes   tetrahedra from order 3 (types 1, 2, 2 + pc, 3), hexahedra from
//             order 1 (types 1, 2, 2 + pc and the three type-3 families).
// Values are vectors in reference coordinates, mapped by the covariant Piola
// transform J^-T in the formulation; curls are taken from the exact first
// derivatives of every component, which this file carries with second-order
// dual numbers (a function's value may itself be a gradient).
#include <array>
#include <cmath>
#include "basis.h"

namespace slx {
namespace {

// value, gradient and Hessian with respect to (ki, eta, phi)
struct D2 {
    double v = 0.0, g[3] = {0, 0, 0}, h[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    D2() = default;
    D2(double x) : v(x) {}  // NOLINT: constants convert implicitly
};
D2 operator+(const D2& a, const D2& b) {
    D2 r(a.v + b.v);
    for (int i = 0; i < 3; i++) {
        r.g[i] = a.g[i] + b.g[i];
        for (int j = 0; j < 3; j++) r.h[i][j] = a.h[i][j] + b.h[i][j];
    }
    return r;
}
D2 operator-(const D2& a, const D2& b) {
    D2 r(a.v - b.v);
    for (int i = 0; i < 3; i++) {
        r.g[i] = a.g[i] - b.g[i];
        for (int j = 0; j < 3; j++) r.h[i][j] = a.h[i][j] - b.h[i][j];
    }
    return r;
}
D2 operator*(const D2& a, const D2& b) {
    D2 r(a.v * b.v);
    for (int i = 0; i < 3; i++) {
        r.g[i] = a.g[i] * b.v + a.v * b.g[i];
        for (int j = 0; j < 3; j++) r.h[i][j] = a.h[i][j] * b.v + a.g[i] * b.g[j] + a.g[j] * b.g[i] + a.v * b.h[i][j];
    }
    return r;
}
D2 operator*(double s, const D2& a) { return D2(s) * a; }

// value and gradient: one component of a vector form function
struct D1 {
    double v = 0.0, g[3] = {0, 0, 0};
};
D1 operator+(const D1& a, const D1& b) {
    D1 r;
    r.v = a.v + b.v;
    for (int i = 0; i < 3; i++) r.g[i] = a.g[i] + b.g[i];
    return r;
}
D1 operator-(const D1& a, const D1& b) {
    D1 r;
    r.v = a.v - b.v;
    for (int i = 0; i < 3; i++) r.g[i] = a.g[i] - b.g[i];
    return r;
}
D1 operator*(const D1& a, const D1& b) {
    D1 r;
    r.v = a.v * b.v;
    for (int i = 0; i < 3; i++) r.g[i] = a.g[i] * b.v + a.v * b.g[i];
    return r;
}
D1 neg(const D1& a) {
    D1 r;
    r.v = -a.v;
    for (int i = 0; i < 3; i++) r.g[i] = -a.g[i];
    return r;
}
D1 val(const D2& a) {
    D1 r;
    r.v = a.v;
    for (int i = 0; i < 3; i++) r.g[i] = a.g[i];
    return r;
}
// d a / d(reference coordinate c) with its gradient
D1 der(const D2& a, int c) {
    D1 r;
    r.v = a.g[c];
    for (int i = 0; i < 3; i++) r.g[i] = a.h[c][i];
    return r;
}

// Legendre families of src/shapefunction/hierarchical/legendre.cpp: l
// (Legendre), L (integrated), ls and Ls (scaled); entries up to index maxn.
std::vector<D2> leg_l(int maxn, const D2& x) {
    std::vector<D2> l(std::max(maxn + 1, 2));
    l[0] = D2(1.0);
    l[1] = x;
    for (int n = 1; n < maxn; n++) l[n + 1] = (1.0 / (n + 1)) * ((double)(2 * n + 1) * l[n] * x - (double)n * l[n - 1]);
    return l;
}
std::vector<D2> leg_L(int maxn, const D2& x) {
    std::vector<D2> L(std::max(maxn + 1, 3));
    L[1] = x;
    L[2] = 0.5 * (x * x - D2(1.0));
    for (int n = 2; n < maxn; n++) L[n + 1] = (1.0 / (n + 1)) * ((double)(2 * n - 1) * x * L[n] - (double)(n - 2) * L[n - 1]);
    return L;
}
std::vector<D2> leg_ls(int maxn, const D2& x, const D2& t) {
    std::vector<D2> ls(std::max(maxn + 1, 2));
    ls[0] = D2(1.0);
    ls[1] = x;
    for (int n = 1; n < maxn; n++) ls[n + 1] = (1.0 / (n + 1)) * ((double)(2 * n + 1) * x * ls[n] - (double)n * t * t * ls[n - 1]);
    return ls;
}
std::vector<D2> leg_Ls(int maxn, const D2& x, const D2& t) {
    std::vector<D2> Ls(std::max(maxn + 1, 3));
    Ls[1] = x;
    Ls[2] = 0.5 * (x * x - t * t);
    for (int n = 2; n < maxn; n++) Ls[n + 1] = (1.0 / (n + 1)) * ((double)(2 * n - 1) * x * Ls[n] - (double)(n - 2) * t * t * Ls[n - 1]);
    return Ls;
}

// Node reorderings bringing an edge / face to its orientation 0 (orientation.cpp:164-180).
const int kEdgeReorder[2][2] = {{0, 1}, {1, 0}};
const int kTriReorder[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 2, 0}, {1, 0, 2}, {2, 0, 1}, {2, 1, 0}};
const int kQuadReorder[8][4] = {{0, 1, 2, 3}, {0, 3, 2, 1}, {1, 2, 3, 0}, {1, 0, 3, 2},
                                {2, 3, 0, 1}, {2, 1, 0, 3}, {3, 0, 1, 2}, {3, 2, 1, 0}};

using Vec = std::array<D1, 3>;
Vec gradient(const D2& f) {
    Vec r;
    for (int c = 0; c < 3; c++) r[c] = der(f, c);
    return r;
}
Vec whitney(const D2& a, const D2& b) {  // grad(a) b - a grad(b)
    Vec r;
    for (int c = 0; c < 3; c++) r[c] = der(a, c) * val(b) - val(a) * der(b, c);
    return r;
}
Vec scaled(const Vec& w, const D1& s) {
    Vec r;
    for (int c = 0; c < 3; c++) r[c] = w[c] * s;
    return r;
}

bool simplex(int type) { return type == TRI || type == TET; }

// Number of functions of order <= `order` on one entity of type `assoc`
// (hcurl*::count(order, dim, num)).
int entity_count(int assoc, int order) {
    if (order < 0) return 0;
    switch (assoc) {
        case LINE: return order + 1;
        case TRI: return order < 2 ? 0 : (order - 1) * (order + 1);
        case QUAD: return 2 * order * (order + 1);
        case TET: return order < 3 ? 0 : (order - 2) * (order - 1) * (order + 1) / 2;
        case HEX: return 3 * order * order * (order + 1);
    }
    return 0;
}

int entity_type(int type, int dim, int num) {
    if (dim == 1) return LINE;
    if (dim == 2) return topo::is_triangular_face(type, num) ? TRI : QUAD;
    return type;
}

void check_type(int type) {
    SLX_CHECK(type == LINE || type == TRI || type == QUAD || type == TET || type == HEX,
              "H(curl) functions are provided on lines, triangles, quadrangles, tetrahedra and hexahedra");
}

// Barycentric 'lambda' and bilinear 'sigma' functions (1-based, index 0 unused).
void lambdas(int type, const D2* x, std::vector<D2>& lam, std::vector<D2>& sig) {
    lam.assign(9, D2(0.0));
    sig.assign(9, D2(0.0));
    if (type == TRI) {
        lam[1] = D2(1.0) - x[0] - x[1];
        lam[2] = x[0];
        lam[3] = x[1];
    } else if (type == TET) {
        lam[1] = D2(1.0) - x[0] - x[1] - x[2];
        lam[2] = x[0];
        lam[3] = x[1];
        lam[4] = x[2];
    } else {
        // reference coordinates in [-1, 1] shifted to [0, 1]
        const D2 a = 0.5 * (x[0] + D2(1.0)), b = 0.5 * (x[1] + D2(1.0)), c = 0.5 * (x[2] + D2(1.0));
        const D2 a1 = D2(1.0) - a, b1 = D2(1.0) - b, c1 = D2(1.0) - c;
        if (type == LINE) {
            lam[1] = sig[1] = a1;
            lam[2] = sig[2] = a;
        } else if (type == QUAD) {
            lam[1] = a1 * b1; lam[2] = a * b1; lam[3] = a * b; lam[4] = a1 * b;
            sig[1] = a1 + b1; sig[2] = a + b1; sig[3] = a + b; sig[4] = a1 + b;
        } else {
            lam[1] = a1 * b1 * c1; lam[2] = a * b1 * c1; lam[3] = a * b * c1; lam[4] = a1 * b * c1;
            lam[5] = a1 * b1 * c;  lam[6] = a * b1 * c;  lam[7] = a * b * c;  lam[8] = a1 * b * c;
            sig[1] = a1 + b1 + c1; sig[2] = a + b1 + c1; sig[3] = a + b + c1; sig[4] = a1 + b + c1;
            sig[5] = a1 + b1 + c;  sig[6] = a + b1 + c;  sig[7] = a + b + c;  sig[8] = a1 + b + c;
        }
    }
}

// Function f (entity orientation `orient`) at one point; x are the reference
// coordinates as dual numbers, lam / sig the element's lambda / sigma.
Vec eval_hcurl(int type, int maxorder, const FormFunction& f, int orient, const D2* x, const std::vector<D2>& lam,
               const std::vector<D2>& sig) {
    if (f.dim == 1) {
        const std::vector<int>& en = topo::edge_nodes(type);
        const int e1 = en[2 * f.num + kEdgeReorder[orient][0]] + 1, e2 = en[2 * f.num + kEdgeReorder[orient][1]] + 1;
        if (simplex(type)) {
            // hcurltriangle.cpp / hcurltetrahedron.cpp edge functions
            if (f.order == 0) return whitney(lam[e1], lam[e2]);
            return gradient(leg_Ls(maxorder + 1, lam[e1] - lam[e2], lam[e1] + lam[e2])[f.order + 1]);
        }
        // hcurlline.cpp / hcurlquadrangle.cpp / hcurlhexahedron.cpp edge functions
        const D2 s = sig[e1] - sig[e2], l = lam[e1] + lam[e2];
        if (f.order == 0) return scaled(gradient(s), val(0.5 * l));
        return gradient(leg_L(maxorder + 1, s)[f.order + 1] * l);
    }
    if (f.dim == 2 && f.assoc_type == TRI) {
        const std::vector<int>& fn = topo::face_nodes(type);
        const int f1 = fn[3 * f.num + kTriReorder[orient][0]] + 1, f2 = fn[3 * f.num + kTriReorder[orient][1]] + 1,
                  f3 = fn[3 * f.num + kTriReorder[orient][2]] + 1;
        const D2 lf = lam[f1] + lam[f2] + lam[f3];
        const std::vector<D2> Ls = leg_Ls(maxorder, lam[f1] - lam[f2], lam[f1] + lam[f2]);
        const std::vector<D2> ls = leg_ls(maxorder, 2.0 * lam[f3] - lf, lf);
        int k = 0;
        for (int i = 0; i <= f.order - 2; i++) {
            const int j = f.order - 2 - i;
            const D2 ui = Ls[i + 2], vj = lam[f3] * ls[j];
            if (k++ == f.l) return gradient(ui * vj);                                    // type 1
            if (k++ == f.l) return whitney(ui, vj);                                      // type 2
            if (i == 0 && k++ == f.l) return scaled(whitney(lam[f1], lam[f2]), val(vj));  // type 3
        }
        fail("internal: H(curl) triangular face function out of range");
    }
    if (f.dim == 2) {
        // quadrangular face (of a quadrangle or a hexahedron)
        const std::vector<int>& fn = topo::face_nodes(type);
        const int off = 3 * topo::count_tri_faces(type) + 4 * (f.num - topo::count_tri_faces(type));
        int q[4];
        for (int m = 0; m < 4; m++) q[m] = fn[off + kQuadReorder[orient][m]] + 1;
        const D2 lf = lam[q[0]] + lam[q[1]] + lam[q[2]] + lam[q[3]];
        const D2 xi = sig[q[0]] - sig[q[1]], eta = sig[q[0]] - sig[q[3]];
        const std::vector<D2> lxi = leg_l(maxorder + 1, xi), leta = leg_l(maxorder + 1, eta);
        const std::vector<D2> Lxi = leg_L(maxorder + 1, xi), Leta = leg_L(maxorder + 1, eta);
        int k = 0;
        for (int i = 0; i <= f.order - 1; i++)
            for (int j = 0; j <= f.order - 1; j++) {
                if (i != f.order - 1 && j != f.order - 1) continue;
                if (k++ == f.l) return gradient(Lxi[i + 2] * Leta[j + 2] * lf);  // type 1
                if (k++ == f.l) {                                                  // type 2
                    Vec r;
                    for (int c = 0; c < 3; c++)
                        r[c] = (val(lxi[i + 1] * Leta[j + 2]) * der(xi, c) - val(Lxi[i + 2] * leta[j + 1]) * der(eta, c)) * val(lf);
                    return r;
                }
                if (i == 0 && k++ == f.l) return scaled(gradient(xi), val(Leta[j + 2] * lf));  // type 3
                if (j == 0 && k++ == f.l) return scaled(gradient(eta), val(Lxi[i + 2] * lf));
            }
        fail("internal: H(curl) quadrangular face function out of range");
    }
    if (type == TET) {
        const std::vector<D2> Ls = leg_Ls(maxorder, lam[1] - lam[2], lam[1] + lam[2]);
        const std::vector<D2> ls = leg_ls(maxorder, 2.0 * lam[3] - (D2(1.0) - lam[4]), D2(1.0) - lam[4]);
        const std::vector<D2> l = leg_l(maxorder, 2.0 * lam[4] - D2(1.0));
        int k = 0;
        for (int i = 0; i <= f.order - 3; i++)
            for (int j = 0; j <= f.order - 3; j++)
                for (int kk = 0; kk <= f.order - 3; kk++) {
                    if (i + j + kk != f.order - 3) continue;
                    const D2 ui = Ls[i + 2], vj = lam[3] * ls[j], wk = lam[4] * l[kk];
                    if (k++ == f.l) return gradient(ui * vj * wk);  // type 1
                    Vec r;
                    if (k++ == f.l) {  // type 2
                        for (int c = 0; c < 3; c++)
                            r[c] = der(ui, c) * val(vj * wk) - val(ui) * der(vj, c) * val(wk) + val(ui * vj) * der(wk, c);
                        return r;
                    }
                    if (k++ == f.l) {  // type 2 + pc
                        for (int c = 0; c < 3; c++)
                            r[c] = der(ui, c) * val(vj * wk) + val(ui) * der(vj, c) * val(wk) - val(ui * vj) * der(wk, c);
                        return r;
                    }
                    if (i == 0 && k++ == f.l) return scaled(whitney(lam[1], lam[2]), val(vj * wk));  // type 3
                }
        fail("internal: H(curl) tetrahedron volume function out of range");
    }
    // hexahedron volume
    const std::vector<D2> La = leg_L(maxorder + 1, x[0]), Lb = leg_L(maxorder + 1, x[1]), Lc = leg_L(maxorder + 1, x[2]);
    int k = 0;
    for (int i = 0; i <= f.order - 1; i++)
        for (int j = 0; j <= f.order - 1; j++)
            for (int kk = 0; kk <= f.order - 1; kk++) {
                if (i != f.order - 1 && j != f.order - 1 && kk != f.order - 1) continue;
                const Vec g = gradient(La[i + 2] * Lb[j + 2] * Lc[kk + 2]);
                if (k++ == f.l) return g;                                // type 1
                if (k++ == f.l) return Vec{g[0], neg(g[1]), g[2]};       // type 2
                if (k++ == f.l) return Vec{g[0], neg(g[1]), neg(g[2])};  // type 2 + pc
                const D1 zero;
                if (i == 0 && k++ == f.l) return Vec{val(Lb[j + 2] * Lc[kk + 2]), zero, zero};  // type 3
                if (j == 0 && k++ == f.l) return Vec{zero, val(La[i + 2] * Lc[kk + 2]), zero};
                if (kk == 0 && k++ == f.l) return Vec{zero, zero, val(La[i + 2] * Lb[j + 2])};
            }
    fail("internal: H(curl) hexahedron volume function out of range");
}

}  // namespace

int hcurl_count(int type, int order, int dim, int num) {
    check_type(type);
    if (dim < 1 || dim > topo::dimension(type)) return 0;
    return entity_count(entity_type(type, dim, num), order);
}

int hcurl_count(int type, int order) {
    check_type(type);
    int n = 0;
    for (int dim = 1; dim <= topo::dimension(type); dim++)
        for (int num = 0; num < topo::count_dim(type, dim); num++) n += hcurl_count(type, order, dim, num);
    return n;
}

std::vector<FormFunction> hcurl_iterate(int type, int order) {
    check_type(type);
    SLX_CHECK(order >= 0, "H(curl) order must be at least 0");
    std::vector<FormFunction> out;
    for (int dim = 1; dim <= topo::dimension(type); dim++)
        for (int num = 0; num < topo::count_dim(type, dim); num++) {
            const int assoc = entity_type(type, dim, num);
            for (int idx = 0; idx < entity_count(assoc, order); idx++) {
                FormFunction f;
                f.dim = dim;
                f.num = num;
                f.index = idx;
                int h = 0;
                while (idx >= entity_count(assoc, h)) h++;
                f.order = h;
                f.l = idx - entity_count(assoc, h - 1);
                f.assoc_type = assoc;
                f.sub = (dim == 2 && assoc == QUAD) ? num - topo::count_tri_faces(type) : num;
                out.push_back(f);
            }
        }
    SLX_CHECK((int)out.size() == hcurl_count(type, order), "H(curl) iterator count mismatch");
    return out;
}

H1Fold hcurl_fold(int type, int order, const std::vector<double>& pts) {
    // the same search as h1_fold on the three components and their reference
    // gradients: quadrangle and hexahedron functions under another edge / face
    // orientation are signed functions of orientation 0 (Legendre parity,
    // exchanged face axes); triangle faces are not, and the search fails
    H1Fold F;
    const std::vector<FormFunction> ffs = hcurl_iterate(type, order);
    const int nff = (int)ffs.size(), np = (int)pts.size() / 3;
    std::vector<D2> lam, sig;
    auto values = [&](const FormFunction& ff, int orient, std::vector<double>& v) {
        v.resize((size_t)12 * np);
        for (int p = 0; p < np; p++) {
            D2 x[3];
            for (int d = 0; d < 3; d++) {
                x[d] = D2(pts[3 * p + d]);
                x[d].g[d] = 1.0;
            }
            lambdas(type, x, lam, sig);
            const Vec r = eval_hcurl(type, order, ff, orient, x, lam, sig);
            for (int c = 0; c < 3; c++) {
                v[12 * (size_t)p + 4 * c] = r[c].v;
                for (int d = 0; d < 3; d++) v[12 * (size_t)p + 4 * c + 1 + d] = r[c].g[d];
            }
        }
    };
    std::vector<std::vector<double>> v0(nff);
    for (int f = 0; f < nff; f++)
        if (ffs[f].dim <= 2) values(ffs[f], 0, v0[f]);
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
    for (int f = 0; f < nff; f++)
        for (size_t v = 0; v < F.map[f].size(); v++)
            for (int g = f + 1; g < nff; g++)
                if (ffs[g].dim == ffs[f].dim && ffs[g].num == ffs[f].num && F.map[g].size() > v &&
                    F.map[g][v].first == F.map[f][v].first)
                    return F;
    F.ok = true;
    return F;
}

void hcurl_tables(int type, int order, int total_orientation, int family, const std::vector<double>& pts,
                  std::vector<double>& out) {
    SLX_CHECK(family == HCURL_VALUE || family == HCURL_CURL, "unknown H(curl) table family " << family);
    const std::vector<FormFunction> ffs = hcurl_iterate(type, order);
    const int nff = (int)ffs.size(), np = (int)pts.size() / 3;
    std::vector<int> eo, fo;
    orientations_from_total(type, total_orientation, eo, fo);
    out.assign((size_t)4 * nff * np, 0.0);
    std::vector<D2> lam, sig;
    for (int p = 0; p < np; p++) {
        D2 x[3];
        for (int d = 0; d < 3; d++) {
            x[d] = D2(pts[3 * p + d]);
            x[d].g[d] = 1.0;
        }
        lambdas(type, x, lam, sig);
        for (int f = 0; f < nff; f++) {
            const FormFunction& ff = ffs[f];
            // a 2D element's own face carries its face orientation too
            const int orient = ff.dim == 1 ? eo[ff.num] : ff.dim == 2 ? fo[ff.num] : 0;
            const Vec v = eval_hcurl(type, order, ff, orient, x, lam, sig);
            double comp[3];
            if (family == HCURL_VALUE) {
                for (int c = 0; c < 3; c++) comp[c] = v[c].v;
            } else {
                comp[0] = v[2].g[1] - v[1].g[2];
                comp[1] = v[0].g[2] - v[2].g[0];
                comp[2] = v[1].g[0] - v[0].g[1];
            }
            // slot 0 is unused; slots 1..3 hold the reference components
            for (int c = 0; c < 3; c++) out[((size_t)(1 + c) * nff + f) * np + p] = comp[c];
        }
    }
}

}  // namespace slx
