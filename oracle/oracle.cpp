// CPU FP64 oracle for the weak-form assembly path.  TEST INFRASTRUCTURE ONLY:
// it is loaded by tests/, by __graft_entry__.smoke() and by the cpu_baseline
// leg of bench.py, as the checker and the CPU baseline; the product never
// calls it.
//
// Restates, on the plan description of include/slx.h:
//   contribution::generate   src/formulation/contribution.cpp:133-306
//     - Jacobian, |det J|, inverse         src/field/jacobian.cpp:4-505
//     - coefficient per term at every GP   contribution.cpp:155-163 (operation::interpolate)
//     - tf*w (multiplycolumns), tf x dof   contribution.cpp:172-187, densemat.cpp:608-636
//     - coef*|detJ|, transpose, GEMM       contribution.cpp:195-211, densemat.cpp:228-311
//     - first term copied, next ones added contribution.cpp:241-245 (getproduct/addproduct)
//     - rhs: minus, VecSetValues ADD       contribution.cpp:293-297, rawvec.cpp:216-254
//     - Dirichlet values in b()            formulation.cpp:361-385, vec.cpp:43-61
//   orientation folding (plan descriptions whose slices carry tf_sign /
//   dof_sign): element matrices on the orientation-0 tables, rows/columns
//   signed when accumulated; pinned against the unfolded description
//   (one batch per total orientation, elementselector.cpp:4-133)
//   rawmat::process                        src/formulation/rawmat.cpp:127-298
//     - per-row bucketing in fragment -> row r -> element order
//     - per-row stable sort by column (ties by position), sequential FP64 sum
//     - split into A (free columns) and D (constrained columns), local numbers
// The fragments are processed in frag_rank order, which is the order in which
// formulation::generate pushes them into the rawmat (formulation.cpp:176-218).
//
// The GEMM inner product runs over the Gauss points in increasing order with
// plain multiply-add, like the reference's non-BLAS path (densemat.cpp:285-307);
// OpenBLAS's summation order for larger shapes is not pinned (parity unpinned
// at the dgemm boundary: values are compared with a relative tolerance).
#include <algorithm>
#include <array>
#include <stdexcept>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <string>
#include <thread>
#include <vector>
#include "../include/slx.h"

namespace {

struct Frag {
    int bucket, slice, rank;
};

bool lagrange2(int type, const double* pts, int np, std::vector<double>& out);
void lagrange1(int type, const double* pts, int np, std::vector<double>& out);

// Geometry functions of a bucket: order 1 on the corner nodes, or order 2 when
// the bucket lists every node of curved elements.
void geometry(const slx_bucket_desc& b, std::vector<double>& out) {
    const int corners = b.elem_type == 1 ? 2 : b.elem_type == 2 ? 3 : b.elem_type == 3 ? 4 : b.elem_type == 4 ? 4 : b.elem_type == 6 ? 6 : 8;
    if (b.n_geo != corners) {
        if (!lagrange2(b.elem_type, b.gp_coords, b.n_gp, out) || (int)out.size() != 4 * b.n_geo * b.n_gp)
            throw std::runtime_error("oracle: unsupported geometry");
        return;
    }
    lagrange1(b.elem_type, b.gp_coords, b.n_gp, out);
}

void jacobian(const slx_plan_desc& d, const slx_bucket_desc& b, const double* xyz, const std::vector<double>& geo, int e, int q,
              double J[3][3], double inv[3][3], double& det) {
    int edim = b.elem_type == 1 ? 1 : (b.elem_type <= 3 ? 2 : 3);
    int md = d.mesh_dim;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) J[r][c] = 0.0;
    for (int r = 0; r < edim; r++)
        for (int c = 0; c < md; c++) {
            double s = 0.0;
            for (int k = 0; k < b.n_geo; k++) {
                int node = b.geo_nodes[(size_t)e * b.n_geo + k];
                s += xyz[3 * (size_t)node + c] * geo[((size_t)(1 + r) * b.n_geo + k) * b.n_gp + q];
            }
            J[r][c] = s;
        }
    if (md == 1) {
        det = J[0][0];
        inv[0][0] = 1.0 / J[0][0];
    } else if (md == 2) {
        det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        inv[0][0] = J[1][1] / det;
        inv[0][1] = -J[0][1] / det;
        inv[1][0] = -J[1][0] / det;
        inv[1][1] = J[0][0] / det;
    } else {
        det = J[0][0] * (J[1][1] * J[2][2] - J[2][1] * J[1][2]) - J[0][1] * (J[1][0] * J[2][2] - J[2][0] * J[1][2]) +
              J[0][2] * (J[1][0] * J[2][1] - J[2][0] * J[1][1]);
        inv[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
        inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
        inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
        inv[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
        inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
        inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
        inv[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
        inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
        inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
    }
}

// Order-2 (curved) geometry functions: the Lagrange interpolants of the nodes
// of lagrange*.cpp getnodecoordinates(2) (gmsh node order), built here by
// inverting the Vandermonde matrix of each type's polynomial space (tensor
// quadratics for lines, quadrangles and hexahedra, total degree 2 for
// triangles and tetrahedra, their product for prisms) -- not from closed forms.
// Returns false for a type without order-2 geometry.
bool lagrange2(int type, const double* pts, int np, std::vector<double>& out) {
    static const double line[] = {-1, 0, 0, 1, 0, 0, 0, 0, 0};
    static const double tri[] = {0, 0, 0, 1, 0, 0, 0, 1, 0, 0.5, 0, 0, 0.5, 0.5, 0, 0, 0.5, 0};
    static const double quad[] = {-1, -1, 0, 1, -1, 0, 1, 1, 0, -1, 1, 0, 0, -1, 0, 1, 0, 0, 0, 1, 0, -1, 0, 0, 0, 0, 0};
    static const double tet[] = {0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1, 0.5, 0, 0, 0.5, 0.5, 0, 0, 0.5, 0, 0, 0, 0.5, 0, 0.5, 0.5, 0.5, 0, 0.5};
    static const double hex[] = {-1, -1, -1, 1, -1, -1, 1, 1, -1, -1, 1, -1, -1, -1, 1, 1, -1, 1, 1, 1, 1, -1, 1, 1, 0, -1, -1,
                                 -1, 0, -1, -1, -1, 0, 1, 0, -1, 1, -1, 0, 0, 1, -1, 1, 1, 0, -1, 1, 0, 0, -1, 1, -1, 0, 1,
                                 1, 0, 1, 0, 1, 1, 0, 0, -1, 0, -1, 0, -1, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};
    static const double prism[] = {0, 0, -1, 1, 0, -1, 0, 1, -1, 0, 0, 1, 1, 0, 1, 0, 1, 1, 0.5, 0, -1, 0, 0.5, -1, 0, 0, 0,
                                   0.5, 0.5, -1, 1, 0, 0, 0, 1, 0, 0.5, 0, 1, 0, 0.5, 1, 0.5, 0.5, 1, 0.5, 0, 0, 0, 0.5, 0,
                                   0.5, 0.5, 0};
    const double* c;
    int nn;
    switch (type) {
        case 1: c = line; nn = 3; break;
        case 2: c = tri; nn = 6; break;
        case 3: c = quad; nn = 9; break;
        case 4: c = tet; nn = 10; break;
        case 5: c = hex; nn = 27; break;
        case 6: c = prism; nn = 18; break;
        default: return false;
    }
    // exponents (a, b, g) of the polynomial space
    std::vector<std::array<int, 3>> mono;
    for (int a = 0; a <= 2; a++)
        for (int b2 = 0; b2 <= 2; b2++)
            for (int g = 0; g <= 2; g++) {
                bool in = false;
                if (type == 1) in = b2 == 0 && g == 0;
                else if (type == 3) in = g == 0;
                else if (type == 5) in = true;
                else if (type == 2) in = g == 0 && a + b2 <= 2;
                else if (type == 4) in = a + b2 + g <= 2;
                else in = a + b2 <= 2;  // prism: triangle quadratics times line quadratics
                if (in) mono.push_back({a, b2, g});
            }
    if ((int)mono.size() != nn) return false;
    // V[k][m] = monomial m at node k; invert (Gauss-Jordan, partial pivoting)
    std::vector<double> V((size_t)nn * nn), I((size_t)nn * nn, 0.0);
    auto pw = [](double x, int e) { return e == 0 ? 1.0 : e == 1 ? x : x * x; };
    for (int k = 0; k < nn; k++) {
        for (int m = 0; m < nn; m++) V[(size_t)k * nn + m] = pw(c[3 * k], mono[m][0]) * pw(c[3 * k + 1], mono[m][1]) * pw(c[3 * k + 2], mono[m][2]);
        I[(size_t)k * nn + k] = 1.0;
    }
    for (int col = 0; col < nn; col++) {
        int piv = col;
        for (int r = col + 1; r < nn; r++)
            if (std::fabs(V[(size_t)r * nn + col]) > std::fabs(V[(size_t)piv * nn + col])) piv = r;
        for (int j = 0; j < nn; j++) {
            std::swap(V[(size_t)col * nn + j], V[(size_t)piv * nn + j]);
            std::swap(I[(size_t)col * nn + j], I[(size_t)piv * nn + j]);
        }
        const double d = V[(size_t)col * nn + col];
        for (int j = 0; j < nn; j++) {
            V[(size_t)col * nn + j] /= d;
            I[(size_t)col * nn + j] /= d;
        }
        for (int r = 0; r < nn; r++)
            if (r != col) {
                const double f = V[(size_t)r * nn + col];
                if (f == 0.0) continue;
                for (int j = 0; j < nn; j++) {
                    V[(size_t)r * nn + j] -= f * V[(size_t)col * nn + j];
                    I[(size_t)r * nn + j] -= f * I[(size_t)col * nn + j];
                }
            }
    }
    // N_k(x) = sum_m mono_m(x) Vinv[m][k]
    auto dpw = [](double x, int e) { return e == 0 ? 0.0 : e == 1 ? 1.0 : 2.0 * x; };
    out.assign((size_t)4 * nn * np, 0.0);
    for (int p = 0; p < np; p++) {
        const double x = pts[3 * p], y = pts[3 * p + 1], z = pts[3 * p + 2];
        for (int m = 0; m < nn; m++) {
            const int a = mono[m][0], b2 = mono[m][1], g = mono[m][2];
            const double v[4] = {pw(x, a) * pw(y, b2) * pw(z, g), dpw(x, a) * pw(y, b2) * pw(z, g), pw(x, a) * dpw(y, b2) * pw(z, g),
                                 pw(x, a) * pw(y, b2) * dpw(z, g)};
            for (int k = 0; k < nn; k++) {
                const double w = I[(size_t)m * nn + k];
                for (int dd = 0; dd < 4; dd++) out[((size_t)dd * nn + k) * np + p] += v[dd] * w;
            }
        }
    }
    return true;
}

// Order-1 Lagrange functions and reference derivatives: out[deriv][node][pt].
void lagrange1(int type, const double* pts, int np, std::vector<double>& out) {
    int nn = type == 1 ? 2 : type == 2 ? 3 : type == 3 ? 4 : type == 4 ? 4 : type == 6 ? 6 : 8;
    out.assign((size_t)4 * nn * np, 0.0);
    for (int p = 0; p < np; p++) {
        double k = pts[3 * p], e = pts[3 * p + 1], f = pts[3 * p + 2];
        double v[8][4];
        if (type == 1) {
            double x = 0.5 * (k + 1);
            double vv[2][4] = {{1 - x, -0.5, 0, 0}, {x, 0.5, 0, 0}};
            std::memcpy(v, vv, sizeof(vv));
        } else if (type == 2) {
            double vv[3][4] = {{1 - k - e, -1, -1, 0}, {k, 1, 0, 0}, {e, 0, 1, 0}};
            std::memcpy(v, vv, sizeof(vv));
        } else if (type == 4) {
            double vv[4][4] = {{1 - k - e - f, -1, -1, -1}, {k, 1, 0, 0}, {e, 0, 1, 0}, {f, 0, 0, 1}};
            std::memcpy(v, vv, sizeof(vv));
        } else if (type == 3) {
            double sx[4] = {-1, 1, 1, -1}, sy[4] = {-1, -1, 1, 1};
            for (int n = 0; n < 4; n++) {
                v[n][0] = 0.25 * (1 + sx[n] * k) * (1 + sy[n] * e);
                v[n][1] = 0.25 * sx[n] * (1 + sy[n] * e);
                v[n][2] = 0.25 * sy[n] * (1 + sx[n] * k);
                v[n][3] = 0;
            }
        } else if (type == 6) {
            // prism: triangle barycentrics (ki, eta) times (1 -+ phi) / 2
            // (lagrangeprism.cpp node order: 0-2 at phi = -1, 3-5 at phi = +1)
            double t[3][3] = {{1 - k - e, -1, -1}, {k, 1, 0}, {e, 0, 1}};
            for (int n = 0; n < 6; n++) {
                const double* tn = t[n % 3];
                const double s = n < 3 ? -1.0 : 1.0, mu = 0.5 * (1 + s * f);
                v[n][0] = tn[0] * mu;
                v[n][1] = tn[1] * mu;
                v[n][2] = tn[2] * mu;
                v[n][3] = 0.5 * s * tn[0];
            }
        } else {
            double sx[8] = {-1, 1, 1, -1, -1, 1, 1, -1}, sy[8] = {-1, -1, 1, 1, -1, -1, 1, 1}, sz[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
            for (int n = 0; n < 8; n++) {
                v[n][0] = 0.125 * (1 + sx[n] * k) * (1 + sy[n] * e) * (1 + sz[n] * f);
                v[n][1] = 0.125 * sx[n] * (1 + sy[n] * e) * (1 + sz[n] * f);
                v[n][2] = 0.125 * sy[n] * (1 + sx[n] * k) * (1 + sz[n] * f);
                v[n][3] = 0.125 * sz[n] * (1 + sx[n] * k) * (1 + sy[n] * e);
            }
        }
        for (int n = 0; n < nn; n++)
            for (int dd = 0; dd < 4; dd++) out[((size_t)dd * nn + n) * np + p] = v[n][dd];
    }
}

// operation::interpolate of every coefficient at one (element, GP):
// evaluates the program's expression DAG in FP64.
void coefficients(const slx_plan_desc& d, const slx_bucket_desc& b, const slx_numeric_desc& in, const std::vector<double>& geo,
                  int e, int q, const double J[3][3], const double inv[3][3], double det, std::vector<double>& coef) {
    std::vector<double> reg(std::max(b.n_regs, 1));
    int edim = b.elem_type == 1 ? 1 : (b.elem_type <= 3 ? 2 : 3);
    for (int pc = 0; pc < b.n_instr; pc++) {
        const slx_instr& ins = b.prog[pc];
        double va = (ins.a >= 0 && ins.a < b.n_regs) ? reg[ins.a] : 0.0;
        double vb = (ins.b >= 0 && ins.b < b.n_regs) ? reg[ins.b] : 0.0;
        double r = 0.0;
        switch (ins.op) {
            case SLX_OP_CONST: r = d.consts[ins.imm]; break;
            case SLX_OP_PARAM: r = in.params[(size_t)ins.imm * in.param_len + b.elem_ids[e]]; break;
            case SLX_OP_JAC: r = J[ins.a][ins.b]; break;
            case SLX_OP_INVJAC: r = inv[ins.a][ins.b]; break;
            case SLX_OP_DETJ: r = det; break;
            case SLX_OP_COORD: {
                for (int k = 0; k < b.n_geo; k++) {
                    int node = b.geo_nodes[(size_t)e * b.n_geo + k];
                    r += in.node_xyz[3 * (size_t)node + ins.a] * geo[(size_t)k * b.n_gp + q];
                }
                break;
            }
            case SLX_OP_FIELD: {
                int bidx = (int)(&b - d.buckets);
                size_t fk = (size_t)ins.imm * d.n_buckets + bidx;
                int nff = d.field_nff[fk];
                const int32_t* addr = d.field_addr[fk];
                const double* tab = d.field_tables[fk];
                double v[4] = {0, 0, 0, 0};
                for (int k = 0; k < nff; k++) {
                    int a = addr[(size_t)k * b.n_elem + e];
                    if (a < 0) continue;
                    double c = in.field_data[ins.imm][a];
                    for (int dd = 0; dd < 4; dd++) v[dd] += c * tab[((size_t)dd * nff + k) * b.n_gp + q];
                }
                if (ins.a == 0) r = v[0];
                else
                    for (int i = 0; i < edim; i++) r += inv[ins.a - 1][i] * v[1 + i];
                break;
            }
            case SLX_OP_ADD: r = va + vb; break;
            case SLX_OP_MUL: r = va * vb; break;
            case SLX_OP_INV: r = 1.0 / va; break;
            case SLX_OP_POW: r = std::pow(va, vb); break;
            case SLX_OP_ABS: r = std::fabs(va); break;
            case SLX_OP_EXP: r = std::exp(va); break;
            case SLX_OP_LOG: r = std::log(va); break;
            case SLX_OP_SIN: r = std::sin(va); break;
            case SLX_OP_COS: r = std::cos(va); break;
            case SLX_OP_TAN: r = std::tan(va); break;
            case SLX_OP_SQRT: r = std::sqrt(va); break;
            case SLX_OP_OUT: coef[ins.imm] = va; continue;
        }
        reg[ins.dst] = r;
    }
}

}  // namespace

extern "C" {

// Coefficient table of one bucket, as the device lays it out:
// out[(slot * nq + q) * n_elem + e] = coefficient * |det J|, nq = 1 for
// buckets whose coefficients do not vary over the Gauss points.
// Order-2 geometry functions of `type` at np points (reference coordinates,
// 3 per point): out[deriv][node][pt]; returns the node count (0: no order-2
// geometry for the type).
int oracle_lagrange2(int type, const double* pts, int np, double* out) {
    std::vector<double> v;
    if (!lagrange2(type, pts, np, v)) return 0;
    std::copy(v.begin(), v.end(), out);
    return (int)(v.size() / (4 * (size_t)np));
}

int oracle_coefficients(const slx_plan_desc* dp, const slx_numeric_desc* inp, int bucket, double* out) {
    const slx_plan_desc& d = *dp;
    if (bucket < 0 || bucket >= d.n_buckets) return -1;
    const slx_bucket_desc& b = d.buckets[bucket];
    const int nq = b.q_invariant ? 1 : b.n_gp;
    std::vector<double> geo;
    geometry(b, geo);
    std::vector<double> cbuf(std::max(b.n_coefs, 1));
    for (int e = 0; e < b.n_elem; e++)
        for (int q = 0; q < nq; q++) {
            double J[3][3], inv[3][3], det;
            jacobian(d, b, inp->node_xyz, geo, e, q, J, inv, det);
            coefficients(d, b, *inp, geo, e, q, J, inv, det, cbuf);
            for (int c = 0; c < b.n_coefs; c++) out[((size_t)c * nq + q) * b.n_elem + e] = cbuf[c] * std::fabs(det);
        }
    return 0;
}


// Full restatement: returns 0 on success.  Output arrays are sized by the
// caller from oracle_shape(); rhs has n_dofs entries (constrained entries hold
// the Dirichlet values, vec::updateconstraints, src/formulation/vec.cpp:43-61).
// Element matrices of every fragment are also returned when ke_out is not
// NULL, in the Ke layout of the plan (bucket, slice, element, i, j).
int oracle_assemble(const slx_plan_desc* dp, const slx_numeric_desc* inp, int nthreads, int64_t* nnzA_out, int64_t* nnzD_out,
                    int64_t** Arow_out, int32_t** Acol_out, double** Aval_out, int64_t** Drow_out, int32_t** Dcol_out,
                    double** Dval_out, double* rhs, double* ke_out);
void oracle_free(void* p);

int oracle_assemble(const slx_plan_desc* dp, const slx_numeric_desc* inp, int nthreads, int64_t* nnzA_out, int64_t* nnzD_out,
                    int64_t** Arow_out, int32_t** Acol_out, double** Aval_out, int64_t** Drow_out, int32_t** Dcol_out,
                    double** Dval_out, double* rhs, double* ke_out) {
    const slx_plan_desc& d = *dp;
    const slx_numeric_desc& in = *inp;
    const int64_t ndofs = d.n_dofs;
    std::vector<Frag> frags;
    std::vector<std::vector<int64_t>> kebase(d.n_buckets);
    int64_t keoff = 0;
    for (int b = 0; b < d.n_buckets; b++) {
        kebase[b].resize(d.buckets[b].n_slices);
        for (int s = 0; s < d.buckets[b].n_slices; s++) {
            const slx_slice_desc& sd = d.buckets[b].slices[s];
            frags.push_back({b, s, sd.frag_rank});
            kebase[b][s] = keoff;
            keoff += (int64_t)d.buckets[b].n_elem * sd.n_tf * (sd.is_rhs ? 1 : sd.n_dof);
        }
    }
    // element matrices of each fragment: vals[(i*n_dof + j) * n_elem + e] (rawmat row-major)
    // nthreads > 1: contiguous element ranges per thread (every entry is computed
    // by the same operations in the same order, so the result does not depend
    // on the thread count)
    std::vector<std::vector<double>> vals(frags.size());
    const int nth = std::max(1, nthreads);
    auto par = [&](int64_t n, auto&& fn) {
        if (nth <= 1 || n < 64) {
            fn((int64_t)0, n);
            return;
        }
        std::vector<std::thread> th;
        for (int t = 0; t < nth; t++) th.emplace_back(fn, n * t / nth, n * (t + 1) / nth);
        for (auto& x : th) x.join();
    };
    // coefficient table of the bucket whose fragments are being generated:
    // every slot at every (q, e), times |det J|, evaluated once per bucket
    // (the reference evaluates each coefficient once per batch and reuses it,
    // universe.cpp:223-314); cslot[c][q][e]
    int cached_bucket = -1;
    std::vector<std::vector<double>> cslot;
    for (size_t f = 0; f < frags.size(); f++) {
        const slx_bucket_desc& b = d.buckets[frags[f].bucket];
        const slx_slice_desc& sd = b.slices[frags[f].slice];
        int ne = b.n_elem, ng = b.n_gp, nt = sd.n_tf, ndf = sd.is_rhs ? 1 : sd.n_dof;
        if (cached_bucket != frags[f].bucket) {
            std::vector<double> geo;
            geometry(b, geo);
            cslot.assign(b.n_coefs, std::vector<double>((size_t)ne * ng));
            par(ne, [&](int64_t e0, int64_t e1) {
                std::vector<double> cbuf(std::max(b.n_coefs, 1));
                for (int64_t e = e0; e < e1; e++)
                    for (int q = 0; q < ng; q++) {
                        double J[3][3], inv[3][3], det;
                        jacobian(d, b, in.node_xyz, geo, (int)e, q, J, inv, det);
                        coefficients(d, b, in, geo, (int)e, q, J, inv, det, cbuf);
                        for (int c = 0; c < b.n_coefs; c++) cslot[c][(size_t)q * ne + e] = cbuf[c] * std::fabs(det);
                    }
            });
            cached_bucket = frags[f].bucket;
        }
        std::vector<double>& V = vals[f];
        V.assign((size_t)nt * ndf * ne, 0.0);
        const double* tft = b.tables[sd.tf_table];
        const double* dft = sd.is_rhs ? nullptr : b.tables[sd.dof_table];
        int nff_tf = b.table_nff[sd.tf_table], nff_dof = sd.is_rhs ? 1 : b.table_nff[sd.dof_table];
        std::vector<double> phi((size_t)nt * ndf * ng);
        for (int t = 0; t < sd.n_terms; t++) {
            const slx_term& tm = sd.terms[t];
            // Phi[(i*ndf + j)][q] = tf[i][q] * w[q] * dof[j][q]
            for (int i = 0; i < nt; i++)
                for (int j = 0; j < ndf; j++)
                    for (int q = 0; q < ng; q++) {
                        double tv = tft[((size_t)tm.tf_deriv * nff_tf + i) * ng + q] * b.gp_weights[q];
                        double dv = sd.is_rhs ? 1.0 : dft[((size_t)tm.dof_deriv * nff_dof + j) * ng + q];
                        phi[((size_t)i * ndf + j) * ng + q] = tv * dv;
                    }
            // term = Phi * C^T  (M = nt*ndf, K = ng, N = ne): per entry a sum over
            // the Gauss points in increasing order; elements innermost (vectorized)
            const double* Ct = cslot[tm.coef].data();
            par(ne, [&](int64_t e0, int64_t e1) {
                constexpr int EB = 64;
                double acc[EB];
                for (int64_t eb = e0; eb < e1; eb += EB) {
                    const int n = (int)std::min<int64_t>(EB, e1 - eb);
                    for (int r = 0; r < nt * ndf; r++) {
                        for (int x = 0; x < n; x++) acc[x] = 0.0;
                        for (int q = 0; q < ng; q++) {
                            const double ph = phi[(size_t)r * ng + q];
                            const double* cq = Ct + (size_t)q * ne + eb;
                            for (int x = 0; x < n; x++) acc[x] += ph * cq[x];
                        }
                        double* vr = V.data() + (size_t)r * ne + eb;
                        if (t == 0)
                            for (int x = 0; x < n; x++) vr[x] = acc[x];
                        else
                            for (int x = 0; x < n; x++) vr[x] += acc[x];
                    }
                }
            });
        }
        if (sd.is_rhs)
            for (double& v : V) v = -v;
        if (ke_out) {
            for (int e = 0; e < ne; e++)
                for (int i = 0; i < nt; i++)
                    for (int j = 0; j < ndf; j++)
                        ke_out[kebase[frags[f].bucket][frags[f].slice] + ((int64_t)e * nt + i) * ndf + j] = V[((size_t)i * ndf + j) * ne + e];
        }
    }
    std::vector<size_t> order(frags.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return frags[a].rank < frags[b].rank; });
    // rhs: rawvec::setvalues "add" in fragment order, tf-function-major
    if (rhs) {
        std::fill(rhs, rhs + ndofs, 0.0);
        for (size_t oi : order) {
            const slx_bucket_desc& b = d.buckets[frags[oi].bucket];
            const slx_slice_desc& sd = b.slices[frags[oi].slice];
            if (!sd.is_rhs) continue;
            for (int i = 0; i < sd.n_tf; i++)
                for (int e = 0; e < b.n_elem; e++) {
                    int a = sd.tf_addr[(size_t)i * b.n_elem + e];
                    // folded orientations: signed rows (exact)
                    double sg = sd.tf_sign ? (double)sd.tf_sign[(size_t)i * b.n_elem + e] : 1.0;
                    if (a >= 0) rhs[a] += sg * vals[oi][(size_t)i * b.n_elem + e];
                }
        }
        // vec::updateconstraints: constrained entries take the (zero) Dirichlet value
        for (int64_t i = 0; i < ndofs; i++)
            if (d.constrained && d.constrained[i]) rhs[i] = d.dirichlet ? d.dirichlet[i] : 0.0;
    }
    // rawmat::process
    std::vector<char> isc(ndofs, 0);
    for (int64_t i = 0; i < ndofs; i++) isc[i] = d.constrained ? d.constrained[i] : 0;
    std::vector<int64_t> renum(ndofs);
    int64_t nt_ = 0, nf_ = 0;
    for (int64_t i = 0; i < ndofs; i++) renum[i] = isc[i] ? nt_++ : nf_++;
    std::vector<int64_t> maxnnz(ndofs, 0);
    int64_t total = 0;
    for (size_t oi : order) {
        const slx_bucket_desc& b = d.buckets[frags[oi].bucket];
        const slx_slice_desc& sd = b.slices[frags[oi].slice];
        if (sd.is_rhs) continue;
        int nr = sd.n_tf * sd.n_dof, nc = b.n_elem;
        for (int r = 0; r < nr; r++)
            for (int c = 0; c < nc; c++) {
                int cr = sd.tf_addr[(size_t)(r / sd.n_dof) * nc + c], cc = sd.dof_addr[(size_t)(r % sd.n_dof) * nc + c];
                if (cr >= 0 && cc >= 0 && !isc[cr]) { maxnnz[cr]++; total++; }
            }
    }
    std::vector<int64_t> ads(ndofs + 1, 0);
    for (int64_t i = 0; i < ndofs; i++) ads[i + 1] = ads[i] + maxnnz[i];
    std::vector<std::pair<int, double>> pairs(total);
    std::vector<int64_t> fill(ads.begin(), ads.end() - 1);
    for (size_t oi : order) {
        const slx_bucket_desc& b = d.buckets[frags[oi].bucket];
        const slx_slice_desc& sd = b.slices[frags[oi].slice];
        if (sd.is_rhs) continue;
        int nr = sd.n_tf * sd.n_dof, nc = b.n_elem;
        for (int r = 0; r < nr; r++)
            for (int c = 0; c < nc; c++) {
                int cr = sd.tf_addr[(size_t)(r / sd.n_dof) * nc + c], cc = sd.dof_addr[(size_t)(r % sd.n_dof) * nc + c];
                if (cr >= 0 && cc >= 0 && !isc[cr]) {
                    double sg = 1.0;
                    if (sd.tf_sign) sg *= sd.tf_sign[(size_t)(r / sd.n_dof) * nc + c];
                    if (sd.dof_sign) sg *= sd.dof_sign[(size_t)(r % sd.n_dof) * nc + c];
                    pairs[fill[cr]++] = {cc, sg * vals[oi][(size_t)r * nc + c]};
                }
            }
    }
    // processrows: stable per-row sort by column, sequential duplicate sum
    std::vector<int64_t> rowlen(ndofs, 0);
    auto processrows = [&](int64_t r0, int64_t r1) {
        std::vector<int> idx;
        std::vector<std::pair<int, double>> tmp;
        for (int64_t i = r0; i < r1; i++) {
            int64_t n = maxnnz[i];
            std::pair<int, double>* p = pairs.data() + ads[i];
            idx.resize(n);
            std::iota(idx.begin(), idx.end(), 0);
            std::sort(idx.begin(), idx.end(), [&](int a, int b) {
                if (p[a].first != p[b].first) return p[a].first < p[b].first;
                return a < b;
            });
            tmp.assign(p, p + n);
            int pc = -1;
            int64_t ind = -1;
            for (int64_t j = 0; j < n; j++) {
                int cc = tmp[idx[j]].first;
                double cv = tmp[idx[j]].second;
                if (cc == pc) p[ind].second += cv;
                else {
                    ind++;
                    p[ind] = {cc, cv};
                }
                pc = cc;
            }
            rowlen[i] = ind + 1;
        }
    };
    if (nthreads <= 1) processrows(0, ndofs);
    else {
        std::vector<std::thread> th;
        for (int t = 0; t < nthreads; t++) th.emplace_back(processrows, ndofs * t / nthreads, ndofs * (t + 1) / nthreads);
        for (auto& x : th) x.join();
    }
    // A / D split
    int64_t nnzA = 0, nnzD = 0;
    for (int64_t i = 0; i < ndofs; i++) {
        if (isc[i]) continue;
        for (int64_t j = 0; j < rowlen[i]; j++) (isc[pairs[ads[i] + j].first] ? nnzD : nnzA)++;
    }
    int64_t nfree = nf_;
    int64_t* Arow = (int64_t*)std::malloc(sizeof(int64_t) * (nfree + 1));
    int64_t* Drow = (int64_t*)std::malloc(sizeof(int64_t) * (nfree + 1));
    int32_t* Acol = (int32_t*)std::malloc(sizeof(int32_t) * std::max<int64_t>(nnzA, 1));
    int32_t* Dcol = (int32_t*)std::malloc(sizeof(int32_t) * std::max<int64_t>(nnzD, 1));
    double* Aval = (double*)std::malloc(sizeof(double) * std::max<int64_t>(nnzA, 1));
    double* Dval = (double*)std::malloc(sizeof(double) * std::max<int64_t>(nnzD, 1));
    Arow[0] = 0;
    Drow[0] = 0;
    int64_t index = 0;
    for (int64_t i = 0; i < ndofs; i++) {
        if (isc[i]) continue;
        Arow[index + 1] = Arow[index];
        Drow[index + 1] = Drow[index];
        for (int64_t j = 0; j < rowlen[i]; j++) {
            int cc = pairs[ads[i] + j].first;
            double cv = pairs[ads[i] + j].second;
            if (isc[cc]) {
                Dcol[Drow[index + 1]] = (int32_t)renum[cc];
                Dval[Drow[index + 1]] = cv;
                Drow[index + 1]++;
            } else {
                Acol[Arow[index + 1]] = (int32_t)renum[cc];
                Aval[Arow[index + 1]] = cv;
                Arow[index + 1]++;
            }
        }
        index++;
    }
    *nnzA_out = nnzA;
    *nnzD_out = nnzD;
    *Arow_out = Arow;
    *Acol_out = Acol;
    *Aval_out = Aval;
    *Drow_out = Drow;
    *Dcol_out = Dcol;
    *Dval_out = Dval;
    return 0;
}

void oracle_free(void* p) { std::free(p); }

}  // extern "C"
