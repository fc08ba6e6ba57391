// Element kernels (VALU path), the Gauss-point coefficient kernel and the
// deterministic CSR reduction.
//
// coef_kernel: one thread per (element, Gauss point): FP64 Jacobian,
//   determinant and inverse (jacobian.cpp:4-505), then the coefficient program
//   (operation::interpolate of every term, contribution.cpp:155-163), times
//   |det J| (contribution.cpp:147, 200), into the table
//   ctab[(slot * nq + q) * n_elem + e]  (element fastest: coalesced for kernels
//   that map lanes to elements), or with qfast ctab[(slot * n_elem + e) * nq + q]
//   (Gauss point fastest: contiguous per element, for the factorized kernel).
// elem_cols: lanes = (element, dof column); per slice, TI tf rows at a time:
//   Ke[i][j] = sum_q sum_t tf_t[i][q] w_q C_t(e,q) dof_t[j][q]
//   (contribution.cpp:172-246; the per-term dgemm of densemat.cpp:228-311 with
//   the terms summed as in addproduct), right-hand-side slices write -sum
//   (contribution.cpp:296).
// reduce_rows: a few lanes per CSR row walk the row's element-row segments
//   (4 entries per lane, aligned vector loads) and add each entry into its
//   column slot in LDS (FP64), then write the A and D blocks
//   (rawmat.cpp:67-125, 238-289).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdlib>
#include "device.h"
#include "coef_device.h"

namespace slx {

typedef float float4v __attribute__((ext_vector_type(4)));
typedef double double2v __attribute__((ext_vector_type(2)));
typedef float float2v __attribute__((ext_vector_type(2)));
typedef unsigned int uint2v __attribute__((ext_vector_type(2)));
typedef unsigned int uint4v __attribute__((ext_vector_type(4)));
typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));

// Coefficient program, evaluated for V (element, Gauss point) pairs per
// thread: every interpreted instruction (uniform across the wave) is applied
// to V independent values, so one register-file access moves a V-vector and
// the instruction fetch and dispatch are shared.  Outputs go to
// ctab[(slot * nq + q) * n_elem + e] multiplied by |det J|.
template <typename T, int V>
__device__ void run_program(const DevBucket& b, const DevInputs& in, const int (&e)[V], const int (&q)[V], const Geo (&g)[V],
                            const bool (&on)[V], int nq, int qfast, T* __restrict__ ctab) {
    T reg[SLX_MAX_REGS][V];
    const int64_t stride = (int64_t)nq * b.n_elem;
    for (int pc = 0; pc < b.n_instr; pc++) {
        const slx_instr ins = b.prog[pc];
        const int ia = ins.a < SLX_MAX_REGS ? ins.a : 0, ib = ins.b < SLX_MAX_REGS ? ins.b : 0;
        T r[V];
        switch (ins.op) {
            case SLX_OP_CONST: {
                const T c = (T)in.consts[ins.imm];
#pragma unroll
                for (int v = 0; v < V; v++) r[v] = c;
                break;
            }
            case SLX_OP_PARAM:
#pragma unroll
                for (int v = 0; v < V; v++) r[v] = (T)in.params[(size_t)ins.imm * in.param_len + b.elem_ids[e[v]]];
                break;
            case SLX_OP_JAC:
#pragma unroll
                for (int v = 0; v < V; v++) r[v] = (T)pick9(g[v].J, 3 * ins.a + ins.b);
                break;
            case SLX_OP_INVJAC:
#pragma unroll
                for (int v = 0; v < V; v++) r[v] = (T)pick9(g[v].inv, 3 * ins.a + ins.b);
                break;
            case SLX_OP_DETJ:
#pragma unroll
                for (int v = 0; v < V; v++) r[v] = (T)g[v].det;
                break;
            case SLX_OP_COORD:
#pragma unroll
                for (int v = 0; v < V; v++) {
                    double s = 0.0;
                    for (int k = 0; k < b.n_geo; k++) {
                        const int node = b.geo_nodes[(size_t)e[v] * b.n_geo + k];
                        s += in.xyz[3 * (size_t)node + ins.a] * b.geo_tab[(size_t)k * b.n_gp + q[v]];
                    }
                    r[v] = (T)s;
                }
                break;
            case SLX_OP_FIELD:
#pragma unroll
                for (int v = 0; v < V; v++) r[v] = (T)field_value(b, in, ins.imm, ins.a, e[v], q[v], g[v]);
                break;
            case SLX_OP_ADD:
#pragma unroll
                for (int v = 0; v < V; v++) r[v] = reg[ia][v] + reg[ib][v];
                break;
            case SLX_OP_MUL:
#pragma unroll
                for (int v = 0; v < V; v++) r[v] = reg[ia][v] * reg[ib][v];
                break;
            case SLX_OP_POW:
#pragma unroll
                for (int v = 0; v < V; v++) r[v] = pow(reg[ia][v], reg[ib][v]);
                break;
            case SLX_OP_OUT:
#pragma unroll
                for (int v = 0; v < V; v++)
                    if (on[v])
                        ctab[(size_t)ins.imm * stride + (qfast ? (size_t)e[v] * nq + q[v] : (size_t)q[v] * b.n_elem + e[v])] =
                            reg[ia][v] * (T)fabs(g[v].det);
                continue;
            default:
#pragma unroll
                for (int v = 0; v < V; v++) r[v] = apply_unary<T>(ins.op, reg[ia][v]);
                break;
        }
#pragma unroll
        for (int v = 0; v < V; v++) reg[ins.dst][v] = r[v];
    }
}

// Coefficient table of one bucket (nq = 1 when the coefficients do not vary
// over the Gauss points); each thread takes V consecutive (element, Gauss
// point) pairs, element fastest.
template <typename T, int V>
__global__ void __launch_bounds__(256) coef_kernel(DevBucket b, DevInputs in, int nq, int qfast, T* __restrict__ ctab) {
    const int64_t total = (int64_t)b.n_elem * nq;
    const int64_t idx0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * V;
    if (idx0 >= total) return;
    int e[V], q[V];
    bool on[V];
    Geo g[V];
#pragma unroll
    for (int v = 0; v < V; v++) {
        const int64_t idx = idx0 + v;
        on[v] = idx < total;
        const int64_t id = on[v] ? idx : idx0;
        e[v] = qfast ? (int)(id / nq) : (int)(id % b.n_elem);
        q[v] = qfast ? (int)(id % nq) : (int)(id / b.n_elem);
        geometry(b, in.xyz, e[v], q[v], g[v]);
    }
    run_program<T, V>(b, in, e, q, g, on, nq, qfast, ctab);
}

// Element matrices, one lane per (element, dof column): `jc` consecutive
// lanes share an element and own one column j each, so the rows written to
// the element-row segments are contiguous runs.  Shape tables of the tf
// functions are read through the constant address space (scalar loads, Gauss
// point major); the dof table value of the lane's column is a vector load;
// coefficients come from the coefficient table.  Per slice:
//   D_rs    = sum of the slice's term coefficients with derivative pair (r, s)
//   h_r     = w_q sum_s D_rs dof_s[j][q]
//   Ke[i][j] += sum_r tf_r[i][q] h_r        (TI rows per pass)
// which is the reference's sum over terms and Gauss points regrouped.
template <typename CT, int TI, bool QV, bool COOP>
__global__ void __launch_bounds__(256) elem_cols(DevBucket b, const CT* __restrict__ ctab, uint32_t slice_mask, int jc, int qfast,
                                                 CT* __restrict__ ke) {
    // QV with COOP: the block's coefficients of the current slice, staged once
    // in LDS ([element][term][q]) and read from there by the element's lanes
    // (dynamic shared memory sized by the launcher)
    extern __shared__ __attribute__((aligned(16))) unsigned char elem_lds[];
    CT* cl = reinterpret_cast<CT*>(elem_lds);
    const int epb = blockDim.x / jc;
    const int el = threadIdx.x / jc, jl = threadIdx.x - (threadIdx.x / jc) * jc;
    const int e = blockIdx.x * epb + el;
    const bool live = el < epb && e < b.n_elem;
    const int ec = live ? e : 0;
    const int nq = b.q_inv ? 1 : b.n_gp;
    const int ngp = b.n_gp;
    const size_t cstride = (size_t)b.n_elem;
    const SLX_CONST double* gw = (const SLX_CONST double*)b.gp_w;
    const SLX_CONST DevSlice* dsl = (const SLX_CONST DevSlice*)b.dslices;
    for (int s = 0; s < b.n_slices; s++) {
        if (!((slice_mask >> s) & 1u)) continue;
        const int n_tf = dsl[s].n_tf, n_dof = dsl[s].n_dof, is_rhs = dsl[s].is_rhs, n_terms = dsl[s].n_terms;
        const int tf_pad = dsl[s].tf_pad, dof_pad = dsl[s].dof_pad;
        const SLX_CONST slx_term* terms = (const SLX_CONST slx_term*)dsl[s].terms;
        // shape tables in the kernel's precision (FP32 copies for CT = float)
        const SLX_CONST CT* tfq = sizeof(CT) == 4 ? (const SLX_CONST CT*)dsl[s].tf_tabq32 : (const SLX_CONST CT*)dsl[s].tf_tabq;
        const CT* dfq = sizeof(CT) == 4 ? (const CT*)dsl[s].dof_tabq32 : (const CT*)dsl[s].dof_tabq;
        const int64_t ke_base = dsl[s].ke_base, ke_ld = dsl[s].ke_ld;
        if (QV && COOP) {
            const int per = n_terms * nq;
            __syncthreads();  // the previous slice's reads are done
            for (int x = threadIdx.x; x < epb * per; x += blockDim.x) {
                const int ee = x / per, rem = x - ee * per, t = rem / nq, qq = rem - t * nq;
                const int eg = blockIdx.x * epb + ee;
                const int eu = eg < b.n_elem ? eg : 0;
                const int cs = terms[t].coef;
                cl[x] = qfast ? ctab[((size_t)cs * cstride + eu) * nq + qq] : ctab[((size_t)cs * nq + qq) * cstride + eu];
            }
            __syncthreads();
        }
        unsigned rmask = 0, dmask = 0;
        for (int t = 0; t < n_terms; t++) {
            rmask |= 1u << terms[t].tf_deriv;
            dmask |= 1u << terms[t].dof_deriv;
        }
        CT D[4][4];
        auto load_D = [&](int qq) {
#pragma unroll
            for (int r = 0; r < 4; r++)
#pragma unroll
                for (int d = 0; d < 4; d++) D[r][d] = 0;
            for (int t = 0; t < n_terms; t++) {
                const int tr = terms[t].tf_deriv, td = terms[t].dof_deriv;
                const CT c = qfast ? ctab[((size_t)terms[t].coef * cstride + ec) * nq + qq] : ctab[((size_t)terms[t].coef * nq + qq) * cstride + ec];
#pragma unroll
                for (int r = 0; r < 4; r++)
#pragma unroll
                    for (int d = 0; d < 4; d++)
                        if (r == tr && d == td) D[r][d] += c;
            }
        };
        if (nq == 1) load_D(0);
        for (int j0 = 0; j0 < n_dof; j0 += jc) {
            const int j = j0 + jl;
            const bool jon = live && j < n_dof;
            const int jr = jon ? j : 0;  // table column read by this lane (in range)
            for (int i0 = 0; i0 < n_tf; i0 += TI) {
                CT acc[TI];
#pragma unroll
                for (int ii = 0; ii < TI; ii++) acc[ii] = 0;
                for (int q = 0; q < ngp; q++) {
                    if (!QV && nq != 1) load_D(q);
                    const CT w = (CT)gw[q];
                    CT dv[4];
#pragma unroll
                    for (int d = 0; d < 4; d++)
                        dv[d] = (!is_rhs && ((dmask >> d) & 1u)) ? (CT)dfq[((size_t)q * 4 + d) * dof_pad + jr] : (CT)(d == 0 ? 1 : 0);
                    if (QV) {
                        // Gauss-point-varying coefficients: h_r = sum over the terms with
                        // tf derivative r of C_t(e, q) dof_{s_t}[j][q], term by term (no
                        // per-point 4 x 4 D assembly)
                        CT h[4] = {0, 0, 0, 0};
                        for (int t = 0; t < n_terms; t++) {
                            const int tr = terms[t].tf_deriv, td = terms[t].dof_deriv;
                            const CT c = COOP ? cl[((size_t)(live ? el : 0) * n_terms + t) * nq + q]
                                              : qfast ? ctab[((size_t)terms[t].coef * cstride + ec) * nq + q]
                                                      : ctab[((size_t)terms[t].coef * nq + q) * cstride + ec];
                            CT dsel = dv[0];
#pragma unroll
                            for (int d = 1; d < 4; d++)
                                if (d == td) dsel = dv[d];
                            const CT x = is_rhs ? c : c * dsel;
#pragma unroll
                            for (int r = 0; r < 4; r++)
                                if (r == tr) h[r] += x;
                        }
#pragma unroll
                        for (int r = 0; r < 4; r++) {
                            if (!((rmask >> r) & 1u)) continue;
                            const CT hw = h[r] * w;
                            const SLX_CONST CT* tv = tfq + ((size_t)q * 4 + r) * tf_pad + i0;
#pragma unroll
                            for (int ii = 0; ii < TI; ii++) acc[ii] += tv[ii] * hw;
                        }
                        continue;
                    }
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        if (!((rmask >> r) & 1u)) continue;
                        CT h = 0;
                        if (is_rhs) {
#pragma unroll
                            for (int d = 0; d < 4; d++) h += D[r][d];
                        } else {
#pragma unroll
                            for (int d = 0; d < 4; d++) h += D[r][d] * dv[d];
                        }
                        h *= w;
                        const SLX_CONST CT* tv = tfq + ((size_t)q * 4 + r) * tf_pad + i0;
#pragma unroll
                        for (int ii = 0; ii < TI; ii++) acc[ii] += tv[ii] * h;
                    }
                }
                if (jon) {
                    const int ni = min(TI, n_tf - i0);
                    CT* out = ke + ke_base + ((int64_t)e * n_tf + i0) * ke_ld + j;
#pragma unroll
                    for (int ii = 0; ii < TI; ii++)
                        if (ii < ni) out[(int64_t)ii * ke_ld] = is_rhs ? -acc[ii] : acc[ii];
                }
            }
        }
    }
}

// Deterministic reduction of the element-row segments into the CSR A/D
// values.  LPR lanes own one row; lane l holds entries 4l..4l+3 of the
// current segment (one aligned 4-entry vector load of values and one of
// slots), so within a segment every slot differs and the FP64 additions into
// the row's LDS accumulator never collide.  Segments are consumed in their
// fixed order, eight loaded ahead, so every value is summed in the same order
// on every run.  Rows are taken in the plan's length order (rows of one
// launch class have similar lengths; the LDS accumulator is sized per class).
template <typename T, int N>
struct VecN {
    T x[N];
};

// N consecutive entries (N = 4 or 2) with one aligned vector load
template <typename KT, int N>
__device__ inline VecN<KT, N> load_vals(const KT* p) {
    VecN<KT, N> r;
    if (sizeof(KT) == 4 && N == 4) {
        const float4v v = *reinterpret_cast<const float4v*>(p);
#pragma unroll
        for (int c = 0; c < N; c++) r.x[c] = (KT)v[c];
    } else if (sizeof(KT) == 4) {
        const float2v v = *reinterpret_cast<const float2v*>(p);
#pragma unroll
        for (int c = 0; c < N; c++) r.x[c] = (KT)v[c];
    } else {
#pragma unroll
        for (int c = 0; c < N; c += 2) {
            const double2v a = *reinterpret_cast<const double2v*>(p + c);
            r.x[c] = (KT)a[0];
            r.x[c + 1] = (KT)a[1];
        }
    }
    return r;
}

template <typename ST, int N>
__device__ inline VecN<ST, N> load_slots(const ST* p) {
    VecN<ST, N> r;
    if (sizeof(ST) * N == 4) {
        const uint32_t w = *reinterpret_cast<const uint32_t*>(p);
#pragma unroll
        for (int c = 0; c < N; c++) r.x[c] = (ST)((w >> (8 * sizeof(ST) * c)) & (sizeof(ST) == 1 ? 0xFFu : 0xFFFFu));
    } else if (sizeof(ST) * N == 2) {
        const uint16_t w = *reinterpret_cast<const uint16_t*>(p);
        r.x[0] = (ST)(w & 0xFFu);
        r.x[1] = (ST)(w >> 8);
    } else if (sizeof(ST) * N == 8) {
        const uint2v w = *reinterpret_cast<const uint2v*>(p);
        if (sizeof(ST) == 2) {
            r.x[0] = (ST)(w[0] & 0xFFFFu);
            r.x[1] = (ST)(w[0] >> 16);
            r.x[2] = (ST)(w[1] & 0xFFFFu);
            r.x[3] = (ST)(w[1] >> 16);
        } else {
            r.x[0] = (ST)w[0];
            r.x[1] = (ST)w[1];
        }
    } else {
        const uint4v w = *reinterpret_cast<const uint4v*>(p);
#pragma unroll
        for (int c = 0; c < N; c++) r.x[c] = (ST)w[c];
    }
    return r;
}

#define SLX_RED_BATCH 8

struct RowDesc {
    int64_t a0, d0, k0, sp;
    int lenA, lenD, nk;
};

__device__ inline RowDesc load_rowdesc(const uint64_t* rdesc, int64_t pos) {
    const uint64_t M48 = 0xFFFFFFFFFFFFull;
    const u64x2 w01 = *reinterpret_cast<const u64x2*>(rdesc + 4 * pos);
    const u64x2 w23 = *reinterpret_cast<const u64x2*>(rdesc + 4 * pos + 2);
    RowDesc r;
    r.a0 = (int64_t)(w01[0] & M48);
    r.lenA = (int)(w01[0] >> 48);
    r.d0 = (int64_t)(w01[1] & M48);
    r.lenD = (int)(w01[1] >> 48);
    r.k0 = (int64_t)(w23[0] & M48);
    r.nk = (int)(w23[0] >> 48);
    r.sp = (int64_t)w23[1];
    return r;
}

// Signed slots (folded orientations): the top bit negates the entry.
template <typename ST, bool SG>
__device__ inline void slot_add(double* acc, ST sl, double v) {
    const ST SKIP = (ST)~(ST)0;
    if (sl == SKIP) return;
    if (SG) {
        const ST top = (ST)(ST(1) << (8 * sizeof(ST) - 1));
        if (sl & top) v = -v;
        sl = (ST)(sl & (ST)~top);
    }
    (void)atomicAdd(&acc[sl], v);
}

template <typename KT, typename ST, int LPR, bool SG, int EPL>
__global__ void __launch_bounds__(256) reduce_rows(DevReduce rd, const KT* __restrict__ ke, int64_t pos0, int64_t pos1, int acc_len,
                                                    double* __restrict__ Aval, double* __restrict__ Dval) {
    extern __shared__ unsigned char smem_raw[];
    double* accall = reinterpret_cast<double*>(smem_raw);
    const int rpb = blockDim.x / LPR;
    const int gid = threadIdx.x / LPR, l = threadIdx.x % LPR;
    const int64_t pos = pos0 + (int64_t)blockIdx.x * rpb + gid;
    if (pos >= pos1) return;
    double* acc = accall + (size_t)gid * (acc_len | 1);  // odd stride: the rows' accumulators start on different banks
    const ST* slots = reinterpret_cast<const ST*>(rd.slots);
    const ST SKIP = (ST)~(ST)0;
    const RowDesc rw = load_rowdesc(rd.rdesc, pos);
    const int64_t a0 = rw.a0, d0 = rw.d0, k0 = rw.k0;
    const int lenA = rw.lenA, lenD = rw.lenD, nk = rw.nk;
    int64_t sp = rw.sp;
    for (int x = l; x < lenA + lenD; x += LPR) acc[x] = 0.0;
    // segment descriptors: the row's lanes load a batch together (DPL each) and
    // pass them around with shuffles.  Software pipeline: the values and slots
    // of batch b + 1 and the descriptors of batch b + 2 are in flight while
    // batch b is summed.
    constexpr int DPL = LPR >= SLX_RED_BATCH ? 1 : SLX_RED_BATCH / LPR;
    auto load_desc = [&](int kb, uint64_t (&d)[DPL]) {
#pragma unroll
        for (int t = 0; t < DPL; t++) {
            const int i = kb + l * DPL + t;
            d[t] = (l * DPL + t < SLX_RED_BATCH && i < nk) ? rd.erows[k0 + i] : 0ull;
        }
    };
    auto load_batch = [&](const uint64_t (&d)[DPL], VecN<KT, EPL> (&v)[SLX_RED_BATCH], VecN<ST, EPL> (&sl)[SLX_RED_BATCH]) {
#pragma unroll
        for (int u = 0; u < SLX_RED_BATCH; u++) {
            const uint64_t dsc = __shfl(d[u % DPL], u / DPL, LPR);
            const int n = (int)(dsc >> 48);
            const int64_t off = (int64_t)(dsc & 0xFFFFFFFFFFFFull);
            if (EPL * l < n) {
                v[u] = load_vals<KT, EPL>(ke + off + EPL * l);
                sl[u] = load_slots<ST, EPL>(slots + sp + EPL * l);
            } else {
#pragma unroll
                for (int c = 0; c < EPL; c++) {
                    v[u].x[c] = (KT)0;
                    sl[u].x[c] = SKIP;
                }
            }
            sp += (n + 3) & ~3;
        }
    };
    uint64_t dsc_next[DPL];
    VecN<KT, EPL> v[SLX_RED_BATCH];
    VecN<ST, EPL> sl[SLX_RED_BATCH];
    {
        uint64_t d0_[DPL];
        load_desc(0, d0_);
        load_desc(SLX_RED_BATCH, dsc_next);
        load_batch(d0_, v, sl);
    }
    for (int kb = 0; kb < nk; kb += SLX_RED_BATCH) {
        VecN<KT, EPL> vn[SLX_RED_BATCH];
        VecN<ST, EPL> sn[SLX_RED_BATCH];
        const bool more = kb + SLX_RED_BATCH < nk;
        if (more) {
            load_batch(dsc_next, vn, sn);
            load_desc(kb + 2 * SLX_RED_BATCH, dsc_next);
        }
        // LDS FP64 add without return (ds_add_f64): no read-after-write wait per
        // update; one wave's LDS operations complete in issue order and the
        // lanes of one segment never share a slot, so the order of the
        // additions into every slot is the program order (deterministic).
#pragma unroll
        for (int u = 0; u < SLX_RED_BATCH; u++)
#pragma unroll
            for (int c = 0; c < EPL; c++) slot_add<ST, SG>(acc, sl[u].x[c], (double)v[u].x[c]);
        if (more) {
#pragma unroll
            for (int u = 0; u < SLX_RED_BATCH; u++) {
                v[u] = vn[u];
                sl[u] = sn[u];
            }
        }
    }
    for (int x = l; x < lenA; x += LPR) Aval[a0 + x] = acc[x];
    for (int x = l; x < lenD; x += LPR) Dval[d0 + x] = acc[lenA + x];
}

// Segments longer than 64 entries: a wave per row, lanes stride over the segment.
// U > 0 (segments of at most 64 U entries): lane l takes entries l + 64 u, the
// next segment's values and slots are in flight while the current one is
// summed, and the descriptors come 64 at a time (one per lane, shuffled).
template <typename KT, typename ST, bool SG, int U>
__global__ void __launch_bounds__(256) reduce_matrix_long(DevReduce rd, const KT* __restrict__ ke, double* __restrict__ Aval,
                                                           double* __restrict__ Dval) {
    extern __shared__ unsigned char smem_raw[];
    double* accall = reinterpret_cast<double*>(smem_raw);
    const int group = 64;
    const int gid = threadIdx.x / group, lane = threadIdx.x % group;
    const int64_t pos = (int64_t)blockIdx.x * (blockDim.x / group) + gid;
    if (pos >= rd.nrows) return;
    double* acc = accall + (size_t)gid * rd.max_row_len;
    const ST* slots = reinterpret_cast<const ST*>(rd.slots);
    const RowDesc rw = load_rowdesc(rd.rdesc, pos);
    const int64_t a0 = rw.a0, d0 = rw.d0;
    const int lenA = rw.lenA, lenD = rw.lenD;
    for (int x = lane; x < lenA + lenD; x += group) acc[x] = 0.0;
    int64_t sp = rw.sp;
    if (U == 0) {
        for (int64_t k = rw.k0; k < rw.k0 + rw.nk; k++) {
            const uint64_t dsc = rd.erows[k];
            const int n = (int)(dsc >> 48);
            const int64_t off = (int64_t)(dsc & 0xFFFFFFFFFFFFull);
            for (int j = lane; j < n; j += group) slot_add<ST, SG>(acc, slots[sp + j], (double)ke[off + j]);
            sp += (n + 3) & ~3;
        }
    } else {
        constexpr int UU = U > 0 ? U : 1;
        const ST SKIP = (ST)~(ST)0;
        const int nk = rw.nk;
        uint64_t dl = lane < nk ? rd.erows[rw.k0 + lane] : 0ull;
        auto load_seg = [&](int k, KT (&v)[UU], ST (&sl)[UU]) {
            if ((k & 63) == 0 && k > 0) dl = (k + lane < nk) ? rd.erows[rw.k0 + k + lane] : 0ull;
            const uint64_t dsc = __shfl(dl, k & 63, 64);
            const int n = (int)(dsc >> 48);
            const int64_t off = (int64_t)(dsc & 0xFFFFFFFFFFFFull);
#pragma unroll
            for (int u = 0; u < UU; u++) {
                const int j = lane + 64 * u;
                v[u] = j < n ? ke[off + j] : (KT)0;
                sl[u] = j < n ? slots[sp + j] : SKIP;
            }
            sp += (n + 3) & ~3;
        };
        KT v[UU], vn[UU];
        ST sl[UU], sn[UU];
        if (nk > 0) load_seg(0, v, sl);
        for (int k = 0; k < nk; k++) {
            const bool more = k + 1 < nk;
            if (more) load_seg(k + 1, vn, sn);
#pragma unroll
            for (int u = 0; u < UU; u++) slot_add<ST, SG>(acc, sl[u], (double)v[u]);
            if (more) {
#pragma unroll
                for (int u = 0; u < UU; u++) {
                    v[u] = vn[u];
                    sl[u] = sn[u];
                }
            }
        }
    }
    for (int x = lane; x < lenA; x += group) Aval[a0 + x] = acc[x];
    for (int x = lane; x < lenD; x += group) Dval[d0 + x] = acc[lenA + x];
}

// rhs: b[dof] = sum of the element-row values in accumulation order
// (rawvec::setvalues "add", rawvec.cpp:216-254); constrained entries are 0.
template <typename KT>
__global__ void reduce_rhs(DevReduce rd, const KT* __restrict__ ke, double* __restrict__ rhs) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rd.nrows) return;
    double s = 0.0;
    const int64_t k1 = rd.rhs_erow_ptr[r + 1];
    int64_t k = rd.rhs_erow_ptr[r];
    // four entries' loads in flight at once; the sum keeps the entry order
    for (; k + 4 <= k1; k += 4) {
        uint64_t pk[4];
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) pk[u] = rd.rhs_erows[k + u];
#pragma unroll
        for (int u = 0; u < 4; u++) v[u] = (double)ke[(int64_t)(pk[u] & 0xFFFFFFFFFFFFULL)];
#pragma unroll
        for (int u = 0; u < 4; u++) s += (pk[u] >> 63) ? -v[u] : v[u];
    }
    for (; k < k1; k++) {
        uint64_t pk = rd.rhs_erows[k];
        const double v = (double)ke[(int64_t)(pk & 0xFFFFFFFFFFFFULL)];
        s += (pk >> 63) ? -v : v;
    }
    rhs[rd.free_dofs[r]] = s;
}

// ---------------------------------------------------------------- launchers --

hipError_t launch_coef(const DevBucket& b, const DevInputs& in, bool fp64, void* ctab, hipStream_t st, int qfast) {
    if (b.n_elem == 0 || b.n_coefs == 0) return hipSuccess;
    const int nq = b.q_inv ? 1 : b.n_gp;
    const int64_t n = (int64_t)b.n_elem * nq;
    if (fp64) {
        const unsigned grid = (unsigned)((n + 2 * 256 - 1) / (2 * 256));
        hipLaunchKernelGGL((coef_kernel<double, 2>), dim3(grid), dim3(256), 0, st, b, in, nq, qfast, (double*)ctab);
    } else {
        const unsigned grid = (unsigned)((n + 4 * 256 - 1) / (4 * 256));
        hipLaunchKernelGGL((coef_kernel<float, 4>), dim3(grid), dim3(256), 0, st, b, in, nq, qfast, (float*)ctab);
    }
    return hipGetLastError();
}

template <typename CT, int TI>
static void launch_cols(const DevBucket& b, const void* ctab, uint32_t mask, int jc, int qfast, void* ke, hipStream_t st) {
    const int epb = 256 / jc;
    const unsigned grid = (unsigned)((b.n_elem + epb - 1) / epb);
    if (b.q_inv) {
        hipLaunchKernelGGL((elem_cols<CT, TI, false, false>), dim3(grid), dim3(256), 0, st, b, (const CT*)ctab, mask, jc, qfast, (CT*)ke);
    } else {
        // stage the block's coefficients in LDS when they fit under the limit
        // (SLX_TUNE_ELEM_LDS_MAX bytes): a larger stage lowers the blocks a CU
        // holds at once
        int maxt = 0;
        for (int s = 0; s < b.n_slices; s++)
            if ((mask >> s) & 1u) maxt = max(maxt, b.slices[s].n_terms);
        const size_t bytes = (size_t)epb * maxt * b.n_gp * sizeof(CT);
        size_t lim = 48 * 1024;
        if (const char* t = std::getenv("SLX_TUNE_ELEM_LDS_MAX")) lim = (size_t)std::min(48L * 1024, std::max(0L, std::atol(t)));
        if (bytes <= lim)
            hipLaunchKernelGGL((elem_cols<CT, TI, true, true>), dim3(grid), dim3(256), bytes, st, b, (const CT*)ctab, mask, jc, qfast, (CT*)ke);
        else
            hipLaunchKernelGGL((elem_cols<CT, TI, true, false>), dim3(grid), dim3(256), 0, st, b, (const CT*)ctab, mask, jc, qfast, (CT*)ke);
    }
}

template <typename CT>
static void launch_cols_any(const DevBucket& b, const void* ctab, uint32_t mask, int qfast, void* ke, hipStream_t st) {
    int ntf = 1, ndof = 1;
    for (int s = 0; s < b.n_slices; s++)
        if ((mask >> s) & 1u) {
            ntf = max(ntf, b.slices[s].n_tf);
            ndof = max(ndof, b.slices[s].n_dof);
        }
    const int jc = min(ndof, 64);
    if (ntf <= 4) launch_cols<CT, 4>(b, ctab, mask, jc, qfast, ke, st);
    else if (ntf <= 8) launch_cols<CT, 8>(b, ctab, mask, jc, qfast, ke, st);
    else if (ntf <= 10) launch_cols<CT, 10>(b, ctab, mask, jc, qfast, ke, st);
    else launch_cols<CT, 16>(b, ctab, mask, jc, qfast, ke, st);
}

// Right-hand-side slices with many tf functions, as a blocked product over
// tiles of RB rows x EB elements:
//   b_e[i] = -sum_q sum_r tf_r[i][q] (w_q sum_{t: tf derivative r} C_t(e, q))
// (contribution.cpp:293-297).  Gauss points go in chunks of RQ: the tile's
// table rows and the elements' weighted coefficients are staged in LDS, and
// every thread accumulates 4 consecutive rows of 4 elements, so each table
// value and each coefficient is read from memory once per tile.
template <typename CT>
struct RhsVec4 { typedef CT type __attribute__((ext_vector_type(4))); };

template <typename CT>
__global__ void __launch_bounds__(256) rhs_tiles(DevBucket b, const CT* __restrict__ ctab, int s, int qfast, CT* __restrict__ ke) {
    constexpr int RB = 64, EB = 64, RQ = 16;
    typedef typename RhsVec4<CT>::type CT4;
    __shared__ __attribute__((aligned(32))) CT Ts[RQ][RB];  // table values of the chunk, [Gauss point][row]
    __shared__ __attribute__((aligned(32))) CT Cs[RQ][EB];  // weighted coefficients, [Gauss point][element]
    const SLX_CONST DevSlice* dsl = (const SLX_CONST DevSlice*)b.dslices;
    const int n_tf = dsl[s].n_tf, n_terms = dsl[s].n_terms, tf_pad = dsl[s].tf_pad;
    const SLX_CONST slx_term* terms = (const SLX_CONST slx_term*)dsl[s].terms;
    const SLX_CONST double* gw = (const SLX_CONST double*)b.gp_w;
    const CT* tab;
    if constexpr (sizeof(CT) == 4) tab = (const CT*)dsl[s].tf_tabq32;
    else tab = (const CT*)dsl[s].tf_tabq;
    const int ngp = b.n_gp, nq = b.q_inv ? 1 : b.n_gp, n_elem = b.n_elem;
    const size_t ne = (size_t)n_elem;
    const int e0 = blockIdx.x * EB, i0 = blockIdx.y * RB;
    const int tid = threadIdx.x, ri = tid & 15, ej = tid >> 4;  // rows i0 + 4 ri + 0..3, elements e0 + 4 ej + 0..3
    CT acc[4][4];
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 4; y++) acc[x][y] = 0;
    for (int r = 0; r < 4; r++) {
        bool used = false;
        for (int t = 0; t < n_terms; t++) used |= terms[t].tf_deriv == r;
        if (!used) continue;
        for (int q0 = 0; q0 < ngp; q0 += RQ) {
#pragma unroll
            for (int u = 0; u < RQ * RB / 256; u++) {
                const int x = tid + 256 * u, k = x / RB, il = x % RB, q = q0 + k, i = i0 + il;
                Ts[k][il] = (q < ngp && i < n_tf) ? tab[((size_t)q * 4 + r) * tf_pad + i] : (CT)0;
            }
#pragma unroll
            for (int u = 0; u < RQ * EB / 256; u++) {
                const int x = tid + 256 * u;
                // Gauss point fastest where the coefficient table is [slot][element][point]
                const int k = qfast ? x % RQ : x / EB, el = qfast ? x / RQ : x % EB, q = q0 + k, e = e0 + el;
                CT c = 0;
                if (q < ngp && e < n_elem) {
                    const int qq = nq == 1 ? 0 : q;
                    for (int t = 0; t < n_terms; t++)
                        if (terms[t].tf_deriv == r) {
                            const int cs = terms[t].coef;
                            c += qfast ? ctab[((size_t)cs * ne + e) * nq + qq] : ctab[((size_t)cs * nq + qq) * ne + e];
                        }
                    c *= (CT)gw[q];
                }
                Cs[k][el] = c;
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < RQ; k++) {
                const CT4 a = *reinterpret_cast<const CT4*>(&Ts[k][4 * ri]);
                const CT4 c = *reinterpret_cast<const CT4*>(&Cs[k][4 * ej]);
#pragma unroll
                for (int x = 0; x < 4; x++)
#pragma unroll
                    for (int y = 0; y < 4; y++) acc[x][y] += c[x] * a[y];
            }
            __syncthreads();
        }
    }
    const int64_t base = dsl[s].ke_base, ld = dsl[s].ke_ld;
#pragma unroll
    for (int x = 0; x < 4; x++) {
        const int e = e0 + 4 * ej + x;
        if (e >= n_elem) continue;
#pragma unroll
        for (int y = 0; y < 4; y++) {
            const int i = i0 + 4 * ri + y;
            if (i < n_tf) ke[base + ((int64_t)e * n_tf + i) * ld] = -acc[x][y];
        }
    }
}

hipError_t launch_elem(const DevBucket& b, const void* ctab, int precision, void* ke, hipStream_t st, uint32_t slice_mask, int qfast) {
    if (b.n_elem == 0 || slice_mask == 0) return hipSuccess;
    // right-hand sides with more tf functions than one column pass of elem_cols
    // covers go to the tiled kernel
    uint32_t rhs_mask = 0;
    for (int s = 0; s < b.n_slices; s++)
        if (((slice_mask >> s) & 1u) && b.slices[s].is_rhs && b.slices[s].n_tf > 16) rhs_mask |= 1u << s;
    const uint32_t mat_mask = slice_mask & ~rhs_mask;
    if (mat_mask) {
        if (precision == SLX_PREC_FP64)
            launch_cols_any<double>(b, ctab, mat_mask, qfast, ke, st);
        else
            launch_cols_any<float>(b, ctab, mat_mask, qfast, ke, st);
    }
    for (int s = 0; s < b.n_slices; s++) {
        if (!((rhs_mask >> s) & 1u)) continue;
        const dim3 grid((unsigned)((b.n_elem + 63) / 64), (unsigned)((b.slices[s].n_tf + 63) / 64));
        if (precision == SLX_PREC_FP64)
            hipLaunchKernelGGL((rhs_tiles<double>), grid, dim3(256), 0, st, b, (const double*)ctab, s, qfast, (double*)ke);
        else
            hipLaunchKernelGGL((rhs_tiles<float>), grid, dim3(256), 0, st, b, (const float*)ctab, s, qfast, (float*)ke);
    }
    return hipGetLastError();
}

template <typename KT, typename ST, int LPR, bool SG, int EPL>
static hipError_t launch_rows(const DevReduce& rd, const void* ke, double* Aval, double* Dval, hipStream_t st, const int64_t* cb,
                              const int64_t* ce, const int* cl, int nc) {
    for (int c = 0; c < nc; c++) {
        const int64_t nrows = ce[c] - cb[c];
        if (nrows <= 0) continue;
        int threads = 256;
        size_t lds = (size_t)(threads / LPR) * (cl[c] | 1) * sizeof(double);
        while (lds > 64 * 1024 && threads > 64) {
            threads /= 2;
            lds = (size_t)(threads / LPR) * (cl[c] | 1) * sizeof(double);
        }
        if (lds > 64 * 1024) return hipErrorInvalidValue;
        const int rpb = threads / LPR;
        const int64_t grid = (nrows + rpb - 1) / rpb;
        hipLaunchKernelGGL((reduce_rows<KT, ST, LPR, SG, EPL>), dim3((unsigned)grid), dim3(threads), lds, st, rd, (const KT*)ke, cb[c], ce[c], cl[c],
                           Aval, Dval);
    }
    return hipGetLastError();
}

template <typename KT, typename ST, bool SG>
static hipError_t launch_reduce_t(const DevReduce& rd, int max_erow_len, const void* ke, double* Aval, double* Dval, hipStream_t st,
                                  const int64_t* cb, const int64_t* ce, const int* cl, int nc) {
    if (max_erow_len > 64) {
        int threads = 256, gpb = threads / 64;
        size_t lds = (size_t)gpb * rd.max_row_len * sizeof(double);
        while (lds > 64 * 1024 && threads > 64) {
            threads /= 2;
            gpb = threads / 64;
            lds = (size_t)gpb * rd.max_row_len * sizeof(double);
        }
        if (lds > 64 * 1024) return hipErrorInvalidValue;
        const int64_t grid = (rd.nrows + gpb - 1) / gpb;
        if (max_erow_len <= 256)
            hipLaunchKernelGGL((reduce_matrix_long<KT, ST, SG, 4>), dim3((unsigned)grid), dim3(threads), lds, st, rd, (const KT*)ke, Aval, Dval);
        else
            hipLaunchKernelGGL((reduce_matrix_long<KT, ST, SG, 0>), dim3((unsigned)grid), dim3(threads), lds, st, rd, (const KT*)ke, Aval, Dval);
        return hipGetLastError();
    }
    // lanes per row: 4 entries per lane (2 entries per lane with twice the
    // lanes measured slower on the elasticity rows)
    if (max_erow_len <= 4) return launch_rows<KT, ST, 1, SG, 4>(rd, ke, Aval, Dval, st, cb, ce, cl, nc);
    if (max_erow_len <= 8) return launch_rows<KT, ST, 2, SG, 4>(rd, ke, Aval, Dval, st, cb, ce, cl, nc);
    if (max_erow_len <= 16) return launch_rows<KT, ST, 4, SG, 4>(rd, ke, Aval, Dval, st, cb, ce, cl, nc);
    if (max_erow_len <= 32) return launch_rows<KT, ST, 8, SG, 4>(rd, ke, Aval, Dval, st, cb, ce, cl, nc);
    return launch_rows<KT, ST, 16, SG, 4>(rd, ke, Aval, Dval, st, cb, ce, cl, nc);
}

template <typename KT>
static hipError_t launch_reduce_k(const DevReduce& rd, int max_erow_len, const void* ke, double* Aval, double* Dval, hipStream_t st,
                                  const int64_t* cb, const int64_t* ce, const int* cl, int nc) {
    if (rd.slot_signed) {
        if (rd.slot_bytes == 1) return launch_reduce_t<KT, uint8_t, true>(rd, max_erow_len, ke, Aval, Dval, st, cb, ce, cl, nc);
        if (rd.slot_bytes == 2) return launch_reduce_t<KT, uint16_t, true>(rd, max_erow_len, ke, Aval, Dval, st, cb, ce, cl, nc);
        return launch_reduce_t<KT, uint32_t, true>(rd, max_erow_len, ke, Aval, Dval, st, cb, ce, cl, nc);
    }
    if (rd.slot_bytes == 1) return launch_reduce_t<KT, uint8_t, false>(rd, max_erow_len, ke, Aval, Dval, st, cb, ce, cl, nc);
    if (rd.slot_bytes == 2) return launch_reduce_t<KT, uint16_t, false>(rd, max_erow_len, ke, Aval, Dval, st, cb, ce, cl, nc);
    return launch_reduce_t<KT, uint32_t, false>(rd, max_erow_len, ke, Aval, Dval, st, cb, ce, cl, nc);
}

hipError_t launch_reduce(const DevReduce& rd, int max_erow_len, bool ke_fp64, const void* ke, double* Aval, double* Dval,
                         double* rhs, bool has_matrix, bool has_rhs, hipStream_t st, const int64_t* class_begin,
                         const int64_t* class_end, const int* class_len, int n_classes) {
    if (rd.nrows == 0) return hipSuccess;
    if (has_matrix) {
        hipError_t err = ke_fp64 ? launch_reduce_k<double>(rd, max_erow_len, ke, Aval, Dval, st, class_begin, class_end, class_len, n_classes)
                                 : launch_reduce_k<float>(rd, max_erow_len, ke, Aval, Dval, st, class_begin, class_end, class_len, n_classes);
        if (err != hipSuccess) return err;
    }
    if (has_rhs) {
        int64_t grid = (rd.nrows + 255) / 256;
        if (ke_fp64)
            hipLaunchKernelGGL((reduce_rhs<double>), dim3((unsigned)grid), dim3(256), 0, st, rd, (const double*)ke, rhs);
        else
            hipLaunchKernelGGL((reduce_rhs<float>), dim3((unsigned)grid), dim3(256), 0, st, rd, (const float*)ke, rhs);
    }
    return hipGetLastError();
}

}  // namespace slx
