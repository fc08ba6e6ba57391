// Element kernels (VALU path) and the deterministic CSR reduction.
//
// elem_kernel: one workgroup per tile of elements of a bucket.
//   phase 1  per (element, Gauss point): FP64 Jacobian, determinant and
//            inverse (jacobian.cpp:4-505), then the coefficient program
//            (operation::interpolate of every term, contribution.cpp:155-163),
//            times |det J| (contribution.cpp:147, 200) -> LDS table C.
//   phase 2  per (element, slice, tf function i, block of dof functions j):
//            Ke[i][j] = sum_q sum_t tf_t[i][q] w_q C_t(e,q) dof_t[j][q]
//            (contribution.cpp:172-246; the per-term dgemm of densemat.cpp:228-311
//            with the terms summed as in addproduct).  Right-hand-side slices
//            write -sum (contribution.cpp:296).
// reduce_matrix: one lane group per CSR row walks the row's element rows in
//   the cached accumulation order and adds each entry into its column slot
//   in LDS (FP64), then writes the A and D blocks (rawmat.cpp:67-125, 238-289).
#include <hip/hip_runtime.h>
#include "device.h"

namespace slx {

template <typename T>
__device__ inline T tabv(const double* tab, int deriv, int nff, int f, int ngp, int q) {
    return (T)tab[((size_t)deriv * nff + f) * ngp + q];
}

struct Geo {
    double J[3][3], inv[3][3], det;
};

// J[r][c] = sum_k x_k^c dN_k/dki_r (rawfield.cpp:2442-2502); det and inverse
// as jacobian.cpp:308-349, 396-505.
__device__ inline void geometry(const DevBucket& b, const double* xyz, int e, int q, Geo& g) {
    double x[SLX_MAX_GEO][3];
    for (int k = 0; k < b.n_geo; k++) {
        int node = b.geo_nodes[(size_t)e * b.n_geo + k];
        x[k][0] = xyz[3 * (size_t)node];
        x[k][1] = xyz[3 * (size_t)node + 1];
        x[k][2] = xyz[3 * (size_t)node + 2];
    }
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) g.J[r][c] = 0.0;
    for (int r = 0; r < b.elemdim; r++)
        for (int c = 0; c < b.meshdim; c++) {
            double s = 0.0;
            for (int k = 0; k < b.n_geo; k++) s += x[k][c] * b.geo_tab[((size_t)(1 + r) * b.n_geo + k) * b.n_gp + q];
            g.J[r][c] = s;
        }
    if (b.meshdim == 1) {
        g.det = g.J[0][0];
        g.inv[0][0] = 1.0 / g.J[0][0];
    } else if (b.meshdim == 2) {
        const double(*J)[3] = g.J;
        g.det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        g.inv[0][0] = J[1][1] / g.det;
        g.inv[0][1] = -J[0][1] / g.det;
        g.inv[1][0] = -J[1][0] / g.det;
        g.inv[1][1] = J[0][0] / g.det;
    } else {
        const double(*J)[3] = g.J;
        g.det = J[0][0] * (J[1][1] * J[2][2] - J[2][1] * J[1][2]) - J[0][1] * (J[1][0] * J[2][2] - J[2][0] * J[1][2]) +
                J[0][2] * (J[1][0] * J[2][1] - J[2][0] * J[1][1]);
        double d = g.det;
        g.inv[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / d;
        g.inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / d;
        g.inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / d;
        g.inv[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / d;
        g.inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / d;
        g.inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / d;
        g.inv[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / d;
        g.inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / d;
        g.inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / d;
    }
}

// Value of field f (space derivative sd) at Gauss point q of element e
// (opfield / rawfield::interpolate: coefficients x form functions, space
// derivatives through the inverse Jacobian).
__device__ inline double field_value(const DevBucket& b, const DevInputs& in, int f, int sd, int e, int q, const Geo& g) {
    const DevField& fl = b.fields[f];
    const double* data = in.field_data[f];
    double v[4] = {0, 0, 0, 0};
    int nd = sd == 0 ? 1 : 1 + b.elemdim;
    for (int k = 0; k < fl.nff; k++) {
        int a = fl.addr[(size_t)k * b.n_elem + e];
        if (a < 0) continue;
        double c = data[a];
        for (int d = 0; d < nd; d++) v[d] += c * fl.tab[((size_t)d * fl.nff + k) * b.n_gp + q];
    }
    if (sd == 0) return v[0];
    double s = 0.0;
    for (int i = 0; i < b.elemdim; i++) s += g.inv[sd - 1][i] * v[1 + i];
    return s;
}

template <typename T>
__device__ inline T apply_unary(int op, T x) {
    switch (op) {
        case SLX_OP_INV: return (T)1 / x;
        case SLX_OP_ABS: return fabs(x);
        case SLX_OP_EXP: return exp(x);
        case SLX_OP_LOG: return log(x);
        case SLX_OP_SIN: return sin(x);
        case SLX_OP_COS: return cos(x);
        case SLX_OP_TAN: return tan(x);
        case SLX_OP_SQRT: return sqrt(x);
    }
    return x;
}

// Coefficient program for one (element, Gauss point); outputs go to
// out[slot] multiplied by |det J|.
template <typename T>
__device__ void run_program(const DevBucket& b, const DevInputs& in, int e, int q, const Geo& g, T* out, int out_stride) {
    T reg[SLX_MAX_REGS];
    T adet = (T)fabs(g.det);
    for (int pc = 0; pc < b.n_instr; pc++) {
        slx_instr ins = b.prog[pc];
        T va = reg[ins.a < SLX_MAX_REGS ? ins.a : 0], vb = reg[ins.b < SLX_MAX_REGS ? ins.b : 0];
        T r = 0;
        switch (ins.op) {
            case SLX_OP_CONST: r = (T)in.consts[ins.imm]; break;
            case SLX_OP_PARAM: r = (T)in.params[(size_t)ins.imm * in.param_len + b.elem_ids[e]]; break;
            case SLX_OP_JAC: r = (T)g.J[ins.a][ins.b]; break;
            case SLX_OP_INVJAC: r = (T)g.inv[ins.a][ins.b]; break;
            case SLX_OP_DETJ: r = (T)g.det; break;
            case SLX_OP_COORD: {
                double s = 0.0;
                for (int k = 0; k < b.n_geo; k++) {
                    int node = b.geo_nodes[(size_t)e * b.n_geo + k];
                    s += in.xyz[3 * (size_t)node + ins.a] * b.geo_tab[((size_t)0 * b.n_geo + k) * b.n_gp + q];
                }
                r = (T)s;
                break;
            }
            case SLX_OP_FIELD: r = (T)field_value(b, in, ins.imm, ins.a, e, q, g); break;
            case SLX_OP_ADD: r = va + vb; break;
            case SLX_OP_MUL: r = va * vb; break;
            case SLX_OP_POW: r = pow(va, vb); break;
            case SLX_OP_OUT: out[(size_t)ins.imm * out_stride] = va * adet; continue;
            default: r = apply_unary<T>(ins.op, va); break;
        }
        reg[ins.dst] = r;
    }
}

// Work item decomposition of phase 2: (slice, element, row i, column block).
#define SLX_JBLK 16

template <typename CT, typename KT>
__global__ void __launch_bounds__(256) elem_kernel(DevBucket b, DevInputs in, int ne_tile, uint32_t slice_mask, KT* __restrict__ ke) {
    extern __shared__ unsigned char smem_raw[];
    CT* C = reinterpret_cast<CT*>(smem_raw);  // [ne_tile][nq][n_coefs]
    const int nq = b.q_inv ? 1 : b.n_gp;
    const int e0 = blockIdx.x * ne_tile;
    const int nloc = min(ne_tile, b.n_elem - e0);
    for (int idx = threadIdx.x; idx < nloc * nq; idx += blockDim.x) {
        int el = idx / nq, q = idx % nq;
        Geo g;
        geometry(b, in.xyz, e0 + el, q, g);
        run_program<CT>(b, in, e0 + el, q, g, C + ((size_t)el * nq + q) * b.n_coefs, 1);
    }
    __syncthreads();
    // phase 2
    int items_before[SLX_MAX_SLICES + 1];
    items_before[0] = 0;
    for (int s = 0; s < b.n_slices; s++) {
        int nj = (b.slices[s].n_dof + SLX_JBLK - 1) / SLX_JBLK;
        int n = ((slice_mask >> s) & 1u) ? nloc * b.slices[s].n_tf * nj : 0;
        items_before[s + 1] = items_before[s] + n;
    }
    int total = items_before[b.n_slices];
    for (int it = threadIdx.x; it < total; it += blockDim.x) {
        int s = 0;
        while (it >= items_before[s + 1]) s++;
        const DevSlice& sl = b.slices[s];
        int nj = (sl.n_dof + SLX_JBLK - 1) / SLX_JBLK;
        int local = it - items_before[s];
        int jb = local % nj;
        int rest = local / nj;
        int i = rest % sl.n_tf;
        int el = rest / sl.n_tf;
        int j0 = jb * SLX_JBLK;
        int jn = min(SLX_JBLK, sl.n_dof - j0);
        CT acc[SLX_JBLK];
        for (int jj = 0; jj < SLX_JBLK; jj++) acc[jj] = 0;
        const CT* Ce = C + (size_t)el * nq * b.n_coefs;
        for (int q = 0; q < b.n_gp; q++) {
            const CT* Cq = Ce + (size_t)(b.q_inv ? 0 : q) * b.n_coefs;
            CT w = (CT)b.gp_w[q];
            for (int t = 0; t < sl.n_terms; t++) {
                slx_term tm = sl.terms[t];
                CT c = Cq[tm.coef] * w * tabv<CT>(sl.tf_tab, tm.tf_deriv, sl.n_tf, i, b.n_gp, q);
                if (sl.is_rhs) {
                    acc[0] += c;
                } else {
                    for (int jj = 0; jj < jn; jj++) acc[jj] += c * tabv<CT>(sl.dof_tab, tm.dof_deriv, sl.n_dof, j0 + jj, b.n_gp, q);
                }
            }
        }
        size_t base = sl.ke_base + ((size_t)(e0 + el) * sl.n_tf + i) * sl.n_dof + j0;
        if (sl.is_rhs)
            ke[base] = (KT)(-acc[0]);
        else
            for (int jj = 0; jj < jn; jj++) ke[base + jj] = (KT)acc[jj];
    }
}

// Deterministic reduction of the element rows into the CSR A/D values.
template <typename KT>
__global__ void __launch_bounds__(256) reduce_matrix(DevReduce rd, const KT* __restrict__ ke, int group, double* __restrict__ Aval,
                                                      double* __restrict__ Dval) {
    extern __shared__ unsigned char smem_raw[];
    double* accall = reinterpret_cast<double*>(smem_raw);
    const int groups_per_block = blockDim.x / group;
    const int gid = threadIdx.x / group, lane = threadIdx.x % group;
    const int64_t r = (int64_t)blockIdx.x * groups_per_block + gid;
    double* acc = accall + (size_t)gid * rd.max_row_len;
    if (r < rd.nrows) {
        int64_t a0 = rd.Arow[r], d0 = rd.Drow[r];
        int lenA = (int)(rd.Arow[r + 1] - a0), lenD = (int)(rd.Drow[r + 1] - d0);
        int len = lenA + lenD;
        for (int x = lane; x < len; x += group) acc[x] = 0.0;
        // Lanes of a group belong to one wavefront and step through the element
        // rows in lockstep: within an element row all slots differ, so every
        // slot receives its contributions in element-row order.
        __builtin_amdgcn_wave_barrier();
        int64_t sp = rd.slot_ptr[r];
        for (int64_t k = rd.erow_ptr[r]; k < rd.erow_ptr[r + 1]; k++) {
            uint64_t pk = rd.erows[k];
            int n = (int)(pk >> 48);
            int64_t off = (int64_t)(pk & 0xFFFFFFFFFFFFULL);
            for (int j = lane; j < n; j += group) {
                int sl = rd.slots[sp + j];
                if (sl >= 0) acc[sl] += (double)ke[off + j];
            }
            sp += n;
            __builtin_amdgcn_wave_barrier();
        }
        for (int x = lane; x < lenA; x += group) Aval[a0 + x] = acc[x];
        for (int x = lane; x < lenD; x += group) Dval[d0 + x] = acc[lenA + x];
    }
}

// rhs: b[dof] = sum of the element-row values in accumulation order
// (rawvec::setvalues "add", rawvec.cpp:216-254); constrained entries are 0.
template <typename KT>
__global__ void reduce_rhs(DevReduce rd, const KT* __restrict__ ke, double* __restrict__ rhs) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rd.nrows) return;
    double s = 0.0;
    for (int64_t k = rd.rhs_erow_ptr[r]; k < rd.rhs_erow_ptr[r + 1]; k++) {
        uint64_t pk = rd.rhs_erows[k];
        s += (double)ke[(int64_t)(pk & 0xFFFFFFFFFFFFULL)];
    }
    rhs[rd.free_dofs[r]] = s;
}

// FP64 element kernel, factorized and register-blocked ("lanes = elements").
// A workgroup takes 64 elements; every wave walks work combos (slice, row
// tile, column tile) that are uniform across its lanes, so the shape-function
// tables are read through the scalar unit while each lane owns one element:
//   H_r[j] = w_q sum_{t: r_t = r} C_t(e,q) dof_{s_t}[j][q]
//   Ke[i][j] += sum_r tf_r[i][q] H_r[j]
// which is the reference's sum over terms and Gauss points regrouped
// (B^T D B with D gathered per tf derivative).
template <int TI, int TJ>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) elem64_tiled(DevBucket b, DevInputs in, uint32_t slice_mask, double* __restrict__ ke) {
    extern __shared__ unsigned char smem_raw[];
    double* C = reinterpret_cast<double*>(smem_raw);  // [64][nq][n_coefs]
    const int NE = 64;
    const int nq = b.q_inv ? 1 : b.n_gp;
    const int e0 = blockIdx.x * NE;
    const int nloc = min(NE, b.n_elem - e0);
    for (int idx = threadIdx.x; idx < nloc * nq; idx += blockDim.x) {
        int el = idx % nloc, q = idx / nloc;
        Geo g;
        geometry(b, in.xyz, e0 + el, q, g);
        run_program<double>(b, in, e0 + el, q, g, C + ((size_t)el * nq + q) * b.n_coefs, 1);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // combos: for each masked slice, row tiles x column tiles
    int ncombo = 0;
    for (int s = 0; s < b.n_slices; s++)
        if ((slice_mask >> s) & 1u) ncombo += ((b.slices[s].n_tf + TI - 1) / TI) * ((b.slices[s].n_dof + TJ - 1) / TJ);
    for (int c = wave; c < ncombo; c += 4) {
        // decode the combo (uniform)
        int s = 0, rem = c;
        for (; s < b.n_slices; s++) {
            if (!((slice_mask >> s) & 1u)) continue;
            int n = ((b.slices[s].n_tf + TI - 1) / TI) * ((b.slices[s].n_dof + TJ - 1) / TJ);
            if (rem < n) break;
            rem -= n;
        }
        const DevSlice& sl = b.slices[s];
        const int njt = (sl.n_dof + TJ - 1) / TJ;
        const int i0 = (rem / njt) * TI, j0 = (rem % njt) * TJ;
        const int ni = min(TI, sl.n_tf - i0), nj = min(TJ, sl.n_dof - j0);
        double acc[TI][TJ];
#pragma unroll
        for (int ii = 0; ii < TI; ii++)
#pragma unroll
            for (int jj = 0; jj < TJ; jj++) acc[ii][jj] = 0.0;
        const int el = lane;
        const double* Ce = C + (size_t)min(el, NE - 1) * nq * b.n_coefs;
        for (int q = 0; q < b.n_gp; q++) {
            const double* Cq = Ce + (size_t)(b.q_inv ? 0 : q) * b.n_coefs;
            const double w = b.gp_w[q];
            double h[4][TJ];
#pragma unroll
            for (int r = 0; r < 4; r++)
#pragma unroll
                for (int jj = 0; jj < TJ; jj++) h[r][jj] = 0.0;
            unsigned rmask = 0;
            for (int t = 0; t < sl.n_terms; t++) {
                const slx_term tm = sl.terms[t];
                const double cw = Cq[tm.coef] * w;
                rmask |= 1u << tm.tf_deriv;
                if (sl.is_rhs) {
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        if (r == tm.tf_deriv) h[r][0] += cw;
                } else {
                    const double* dt = sl.dof_tab + ((size_t)tm.dof_deriv * sl.n_dof + j0) * b.n_gp + q;
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        if (r == tm.tf_deriv)
#pragma unroll
                            for (int jj = 0; jj < TJ; jj++) h[r][jj] += jj < nj ? cw * dt[(size_t)jj * b.n_gp] : 0.0;
                }
            }
#pragma unroll
            for (int r = 0; r < 4; r++) {
                if (!((rmask >> r) & 1u)) continue;
                const double* tt = sl.tf_tab + ((size_t)r * sl.n_tf + i0) * b.n_gp + q;
#pragma unroll
                for (int ii = 0; ii < TI; ii++) {
                    const double tv = ii < ni ? tt[(size_t)ii * b.n_gp] : 0.0;
#pragma unroll
                    for (int jj = 0; jj < TJ; jj++) acc[ii][jj] += tv * h[r][jj];
                }
            }
        }
        if (el < nloc) {
            const size_t base = sl.ke_base + ((size_t)(e0 + el) * sl.n_tf + i0) * sl.n_dof + j0;
#pragma unroll
            for (int ii = 0; ii < TI; ii++)
#pragma unroll
                for (int jj = 0; jj < TJ; jj++)
                    if (ii < ni && jj < nj) ke[base + (size_t)ii * sl.n_dof + jj] = sl.is_rhs ? -acc[ii][jj] : acc[ii][jj];
        }
    }
}

// ---------------------------------------------------------------- launchers --

int elem_tile_size(const DevBucket& b, size_t elem_bytes, size_t lds_budget) {
    int nq = b.q_inv ? 1 : b.n_gp;
    size_t per = (size_t)nq * b.n_coefs * elem_bytes;
    int t = (int)(lds_budget / (per ? per : 1));
    if (t > 64) t = 64;
    if (t < 1) t = 1;
    return t;
}

template <int T>
static hipError_t launch_tiled64(const DevBucket& b, const DevInputs& in, uint32_t mask, double* ke, hipStream_t st, size_t lds) {
    int grid = (b.n_elem + 63) / 64;
    hipLaunchKernelGGL((elem64_tiled<T, T>), dim3(grid), dim3(256), lds, st, b, in, mask, ke);
    return hipGetLastError();
}

hipError_t launch_elem(const DevBucket& b, const DevInputs& in, int precision, void* ke, hipStream_t st, uint32_t slice_mask) {
    if (b.n_elem == 0 || slice_mask == 0) return hipSuccess;
    if (precision == SLX_PREC_FP64) {
        // lanes-as-elements kernel when the 64-element coefficient table fits
        int nq = b.q_inv ? 1 : b.n_gp;
        size_t lds64 = (size_t)64 * nq * b.n_coefs * sizeof(double);
        int nmax = 0;
        for (int s = 0; s < b.n_slices; s++)
            if ((slice_mask >> s) & 1u) nmax = max(nmax, max(b.slices[s].n_tf, b.slices[s].n_dof));
        if (lds64 <= 48 * 1024 && b.n_coefs > 0) {
            if (nmax <= 4) return launch_tiled64<4>(b, in, slice_mask, (double*)ke, st, lds64);
            if (nmax <= 6) return launch_tiled64<6>(b, in, slice_mask, (double*)ke, st, lds64);
            if (nmax <= 10) return launch_tiled64<10>(b, in, slice_mask, (double*)ke, st, lds64);
            return launch_tiled64<8>(b, in, slice_mask, (double*)ke, st, lds64);
        }
        int tile = elem_tile_size(b, sizeof(double), 48 * 1024);
        size_t lds = (size_t)tile * nq * b.n_coefs * sizeof(double);
        if (lds > 64 * 1024) return hipErrorInvalidValue;
        int grid = (b.n_elem + tile - 1) / tile;
        hipLaunchKernelGGL((elem_kernel<double, double>), dim3(grid), dim3(256), lds, st, b, in, tile, slice_mask, (double*)ke);
    } else {
        int tile = elem_tile_size(b, sizeof(float), 48 * 1024);
        int nq = b.q_inv ? 1 : b.n_gp;
        size_t lds = (size_t)tile * nq * b.n_coefs * sizeof(float);
        if (lds > 64 * 1024) return hipErrorInvalidValue;
        int grid = (b.n_elem + tile - 1) / tile;
        hipLaunchKernelGGL((elem_kernel<float, float>), dim3(grid), dim3(256), lds, st, b, in, tile, slice_mask, (float*)ke);
    }
    return hipGetLastError();
}

int reduce_group_size(int max_erow_len) {
    int g = 1;
    while (g < max_erow_len && g < 64) g *= 2;
    return g;
}

hipError_t launch_reduce(const DevReduce& rd, int max_erow_len, bool ke_fp64, const void* ke, double* Aval, double* Dval,
                         double* rhs, bool has_matrix, bool has_rhs, hipStream_t st) {
    if (rd.nrows == 0) return hipSuccess;
    if (has_matrix) {
        int g = reduce_group_size(max_erow_len);
        int threads = 256, gpb = threads / g;
        size_t lds = (size_t)gpb * rd.max_row_len * sizeof(double);
        while (lds > 64 * 1024 && threads > g) {
            threads /= 2;
            gpb = threads / g;
            lds = (size_t)gpb * rd.max_row_len * sizeof(double);
        }
        if (lds > 64 * 1024) return hipErrorInvalidValue;
        int64_t grid = (rd.nrows + gpb - 1) / gpb;
        if (ke_fp64)
            hipLaunchKernelGGL((reduce_matrix<double>), dim3((unsigned)grid), dim3(threads), lds, st, rd, (const double*)ke, g, Aval, Dval);
        else
            hipLaunchKernelGGL((reduce_matrix<float>), dim3((unsigned)grid), dim3(threads), lds, st, rd, (const float*)ke, g, Aval, Dval);
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
