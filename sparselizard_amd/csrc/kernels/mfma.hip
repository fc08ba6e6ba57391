// Mixed-precision element contraction on MFMA (gfx950 v_mfma_f32_16x16x32_f16).
//
// For the matrix slices of a bucket the reference computes, per term t,
//   Ke_t[(i,j), e] = sum_q tf_t[i][q] w_q dof_t[j][q] * C_t(e,q)
// (contribution.cpp:172-211) and sums the terms (addproduct).  Stacking the
// terms' derivative pairs p = (r, s) and the Gauss points q into one K index
// k = p * n_gp + q gives one GEMM per slice:
//   Ke[m = i*n_dof + j, e] = sum_k Phi[m][k] * C[k][e]
// with Phi[m][(p,q)] = tf_r[i][q] w_q dof_s[j][q] constant per bucket.
//
// Precision: Phi and C are split into FP16 hi + lo parts with
//   x = hi + 2^-11 lo',   lo' = fp16((x - hi) * 2^11),
// Phi scaled once by a power of two, C scaled per (slice, element) column by
// a power of two so that both parts stay in the FP16 normal range.  Three
// MFMA products (hi*hi into one accumulator, hi*lo' + lo'*hi into a second)
// recover ~22 bits; accumulation is FP32; the reduction into the CSR is FP64.
// The Jacobian, its inverse and determinant stay FP64.
//
// Layout of v_mfma_f32_16x16x32_f16 operands (wave64, l = lane):
//   A (16 x 32): lane l holds A[l % 16][8 (l / 16) + 0..7]
//   B (32 x 16): lane l holds B[8 (l / 16) + 0..7][l % 16]
//   D (16 x 16): lane l holds D[4 (l / 16) + 0..3][l % 16]
#include <hip/hip_runtime.h>
#include "device.h"
#include "mfma.h"

namespace slx {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float float4v __attribute__((ext_vector_type(4)));

// Geometry and coefficient program are shared with the VALU kernels.
struct GeoM {
    double J[3][3], inv[3][3], det;
};

__device__ inline void geometry_m(const DevBucket& b, const double* xyz, int e, int q, GeoM& g) {
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
    const double(*J)[3] = g.J;
    if (b.meshdim == 1) {
        g.det = J[0][0];
        g.inv[0][0] = 1.0 / J[0][0];
    } else if (b.meshdim == 2) {
        g.det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        g.inv[0][0] = J[1][1] / g.det;
        g.inv[0][1] = -J[0][1] / g.det;
        g.inv[1][0] = -J[1][0] / g.det;
        g.inv[1][1] = J[0][0] / g.det;
    } else {
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

__device__ inline double field_value_m(const DevBucket& b, const DevInputs& in, int f, int sd, int e, int q, const GeoM& g) {
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

// FP32 coefficient program; OUT slots are routed through `route`
// (slot -> index into the tile's coefficient table, -1 = not for this kernel).
__device__ void run_program_f32(const DevBucket& b, const DevInputs& in, int e, int q, const GeoM& g, const int* route, float* Cf,
                                int cf_stride_slot_unused) {
    float reg[SLX_MAX_REGS];
    float adet = (float)fabs(g.det);
    for (int pc = 0; pc < b.n_instr; pc++) {
        slx_instr ins = b.prog[pc];
        float va = reg[ins.a], vb = reg[ins.b];
        float r = 0.f;
        switch (ins.op) {
            case SLX_OP_CONST: r = (float)in.consts[ins.imm]; break;
            case SLX_OP_PARAM: r = (float)in.params[(size_t)ins.imm * in.param_len + b.elem_ids[e]]; break;
            case SLX_OP_JAC: r = (float)g.J[ins.a][ins.b]; break;
            case SLX_OP_INVJAC: r = (float)g.inv[ins.a][ins.b]; break;
            case SLX_OP_DETJ: r = (float)g.det; break;
            case SLX_OP_COORD: {
                double s = 0.0;
                for (int k = 0; k < b.n_geo; k++) {
                    int node = b.geo_nodes[(size_t)e * b.n_geo + k];
                    s += in.xyz[3 * (size_t)node + ins.a] * b.geo_tab[((size_t)0 * b.n_geo + k) * b.n_gp + q];
                }
                r = (float)s;
                break;
            }
            case SLX_OP_FIELD: r = (float)field_value_m(b, in, ins.imm, ins.a, e, q, g); break;
            case SLX_OP_ADD: r = va + vb; break;
            case SLX_OP_MUL: r = va * vb; break;
            case SLX_OP_POW: r = powf(va, vb); break;
            case SLX_OP_INV: r = 1.f / va; break;
            case SLX_OP_ABS: r = fabsf(va); break;
            case SLX_OP_EXP: r = expf(va); break;
            case SLX_OP_LOG: r = logf(va); break;
            case SLX_OP_SIN: r = sinf(va); break;
            case SLX_OP_COS: r = cosf(va); break;
            case SLX_OP_TAN: r = tanf(va); break;
            case SLX_OP_SQRT: r = sqrtf(va); break;
            case SLX_OP_OUT: {
                int dst = route[ins.imm];
                if (dst >= 0) Cf[dst] = va * adet;
                continue;
            }
        }
        reg[ins.dst] = r;
    }
}

template <int MTW, int MAXK>
__global__ void __launch_bounds__(256) mfma_f16split_kernel(DevBucket b, MfmaBucket mb, DevInputs in, float* __restrict__ ke) {
    extern __shared__ unsigned char smem_raw[];
    const int NE = mb.ne_tile;
    const int ns = mb.n_mslices, np = mb.n_pairs, nq = mb.nq, ngp = b.n_gp;
    const int colvals = np * nq;  // coefficient values per (element, slice) column
    float* Cf = reinterpret_cast<float*>(smem_raw);                         // [NE][ns][np][nq]
    _Float16* Cs = reinterpret_cast<_Float16*>(Cf + (size_t)NE * ns * colvals);  // [NE][ns][np][nq][2]
    float* colsc = reinterpret_cast<float*>(Cs + (size_t)NE * ns * colvals * 2);  // [ns][NE]
    int* route = reinterpret_cast<int*>(colsc + (size_t)ns * NE);                 // [n_coefs]
    for (int c = threadIdx.x; c < b.n_coefs; c += blockDim.x) route[c] = mb.route[c];

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int lrow = lane & 15, lgrp = lane >> 4;
    // A fragments of this wave's M tiles (interleaved: tile = wave + 4 i)
    half8 Ahi[MTW][MAXK], Alo[MTW][MAXK];
#pragma unroll
    for (int i = 0; i < MTW; i++) {
        int mt = wave + 4 * i;
#pragma unroll
        for (int ks = 0; ks < MAXK; ks++) {
            if (mt < mb.mtiles && ks < mb.ksteps) {
                const half8* src = reinterpret_cast<const half8*>(mb.afrag) + ((size_t)(mt * mb.ksteps + ks) * 2) * 64;
                Ahi[i][ks] = src[lane];
                Alo[i][ks] = src[64 + lane];
            } else {
                for (int v = 0; v < 8; v++) {
                    Ahi[i][ks][v] = (_Float16)0.f;
                    Alo[i][ks][v] = (_Float16)0.f;
                }
            }
        }
    }
    const int ntiles_e = (b.n_elem + NE - 1) / NE;
    for (int tile = blockIdx.x; tile < ntiles_e; tile += gridDim.x) {
        const int e0 = tile * NE;
        const int nloc = min(NE, b.n_elem - e0);
        __syncthreads();
        for (int x = threadIdx.x; x < NE * ns * colvals; x += blockDim.x) Cf[x] = 0.f;
        __syncthreads();
        // phase 1a: coefficients of every (element, Gauss point) -> Cf
        for (int idx = threadIdx.x; idx < nloc * nq; idx += blockDim.x) {
            int el = idx / nq, q = idx % nq;
            GeoM g;
            geometry_m(b, in.xyz, e0 + el, q, g);
            // route[slot] = (s * np + p) * nq; add the element / Gauss point offsets
            float* base = Cf + (size_t)el * ns * colvals + q;
            run_program_f32(b, in, e0 + el, q, g, route, base, 0);
        }
        __syncthreads();
        // phase 1b: per (element, slice) power-of-two scaling and hi/lo split
        for (int idx = threadIdx.x; idx < NE * ns; idx += blockDim.x) {
            int el = idx / ns, s = idx % ns;
            const float* src = Cf + ((size_t)el * ns + s) * colvals;
            _Float16* dst = Cs + ((size_t)el * ns + s) * colvals * 2;
            float mx = 0.f;
            if (el < nloc)
                for (int v = 0; v < colvals; v++) mx = fmaxf(mx, fabsf(src[v]));
            int ex = 0;
            if (mx > 0.f) (void)frexpf(mx, &ex);  // mx in [2^(ex-1), 2^ex)
            float sc = ldexpf(1.f, -ex);          // scaled max in [0.5, 1)
            colsc[s * NE + el] = ldexpf(1.f, ex);
            for (int v = 0; v < colvals; v++) {
                float x = el < nloc ? src[v] * sc : 0.f;
                _Float16 hi = (_Float16)x;
                _Float16 lo = (_Float16)((x - (float)hi) * 2048.f);
                dst[2 * v] = hi;
                dst[2 * v + 1] = lo;
            }
        }
        __syncthreads();
        // phase 2: per slice and 16-element column tile, the wave's M tiles
        const int nct = (NE + 15) / 16;
        for (int s = 0; s < ns; s++) {
            for (int ct = 0; ct < nct; ct++) {
                const int el = ct * 16 + lrow;  // this lane's B / D column
                float4v acc_hh[MTW], acc_x[MTW];
#pragma unroll
                for (int i = 0; i < MTW; i++)
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        acc_hh[i][v] = 0.f;
                        acc_x[i][v] = 0.f;
                    }
                const _Float16* colbase = Cs + ((size_t)el * ns + s) * colvals * 2;
#pragma unroll
                for (int ks = 0; ks < MAXK; ks++) {
                    if (ks < mb.ksteps) {
                    half8 bhi, blo;
                    int k0 = ks * 32 + 8 * lgrp;
#pragma unroll
                    for (int v = 0; v < 8; v++) {
                        int k = k0 + v;
                        _Float16 h = (_Float16)0.f, l = (_Float16)0.f;
                        if (k < mb.K && el < NE) {
                            int p = k / ngp, q = k - p * ngp;
                            int off = (p * nq + (nq == 1 ? 0 : q)) * 2;
                            h = colbase[off];
                            l = colbase[off + 1];
                        }
                        bhi[v] = h;
                        blo[v] = l;
                    }
#pragma unroll
                    for (int i = 0; i < MTW; i++) {
                        acc_hh[i] = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ahi[i][ks], bhi, acc_hh[i], 0, 0, 0);
                        acc_x[i] = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ahi[i][ks], blo, acc_x[i], 0, 0, 0);
                        acc_x[i] = __builtin_amdgcn_mfma_f32_16x16x32_f16(Alo[i][ks], bhi, acc_x[i], 0, 0, 0);
                    }
                    }
                }
                if (el < nloc) {
                    const float colscale = colsc[s * NE + el] * mb.a_unscale;
                    const int64_t base = mb.ke_base[s] + (int64_t)(e0 + el) * mb.M;
#pragma unroll
                    for (int i = 0; i < MTW; i++) {
                        int mt = wave + 4 * i;
                        if (mt >= mb.mtiles) continue;
#pragma unroll
                        for (int v = 0; v < 4; v++) {
                            int row = mt * 16 + 4 * lgrp + v;
                            if (row < mb.M) ke[base + row] = (acc_hh[i][v] + acc_x[i][v] * (1.f / 2048.f)) * colscale;
                        }
                    }
                }
            }
        }
    }
}

template <int MTW, int MAXK>
static hipError_t launch_variant(const DevBucket& b, const MfmaBucket& mb, const DevInputs& in, float* ke, hipStream_t st) {
    size_t lds = mfma_lds_bytes(mb, b.n_coefs);
    int tiles = (b.n_elem + mb.ne_tile - 1) / mb.ne_tile;
    int grid = tiles < mb.max_blocks ? tiles : mb.max_blocks;
    hipLaunchKernelGGL((mfma_f16split_kernel<MTW, MAXK>), dim3(grid), dim3(256), lds, st, b, mb, in, ke);
    return hipGetLastError();
}

size_t mfma_lds_bytes(const MfmaBucket& mb, int n_coefs) {
    size_t cv = (size_t)mb.ne_tile * mb.n_mslices * mb.n_pairs * mb.nq;
    return cv * 4 + cv * 4 + (size_t)mb.n_mslices * mb.ne_tile * 4 + (size_t)n_coefs * 4;
}

bool mfma_variant_fits(int mtiles, int ksteps, int* mtw, int* maxk) {
    static const int variants[][2] = {{1, 2}, {1, 4}, {1, 8}, {2, 4}, {2, 7}, {2, 8}, {4, 4}, {4, 8}};
    int need_mtw = (mtiles + 3) / 4;
    for (auto& v : variants)
        if (v[0] >= need_mtw && v[1] >= ksteps) {
            *mtw = v[0];
            *maxk = v[1];
            return true;
        }
    return false;
}

hipError_t launch_mfma_f16split(const DevBucket& b, const MfmaBucket& mb, const DevInputs& in, float* ke, hipStream_t st) {
    if (b.n_elem == 0) return hipSuccess;
    if (mb.mtw == 1 && mb.maxk == 2) return launch_variant<1, 2>(b, mb, in, ke, st);
    if (mb.mtw == 1 && mb.maxk == 4) return launch_variant<1, 4>(b, mb, in, ke, st);
    if (mb.mtw == 1 && mb.maxk == 8) return launch_variant<1, 8>(b, mb, in, ke, st);
    if (mb.mtw == 2 && mb.maxk == 4) return launch_variant<2, 4>(b, mb, in, ke, st);
    if (mb.mtw == 2 && mb.maxk == 7) return launch_variant<2, 7>(b, mb, in, ke, st);
    if (mb.mtw == 2 && mb.maxk == 8) return launch_variant<2, 8>(b, mb, in, ke, st);
    if (mb.mtw == 4 && mb.maxk == 4) return launch_variant<4, 4>(b, mb, in, ke, st);
    if (mb.mtw == 4 && mb.maxk == 8) return launch_variant<4, 8>(b, mb, in, ke, st);
    return hipErrorInvalidValue;
}

// In-bounds probe of the 16x16x32 f16 operand layout: D = A * B for
// A[m][k] = m + 16 k / 64, B[k][n] = (k == n) + (k == n + 16), written as
// out[m * 16 + n] (256 floats).  Used by the layout self-test.
__global__ void mfma_layout_probe(float* out) {
    const int lane = threadIdx.x & 63;
    half8 a, bb;
    for (int v = 0; v < 8; v++) {
        int m = lane % 16, k = 8 * (lane / 16) + v;
        a[v] = (_Float16)(m + k / 64.f);
        int n = lane % 16;
        bb[v] = (_Float16)((k == n) || (k == n + 16) ? 1.f : 0.f);
    }
    float4v c = {0.f, 0.f, 0.f, 0.f};
    c = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, bb, c, 0, 0, 0);
    for (int v = 0; v < 4; v++) {
        int m = 4 * (lane / 16) + v, n = lane % 16;
        out[m * 16 + n] = c[v];
    }
}

hipError_t launch_mfma_probe(float* out, hipStream_t st) {
    hipLaunchKernelGGL(mfma_layout_probe, dim3(1), dim3(64), 0, st, out);
    return hipGetLastError();
}

}  // namespace slx
