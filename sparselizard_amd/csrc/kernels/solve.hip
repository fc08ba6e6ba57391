// Device solve of an assembled system (kernels).
//
// x_A = A^-1 (b_A - D x_D) with x_D the Dirichlet values (mat::eliminate,
// src/formulation/mat.cpp:78-103; the reference then factorizes with MUMPS,
// src/expression/operation/sl.cpp:1565-1627).  Here: FP64 iterative
// refinement (FP64 CSR residuals) around a Jacobi-preconditioned FP32
// conjugate-gradient inner solve, for symmetric positive definite A.
//
// One inner iteration is three streaming kernels and two one-block sums:
//   K1  q = A p (FP32 values),             partials of p.q
//   F   pq = ordered sum of the partials
//   K2  d += a p, r -= a q, z = r / diag,  partials of r.z and r.r  (a = rz / pq)
//   F   rz' and rr, beta = rz' / rz
//   K3  p = z + beta p
// Every reduction is deterministic: a fixed grid of partial sums, summed in a
// fixed order by one block, and fixed lanes per row for the products.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <vector>
#include "solve.h"

namespace slx {

enum { SC_RZ = 0, SC_PQ = 1, SC_RZN = 2, SC_RR = 3, SC_BETA = 4, SC_DOT = 5 };

// block sum of one double per thread (256 threads); result valid in thread 0
__device__ inline double block_sum(double v, double* sh) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) s = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    __syncthreads();
    return s;
}

// y[r] = c0 f[r] - c1 (A x)[r] over one CSR matrix, LPR lanes per row (the
// lanes of a row group share r, so the width-LPR shuffles stay inside
// active lanes); with part != null the block also writes its partial sum of
// x[r] y[r] (square A)
template <int LPR, typename VT, typename XT, typename YT>
__global__ void __launch_bounds__(256) spmv_kernel(int64_t nrows, const int64_t* __restrict__ rowp, const int32_t* __restrict__ col,
                                                   const VT* __restrict__ val, const XT* __restrict__ x, const YT* f, YT c0, YT c1,
                                                   YT* y, double* __restrict__ part) {
    __shared__ double sh[4];
    const int l = threadIdx.x % LPR;
    const int64_t rstride = (int64_t)gridDim.x * (256 / LPR);
    double acc = 0.0;
    for (int64_t r = (int64_t)blockIdx.x * (256 / LPR) + threadIdx.x / LPR; r < nrows; r += rstride) {
        YT s = 0;
        const int64_t e = rowp[r + 1];
        for (int64_t k = rowp[r] + l; k < e; k += LPR) s += (YT)val[k] * (YT)x[col[k]];
#pragma unroll
        for (int o = LPR / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, LPR);
        if (l == 0) {
            const YT v = (f ? c0 * f[r] : (YT)0) - c1 * s;
            y[r] = v;
            if (part) acc += (double)x[r] * (double)v;
        }
    }
    if (part) {
        const double t = block_sum(acc, sh);
        if (threadIdx.x == 0) part[blockIdx.x] = t;
    }
}

// lanes per row from the mean row length (a few products per lane)
static int spmv_lpr(int64_t nrows, int64_t nnz) {
    const double avg = nrows ? (double)nnz / (double)nrows : 0.0;
    if (avg >= 96) return 32;
    if (avg >= 40) return 16;
    if (avg >= 16) return 8;
    if (avg >= 6) return 4;
    return 2;
}

template <typename VT, typename XT, typename YT>
static void spmv(int64_t nrows, int64_t nnz, const int64_t* rowp, const int32_t* col, const VT* val, const XT* x, const YT* f, YT c0,
                 YT c1, YT* y, double* part, hipStream_t st) {
    if (nrows <= 0) return;
    const int lpr = spmv_lpr(nrows, nnz);
    // the fixed grid when partial sums are wanted, else one row group per thread group
    const int64_t want = (nrows * lpr + 255) / 256;
    const unsigned grid = part ? SLX_SOLVE_BLOCKS : (unsigned)std::min<int64_t>(want, 1 << 20);
#define SLX_SPMV_CASE(L)                                                                                                          \
    case L:                                                                                                                       \
        hipLaunchKernelGGL((spmv_kernel<L, VT, XT, YT>), dim3(grid), dim3(256), 0, st, nrows, rowp, col, val, x, f, c0, c1, y, part); \
        break;
    switch (lpr) {
        SLX_SPMV_CASE(2)
        SLX_SPMV_CASE(4)
        SLX_SPMV_CASE(8)
        SLX_SPMV_CASE(16)
        SLX_SPMV_CASE(32)
    }
#undef SLX_SPMV_CASE
}

// q = A p over FP32 values, streamed by row blocks: a block loads the
// nonzeros of its rows with coalesced loads, keeps the products in LDS
// and reduces each row from there (lanes per row from the rows in the block);
// a row longer than SLX_STREAM_NNZ is summed by the whole block.  The fixed
// grid walks the row blocks; each block writes its partial sum of p.q.
// Workgroups are dispatched round-robin over the 8 XCDs (workgroup w on XCD
// w % 8): each XCD takes one contiguous eighth of the row blocks, so the
// entries of x its rows gather stay within a range its own L2 can hold.
__global__ void __launch_bounds__(256) spmv_stream_f32(int64_t nblk, const int64_t* __restrict__ rblk, const int64_t* __restrict__ rowp,
                                                       const int32_t* __restrict__ col, const float* __restrict__ val, int64_t nnz,
                                                       const float* __restrict__ x, float* __restrict__ y, double* __restrict__ part) {
    __shared__ float prod[SLX_STREAM_NNZ];
    __shared__ double sh[4];
    double acc = 0.0;
    const int xcd = blockIdx.x % SLX_XCDS;
    const int64_t b_end = nblk * (xcd + 1) / SLX_XCDS;
    for (int64_t b = nblk * xcd / SLX_XCDS + blockIdx.x / SLX_XCDS; b < b_end; b += gridDim.x / SLX_XCDS) {
        const int64_t r0 = rblk[b], r1 = rblk[b + 1];
        const int64_t k0 = rowp[r0], k1 = rowp[r1];
        if (k1 - k0 > SLX_STREAM_NNZ) {
            // one long row (block-uniform branch)
            double s = 0.0;
            for (int64_t k = k0 + threadIdx.x; k < k1; k += 256) s += (double)(val[k] * x[col[k]]);
            const double t = block_sum(s, sh);
            if (threadIdx.x == 0) {
                y[r0] = (float)t;
                acc += (double)x[r0] * (double)(float)t;
            }
            continue;
        }
        // consecutive lanes take consecutive nonzeros: the x entries one
        // wave gathers then share cache lines more often than with 16-byte
        // loads per lane (measured: 1.54 vs 1.75 ms per product, 64^3 P2 elasticity)
        const int64_t kb = k0;
        for (int64_t k = k0 + threadIdx.x; k < k1; k += 256) prod[k - k0] = val[k] * x[col[k]];
        __syncthreads();
        const int rows = (int)(r1 - r0);
        int lpr = 32;
        while (lpr > 1 && rows * lpr > 256) lpr >>= 1;
        const int g = threadIdx.x / lpr, l = threadIdx.x % lpr;
        for (int rr = g; rr < rows; rr += 256 / lpr) {
            const int a = (int)(rowp[r0 + rr] - kb), e = (int)(rowp[r0 + rr + 1] - kb);
            float s = 0.f;
            for (int k = a + l; k < e; k += lpr) s += prod[k];
            for (int o = lpr / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, lpr);
            if (l == 0) {
                y[r0 + rr] = s;
                acc += (double)x[r0 + rr] * (double)s;
            }
        }
        __syncthreads();
    }
    const double t = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

void stream_row_blocks(const int64_t* rowp, int64_t nrows, std::vector<int64_t>& out) {
    out.clear();
    out.push_back(0);
    int64_t r = 0;
    while (r < nrows) {
        int64_t e = r;
        while (e < nrows && e - r < 256 && rowp[e + 1] - rowp[r] <= SLX_STREAM_NNZ) e++;
        if (e == r) e = r + 1;
        out.push_back(e);
        r = e;
    }
}

__global__ void gather_kernel(int64_t n, const int64_t* __restrict__ idx, const double* __restrict__ src, double* __restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

__global__ void scatter_kernel(int64_t n, const int64_t* __restrict__ idx, const double* __restrict__ src, double* __restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[idx[i]] = src[i];
}

// Jacobi preconditioner (FP32) and the FP32 copy of the matrix values
__global__ void prep_kernel(int64_t nrows, const int64_t* __restrict__ rowp, const int32_t* __restrict__ col, const double* __restrict__ val,
                            float* __restrict__ val32, float* __restrict__ dinv) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    double d = 0.0;
    for (int64_t k = rowp[r]; k < rowp[r + 1]; k++) {
        val32[k] = (float)val[k];
        if (col[k] == r) d = val[k];
    }
    dinv[r] = d != 0.0 ? (float)(1.0 / d) : 1.0f;
}

// partials of a.b (FP64 accumulation) over the fixed grid
template <typename T>
__global__ void __launch_bounds__(256) dot_partial(int64_t n, const T* __restrict__ a, const T* __restrict__ b, double* __restrict__ part) {
    __shared__ double sh[4];
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) s += (double)a[i] * (double)b[i];
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// sc[out0] (and sc[out1]) = ordered sums of one (two) partial arrays; mode 1
// also sets beta = sc[SC_RZN] / sc[SC_RZ] and then sc[SC_RZ] = sc[SC_RZN]
__global__ void __launch_bounds__(256) sum_partials(const double* __restrict__ part, int nsum, int out0, int out1, int mode,
                                                    double* __restrict__ sc) {
    __shared__ double sh[4];
    double t[2] = {0.0, 0.0};
    for (int j = 0; j < nsum; j++) {
        double s = 0.0;
        for (int i = threadIdx.x; i < SLX_SOLVE_BLOCKS; i += 256) s += part[j * SLX_SOLVE_BLOCKS + i];
        t[j] = block_sum(s, sh);
    }
    if (threadIdx.x == 0) {
        sc[out0] = t[0];
        if (nsum > 1) sc[out1] = t[1];
        if (mode == 1) {
            sc[SC_BETA] = sc[SC_RZ] != 0.0 ? sc[SC_RZN] / sc[SC_RZ] : 0.0;
            sc[SC_RZ] = sc[SC_RZN];
        }
    }
}

// inner start: r32 = r, z = r / diag, p = z, d = 0; partials of r.z and r.r
__global__ void __launch_bounds__(256) pcg_start(int64_t n, const double* __restrict__ r, const float* __restrict__ dinv,
                                                 float* __restrict__ r32, float* __restrict__ z, float* __restrict__ p,
                                                 float* __restrict__ d, double* __restrict__ part) {
    __shared__ double sh[4];
    double rz = 0.0, rr = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const float rv = (float)r[i];
        const float zv = dinv[i] * rv;
        r32[i] = rv;
        z[i] = zv;
        p[i] = zv;
        d[i] = 0.f;
        rz += (double)rv * (double)zv;
        rr += (double)rv * (double)rv;
    }
    const double a = block_sum(rz, sh);
    const double b = block_sum(rr, sh);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = a;
        part[SLX_SOLVE_BLOCKS + blockIdx.x] = b;
    }
}

// d += a p ; r -= a q ; z = r / diag   (a = rz / pq); partials of r.z and r.r
__global__ void __launch_bounds__(256) pcg_update(int64_t n, const double* __restrict__ sc, float* __restrict__ d, float* __restrict__ r,
                                                  const float* __restrict__ p, const float* __restrict__ q, const float* __restrict__ dinv,
                                                  float* __restrict__ z, double* __restrict__ part) {
    __shared__ double sh[4];
    const float alpha = sc[SC_PQ] != 0.0 ? (float)(sc[SC_RZ] / sc[SC_PQ]) : 0.f;
    double rz = 0.0, rr = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        d[i] += alpha * p[i];
        const float rv = r[i] - alpha * q[i];
        const float zv = dinv[i] * rv;
        r[i] = rv;
        z[i] = zv;
        rz += (double)rv * (double)zv;
        rr += (double)rv * (double)rv;
    }
    const double a = block_sum(rz, sh);
    const double b = block_sum(rr, sh);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = a;
        part[SLX_SOLVE_BLOCKS + blockIdx.x] = b;
    }
}

// p = z + beta p
__global__ void __launch_bounds__(256) pcg_direction(int64_t n, const double* __restrict__ sc, const float* __restrict__ z,
                                                     float* __restrict__ p) {
    const float beta = (float)sc[SC_BETA];
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) p[i] = z[i] + beta * p[i];
}

__global__ void to_f32(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

__global__ void add_f32_to_f64(int64_t n, const float* __restrict__ d, double* __restrict__ x) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] += (double)d[i];
}

static unsigned blocks(int64_t n) { return (unsigned)((n + 255) / 256); }

// ---- GMRES (FP32 vectors, FP64 sums) ----

// z = dinv .* v
__global__ void precond_kernel(int64_t n, const float* __restrict__ dinv, const float* __restrict__ v, float* __restrict__ z) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) z[i] = dinv[i] * v[i];
}

// dst = alpha * src
__global__ void scale_kernel(int64_t n, float alpha, const float* __restrict__ src, float* __restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = alpha * src[i];
}

// partials of V_i . w for i < nv (basis rows of length n)
__global__ void __launch_bounds__(256) multi_dot_partial(int64_t n, const float* __restrict__ V, int nv, const float* __restrict__ w,
                                                         double* __restrict__ part) {
    __shared__ double sh[4];
    for (int i = 0; i < nv; i++) {
        const float* v = V + (size_t)i * n;
        double s = 0.0;
        for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (int64_t)gridDim.x * 256) s += (double)v[k] * (double)w[k];
        const double t = block_sum(s, sh);
        if (threadIdx.x == 0) part[(size_t)i * SLX_SOLVE_BLOCKS + blockIdx.x] = t;
    }
}

// out[i] = ordered sum of partial row i (one block per row)
__global__ void __launch_bounds__(256) sum_rows(const double* __restrict__ part, double* __restrict__ out) {
    __shared__ double sh[4];
    double s = 0.0;
    for (int i = threadIdx.x; i < SLX_SOLVE_BLOCKS; i += 256) s += part[(size_t)blockIdx.x * SLX_SOLVE_BLOCKS + i];
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = t;
}

// w -= sum_i h_i V_i
__global__ void basis_update(int64_t n, const float* __restrict__ V, int nv, const double* __restrict__ h, float* __restrict__ w) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    float acc = w[k];
    for (int i = 0; i < nv; i++) acc -= (float)h[i] * V[(size_t)i * n + k];
    w[k] = acc;
}

// d = dinv .* sum_i y_i V_i   (right preconditioning)
__global__ void basis_combine(int64_t n, const float* __restrict__ V, int nv, const double* __restrict__ y, const float* __restrict__ dinv,
                              float* __restrict__ d) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    float acc = 0.f;
    for (int i = 0; i < nv; i++) acc += (float)y[i] * V[(size_t)i * n + k];
    d[k] = dinv[k] * acc;
}

// position of column c in CSR row r (sorted columns), or -1
__device__ inline int64_t find_col(const int64_t* __restrict__ rowp, const int32_t* __restrict__ col, int64_t r, int64_t c) {
    int64_t lo = rowp[r], hi = rowp[r + 1];
    while (lo < hi) {
        const int64_t mid = (lo + hi) / 2;
        if (col[mid] < c) lo = mid + 1;
        else hi = mid;
    }
    return (lo < rowp[r + 1] && col[lo] == c) ? lo : -1;
}

// any entry a_rc whose transpose partner a_cr is missing or differs by more
// than 1e-5 (|a_rc| + |a_cr|) + 1e-6 (|a_rr| + |a_cc|) sets *flag (every
// writer stores 1); the diagonal term keeps rounding noise (FP64, or the
// 1e-7 of the mixed-precision element matrices) on entries that cancel to
// almost zero from counting as asymmetry.  SLX_SYM_LPR lanes share a row
// (one entry each per step), so the binary searches of a row run side by side.
#define SLX_SYM_LPR 16
__global__ void __launch_bounds__(256) sym_kernel(int64_t n, const int64_t* __restrict__ rowp, const int32_t* __restrict__ col,
                                                  const double* __restrict__ val, int* __restrict__ flag) {
    const int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / SLX_SYM_LPR;
    const int l = threadIdx.x % SLX_SYM_LPR;
    if (r >= n) return;
    const int64_t kd = find_col(rowp, col, r, r);
    const double drr = kd >= 0 ? fabs(val[kd]) : 0.0;
    bool bad = false;
    for (int64_t k = rowp[r] + l; k < rowp[r + 1] && !bad; k += SLX_SYM_LPR) {
        const int32_t c = col[k];
        const int64_t kt = find_col(rowp, col, c, r);
        if (kt < 0) {
            bad = true;
        } else {
            const int64_t kc = find_col(rowp, col, c, c);
            const double dcc = kc >= 0 ? fabs(val[kc]) : 0.0;
            bad = fabs(val[k] - val[kt]) > 1e-5 * (fabs(val[k]) + fabs(val[kt])) + 1e-6 * (drr + dcc);
        }
    }
    if (bad) flag[0] = 1;
}

hipError_t check_symmetric(const SolveSystem& s, int* flag_dev, int* nonsym, hipStream_t st) {
    HIPERR(hipMemsetAsync(flag_dev, 0, sizeof(int), st));
    if (s.nfree > 0)
        hipLaunchKernelGGL(sym_kernel, dim3(blocks(s.nfree * SLX_SYM_LPR)), dim3(256), 0, st, s.nfree, s.Arow, s.Acol, s.Aval, flag_dev);
    HIPERR(hipMemcpyAsync(nonsym, flag_dev, sizeof(int), hipMemcpyDeviceToHost, st));
    return hipStreamSynchronize(st);
}

static hipError_t read_scalar(const double* dsc, int idx, double* h, hipStream_t st);

// One cycle of right-preconditioned GMRES on A d = r32 (FP32 vectors, FP64
// sums, classical Gram-Schmidt applied twice), at most min(restart,
// max_inner) iterations or until the FP32 residual estimate falls below
// inner_tol of its start.  d = 0 when r32 = 0.
static hipError_t gmres_cycle(const SolveSystem& s, const SolveParams& prm, SolveWork& w, int& iters, hipStream_t st) {
    const int64_t n = s.nfree;
    const unsigned G = SLX_SOLVE_BLOCKS;
    const int m = std::max(1, std::min(prm.restart, prm.max_inner));
    double beta2 = 0.0;
    hipLaunchKernelGGL((dot_partial<float>), dim3(G), dim3(256), 0, st, n, w.r32, w.r32, w.part);
    hipLaunchKernelGGL(sum_partials, dim3(1), dim3(256), 0, st, w.part, 1, (int)SC_DOT, 0, 0, w.sc);
    HIPERR(read_scalar(w.sc, SC_DOT, &beta2, st));
    const double beta = sqrt(beta2);
    if (!(beta > 0.0)) {
        HIPERR(hipMemsetAsync(w.d, 0, n * sizeof(float), st));
        return hipSuccess;
    }
    hipLaunchKernelGGL(scale_kernel, dim3(blocks(n)), dim3(256), 0, st, n, (float)(1.0 / beta), w.r32, w.V);
    std::vector<double> H((size_t)(m + 1) * m, 0.0), cs(m, 0.0), sn(m, 0.0), g(m + 1, 0.0), hc(m + 1);
    g[0] = beta;
    int k = 0;
    for (int j = 0; j < m; j++) {
        const float* vj = w.V + (size_t)j * n;
        hipLaunchKernelGGL(precond_kernel, dim3(blocks(n)), dim3(256), 0, st, n, w.dinv, vj, w.z);
        hipLaunchKernelGGL(spmv_stream_f32, dim3(G), dim3(256), 0, st, s.nblk, s.Ablk, s.Arow, s.Acol, w.val32, s.nnzA, w.z, w.q,
                           w.part);
        // two passes of classical Gram-Schmidt against v_0 .. v_j
        for (int pass = 0; pass < 2; pass++) {
            hipLaunchKernelGGL(multi_dot_partial, dim3(G), dim3(256), 0, st, n, w.V, j + 1, w.q, w.mpart);
            hipLaunchKernelGGL(sum_rows, dim3(j + 1), dim3(256), 0, st, w.mpart, w.hv);
            hipLaunchKernelGGL(basis_update, dim3(blocks(n)), dim3(256), 0, st, n, w.V, j + 1, w.hv, w.q);
            HIPERR(hipMemcpyAsync(hc.data(), w.hv, (size_t)(j + 1) * sizeof(double), hipMemcpyDeviceToHost, st));
            HIPERR(hipStreamSynchronize(st));
            for (int i = 0; i <= j; i++) H[(size_t)i * m + j] += hc[i];
        }
        double hn2 = 0.0;
        hipLaunchKernelGGL((dot_partial<float>), dim3(G), dim3(256), 0, st, n, w.q, w.q, w.part);
        hipLaunchKernelGGL(sum_partials, dim3(1), dim3(256), 0, st, w.part, 1, (int)SC_DOT, 0, 0, w.sc);
        HIPERR(read_scalar(w.sc, SC_DOT, &hn2, st));
        const double hn = sqrt(hn2);
        H[(size_t)(j + 1) * m + j] = hn;
        if (hn > 0.0 && j + 1 <= m)
            hipLaunchKernelGGL(scale_kernel, dim3(blocks(n)), dim3(256), 0, st, n, (float)(1.0 / hn), w.q, w.V + (size_t)(j + 1) * n);
        // Givens rotations of column j
        for (int i = 0; i < j; i++) {
            const double a = H[(size_t)i * m + j], b = H[(size_t)(i + 1) * m + j];
            H[(size_t)i * m + j] = cs[i] * a + sn[i] * b;
            H[(size_t)(i + 1) * m + j] = -sn[i] * a + cs[i] * b;
        }
        const double a = H[(size_t)j * m + j], b = H[(size_t)(j + 1) * m + j];
        const double rr = std::hypot(a, b);
        cs[j] = rr > 0.0 ? a / rr : 1.0;
        sn[j] = rr > 0.0 ? b / rr : 0.0;
        H[(size_t)j * m + j] = rr;
        H[(size_t)(j + 1) * m + j] = 0.0;
        g[j + 1] = -sn[j] * g[j];
        g[j] = cs[j] * g[j];
        k = j + 1;
        iters++;
        if (fabs(g[j + 1]) <= prm.inner_tol * beta || !(hn > 0.0)) break;
    }
    // y = H(0:k, 0:k)^-1 g(0:k)
    std::vector<double> y(k, 0.0);
    for (int i = k - 1; i >= 0; i--) {
        double t = g[i];
        for (int c = i + 1; c < k; c++) t -= H[(size_t)i * m + c] * y[c];
        y[i] = H[(size_t)i * m + i] != 0.0 ? t / H[(size_t)i * m + i] : 0.0;
    }
    HIPERR(hipMemcpyAsync(w.hv, y.data(), (size_t)k * sizeof(double), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(basis_combine, dim3(blocks(n)), dim3(256), 0, st, n, w.V, k, w.hv, w.dinv, w.d);
    return hipStreamSynchronize(st);
}

static hipError_t read_scalar(const double* dsc, int idx, double* h, hipStream_t st) {
    HIPERR(hipMemcpyAsync(h, dsc + idx, sizeof(double), hipMemcpyDeviceToHost, st));
    return hipStreamSynchronize(st);
}

hipError_t eliminate_system(const SolveSystem& s, const double* b, double* out, double* xd, hipStream_t st) {
    const int64_t n = s.nfree;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(gather_kernel, dim3(blocks(n)), dim3(256), 0, st, n, s.free_dofs, b, out);
    if (s.ncons > 0) {
        hipLaunchKernelGGL(gather_kernel, dim3(blocks(s.ncons)), dim3(256), 0, st, s.ncons, s.cons_dofs, b, xd);
        spmv<double, double, double>(n, s.nnzD, s.Drow, s.Dcol, s.Dval, xd, out, 1.0, 1.0, out, nullptr, st);
    }
    return hipGetLastError();
}

hipError_t matvec_system(const SolveSystem& s, const double* x, double* y, double* xa, double* xd, double* ya, hipStream_t st) {
    const int64_t n = s.nfree;
    HIPERR(hipMemcpyAsync(y, x, s.n_dofs * sizeof(double), hipMemcpyDeviceToDevice, st));
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(gather_kernel, dim3(blocks(n)), dim3(256), 0, st, n, s.free_dofs, x, xa);
    spmv<double, double, double>(n, s.nnzA, s.Arow, s.Acol, s.Aval, xa, (const double*)nullptr, 0.0, -1.0, ya, nullptr, st);
    if (s.ncons > 0) {
        hipLaunchKernelGGL(gather_kernel, dim3(blocks(s.ncons)), dim3(256), 0, st, s.ncons, s.cons_dofs, x, xd);
        spmv<double, double, double>(n, s.nnzD, s.Drow, s.Dcol, s.Dval, xd, ya, 1.0, -1.0, ya, nullptr, st);
    }
    hipLaunchKernelGGL(scatter_kernel, dim3(blocks(n)), dim3(256), 0, st, n, s.free_dofs, ya, y);
    return hipGetLastError();
}

hipError_t spmv_stream_probe(int64_t nrows, const int64_t* rowp, const int32_t* col, const float* val, int64_t nnz, const int64_t* blk,
                             int64_t nblk, const float* x, float* y, double* part, double* sc, hipStream_t st) {
    hipLaunchKernelGGL(spmv_stream_f32, dim3(SLX_SOLVE_BLOCKS), dim3(256), 0, st, nblk, blk, rowp, col, val, nnz, x, y, part);
    hipLaunchKernelGGL(sum_partials, dim3(1), dim3(256), 0, st, part, 1, (int)SC_PQ, 0, 0, sc);
    return hipGetLastError();
}

hipError_t solve_system(const SolveSystem& s, const SolveParams& prm, SolveWork& w, SolveStats& out, hipStream_t st) {
    const int64_t n = s.nfree;
    out = SolveStats{};
    if (n == 0) return hipSuccess;
    const unsigned G = SLX_SOLVE_BLOCKS;
    // f = b_A - D x_D ; x_D = Dirichlet values from the rhs
    hipLaunchKernelGGL(gather_kernel, dim3(blocks(n)), dim3(256), 0, st, n, s.free_dofs, s.rhs, w.f);
    if (s.ncons > 0) {
        hipLaunchKernelGGL(gather_kernel, dim3(blocks(s.ncons)), dim3(256), 0, st, s.ncons, s.cons_dofs, s.rhs, w.xd);
        spmv<double, double, double>(n, s.nnzD, s.Drow, s.Dcol, s.Dval, w.xd, w.f, 1.0, 1.0, w.f, nullptr, st);
    }
    hipLaunchKernelGGL(prep_kernel, dim3(blocks(n)), dim3(256), 0, st, n, s.Arow, s.Acol, s.Aval, w.val32, w.dinv);
    HIPERR(hipMemsetAsync(w.x, 0, n * sizeof(double), st));
    double h = 0.0;
    hipLaunchKernelGGL((dot_partial<double>), dim3(G), dim3(256), 0, st, n, w.f, w.f, w.part);
    hipLaunchKernelGGL(sum_partials, dim3(1), dim3(256), 0, st, w.part, 1, (int)SC_DOT, 0, 0, w.sc);
    HIPERR(read_scalar(w.sc, SC_DOT, &h, st));
    const double fnorm = sqrt(h);
    // r = f (x = 0)
    HIPERR(hipMemcpyAsync(w.r, w.f, n * sizeof(double), hipMemcpyDeviceToDevice, st));
    double rnorm = fnorm;
    for (int it = 0; it < prm.max_outer; it++) {
        if (rnorm <= prm.rel_tol * fnorm) break;
        out.outer++;
        if (prm.gmres) {
            hipLaunchKernelGGL(to_f32, dim3(blocks(n)), dim3(256), 0, st, n, w.r, w.r32);
            HIPERR(gmres_cycle(s, prm, w, out.inner, st));
        } else {
        // inner FP32 PCG on A d = r
        hipLaunchKernelGGL(pcg_start, dim3(G), dim3(256), 0, st, n, w.r, w.dinv, w.r32, w.z, w.p, w.d, w.part);
        hipLaunchKernelGGL(sum_partials, dim3(1), dim3(256), 0, st, w.part, 2, (int)SC_RZ, (int)SC_RR, 0, w.sc);
        HIPERR(read_scalar(w.sc, SC_RR, &h, st));
        const double r0 = sqrt(h);
        for (int k = 0; k < prm.max_inner; k++) {
            hipLaunchKernelGGL(spmv_stream_f32, dim3(G), dim3(256), 0, st, s.nblk, s.Ablk, s.Arow, s.Acol, w.val32, s.nnzA, w.p, w.q,
                               w.part);
            hipLaunchKernelGGL(sum_partials, dim3(1), dim3(256), 0, st, w.part, 1, (int)SC_PQ, 0, 0, w.sc);
            hipLaunchKernelGGL(pcg_update, dim3(G), dim3(256), 0, st, n, w.sc, w.d, w.r32, w.p, w.q, w.dinv, w.z, w.part);
            hipLaunchKernelGGL(sum_partials, dim3(1), dim3(256), 0, st, w.part, 2, (int)SC_RZN, (int)SC_RR, 1, w.sc);
            out.inner++;
            if ((k + 1) % 8 == 0 || k + 1 == prm.max_inner) {
                HIPERR(read_scalar(w.sc, SC_RR, &h, st));
                if (sqrt(h) <= prm.inner_tol * r0) break;
            }
            hipLaunchKernelGGL(pcg_direction, dim3(G), dim3(256), 0, st, n, w.sc, w.z, w.p);
        }
        }
        hipLaunchKernelGGL(add_f32_to_f64, dim3(blocks(n)), dim3(256), 0, st, n, w.d, w.x);
        // FP64 residual r = f - A x
        spmv<double, double, double>(n, s.nnzA, s.Arow, s.Acol, s.Aval, w.x, w.f, 1.0, 1.0, w.r, nullptr, st);
        hipLaunchKernelGGL((dot_partial<double>), dim3(G), dim3(256), 0, st, n, w.r, w.r, w.part);
        hipLaunchKernelGGL(sum_partials, dim3(1), dim3(256), 0, st, w.part, 1, (int)SC_DOT, 0, 0, w.sc);
        HIPERR(read_scalar(w.sc, SC_DOT, &h, st));
        rnorm = sqrt(h);
    }
    out.rel_residual = fnorm > 0 ? rnorm / fnorm : 0.0;
    // full solution: free entries from x, constrained entries the Dirichlet values
    HIPERR(hipMemcpyAsync(w.xfull, s.rhs, s.n_dofs * sizeof(double), hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(scatter_kernel, dim3(blocks(n)), dim3(256), 0, st, n, s.free_dofs, w.x, w.xfull);
    return hipGetLastError();
}

}  // namespace slx
