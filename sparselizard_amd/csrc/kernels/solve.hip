// Device solve of an assembled system (kernels).
//
// x_A = A^-1 (b_A - D x_D) with x_D the Dirichlet values (mat::eliminate,
// src/formulation/mat.cpp:78-103; the reference then factorizes with MUMPS,
// src/expression/operation/sl.cpp:1565-1627).  Here: FP64 iterative
// refinement (FP64 CSR residuals) around a Jacobi-preconditioned FP32
// conjugate-gradient inner solve, for symmetric positive definite A.
//
// Every reduction is deterministic: fixed lanes per row for the products, a
// fixed grid of partial sums and one ordered final sum for the dots.
#include <hip/hip_runtime.h>
#include "solve.h"

namespace slx {

#define SLX_SPMV_LPR 8

// y[r] = c0 * f[r] - c1 * (A x)[r] for the rows of one CSR matrix (FP64);
// f may be null (then y = -c1 A x, with c1 = -1 giving y = A x)
template <typename VT, typename XT, typename YT>
__global__ void __launch_bounds__(256) spmv_kernel(int64_t nrows, const int64_t* __restrict__ rowp, const int32_t* __restrict__ col,
                                                   const VT* __restrict__ val, const XT* __restrict__ x, const YT* f,
                                                   YT c0, YT c1, YT* y) {
    const int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / SLX_SPMV_LPR;
    const int l = threadIdx.x % SLX_SPMV_LPR;
    YT s = 0;
    if (r < nrows)
        for (int64_t k = rowp[r] + l; k < rowp[r + 1]; k += SLX_SPMV_LPR) s += (YT)val[k] * (YT)x[col[k]];
#pragma unroll
    for (int o = SLX_SPMV_LPR / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, SLX_SPMV_LPR);
    if (r < nrows && l == 0) y[r] = (f ? c0 * f[r] : (YT)0) - c1 * s;
}

template <typename VT, typename XT, typename YT>
static void spmv(int64_t nrows, const int64_t* rowp, const int32_t* col, const VT* val, const XT* x, const YT* f, YT c0, YT c1, YT* y,
                 hipStream_t st) {
    if (nrows <= 0) return;
    const int64_t threads = nrows * SLX_SPMV_LPR;
    hipLaunchKernelGGL((spmv_kernel<VT, XT, YT>), dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, nrows, rowp, col, val, x, f,
                       c0, c1, y);
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

// partial sums of a dot product over a fixed grid, then one ordered sum
#define SLX_DOT_BLOCKS 512
template <typename T>
__global__ void __launch_bounds__(256) dot_partial(int64_t n, const T* __restrict__ a, const T* __restrict__ b, double* __restrict__ part) {
    __shared__ double sh[256];
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        s += (double)a[i] * (double)b[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(256) dot_final(const double* __restrict__ part, double* __restrict__ out) {
    __shared__ double sh[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < SLX_DOT_BLOCKS; i += blockDim.x) s += part[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

template <typename T>
static void dot(int64_t n, const T* a, const T* b, double* part, double* out, hipStream_t st) {
    hipLaunchKernelGGL((dot_partial<T>), dim3(SLX_DOT_BLOCKS), dim3(256), 0, st, n, a, b, part);
    hipLaunchKernelGGL(dot_final, dim3(1), dim3(256), 0, st, part, out);
}

// z = dinv .* r ; with first: p = z
__global__ void precond_kernel(int64_t n, const float* __restrict__ dinv, const float* __restrict__ r, float* __restrict__ z) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) z[i] = dinv[i] * r[i];
}

// d += alpha p ; r -= alpha q   (alpha = rz / pq from device scalars)
__global__ void cg_update(int64_t n, const double* __restrict__ sc, float* __restrict__ d, float* __restrict__ r, const float* __restrict__ p,
                          const float* __restrict__ q) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const float alpha = sc[1] != 0.0 ? (float)(sc[0] / sc[1]) : 0.f;
    if (i < n) {
        d[i] += alpha * p[i];
        r[i] -= alpha * q[i];
    }
}

// p = z + beta p   (beta = rz_new / rz)
__global__ void cg_direction(int64_t n, const double* __restrict__ sc, const float* __restrict__ z, float* __restrict__ p) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const float beta = sc[0] != 0.0 ? (float)(sc[2] / sc[0]) : 0.f;
    if (i < n) p[i] = z[i] + beta * p[i];
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

hipError_t solve_system(const SolveSystem& s, const SolveParams& prm, SolveWork& w, SolveStats& out, hipStream_t st) {
    const int64_t n = s.nfree;
    out = SolveStats{};
    if (n == 0) return hipSuccess;
    // f = b_A - D x_D ; x_D = Dirichlet values from the rhs
    hipLaunchKernelGGL(gather_kernel, dim3(blocks(n)), dim3(256), 0, st, n, s.free_dofs, s.rhs, w.f);
    if (s.ncons > 0) {
        hipLaunchKernelGGL(gather_kernel, dim3(blocks(s.ncons)), dim3(256), 0, st, s.ncons, s.cons_dofs, s.rhs, w.xd);
        spmv<double, double, double>(n, s.Drow, s.Dcol, s.Dval, w.xd, w.f, 1.0, 1.0, w.f, st);
    }
    hipLaunchKernelGGL(prep_kernel, dim3(blocks(n)), dim3(256), 0, st, n, s.Arow, s.Acol, s.Aval, w.val32, w.dinv);
    HIPERR(hipMemsetAsync(w.x, 0, n * sizeof(double), st));
    double h[3];
    dot<double>(n, w.f, w.f, w.part, w.sc + 3, st);
    HIPERR(hipMemcpyAsync(h, w.sc + 3, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPERR(hipStreamSynchronize(st));
    const double fnorm = sqrt(h[0]);
    // r = f (x = 0)
    HIPERR(hipMemcpyAsync(w.r, w.f, n * sizeof(double), hipMemcpyDeviceToDevice, st));
    double rnorm = fnorm;
    for (int it = 0; it < prm.max_outer; it++) {
        out.rel_residual = fnorm > 0 ? rnorm / fnorm : 0.0;
        if (rnorm <= prm.rel_tol * fnorm) break;
        out.outer++;
        // inner FP32 PCG on A d = r
        hipLaunchKernelGGL(to_f32, dim3(blocks(n)), dim3(256), 0, st, n, w.r, w.r32);
        HIPERR(hipMemsetAsync(w.d, 0, n * sizeof(float), st));
        hipLaunchKernelGGL(precond_kernel, dim3(blocks(n)), dim3(256), 0, st, n, w.dinv, w.r32, w.z);
        HIPERR(hipMemcpyAsync(w.p, w.z, n * sizeof(float), hipMemcpyDeviceToDevice, st));
        dot<float>(n, w.r32, w.z, w.part, w.sc + 0, st);  // sc[0] = rz
        dot<float>(n, w.r32, w.r32, w.part, w.sc + 3, st);
        HIPERR(hipMemcpyAsync(h, w.sc + 3, sizeof(double), hipMemcpyDeviceToHost, st));
        HIPERR(hipStreamSynchronize(st));
        const double r0 = sqrt(h[0]);
        for (int k = 0; k < prm.max_inner; k++) {
            spmv<float, float, float>(n, s.Arow, s.Acol, w.val32, w.p, (const float*)nullptr, 0.f, -1.f, w.q, st);  // q = A p
            dot<float>(n, w.p, w.q, w.part, w.sc + 1, st);                                                           // sc[1] = pq
            hipLaunchKernelGGL(cg_update, dim3(blocks(n)), dim3(256), 0, st, n, w.sc, w.d, w.r32, w.p, w.q);
            out.inner++;
            if ((k + 1) % 8 == 0 || k + 1 == prm.max_inner) {
                dot<float>(n, w.r32, w.r32, w.part, w.sc + 3, st);
                HIPERR(hipMemcpyAsync(h, w.sc + 3, sizeof(double), hipMemcpyDeviceToHost, st));
                HIPERR(hipStreamSynchronize(st));
                if (sqrt(h[0]) <= prm.inner_tol * r0) break;
            }
            hipLaunchKernelGGL(precond_kernel, dim3(blocks(n)), dim3(256), 0, st, n, w.dinv, w.r32, w.z);
            dot<float>(n, w.r32, w.z, w.part, w.sc + 2, st);  // sc[2] = rz_new
            hipLaunchKernelGGL(cg_direction, dim3(blocks(n)), dim3(256), 0, st, n, w.sc, w.z, w.p);
            HIPERR(hipMemcpyAsync(w.sc + 0, w.sc + 2, sizeof(double), hipMemcpyDeviceToDevice, st));
        }
        hipLaunchKernelGGL(add_f32_to_f64, dim3(blocks(n)), dim3(256), 0, st, n, w.d, w.x);
        // FP64 residual r = f - A x
        spmv<double, double, double>(n, s.Arow, s.Acol, s.Aval, w.x, w.f, 1.0, 1.0, w.r, st);
        dot<double>(n, w.r, w.r, w.part, w.sc + 3, st);
        HIPERR(hipMemcpyAsync(h, w.sc + 3, sizeof(double), hipMemcpyDeviceToHost, st));
        HIPERR(hipStreamSynchronize(st));
        rnorm = sqrt(h[0]);
    }
    out.rel_residual = fnorm > 0 ? rnorm / fnorm : 0.0;
    // full solution: free entries from x, constrained entries the Dirichlet values
    HIPERR(hipMemcpyAsync(w.xfull, s.rhs, s.n_dofs * sizeof(double), hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(scatter_kernel, dim3(blocks(n)), dim3(256), 0, st, n, s.free_dofs, w.x, w.xfull);
    return hipGetLastError();
}

}  // namespace slx
