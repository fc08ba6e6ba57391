// Device solve of an assembled system: FP64 iterative refinement around a
// Jacobi-preconditioned FP32 conjugate-gradient inner solve (SPD systems).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>

#define HIPERR(x)                              \
    do {                                       \
        hipError_t _e = (x);                   \
        if (_e != hipSuccess) return _e;       \
    } while (0)

#define SLX_SOLVE_BLOCKS 2048  // fixed grid of the partial-sum kernels
#define SLX_STREAM_NNZ 4096    // nonzeros of one row block of the streaming FP32 product
#define SLX_GMRES_MAX 100        // Krylov vectors per GMRES cycle at most
#define SLX_XCDS 8              // XCDs of the device (SLX_SOLVE_BLOCKS is a multiple)

namespace slx {

struct SolveSystem {
    int64_t n_dofs, nfree, ncons, nnzA, nnzD;
    const int64_t *Arow, *Drow;
    const int32_t *Acol, *Dcol;
    const double *Aval, *Dval;
    const double* rhs;          // [n_dofs]: free entries b_A, constrained entries the Dirichlet values
    const int64_t* free_dofs;   // [nfree] dof of each A row
    const int64_t* cons_dofs;   // [ncons] dof of each D column
    const int64_t* Ablk;        // [nblk + 1] row blocks of A (stream_row_blocks)
    int64_t nblk;
};

struct SolveParams {
    double rel_tol;    // stop when ||b_A - D x_D - A x|| <= rel_tol ||b_A - D x_D||
    double inner_tol;  // inner solve: relative reduction of its own FP32 residual
    int max_outer, max_inner;
    int gmres;         // 0: Jacobi-PCG inner solve (SPD A); 1: Jacobi right-preconditioned GMRES
    int restart;       // GMRES: Krylov vectors per refinement (<= SLX_GMRES_MAX)
};

struct SolveStats {
    int outer = 0, inner = 0;
    double rel_residual = 0.0;
};

// device workspace: nfree doubles f, r, x; ncons doubles xd; n_dofs doubles
// xfull; 2 * SLX_SOLVE_BLOCKS + 8 doubles part / sc; nfree floats r32, z, p,
// q, d, dinv; nnzA floats val32 (16-byte aligned)
struct SolveWork {
    double *f, *r, *x, *xd, *xfull, *part, *sc;
    float *r32, *z, *p, *q, *d, *dinv, *val32;
    // GMRES only: Krylov basis [restart + 1][nfree] floats, partial sums
    // [restart + 1][SLX_SOLVE_BLOCKS] and restart + 1 coefficients (doubles)
    float* V;
    double *mpart, *hv;
};

// 1 when A differs from its transpose structurally, or numerically by more
// than 1e-5 relative plus 1e-6 of the two diagonal entries in any entry;
// flag: one int of device memory
hipError_t check_symmetric(const SolveSystem& s, int* flag_dev, int* nonsym, hipStream_t st);

// Row blocks of the streaming product: consecutive rows whose nonzeros fit
// SLX_STREAM_NNZ (at most 256 rows); a longer row is a block of its own.
void stream_row_blocks(const int64_t* rowp, int64_t nrows, std::vector<int64_t>& out);

// out[nfree] = b[free] - D b[constrained]  (b: n_dofs entries; xd: ncons work)
hipError_t eliminate_system(const SolveSystem& s, const double* b, double* out, double* xd, hipStream_t st);
// y[free] = A x[free] + D x[constrained], y[constrained] = x[constrained]
// (x, y: n_dofs entries; xa, ya: nfree work; xd: ncons work)
hipError_t matvec_system(const SolveSystem& s, const double* x, double* y, double* xa, double* xd, double* ya, hipStream_t st);
// y = A x with the streaming FP32 product on any CSR; sc[1] = x.y
hipError_t spmv_stream_probe(int64_t nrows, const int64_t* rowp, const int32_t* col, const float* val, int64_t nnz, const int64_t* blk,
                             int64_t nblk, const float* x, float* y, double* part, double* sc, hipStream_t st);

hipError_t solve_system(const SolveSystem& s, const SolveParams& prm, SolveWork& w, SolveStats& out, hipStream_t st);

}  // namespace slx
