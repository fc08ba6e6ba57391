// Device solve of an assembled system: FP64 iterative refinement around a
// Jacobi-preconditioned FP32 conjugate-gradient inner solve (SPD systems).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#define HIPERR(x)                              \
    do {                                       \
        hipError_t _e = (x);                   \
        if (_e != hipSuccess) return _e;       \
    } while (0)

namespace slx {

struct SolveSystem {
    int64_t n_dofs, nfree, ncons;
    const int64_t *Arow, *Drow;
    const int32_t *Acol, *Dcol;
    const double *Aval, *Dval;
    const double* rhs;          // [n_dofs]: free entries b_A, constrained entries the Dirichlet values
    const int64_t* free_dofs;   // [nfree] dof of each A row
    const int64_t* cons_dofs;   // [ncons] dof of each D column
};

struct SolveParams {
    double rel_tol;    // stop when ||b_A - D x_D - A x|| <= rel_tol ||b_A - D x_D||
    double inner_tol;  // inner PCG: relative reduction of its own FP32 residual
    int max_outer, max_inner;
};

struct SolveStats {
    int outer = 0, inner = 0;
    double rel_residual = 0.0;
};

// device workspace: nfree doubles f, r, x; ncons doubles xd; n_dofs doubles
// xfull; nfree floats r32, z, p, q, d, dinv; nnzA floats val32; 512 + 8
// doubles part / sc
struct SolveWork {
    double *f, *r, *x, *xd, *xfull, *part, *sc;
    float *r32, *z, *p, *q, *d, *dinv, *val32;
};

hipError_t solve_system(const SolveSystem& s, const SolveParams& prm, SolveWork& w, SolveStats& out, hipStream_t st);

}  // namespace slx
