// Factorized MFMA contraction for large element matrices (high-order
// hexahedra and other slices with n_tf * n_dof > 256).  Per element and slice
//   Ke = P^T Q_e,   P[k][i] = tf_a(k)[i][q(k)],
//                   Q_e[k][j] = w_q(k) sum_{terms t with tf derivative a(k)} C_t(e, q(k)) dof_{b_t}[j][q(k)]
// which is the reference's sum over terms and Gauss points of
// tf_t[i][q] w_q C_t(e,q) dof_t[j][q] (contribution.cpp:172-246) regrouped by
// tf derivative.  P is shared by the elements of a bucket (orientation-0
// tables), Q_e is built per element; both run on MFMA with FP32 accumulation
// in one of the arithmetics of kernels/mfma.h (FP16 hi/lo split, BF16
// hi/mid/lo split, plain FP32).
#pragma once
#include <hip/hip_runtime.h>
#include "device.h"

#define SLX_FK_WAVES 8     // consumer (MFMA) waves
#define SLX_FK_BWAVES 4    // producer waves: 256 threads, one per padded dof column
#define SLX_FK_RU 16       // Gauss points per k-step at most
#define SLX_FK_AREG 11     // P chunk words per producer thread (mt * planes * 64 <= 11 * 256)
#define SLX_FK_MAXT 16     // terms of a slice

namespace slx {

// K order of a k-step (32 entries): 32 / segw segments of segw entries, segw
// = 16 for the variants with two column tiles per wave and 8 for the one-tile
// variant (one producer thread per column and segment either way); a segment
// holds segw / na whole Gauss points, entry r of segment s being Gauss point
// ks * qb + s * (segw / na) + r / na with tf derivative r % na (entries past
// the segment's last whole Gauss point are zero).
__host__ __device__ constexpr int factor_segw(int bn) { return bn >= 2 ? 16 : 8; }
__host__ __device__ constexpr int factor_qb(int na, int bn) { return (32 / factor_segw(bn)) * (factor_segw(bn) / na); }

struct FactorSlice {
    int32_t n_tf, n_dof, ngp, nq;   // nq: coefficient points per element (1 when q-invariant)
    int32_t na, qb, nks;            // tf derivatives, Gauss points per k-step (factor_qb), k-steps
    int32_t mt, nt, ntp;            // 16-row / 16-column tiles, padded dof columns (nt * 16)
    int32_t ntq;                    // row stride of tq: dof columns padded to a multiple of 128
    int32_t n_terms, bmask;         // terms; dof derivatives in use (bit b)
    int32_t bm, bn;                 // tiles per wave (rows, columns): waves 2 x 4
    // tinfo: [n_terms] tf-derivative index (into the na derivatives), then
    // [n_terms] dof derivative, then [n_terms] coefficient slot
    const int32_t* tinfo;
    float a_unscale;                // 1 / power-of-two scale of P
    float tqmax;                    // max |tq|
    int32_t arith;                  // MfmaArith of this launch
    const uint32_t* afrag;          // [nks][mt][planes][64][4] words: fragments of the scaled P in this launch's arithmetic
    const float* tq;                // [nks * qb][ntq][4]: w_q dof_b[j][q], zero padded rows and columns
    int64_t ke_base, ke_ld;
    const uint32_t* afrag_x[3];     // host side: the fragments of every arithmetic
    uint32_t arith_ok;              // bit a: arithmetic a fits (LDS, P chunk staging)
    // FP64 (v_mfma_f64_16x16x4_f64): k-steps of 16 (qb64 Gauss points x na
    // derivatives), P fragments [nks64][mt][4 steps][64] doubles (lane l of
    // step j: P^T[mt*16 + l%16][4j + l/16]), w_q dof_b[j][q] as [ngp][ntp][4]
    int32_t qb64, nks64, f64_ok;
    const double* afrag64;
    const double* tq64;
};

size_t factor_lds_bytes(const FactorSlice& f, int arith);
size_t factor64_lds_bytes(const FactorSlice& f);
// Tile blocks per wave (bm row tiles, bn column tiles per pass) for an
// mt x nt tile matrix: `variant` (0..2) when it covers the rows, else the
// smallest variant that does.
bool factor_variant_fits(int mt, int nt, int* bm, int* bn, int variant = -1);
// Contraction of one slice in arithmetic `arith` (MfmaArith).
hipError_t launch_factor(const DevBucket& b, const FactorSlice& f, int arith, const float* ctab, float* ke, int n_cu,
                         hipStream_t st);
// The same contraction in FP64 on v_mfma_f64_16x16x4_f64 (FP64 coefficients
// and element matrices).
hipError_t launch_factor64(const DevBucket& b, const FactorSlice& f, const double* ctab, double* ke, int n_cu, hipStream_t st);

// One slice of a fused launch: every (slice, element) of the launch is a work
// item; items are numbered slice by slice (slices in decreasing per-element
// cost) and taken by the workgroups from a shared counter.
struct FusedSlice {
    FactorSlice f;         // arith / afrag set for the launch's arithmetic
    const void* ctab;      // the bucket's coefficient table (FP32, or FP64 for the FP64 launch)
    int64_t first;         // item number of the slice's first element
    int32_t n_elem;
    int32_t variant;       // 0: (2, 1), 1: (4, 2), 2: (7, 2) tiles per wave
};
int factor_variant_index(int bm, int bn);
// One persistent launch over every slice of `fs` (device array of n_fs
// entries, `total` items): arithmetic `arith` (MfmaArith) or 3 for FP64.
// `counter` (device, 8 bytes) is cleared by the launch's first kernel.
hipError_t launch_factor_fused(const FusedSlice* fs, int n_fs, int64_t total, unsigned long long* counter, int arith, size_t lds,
                               void* ke, int n_cu, hipStream_t st);

}  // namespace slx
