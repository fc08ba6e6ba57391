// Element matrices of small slices (n_tf, n_dof <= 64) whose coefficients
// vary over the Gauss points, one wave per element holding its mtt x ntt
// 16x16 MFMA tiles (mtt, ntt = 1 or 2, or 3 x 3 or 4 x 4):
//   Ke = P^T Q_e,   P[k][i] = tf_a(k)[i][q(k)]  (shared by the bucket's elements),
//                   Q_e[k][j] = w_q(k) sum_{terms t with tf derivative a(k)} C_t(e, q(k)) dof_{b_t}[j][q(k)]
// the reference's sum over terms and Gauss points of tf_t[i][q] w_q C_t(e,q)
// dof_t[j][q] (contribution.cpp:172-246) regrouped by tf derivative, as in the
// factorized kernel (kernels/factor.h) but with the whole element in one wave:
// P and the weighted dof table stay in LDS for the whole launch (up to 2 x 2
// tiles; read through the caches for the larger ones), and each element
// builds its Q_e fragments from its coefficients.  FP16
// hi/lo or BF16 hi/mid/lo split, or plain FP32, with FP32 accumulation
// (kernels/mfma.h, AR_F16, AR_BF16 and AR_F32).
#pragma once
#include <hip/hip_runtime.h>
#include "device.h"

#define SLX_TK_MAXC 16     // k-chunks of 32 entries at most
#define SLX_TK_MAXQ 128    // padded Gauss points at most (the lanes sum the coefficients in two passes)
#define SLX_TK_WAVES 4     // waves (elements in flight) per workgroup
#define SLX_TK_MAXT 16     // terms of a slice

namespace slx {

// K order: chunk c, lane group g (lanes 16 g .. 16 g + 15), entry v = 0..7 of
// the lane's 8 k values (k = 32 c + 8 g + v) is Gauss point (4 c + g) qh + v / na
// with tf derivative v % na, where qh = 8 / na whole Gauss points fill a group
// (entries past qh na are zero).
__host__ __device__ constexpr int tile_qh(int na) { return 8 / na; }

struct TileSlice {
    int32_t n_tf, n_dof, ngp, nq;   // nq: coefficient points per element
    int32_t na, nkc, ngpp;          // tf derivatives, k-chunks, padded Gauss points (nkc * 4 * qh)
    int32_t mtt, ntt;               // 16-row and 16-column tiles (1 or 2 each, or 3 x 3, or 4 x 4)
    int32_t n_terms;
    // tinfo: [n_terms] tf-derivative index (into the na derivatives), then
    // [n_terms] dof derivative, then [n_terms] coefficient slot
    const int32_t* tinfo;
    float a_unscale;                // 1 / power-of-two scale of P
    float tqmax;                    // max |tq|
    const uint32_t* afrag;          // set per launch: afrag_x[arith]
    const uint32_t* afrag_x[3];     // [nkc][mtt][planes][64 lanes][4 words]: lane l of chunk c and row tile mt holds its entries
                                    // v = 0..7, the scaled P^T[16 mt + l % 16][k(c, l / 16, v)], as f16 hi, lo (2 planes),
                                    // bf16 hi, mid, lo (3 planes) or floats (2 words: v = 0..3, 4..7)
    const float* tq;                // [ngpp][16 ntt][4]: w_q dof_b[j][q], zero for q >= ngp or j >= n_dof
    int64_t ke_base, ke_ld;
};

size_t tile_lds_bytes(const TileSlice& t, int planes);
// Element matrices of one slice (coefficient table [slot][element][point]
// when qfast, else [slot][point][element]).
// arith: 0 FP16 split, 1 BF16 split, 2 FP32.
hipError_t launch_tile(const DevBucket& b, const TileSlice& t, int arith, const float* ctab, int qfast, float* ke, int n_cu,
                       hipStream_t st);

}  // namespace slx
