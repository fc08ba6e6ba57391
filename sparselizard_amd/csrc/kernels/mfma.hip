// Mixed-precision element contraction on MFMA (gfx950).
//
// For the matrix slices of a bucket that share one tf and one dof shape
// table the reference computes, per term t,
//   Ke_t[(i,j), e] = sum_q tf_t[i][q] w_q dof_t[j][q] * C_t(e,q)
// (contribution.cpp:172-211) and sums the terms (addproduct).  Stacking the
// terms' derivative pairs p = (r, s) and the Gauss points q into one K index
// k = p * n_gp + q gives one GEMM per slice:
//   Ke[m = i*n_dof + j, e] = sum_k Phi[m][k] * C[k][e]
// with Phi[m][(p,q)] = tf_r[i][q] w_q dof_s[j][q] constant per bucket.
//
// Arithmetic (MfmaArith), FP32 accumulation in every case, FP64 reduction
// into the CSR, Jacobian / inverse / determinant in FP64:
//   AR_F16: Phi and C split into FP16 hi + lo parts,
//       x = hi + 2^-11 lo',   lo' = fp16((x - hi) * 2^11),
//     Phi scaled once by a power of two, C scaled per (slice, element) column
//     by a power of two so both parts stay in the FP16 normal range.  Three
//     products (hi*hi into one accumulator, hi*lo' + lo'*hi into a second)
//     recover ~22 bits (v_mfma_f32_16x16x32_f16).
//   AR_BF16: three BF16 pieces x = hi + mid + lo (each the BF16 rounding of
//     what the previous pieces leave; BF16 has the FP32 exponent range, so the
//     pieces need no scaling beyond the shared one).  Six products (hi*hi into
//     one accumulator; hi*mid, mid*hi, hi*lo, mid*mid, lo*hi into a second)
//     keep every term down to 2^-16 of the product: ~22 bits
//     (v_mfma_f32_16x16x32_bf16).
//   AR_F32: plain FP32 operands, one product (v_mfma_f32_16x16x4_f32).
//
// Operand layouts (wave64, l = lane):
//   16x16x32 f16/bf16:  A lane l holds A[l % 16][8 (l / 16) + 0..7]
//                       B lane l holds B[8 (l / 16) + 0..7][l % 16]
//   16x16x4 f32:        A lane l holds A[l % 16][l / 16]
//                       B lane l holds B[l / 16][l % 16]
//   D (16 x 16):        lane l holds D[4 (l / 16) + 0..3][l % 16]
//   (v_mfma_f64_16x16x4_f64: lane l holds D[l / 16 + 4 v][l % 16], v = 0..3)
// (checked on the GPU by the layout probe at the end of this file).
#include <hip/hip_runtime.h>
#include "device.h"
#include "mfma.h"

namespace slx {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float float4v __attribute__((ext_vector_type(4)));
typedef unsigned int uint4v __attribute__((ext_vector_type(4)));
typedef unsigned int uint2v __attribute__((ext_vector_type(2)));

__device__ inline uint32_t pack_split(float x) {
    _Float16 hi = (_Float16)x;
    _Float16 lo = (_Float16)((x - (float)hi) * 2048.f);
    return (uint32_t)__builtin_bit_cast(uint16_t, hi) | ((uint32_t)__builtin_bit_cast(uint16_t, lo) << 16);
}

// BF16 pieces of x: hi = bf16(x), mid = bf16(x - hi), lo = bf16(x - hi - mid)
// (each difference is exact in FP32).  Words (hi | mid << 16, lo).
__device__ inline void pack_bf3(float x, uint32_t& w0, uint32_t& w1) {
    const __bf16 hi = (__bf16)x;
    const float r1 = x - (float)hi;
    const __bf16 mid = (__bf16)r1;
    const __bf16 lo = (__bf16)(r1 - (float)mid);
    w0 = (uint32_t)__builtin_bit_cast(uint16_t, hi) | ((uint32_t)__builtin_bit_cast(uint16_t, mid) << 16);
    w1 = (uint32_t)__builtin_bit_cast(uint16_t, lo);
}

// Compact column words of pair p (value x already scaled), and the zero pad pair.
template <int AR>
__device__ inline void store_pair(uint32_t* dst, int p, float x) {
    if constexpr (AR == AR_F16) {
        dst[p] = pack_split(x);
    } else if constexpr (AR == AR_BF16) {
        uint32_t w0, w1;
        pack_bf3(x, w0, w1);
        dst[2 * p] = w0;
        dst[2 * p + 1] = w1;
    } else {
        dst[p] = __float_as_uint(x);
    }
}
template <int AR>
__device__ inline void store_pad(uint32_t* dst, int np);
// Padding of a compact column: the zero pad pair; presummed F16 columns: zero
// words up to 16 and, after the last slice of an element, 16 more zero words
// (read by the lanes of k 16..31).
template <int AR>
__device__ inline void presum_pad(uint32_t* dst, int np, int presum, bool last) {
    if (AR == AR_F16 && presum) {
        for (int k = np; k < 16; k++) dst[k] = 0u;
        if (last)
            for (int k = 16; k < 32; k++) dst[k] = 0u;
    } else {
        store_pad<AR>(dst, np);
    }
}
template <int AR>
__device__ inline void store_pad(uint32_t* dst, int np) {
    if constexpr (AR == AR_BF16) {
        dst[2 * np] = 0u;
        dst[2 * np + 1] = 0u;
    } else {
        dst[np] = 0u;
    }
}

// Column preparation (per workgroup tile of NE elements): for every
// (element, slice) column the coefficient values C[p][q] are read from the
// coalesced coefficient table, scaled by a power of two and split.  A column
// lives in LDS as 32-bit words:
//   compact  (coefficients constant over the Gauss points): the words of
//            pair p (F16 one word hi | lo' << 16, BF16 two words, F32 the
//            value).  Phi holds each pair's Gauss-point sum in row k = 8 p
//            (kq = 8; rows 8 p + 1 .. 8 p + 7 are zero), so the eight
//            consecutive k of a lane (16x16x32) or the four k of one 16x16x4
//            step share one pair: the B fragment is a broadcast of one pair's
//            words.  Pair n_pairs is 0 (K padding).
//   presum   (F16, compact columns whose Phi holds each pair's Gauss-point
//            sum in row k = p, kq = 1): 16 words per slice, pairs 0..15, zero
//            past the last pair; a lane's eight k are eight pairs, read as two
//            16-byte LDS loads and split by byte permutes; one K step.  The
//            lanes of k 16..31 read the next slice's words (or 16 zero words
//            after the last slice) against zero rows of Phi.
//   expanded otherwise (F16 only): word k = C[k / kq][k % kq] for k < K,
//            0 up to the padded K, read as two 16-byte LDS loads per K step
//            (kq = n_gp).
// Output: the element rows of every segment group are assembled in LDS for
// 16 elements and written with coalesced 16-byte stores (one contiguous
// range per group).
//
// Work split (one workgroup per CU, persistent over 16-element tiles):
//   consumer waves (one or two M tiles each) run the contraction of tile s
//     from the column buffer Cs[s & 1] into the staging buffer stg[s & 1];
//   producer waves (in two groups) meanwhile prepare the columns of tile
//     s + 1 into Cs[(s + 1) & 1] (SLX_MFMA_PREP_WAVES) and write tile s - 1
//     from stg[(s - 1) & 1] to the element matrices with coalesced 16-byte
//     stores (the remaining waves of the workgroup: the write-out is the
//     longest of the three tasks, so small-fragment variants fill 16 waves).
// One barrier per tile; every buffer is written one tile after its last read.
// WAVES: the largest workgroup of the variant, in waves (12 leaves 168 VGPRs
// per wave: up to 7 consumer waves; 16 leaves 128).
template <int AR, int MTW, int MAXK, bool COMPACT, int G, int WAVES>
__global__ void __launch_bounds__(64 * WAVES) mfma_contract_kernel(DevBucket b, MfmaBucket mb, const float* __restrict__ ctab,
                                                                                  float* __restrict__ ke) {
    static_assert(COMPACT || AR == AR_F16, "per-Gauss-point columns are built for the FP16 split only");
    constexpr int NPL = AR == AR_BF16 ? 3 : 2;  // A fragment planes
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    constexpr int NE = 16;
    const int ns = mb.n_mslices, np = mb.n_pairs, nq = mb.nq;
    const int estride = mb.estride, cstride = mb.cstride;
    const int n_tf = mb.n_tf;
    uint32_t* Csb = reinterpret_cast<uint32_t*>(smem_raw);                     // [2][NE][estride]
    float* colscb = reinterpret_cast<float*>(Csb + (size_t)2 * NE * estride);   // [2][ns][NE]
    float* stgb = colscb + (size_t)2 * ns * NE;                                 // [2][16 * stg_words]
    int* csl = reinterpret_cast<int*>(stgb + (size_t)2 * 16 * mb.stg_words);    // [ns][np] coefficient slots
    float* trash = reinterpret_cast<float*>(csl + ns * np);                      // [64] sink of padding-row stores
    const SLX_CONST int32_t* csrc = (const SLX_CONST int32_t*)mb.csrc;
    const SLX_CONST int32_t* sprmc = (const SLX_CONST int32_t*)mb.sprm;
    const SLX_CONST int64_t* gbase = (const SLX_CONST int64_t*)mb.gbase;
    const SLX_CONST int32_t* gld = (const SLX_CONST int32_t*)mb.gld;
    for (int x = threadIdx.x; x < ns * np; x += blockDim.x) csl[x] = csrc[x];

    const int nw = mb.nw;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int lrow = lane & 15, lgrp = lane >> 4;
    const int ntiles = (b.n_elem + NE - 1) / NE;
    const int nsteps = ntiles > (int)blockIdx.x ? (ntiles - 1 - (int)blockIdx.x) / (int)gridDim.x + 1 : 0;
    const int64_t nel = b.n_elem;
    __syncthreads();  // csl

    if (wave < nw) {
        // ------------------------------------------------------- consumers --
        // A fragments of this wave's M tiles (tile = wave + nw * i)
        uint4v Af[MTW][MAXK][NPL];
        int irow[MTW][4], jrow[MTW][4];
        bool rowok[MTW][4];
#pragma unroll
        for (int i = 0; i < MTW; i++) {
            const int mt = wave + nw * i;
#pragma unroll
            for (int ks = 0; ks < MAXK; ks++) {
#pragma unroll
                for (int pl = 0; pl < NPL; pl++) {
                    if (mt < mb.mtiles) {
                        const uint4v* src = reinterpret_cast<const uint4v*>(mb.afrag) + ((size_t)(mt * MAXK + ks) * NPL + pl) * 64;
                        Af[i][ks][pl] = src[lane];
                    } else {
                        Af[i][ks][pl] = uint4v{0u, 0u, 0u, 0u};
                    }
                }
            }
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = mt * 16 + 4 * lgrp + v;
                irow[i][v] = row / mb.n_dof;
                jrow[i][v] = row - irow[i][v] * mb.n_dof;
                rowok[i][v] = row < mb.M;
            }
        }
        // compact columns: the pair of this lane's eight k per K step (16x16x32)
        // or of the four k of each 16x16x4 step (F32: eight-k block kb of the
        // K step, wave-uniform)
        constexpr int WPP = AR == AR_BF16 ? 2 : 1;
        int pword[MAXK];
        int pw32[MAXK][4];
#pragma unroll
        for (int ks = 0; ks < MAXK; ks++) {
            pword[ks] = WPP * min((ks * 32 + 8 * lgrp) / mb.kq, np);
#pragma unroll
            for (int kb = 0; kb < 4; kb++) pw32[ks][kb] = min((ks * 32 + 8 * kb) / mb.kq, np);
        }
        __syncthreads();  // tile 0 prepared
        if constexpr (COMPACT && G == 1 && MTW == 1 && AR != AR_F32) {
            // One M tile per wave (wave < mtiles), one slice at a time, software
            // pipelined: the MFMAs of slice s and the output stores of slice
            // s - 1 form one basic block (branch-free stores: rows past M go to
            // the sink), so the stores issue while the MFMA pipe is busy.
            const bool presum = AR == AR_F16 && mb.presum;
            auto contract = [&](const uint32_t* colbase, float4v& hh, float4v& xx) {
                if constexpr (AR == AR_F16) {
                    if (presum) {
                        // the words of pairs 8 lgrp .. 8 lgrp + 7 (one K step)
                        const uint4v x0 = *reinterpret_cast<const uint4v*>(colbase + 8 * lgrp);
                        const uint4v x1 = *reinterpret_cast<const uint4v*>(colbase + 8 * lgrp + 4);
                        uint4v hw, lw;
                        hw[0] = __builtin_amdgcn_perm(x0[1], x0[0], 0x05040100u);
                        hw[1] = __builtin_amdgcn_perm(x0[3], x0[2], 0x05040100u);
                        hw[2] = __builtin_amdgcn_perm(x1[1], x1[0], 0x05040100u);
                        hw[3] = __builtin_amdgcn_perm(x1[3], x1[2], 0x05040100u);
                        lw[0] = __builtin_amdgcn_perm(x0[1], x0[0], 0x07060302u);
                        lw[1] = __builtin_amdgcn_perm(x0[3], x0[2], 0x07060302u);
                        lw[2] = __builtin_amdgcn_perm(x1[1], x1[0], 0x07060302u);
                        lw[3] = __builtin_amdgcn_perm(x1[3], x1[2], 0x07060302u);
                        const half8 bhi = __builtin_bit_cast(half8, hw);
                        const half8 blo = __builtin_bit_cast(half8, lw);
                        const half8 ahi = __builtin_bit_cast(half8, Af[0][0][0]);
                        const half8 alo = __builtin_bit_cast(half8, Af[0][0][1]);
                        hh = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi, bhi, float4v{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
                        xx = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi, blo, float4v{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
                        xx = __builtin_amdgcn_mfma_f32_16x16x32_f16(alo, bhi, xx, 0, 0, 0);
                        return;
                    }
                }
                uint32_t cw[MAXK][AR == AR_BF16 ? 2 : (AR == AR_F32 ? 4 : 1)];
#pragma unroll
                for (int ks = 0; ks < MAXK; ks++) {
                    if constexpr (AR == AR_F16) {
                        cw[ks][0] = colbase[pword[ks]];
                    } else if constexpr (AR == AR_BF16) {
                        cw[ks][0] = colbase[pword[ks]];
                        cw[ks][1] = colbase[pword[ks] + 1];
                    } else {
#pragma unroll
                        for (int kb = 0; kb < 4; kb++) cw[ks][kb] = colbase[pw32[ks][kb]];
                    }
                }
                hh = float4v{0.f, 0.f, 0.f, 0.f};
                xx = float4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int ks = 0; ks < MAXK; ks++) {
                    if constexpr (AR == AR_F16) {
                        const uint32_t u = cw[ks][0];
                        const uint32_t h2 = __builtin_amdgcn_perm(u, u, 0x01000100u);
                        const uint32_t l2 = __builtin_amdgcn_perm(u, u, 0x03020302u);
                        const half8 bhi = __builtin_bit_cast(half8, (uint4v){h2, h2, h2, h2});
                        const half8 blo = __builtin_bit_cast(half8, (uint4v){l2, l2, l2, l2});
                        const half8 ahi = __builtin_bit_cast(half8, Af[0][ks][0]);
                        const half8 alo = __builtin_bit_cast(half8, Af[0][ks][1]);
                        hh = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi, bhi, hh, 0, 0, 0);
                        xx = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi, blo, xx, 0, 0, 0);
                        xx = __builtin_amdgcn_mfma_f32_16x16x32_f16(alo, bhi, xx, 0, 0, 0);
                    } else if constexpr (AR == AR_BF16) {
                        const uint32_t u0 = cw[ks][0], u1 = cw[ks][1];
                        const uint32_t h2 = __builtin_amdgcn_perm(u0, u0, 0x01000100u);
                        const uint32_t m2 = __builtin_amdgcn_perm(u0, u0, 0x03020302u);
                        const uint32_t l2 = __builtin_amdgcn_perm(u1, u1, 0x01000100u);
                        const bf16x8 bh = __builtin_bit_cast(bf16x8, (uint4v){h2, h2, h2, h2});
                        const bf16x8 bm = __builtin_bit_cast(bf16x8, (uint4v){m2, m2, m2, m2});
                        const bf16x8 bl = __builtin_bit_cast(bf16x8, (uint4v){l2, l2, l2, l2});
                        const bf16x8 ah = __builtin_bit_cast(bf16x8, Af[0][ks][0]);
                        const bf16x8 am = __builtin_bit_cast(bf16x8, Af[0][ks][1]);
                        const bf16x8 al = __builtin_bit_cast(bf16x8, Af[0][ks][2]);
                        hh = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, hh, 0, 0, 0);
                        xx = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bm, xx, 0, 0, 0);
                        xx = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bh, xx, 0, 0, 0);
                        xx = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, xx, 0, 0, 0);
                        xx = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bm, xx, 0, 0, 0);
                        xx = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, xx, 0, 0, 0);
                    } else {
                        const float4v a0 = __builtin_bit_cast(float4v, Af[0][ks][0]);
                        const float4v a1 = __builtin_bit_cast(float4v, Af[0][ks][1]);
#pragma unroll
                        for (int j = 0; j < 4; j++)
                            hh = __builtin_amdgcn_mfma_f32_16x16x4f32(a0[j], __uint_as_float(cw[ks][j >> 1]), hh, 0, 0, 0);
#pragma unroll
                        for (int j = 0; j < 4; j++)
                            hh = __builtin_amdgcn_mfma_f32_16x16x4f32(a1[j], __uint_as_float(cw[ks][2 + (j >> 1)]), hh, 0, 0, 0);
                    }
                }
            };
            auto store = [&](int s, const float* colsc, float* stg, const float4v& hh, const float4v& xx) {
                const int soff = sprmc[3 * s], ld = sprmc[3 * s + 1], ssrow = sprmc[3 * s + 2];
                const float colscale = colsc[s * NE + lrow];
                const float xscale = AR == AR_F16 ? colscale * (1.f / 2048.f) : colscale;
                float* out = stg + soff + __umul24(lrow, ssrow);  // 24-bit products: full-rate, factors < 2^24
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const float val = AR == AR_F32 ? hh[v] * colscale : hh[v] * colscale + xx[v] * xscale;
                    float* dst = rowok[0][v] ? out + __umul24(irow[0][v], ld) + jrow[0][v] : trash + lane;
                    *dst = val;
                }
            };
            for (int st = 0; st < nsteps; st++) {
                const uint32_t* colb = Csb + (size_t)(st & 1) * NE * estride + (size_t)lrow * estride;
                const float* colsc = colscb + (size_t)(st & 1) * ns * NE;
                float* stg = stgb + (size_t)(st & 1) * 16 * mb.stg_words;
                float4v hh0, xx0;
                contract(colb, hh0, xx0);
                for (int s = 1; s < ns; s++) {
                    float4v hh1, xx1;
                    contract(colb + (size_t)s * cstride, hh1, xx1);
                    store(s - 1, colsc, stg, hh0, xx0);
                    hh0 = hh1;
                    xx0 = xx1;
                }
                store(ns - 1, colsc, stg, hh0, xx0);
                __syncthreads();  // tile st staged; tile st + 1 prepared
            }
        } else
        for (int st = 0; st < nsteps; st++) {
            const uint32_t* Cs = Csb + (size_t)(st & 1) * NE * estride;
            const float* colsc = colscb + (size_t)(st & 1) * ns * NE;
            float* stg = stgb + (size_t)(st & 1) * 16 * mb.stg_words;
            // slices G at a time (G divides the slice count): 2 G independent
            // accumulator chains per M tile (main product, and correction
            // products) share the A fragments
            for (int s0 = 0; s0 < ns; s0 += G) {
                float4v acc_hh[G][MTW], acc_x[G][MTW];
#pragma unroll
                for (int c = 0; c < G; c++)
#pragma unroll
                    for (int i = 0; i < MTW; i++)
#pragma unroll
                        for (int v = 0; v < 4; v++) {
                            acc_hh[c][i][v] = 0.f;
                            acc_x[c][i][v] = 0.f;
                        }
                const uint32_t* colbase0 = Cs + (size_t)lrow * estride + (size_t)s0 * cstride;
                uint32_t cw[G][(COMPACT && AR == AR_F16) ? MAXK : 1];
                if constexpr (COMPACT && AR == AR_F16) {
#pragma unroll
                    for (int c = 0; c < G; c++)
#pragma unroll
                        for (int ks = 0; ks < MAXK; ks++) cw[c][ks] = colbase0[c * cstride + pword[ks]];
                }
#pragma unroll
                for (int ks = 0; ks < MAXK; ks++) {
#pragma unroll
                    for (int c = 0; c < G; c++) {
                        const uint32_t* colbase = colbase0 + c * cstride;
                        if constexpr (AR == AR_F16) {
                            uint4v hw, lw;
                            if constexpr (COMPACT) {
                                // one packed word (hi | lo' << 16) -> both halves broadcast
                                // to all eight k of the lane (byte permutes)
                                const uint32_t u = cw[c][ks];
                                const uint32_t h2 = __builtin_amdgcn_perm(u, u, 0x01000100u);
                                const uint32_t l2 = __builtin_amdgcn_perm(u, u, 0x03020302u);
                                hw = (uint4v){h2, h2, h2, h2};
                                lw = (uint4v){l2, l2, l2, l2};
                            } else {
                                const uint4v x0 = *reinterpret_cast<const uint4v*>(colbase + ks * 32 + 8 * lgrp);
                                const uint4v x1 = *reinterpret_cast<const uint4v*>(colbase + ks * 32 + 8 * lgrp + 4);
                                hw[0] = __builtin_amdgcn_perm(x0[1], x0[0], 0x05040100u);
                                hw[1] = __builtin_amdgcn_perm(x0[3], x0[2], 0x05040100u);
                                hw[2] = __builtin_amdgcn_perm(x1[1], x1[0], 0x05040100u);
                                hw[3] = __builtin_amdgcn_perm(x1[3], x1[2], 0x05040100u);
                                lw[0] = __builtin_amdgcn_perm(x0[1], x0[0], 0x07060302u);
                                lw[1] = __builtin_amdgcn_perm(x0[3], x0[2], 0x07060302u);
                                lw[2] = __builtin_amdgcn_perm(x1[1], x1[0], 0x07060302u);
                                lw[3] = __builtin_amdgcn_perm(x1[3], x1[2], 0x07060302u);
                            }
                            const half8 bhi = __builtin_bit_cast(half8, hw);
                            const half8 blo = __builtin_bit_cast(half8, lw);
#pragma unroll
                            for (int i = 0; i < MTW; i++) {
                                if (i > 0 && wave + nw * i >= mb.mtiles) continue;  // wave-uniform: no work for an empty tile
                                const half8 ahi = __builtin_bit_cast(half8, Af[i][ks][0]);
                                const half8 alo = __builtin_bit_cast(half8, Af[i][ks][1]);
                                acc_hh[c][i] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi, bhi, acc_hh[c][i], 0, 0, 0);
                                acc_x[c][i] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi, blo, acc_x[c][i], 0, 0, 0);
                                acc_x[c][i] = __builtin_amdgcn_mfma_f32_16x16x32_f16(alo, bhi, acc_x[c][i], 0, 0, 0);
                            }
                        } else if constexpr (AR == AR_BF16) {
                            // two words (hi | mid << 16, lo) -> the three pieces broadcast
                            const uint32_t u0 = colbase[pword[ks]], u1 = colbase[pword[ks] + 1];
                            const uint32_t h2 = __builtin_amdgcn_perm(u0, u0, 0x01000100u);
                            const uint32_t m2 = __builtin_amdgcn_perm(u0, u0, 0x03020302u);
                            const uint32_t l2 = __builtin_amdgcn_perm(u1, u1, 0x01000100u);
                            const bf16x8 bh = __builtin_bit_cast(bf16x8, (uint4v){h2, h2, h2, h2});
                            const bf16x8 bm = __builtin_bit_cast(bf16x8, (uint4v){m2, m2, m2, m2});
                            const bf16x8 bl = __builtin_bit_cast(bf16x8, (uint4v){l2, l2, l2, l2});
#pragma unroll
                            for (int i = 0; i < MTW; i++) {
                                if (i > 0 && wave + nw * i >= mb.mtiles) continue;
                                const bf16x8 ah = __builtin_bit_cast(bf16x8, Af[i][ks][0]);
                                const bf16x8 am = __builtin_bit_cast(bf16x8, Af[i][ks][1]);
                                const bf16x8 al = __builtin_bit_cast(bf16x8, Af[i][ks][2]);
                                acc_hh[c][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc_hh[c][i], 0, 0, 0);
                                acc_x[c][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bm, acc_x[c][i], 0, 0, 0);
                                acc_x[c][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bh, acc_x[c][i], 0, 0, 0);
                                acc_x[c][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc_x[c][i], 0, 0, 0);
                                acc_x[c][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bm, acc_x[c][i], 0, 0, 0);
                                acc_x[c][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc_x[c][i], 0, 0, 0);
                            }
                        } else {
                            // F32: step j (k = 32 ks + 4 j + lane / 16) reads the pair of
                            // eight-k block j / 2
                            float bv[4];
#pragma unroll
                            for (int kb = 0; kb < 4; kb++) bv[kb] = __uint_as_float(colbase[pw32[ks][kb]]);
#pragma unroll
                            for (int i = 0; i < MTW; i++) {
                                if (i > 0 && wave + nw * i >= mb.mtiles) continue;
                                const float4v a0 = __builtin_bit_cast(float4v, Af[i][ks][0]);
                                const float4v a1 = __builtin_bit_cast(float4v, Af[i][ks][1]);
#pragma unroll
                                for (int j = 0; j < 4; j++)
                                    acc_hh[c][i] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0[j], bv[j >> 1], acc_hh[c][i], 0, 0, 0);
#pragma unroll
                                for (int j = 0; j < 4; j++)
                                    acc_hh[c][i] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1[j], bv[2 + (j >> 1)], acc_hh[c][i], 0, 0, 0);
                            }
                        }
                    }
                }
                // the output stores stay below the MFMAs (hoisting their loads and
                // addresses above them costs registers)
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int c = 0; c < G; c++) {
                    const int s = s0 + c;
                    // the group's LDS block: [16][n_tf * ld + 4] words per tile;
                    // the slice's three constants: independent scalar loads
                    const int soff = sprmc[3 * s], ld = sprmc[3 * s + 1], ssrow = sprmc[3 * s + 2];
                    const float colscale = colsc[s * NE + lrow];
                    const float xscale = AR == AR_F16 ? colscale * (1.f / 2048.f) : colscale;
                    float* out = stg + soff + __umul24(lrow, ssrow);
#pragma unroll
                    for (int i = 0; i < MTW; i++) {
                        const int mt = wave + nw * i;
                        if (mt >= mb.mtiles) continue;
                        // wave-uniform: every row of a full tile is a matrix row
                        const bool full = (mt + 1) * 16 <= mb.M;
#pragma unroll
                        for (int v = 0; v < 4; v++) {
                            if (!full && !rowok[i][v]) continue;
                            const float val = AR == AR_F32 ? acc_hh[c][i][v] * colscale
                                                           : acc_hh[c][i][v] * colscale + acc_x[c][i][v] * xscale;
                            out[__umul24(irow[i][v], ld) + jrow[i][v]] = val;
                        }
                    }
                }
            }
            __syncthreads();  // tile st staged; tile st + 1 prepared
        }
    } else {
        // ------------------------------------------------------- producers --
        // two producer groups: the first prepares columns, the second writes out
        constexpr int prep_threads = 64 * SLX_MFMA_PREP_WAVES;
        const int write_threads = (int)blockDim.x - 64 * nw - prep_threads;
        const int pt0 = threadIdx.x - 64 * nw;
        const bool is_prep = pt0 < prep_threads;  // wave-uniform
        const int pt = is_prep ? pt0 : pt0 - prep_threads;
        // columns of a tile: one thread per (element, slice), element fastest:
        // power-of-two scale and the split words
        auto prepare = [&](int tile, int buf) {
            const int e0 = tile * NE;
            const int nloc = min(NE, b.n_elem - e0);
            uint32_t* Cs = Csb + (size_t)buf * NE * estride;
            float* colsc = colscb + (size_t)buf * ns * NE;
            for (int col = pt; col < NE * ns; col += prep_threads) {
                const int el = col % NE, s = col / NE;
                const bool valid = el < nloc;
                const int64_t e = e0 + (valid ? el : 0);
                uint32_t* dst = Cs + (size_t)el * estride + (size_t)s * cstride;
                float mx = 0.f;
                if constexpr (COMPACT) {
                    // every coefficient of the column loaded at once (independent loads)
                    float cv[SLX_MFMA_MAXP];
#pragma unroll
                    for (int p = 0; p < SLX_MFMA_MAXP; p++) {
                        const int c = p < np ? csl[s * np + p] : -1;
                        cv[p] = (c >= 0 && valid) ? ctab[(int64_t)c * nel + e] : 0.f;
                        mx = fmaxf(mx, fabsf(cv[p]));
                    }
                    int ex = 0;
                    if (mx > 0.f) (void)frexpf(mx, &ex);  // mx in [2^(ex-1), 2^ex)
                    const float sc = ldexpf(1.f, -ex);
                    colsc[s * NE + el] = ldexpf(1.f, ex) * mb.a_unscale;
#pragma unroll
                    for (int p = 0; p < SLX_MFMA_MAXP; p++)
                        if (p < np) store_pair<AR>(dst, p, cv[p] * sc);
                    presum_pad<AR>(dst, np, mb.presum, s == ns - 1);
                } else {
                    // Gauss-point values in batches of 8 independent loads
                    for (int p = 0; p < np; p++) {
                        const int c = csl[s * np + p];
                        if (c < 0 || !valid) continue;
                        for (int q0 = 0; q0 < nq; q0 += 8) {
#pragma unroll
                            for (int u = 0; u < 8; u++)
                                if (q0 + u < nq) mx = fmaxf(mx, fabsf(ctab[((int64_t)c * nq + q0 + u) * nel + e]));
                        }
                    }
                    int ex = 0;
                    if (mx > 0.f) (void)frexpf(mx, &ex);  // mx in [2^(ex-1), 2^ex)
                    const float sc = ldexpf(1.f, -ex);
                    colsc[s * NE + el] = ldexpf(1.f, ex) * mb.a_unscale;
                    const int kq = mb.kq;  // the Gauss points of a pair
                    for (int p = 0; p < np; p++) {
                        const int c = csl[s * np + p];
                        for (int q0 = 0; q0 < kq; q0 += 8) {
#pragma unroll
                            for (int u = 0; u < 8; u++) {
                                const int q = q0 + u;
                                if (q >= kq) break;
                                const float x = (c >= 0 && valid) ? ctab[((int64_t)c * nq + (nq == 1 ? 0 : q)) * nel + e] * sc : 0.f;
                                dst[p * kq + q] = pack_split(x);
                            }
                        }
                    }
                    for (int k = np * kq; k < cstride; k++) dst[k] = 0u;
                }
            }
        };
        // each group's rows of the tile's elements are one contiguous range
        auto write_out = [&](int tile, int buf) {
            const int e0 = tile * NE;
            const int nel_here = min(NE, b.n_elem - e0);
            const float* gsrc = stgb + (size_t)buf * 16 * mb.stg_words;
            for (int g = 0; g < mb.n_groups; g++) {
                const int erow = n_tf * gld[g], srow = erow + 4;
                float* dst = ke + gbase[g] + (int64_t)e0 * erow;
                if (((uintptr_t)dst & 15u) == 0 && (erow & 3) == 0) {
                    const int per = erow / 4;  // float4 per element row
                    float4v* d4 = reinterpret_cast<float4v*>(dst);
                    int eo = pt / per, w = pt - eo * per;
                    const int step_e = write_threads / per, step_w = write_threads - step_e * per;
                    const int nx = nel_here * per;
                    // four LDS reads in flight before their four stores
                    constexpr int WU = 4;
                    for (int x = pt; x < nx; x += WU * write_threads) {
                        float4v v[WU];
#pragma unroll
                        for (int u = 0; u < WU; u++) {
                            if (x + u * write_threads < nx) v[u] = *reinterpret_cast<const float4v*>(gsrc + (size_t)eo * srow + 4 * w);
                            eo += step_e;
                            w += step_w;
                            if (w >= per) {
                                w -= per;
                                eo++;
                            }
                        }
#pragma unroll
                        for (int u = 0; u < WU; u++)
                            if (x + u * write_threads < nx) d4[x + u * write_threads] = v[u];
                    }
                } else {
                    for (int x = pt; x < nel_here * erow; x += write_threads) {
                        const int eo = x / erow, w = x - eo * erow;
                        dst[x] = gsrc[(size_t)eo * srow + w];
                    }
                }
                gsrc += (size_t)16 * srow;
            }
        };
        // compact columns: one column per producer thread, its coefficients
        // loaded two tiles ahead (the loads of tile s + 2 are in flight while
        // tile s + 1 is packed and tile s - 1 written)
        float cvn[SLX_MFMA_MAXP];
        auto load_cols = [&](int tile) {
            const int e0 = tile * NE, nloc = min(NE, b.n_elem - e0);
            const int el = pt % NE, s = pt / NE;
            const bool valid = pt < NE * ns && el < nloc;
            const int64_t e = e0 + (valid ? el : 0);
#pragma unroll
            for (int p = 0; p < SLX_MFMA_MAXP; p++) {
                const int c = (valid && p < np) ? csl[s * np + p] : -1;
                cvn[p] = c >= 0 ? ctab[(int64_t)c * nel + e] : 0.f;
            }
        };
        auto finish_cols = [&](int buf) {
            if (pt >= NE * ns) return;
            const int el = pt % NE, s = pt / NE;
            uint32_t* dst = Csb + (size_t)buf * NE * estride + (size_t)el * estride + (size_t)s * cstride;
            float* colsc = colscb + (size_t)buf * ns * NE;
            float mx = 0.f;
#pragma unroll
            for (int p = 0; p < SLX_MFMA_MAXP; p++) mx = fmaxf(mx, fabsf(cvn[p]));
            int ex = 0;
            if (mx > 0.f) (void)frexpf(mx, &ex);  // mx in [2^(ex-1), 2^ex)
            const float sc = ldexpf(1.f, -ex);
            colsc[s * NE + el] = ldexpf(1.f, ex) * mb.a_unscale;
#pragma unroll
            for (int p = 0; p < SLX_MFMA_MAXP; p++)
                if (p < np) store_pair<AR>(dst, p, cvn[p] * sc);
            presum_pad<AR>(dst, np, mb.presum, s == ns - 1);
        };
        const bool pipelined = COMPACT && NE * ns <= prep_threads;
        auto tile_of = [&](int st) { return (int)blockIdx.x + st * (int)gridDim.x; };
        if (is_prep) {
            if (pipelined) {
                if (nsteps > 0) {
                    load_cols(tile_of(0));
                    finish_cols(0);
                }
                if (nsteps > 1) load_cols(tile_of(1));
            } else if (nsteps > 0) {
                prepare(tile_of(0), 0);
            }
        }
        __syncthreads();  // tile 0 prepared
        for (int st = 0; st < nsteps; st++) {
            if (is_prep) {
                if (st + 1 < nsteps) {
                    if (pipelined) {
                        finish_cols((st + 1) & 1);
                        if (st + 2 < nsteps) load_cols(tile_of(st + 2));
                    } else {
                        prepare(tile_of(st + 1), (st + 1) & 1);
                    }
                }
            } else if (st > 0) {
                write_out(tile_of(st - 1), (st - 1) & 1);
            }
            __syncthreads();  // tile st staged; tile st + 1 prepared
        }
        if (!is_prep && nsteps > 0) write_out(tile_of(nsteps - 1), (nsteps - 1) & 1);
    }
}

template <int AR, int MTW, int MAXK, bool COMPACT, int G>
static void launch_waves(const dim3& gd, const dim3& bd, size_t lds, hipStream_t st, const DevBucket& b, const MfmaBucket& mb,
                         const float* ctab, float* ke) {
    if (bd.x <= 64 * 12)
        hipLaunchKernelGGL((mfma_contract_kernel<AR, MTW, MAXK, COMPACT, G, 12>), gd, bd, lds, st, b, mb, ctab, ke);
    else
        hipLaunchKernelGGL((mfma_contract_kernel<AR, MTW, MAXK, COMPACT, G, 16>), gd, bd, lds, st, b, mb, ctab, ke);
}

template <int AR, int MTW, int MAXK>
static hipError_t launch_variant(const DevBucket& b, const MfmaBucket& mb, const float* ctab, float* ke, hipStream_t st) {
    const size_t lds = mfma_lds_bytes(mb);
    const int tiles = (b.n_elem + 15) / 16;
    const int grid = tiles < mb.max_blocks ? tiles : mb.max_blocks;
    // F16 with small A fragments (at most 32 VGPRs): a 16-wave workgroup,
    // the waves beyond consumers and column producers write out (measured:
    // 2.00 -> 1.84 ms at 64^3 P2 elasticity; FP32 MFMA is slower that way);
    // else two writers
    const int waves = AR == AR_F16 && MTW * MAXK <= 4 ? 16 : mb.nw + SLX_MFMA_PREP_WAVES + SLX_MFMA_MIN_WRITE_WAVES;
    const dim3 gd(grid), bd(64 * waves);
    if constexpr (AR != AR_F16) {
        // compact columns, one M tile per wave, one slice at a time
        if (!mb.compact || MTW != 1) return hipErrorInvalidValue;
        launch_waves<AR, 1, MAXK, true, 1>(gd, bd, lds, st, b, mb, ctab, ke);
    } else if (!mb.compact || MTW != 1 || mb.sgroup == 1) {
        if (mb.compact)
            launch_waves<AR, MTW, MAXK, true, 1>(gd, bd, lds, st, b, mb, ctab, ke);
        else
            launch_waves<AR, MTW, MAXK, false, 1>(gd, bd, lds, st, b, mb, ctab, ke);
    } else if (mb.sgroup == 2) {
        launch_waves<AR, MTW, MAXK, true, (MTW == 1 ? 2 : 1)>(gd, bd, lds, st, b, mb, ctab, ke);
    } else {
        launch_waves<AR, MTW, MAXK, true, (MTW == 1 ? 3 : 1)>(gd, bd, lds, st, b, mb, ctab, ke);
    }
    return hipGetLastError();
}

size_t mfma_lds_bytes(const MfmaBucket& mb) {
    return (size_t)2 * 16 * mb.estride * 4 + (size_t)2 * mb.n_mslices * 16 * 4 + (size_t)2 * 16 * mb.stg_words * 4 +
           (size_t)mb.n_mslices * mb.n_pairs * 4 + 64 * 4;
}

bool mfma_variant_fits(int mtiles, int ksteps, int* nw, int* mtw, int* maxk, int prefer_mtw) {
    if (ksteps < 1 || ksteps > 8 || mtiles < 1) return false;
    int w = mtiles <= 8 ? mtiles : 8;
    if (prefer_mtw == 2 && mtiles > 1) w = (mtiles + 1) / 2 < 8 ? (mtiles + 1) / 2 : 8;
    int t = (mtiles + w - 1) / w;
    if (t > 2) return false;
    *nw = w;
    *mtw = t;
    *maxk = ksteps;
    return true;
}

MfmaBucket mfma_for(const MfmaBucket& mb0, int arith) {
    MfmaBucket mb = mb0;
    mb.arith = arith;
    mb.afrag = mb0.afrag_x[arith];
    if (arith != AR_F16) {
        // BF16 / F32: compact columns only, one M tile per wave, one slice per group
        if (mb0.presum) {
            // the F16 launch reads presummed columns; these take the compact
            // layout (eight K rows per pair)
            mb.presum = 0;
            mb.kq = mb0.ck_kq;
            mb.K = mb0.ck_K;
            mb.ksteps = mb.maxk = mb0.ck_ksteps;
        }
        mb.nw = mb.mtiles;
        mb.mtw = 1;
        mb.sgroup = 1;
        mb.cstride = mfma_words_per_pair(arith) * (mb.n_pairs + 1);
        mb.estride = (mb.n_mslices * mb.cstride) | 1;  // odd: conflict-free 32-bit reads across elements
    }
    return mb;
}

template <int AR, int MTW>
static hipError_t launch_mtw(const DevBucket& b, const MfmaBucket& mb, const float* ctab, float* ke, hipStream_t st) {
    switch (mb.maxk) {
        case 1: return launch_variant<AR, MTW, 1>(b, mb, ctab, ke, st);
        case 2: return launch_variant<AR, MTW, 2>(b, mb, ctab, ke, st);
        case 3: return launch_variant<AR, MTW, 3>(b, mb, ctab, ke, st);
        case 4: return launch_variant<AR, MTW, 4>(b, mb, ctab, ke, st);
        case 5: return launch_variant<AR, MTW, 5>(b, mb, ctab, ke, st);
        case 6: return launch_variant<AR, MTW, 6>(b, mb, ctab, ke, st);
        case 7: return launch_variant<AR, MTW, 7>(b, mb, ctab, ke, st);
        case 8: return launch_variant<AR, MTW, 8>(b, mb, ctab, ke, st);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_mfma(const DevBucket& b, const MfmaBucket& mb, const float* ctab, float* ke, hipStream_t st) {
    if (b.n_elem == 0) return hipSuccess;
    if (!((mb.arith_ok >> mb.arith) & 1u) || !mb.afrag) return hipErrorInvalidValue;
    if (mb.arith == AR_BF16) return mb.mtw == 1 ? launch_mtw<AR_BF16, 1>(b, mb, ctab, ke, st) : hipErrorInvalidValue;
    if (mb.arith == AR_F32) return mb.mtw == 1 ? launch_mtw<AR_F32, 1>(b, mb, ctab, ke, st) : hipErrorInvalidValue;
    if (mb.mtw == 1) return launch_mtw<AR_F16, 1>(b, mb, ctab, ke, st);
    if (mb.mtw == 2) return launch_mtw<AR_F16, 2>(b, mb, ctab, ke, st);
    return hipErrorInvalidValue;
}

// In-bounds probes of the operand layouts: one wave, one MFMA, out[m * 16 + n]
// (256 floats) = D[m][n] written with the assumed D layout.
//   kind 0 (f16 16x16x32): A[m][k] = m + k / 64, B[k][n] = (k == n) + (k == n + 16)
//   kind 1 (bf16 16x16x32), 2 (f32 16x16x4), 3 (f64 16x16x4):
//     A[m][k] = (m + 2 k) % 7 - 3, B[k][n] = (3 k + n) % 5 - 2 (small integers: exact)
__global__ void mfma_layout_probe(int kind, float* out) {
    const int lane = threadIdx.x & 63;
    const int lrow = lane % 16, lgrp = lane / 16;
    auto ia = [](int m, int k) { return (float)((m + 2 * k) % 7 - 3); };
    auto ib = [](int k, int n) { return (float)((3 * k + n) % 5 - 2); };
    float4v c = {0.f, 0.f, 0.f, 0.f};
    if (kind == 0) {
        half8 a, bb;
        for (int v = 0; v < 8; v++) {
            const int k = 8 * lgrp + v;
            a[v] = (_Float16)(lrow + k / 64.f);
            bb[v] = (_Float16)((k == lrow) || (k == lrow + 16) ? 1.f : 0.f);
        }
        c = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, bb, c, 0, 0, 0);
    } else if (kind == 1) {
        bf16x8 a, bb;
        for (int v = 0; v < 8; v++) {
            const int k = 8 * lgrp + v;
            a[v] = (__bf16)ia(lrow, k);
            bb[v] = (__bf16)ib(k, lrow);
        }
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bb, c, 0, 0, 0);
    } else if (kind == 2) {
        c = __builtin_amdgcn_mfma_f32_16x16x4f32(ia(lrow, lgrp), ib(lgrp, lrow), c, 0, 0, 0);
    } else {
        typedef double double4v __attribute__((ext_vector_type(4)));
        double4v d = {0.0, 0.0, 0.0, 0.0};
        d = __builtin_amdgcn_mfma_f64_16x16x4f64((double)ia(lrow, lgrp), (double)ib(lgrp, lrow), d, 0, 0, 0);
        for (int v = 0; v < 4; v++) c[v] = (float)d[v];
    }
    // D layout: lane l holds D[4 (l / 16) + v][l % 16] for the f32-accumulating
    // kinds, D[l / 16 + 4 v][l % 16] for v_mfma_f64_16x16x4_f64
    for (int v = 0; v < 4; v++) {
        const int m = kind == 3 ? lgrp + 4 * v : 4 * lgrp + v, n = lrow;
        out[m * 16 + n] = c[v];
    }
}

hipError_t launch_mfma_probe(int kind, float* out, hipStream_t st) {
    if (kind < 0 || kind > 3) return hipErrorInvalidValue;
    hipLaunchKernelGGL(mfma_layout_probe, dim3(1), dim3(64), 0, st, kind, out);
    return hipGetLastError();
}

}  // namespace slx
