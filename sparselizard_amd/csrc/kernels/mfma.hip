// Mixed-precision element contraction on MFMA (gfx950 v_mfma_f32_16x16x32_f16).
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
typedef unsigned int uint4v __attribute__((ext_vector_type(4)));

// Column preparation (per workgroup tile of NE elements): for every
// (element, slice) column the coefficient values C[p][q] are read from the
// coalesced coefficient table, scaled by a power of two and split.  A column
// lives in LDS as packed 32-bit words (hi | lo' << 16):
//   compact  (coefficients constant over the Gauss points): word p = C[p].
//            The K index is k = p * kq + q with the Gauss points of a pair
//            padded to kq = n_gp rounded up to 8 (zero rows of Phi), so the
//            eight consecutive k of a lane share one pair: the B fragment is
//            one LDS word broadcast.  Word n_pairs is 0 (K padding).
//   expanded otherwise: word k = C[k / n_gp][k % n_gp] for k < K, 0 up to
//            the padded K, read as two 16-byte LDS loads per K step.
// Output: the element rows of every segment group are assembled in LDS for
// 16 elements and written with coalesced 16-byte stores (one contiguous
// range per group).
__device__ inline uint32_t pack_split(float x) {
    _Float16 hi = (_Float16)x;
    _Float16 lo = (_Float16)((x - (float)hi) * 2048.f);
    return (uint32_t)__builtin_bit_cast(uint16_t, hi) | ((uint32_t)__builtin_bit_cast(uint16_t, lo) << 16);
}

// Work split (one workgroup per CU, persistent over 16-element tiles):
//   consumer waves (one or two M tiles each) run the contraction of tile s
//     from the column buffer Cs[s & 1] into the staging buffer stg[s & 1];
//   producer waves (SLX_MFMA_PWAVES, in two groups) meanwhile prepare the
//     columns of tile s + 1 into Cs[(s + 1) & 1] and write tile s - 1 from
//     stg[(s - 1) & 1] to the element matrices with coalesced 16-byte stores.
// One barrier per tile; every buffer is written one tile after its last read.
template <int MTW, int MAXK, bool COMPACT, int G>
__global__ void __launch_bounds__(64 * (8 + SLX_MFMA_PWAVES)) mfma_f16split_kernel(DevBucket b, MfmaBucket mb, const float* __restrict__ ctab,
                                                                                  float* __restrict__ ke) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    constexpr int NE = 16;
    const int ns = mb.n_mslices, np = mb.n_pairs, nq = mb.nq, ngp = b.n_gp;
    const int estride = mb.estride, cstride = mb.cstride;
    const int n_tf = mb.n_tf;
    uint32_t* Csb = reinterpret_cast<uint32_t*>(smem_raw);                     // [2][NE][estride]
    float* colscb = reinterpret_cast<float*>(Csb + (size_t)2 * NE * estride);   // [2][ns][NE]
    float* stgb = colscb + (size_t)2 * ns * NE;                                 // [2][16 * stg_words]
    int* csl = reinterpret_cast<int*>(stgb + (size_t)2 * 16 * mb.stg_words);    // [ns][np] coefficient slots
    const SLX_CONST int32_t* csrc = (const SLX_CONST int32_t*)mb.csrc;
    const SLX_CONST int64_t* gbase = (const SLX_CONST int64_t*)mb.gbase;
    const SLX_CONST int32_t* gld = (const SLX_CONST int32_t*)mb.gld;
    const SLX_CONST int32_t* sgrp = (const SLX_CONST int32_t*)mb.sgrp;
    const SLX_CONST int32_t* scol = (const SLX_CONST int32_t*)mb.scol;
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
        half8 Ahi[MTW][MAXK], Alo[MTW][MAXK];
        int irow[MTW][4], jrow[MTW][4];
        bool rowok[MTW][4];
#pragma unroll
        for (int i = 0; i < MTW; i++) {
            const int mt = wave + nw * i;
#pragma unroll
            for (int ks = 0; ks < MAXK; ks++) {
                if (mt < mb.mtiles) {
                    const half8* src = reinterpret_cast<const half8*>(mb.afrag) + ((size_t)(mt * MAXK + ks) * 2) * 64;
                    Ahi[i][ks] = src[lane];
                    Alo[i][ks] = src[64 + lane];
                } else {
#pragma unroll
                    for (int v = 0; v < 8; v++) {
                        Ahi[i][ks][v] = (_Float16)0.f;
                        Alo[i][ks][v] = (_Float16)0.f;
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
        // compact columns: the pair word of this lane's eight k per K step
        int pword[MAXK];
#pragma unroll
        for (int ks = 0; ks < MAXK; ks++) pword[ks] = min((ks * 32 + 8 * lgrp) / mb.kq, np);
        __syncthreads();  // tile 0 prepared
        for (int st = 0; st < nsteps; st++) {
            const uint32_t* Cs = Csb + (size_t)(st & 1) * NE * estride;
            const float* colsc = colscb + (size_t)(st & 1) * ns * NE;
            float* stg = stgb + (size_t)(st & 1) * 16 * mb.stg_words;
            // slices G at a time (G divides the slice count): 2 G independent
            // accumulator chains per M tile (hi*hi, and hi*lo' + lo'*hi) share the
            // A fragments; every LDS word of the group is read before the MFMAs
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
                uint32_t cw[G][COMPACT ? MAXK : 1];
                if (COMPACT) {
#pragma unroll
                    for (int c = 0; c < G; c++)
#pragma unroll
                        for (int ks = 0; ks < MAXK; ks++) cw[c][ks] = colbase0[c * cstride + pword[ks]];
                }
#pragma unroll
                for (int ks = 0; ks < MAXK; ks++) {
#pragma unroll
                    for (int c = 0; c < G; c++) {
                        uint4v hw, lw;
                        if (COMPACT) {
                            // one packed word (hi | lo' << 16) -> both halves broadcast
                            // to all eight k of the lane (byte permutes)
                            const uint32_t u = cw[c][COMPACT ? ks : 0];
                            const uint32_t h2 = __builtin_amdgcn_perm(u, u, 0x01000100u);
                            const uint32_t l2 = __builtin_amdgcn_perm(u, u, 0x03020302u);
                            hw = (uint4v){h2, h2, h2, h2};
                            lw = (uint4v){l2, l2, l2, l2};
                        } else {
                            const uint32_t* colbase = colbase0 + c * cstride;
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
                            acc_hh[c][i] = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ahi[i][ks], bhi, acc_hh[c][i], 0, 0, 0);
                            acc_x[c][i] = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ahi[i][ks], blo, acc_x[c][i], 0, 0, 0);
                            acc_x[c][i] = __builtin_amdgcn_mfma_f32_16x16x32_f16(Alo[i][ks], bhi, acc_x[c][i], 0, 0, 0);
                        }
                    }
                }
#pragma unroll
                for (int c = 0; c < G; c++) {
                    const int s = s0 + c;
                    const int g = sgrp[s];
                    float* gstg = stg;  // LDS block of group g: [16][n_tf * ld + 4]
                    for (int g2 = 0; g2 < g; g2++) gstg += (size_t)16 * (n_tf * gld[g2] + 4);
                    const int ld = gld[g], srow = n_tf * ld + 4;
                    const float colscale = colsc[s * NE + lrow];
                    const float xscale = colscale * (1.f / 2048.f);
                    float* out = gstg + (size_t)lrow * srow + scol[s];
#pragma unroll
                    for (int i = 0; i < MTW; i++) {
                        const int mt = wave + nw * i;
                        if (mt >= mb.mtiles) continue;
#pragma unroll
                        for (int v = 0; v < 4; v++)
                            if (rowok[i][v]) out[irow[i][v] * ld + jrow[i][v]] = acc_hh[c][i][v] * colscale + acc_x[c][i][v] * xscale;
                    }
                }
            }
            __syncthreads();  // tile st staged; tile st + 1 prepared
        }
    } else {
        // ------------------------------------------------------- producers --
        // two producer groups: the first prepares columns, the second writes out
        const int np_threads = 32 * SLX_MFMA_PWAVES;
        const int pt0 = threadIdx.x - 64 * nw;
        const bool is_prep = pt0 < np_threads;  // wave-uniform
        const int pt = is_prep ? pt0 : pt0 - np_threads;
        // columns of a tile: one thread per (element, slice), element fastest:
        // power-of-two scale and packed hi/lo' words
        auto prepare = [&](int tile, int buf) {
            const int e0 = tile * NE;
            const int nloc = min(NE, b.n_elem - e0);
            uint32_t* Cs = Csb + (size_t)buf * NE * estride;
            float* colsc = colscb + (size_t)buf * ns * NE;
            for (int col = pt; col < NE * ns; col += np_threads) {
                const int el = col % NE, s = col / NE;
                const bool valid = el < nloc;
                const int64_t e = e0 + (valid ? el : 0);
                uint32_t* dst = Cs + (size_t)el * estride + (size_t)s * cstride;
                float mx = 0.f;
                if (COMPACT) {
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
                        if (p < np) dst[p] = pack_split(cv[p] * sc);
                    dst[np] = 0u;
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
                    for (int p = 0; p < np; p++) {
                        const int c = csl[s * np + p];
                        for (int q0 = 0; q0 < ngp; q0 += 8) {
#pragma unroll
                            for (int u = 0; u < 8; u++) {
                                const int q = q0 + u;
                                if (q >= ngp) break;
                                const float x = (c >= 0 && valid) ? ctab[((int64_t)c * nq + (nq == 1 ? 0 : q)) * nel + e] * sc : 0.f;
                                dst[p * ngp + q] = pack_split(x);
                            }
                        }
                    }
                    for (int k = np * ngp; k < cstride; k++) dst[k] = 0u;
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
                    const int step_e = np_threads / per, step_w = np_threads - step_e * per;
                    for (int x = pt; x < nel_here * per; x += np_threads) {
                        d4[x] = *reinterpret_cast<const float4v*>(gsrc + (size_t)eo * srow + 4 * w);
                        eo += step_e;
                        w += step_w;
                        if (w >= per) {
                            w -= per;
                            eo++;
                        }
                    }
                } else {
                    for (int x = pt; x < nel_here * erow; x += np_threads) {
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
                if (p < np) dst[p] = pack_split(cvn[p] * sc);
            dst[np] = 0u;
        };
        const bool pipelined = COMPACT && NE * ns <= np_threads;
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

template <int MTW, int MAXK>
static hipError_t launch_variant(const DevBucket& b, const MfmaBucket& mb, const float* ctab, float* ke, hipStream_t st) {
    const size_t lds = mfma_lds_bytes(mb);
    const int tiles = (b.n_elem + 15) / 16;
    const int grid = tiles < mb.max_blocks ? tiles : mb.max_blocks;
    const dim3 gd(grid), bd(64 * (mb.nw + SLX_MFMA_PWAVES));
    if (!mb.compact || MTW != 1 || mb.sgroup == 1) {
        if (mb.compact)
            hipLaunchKernelGGL((mfma_f16split_kernel<MTW, MAXK, true, 1>), gd, bd, lds, st, b, mb, ctab, ke);
        else
            hipLaunchKernelGGL((mfma_f16split_kernel<MTW, MAXK, false, 1>), gd, bd, lds, st, b, mb, ctab, ke);
    } else if (mb.sgroup == 2) {
        hipLaunchKernelGGL((mfma_f16split_kernel<MTW, MAXK, true, (MTW == 1 ? 2 : 1)>), gd, bd, lds, st, b, mb, ctab, ke);
    } else {
        hipLaunchKernelGGL((mfma_f16split_kernel<MTW, MAXK, true, (MTW == 1 ? 3 : 1)>), gd, bd, lds, st, b, mb, ctab, ke);
    }
    return hipGetLastError();
}

size_t mfma_lds_bytes(const MfmaBucket& mb) {
    return (size_t)2 * 16 * mb.estride * 4 + (size_t)2 * mb.n_mslices * 16 * 4 + (size_t)2 * 16 * mb.stg_words * 4 +
           (size_t)mb.n_mslices * mb.n_pairs * 4;
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

template <int MTW>
static hipError_t launch_mtw(const DevBucket& b, const MfmaBucket& mb, const float* ctab, float* ke, hipStream_t st) {
    switch (mb.maxk) {
        case 1: return launch_variant<MTW, 1>(b, mb, ctab, ke, st);
        case 2: return launch_variant<MTW, 2>(b, mb, ctab, ke, st);
        case 3: return launch_variant<MTW, 3>(b, mb, ctab, ke, st);
        case 4: return launch_variant<MTW, 4>(b, mb, ctab, ke, st);
        case 5: return launch_variant<MTW, 5>(b, mb, ctab, ke, st);
        case 6: return launch_variant<MTW, 6>(b, mb, ctab, ke, st);
        case 7: return launch_variant<MTW, 7>(b, mb, ctab, ke, st);
        case 8: return launch_variant<MTW, 8>(b, mb, ctab, ke, st);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_mfma_f16split(const DevBucket& b, const MfmaBucket& mb, const float* ctab, float* ke, hipStream_t st) {
    if (b.n_elem == 0) return hipSuccess;
    if (mb.mtw == 1) return launch_mtw<1>(b, mb, ctab, ke, st);
    if (mb.mtw == 2) return launch_mtw<2>(b, mb, ctab, ke, st);
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
