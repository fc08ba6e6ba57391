// Host/device description of the MFMA contraction of one bucket.
#pragma once
#include <hip/hip_runtime.h>
#include "device.h"

#define SLX_MFMA_MAXP 16   // derivative pairs of a batched-MFMA bucket
#define SLX_MFMA_PREP_WAVES 3   // producer waves of the batched-MFMA workgroup that prepare columns
// producer waves that write element matrices out: the rest of a 16-wave
// workgroup when the consumers' A fragments are small (at least 5), else 2
#define SLX_MFMA_MIN_WRITE_WAVES 2

namespace slx {

// MFMA arithmetic of the contraction kernels (FP32 accumulation in all three):
//   AR_F16   v_mfma_f32_16x16x32_f16, operands split x = hi + 2^-11 lo' (3 products)
//   AR_BF16  v_mfma_f32_16x16x32_bf16, operands split x = hi + mid + lo (6 products)
//   AR_F32   v_mfma_f32_16x16x4_f32, plain FP32 operands (1 product)
enum MfmaArith { AR_F16 = 0, AR_BF16 = 1, AR_F32 = 2, AR_COUNT = 3 };

// 16-byte fragment planes per (M tile, K step of 32, lane) of the A operand:
// F16 hi, lo'; BF16 hi, mid, lo; F32 k = 4 j + lane / 16 for j = 0..3, 4..7.
inline int mfma_planes(int arith) { return arith == AR_BF16 ? 3 : 2; }
// LDS words per derivative pair of a compact column: F16 (hi | lo' << 16),
// BF16 (hi | mid << 16, lo), F32 the value.
inline int mfma_words_per_pair(int arith) { return arith == AR_BF16 ? 2 : 1; }

struct MfmaBucket {
    int32_t M, K;               // rows (n_tf * n_dof) and K = n_pairs * kq
    int32_t kq;                 // K entries per derivative pair (n_gp; 8 when compact: the Gauss-point sum, 7 zero rows)
    int32_t n_tf, n_dof;
    int32_t mtiles, ksteps;     // ceil(M / 16), ceil(K / 32)
    int32_t n_pairs, nq;        // derivative pairs; 1 when the coefficients do not vary over Gauss points
    int32_t ne_tile;            // elements per workgroup tile (16)
    int32_t n_mslices;          // matrix slices handled
    int32_t nw, mtw, maxk;      // waves per workgroup, M tiles per wave, K steps (template, = ksteps)
    int32_t compact;            // compact (constant-coefficient) column layout
    int32_t sgroup;             // slices per contraction group (divides n_mslices; 1 unless compact, one M tile per wave)
    int32_t cstride, estride;   // LDS words per column / per element (all slices)
    int32_t max_blocks;         // persistent grid size
    int32_t arith;              // MfmaArith of this launch (mfma_for)
    float a_unscale;            // 2^-kA of the scaled Phi
    const void* afrag;          // [mtiles][ksteps][planes][64 lanes] x 16 bytes, of this launch's arithmetic
    const int32_t* csrc;        // [n_mslices][n_pairs] coefficient slot of (slice, pair), or -1
    int32_t n_groups, max_gld;  // segment groups of the slices, longest group segment
    int32_t stg_words;          // LDS staging words per element of all groups: sum (n_tf * ld + 4)
    const int64_t* gbase;       // [n_groups] Ke base of each group
    const int32_t* gld;         // [n_groups] segment length of each group
    const int32_t* sgrp;        // [n_mslices] group of each slice
    const int32_t* scol;        // [n_mslices] column offset of each slice in its group
    int32_t slice_idx[SLX_MAX_SLICES];
    // [n_mslices][3] per slice, for the consumers' output stores into the LDS
    // staging of a tile: word offset of the slice's first column in its
    // group's block, the group's segment length, the staging words per element
    const int32_t* sprm;
    // host side: the fragments of every arithmetic and which arithmetics fit
    const void* afrag_x[AR_COUNT];
    uint32_t arith_ok;          // bit a: arithmetic a has a kernel variant for this bucket
    // presum: the F16 launch reads presummed columns (one K row per pair, kq =
    // 1, 16 words per slice: a lane reads the words of eight pairs); ck_*: the
    // compact layout (kq = 8) of the BF16 and F32 launches
    int32_t presum;
    int32_t ck_kq, ck_K, ck_ksteps;
};

// The launch description of one arithmetic (fragments, column strides, waves).
MfmaBucket mfma_for(const MfmaBucket& mb, int arith);
size_t mfma_lds_bytes(const MfmaBucket& mb);
// Waves per workgroup and M tiles per wave: one tile per wave up to 8 waves,
// or (prefer_mtw == 2) two tiles per wave so each B fragment feeds 6 MFMAs.
bool mfma_variant_fits(int mtiles, int ksteps, int* nw, int* mtw, int* maxk, int prefer_mtw);
// Contraction of the bucket's matrix slices in the arithmetic of mb (from mfma_for).
hipError_t launch_mfma(const DevBucket& b, const MfmaBucket& mb, const float* ctab, float* ke, hipStream_t st);
// One MFMA of the given arithmetic (AR_*, or 3 for v_mfma_f64_16x16x4_f64) on
// fixed operands: out = the 16x16 result, row-major (operand layout check).
hipError_t launch_mfma_probe(int kind, float* out, hipStream_t st);

}  // namespace slx
