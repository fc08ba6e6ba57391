// Host/device description of the MFMA contraction of one bucket.
#pragma once
#include <hip/hip_runtime.h>
#include "device.h"

#define SLX_MFMA_MAXP 16   // derivative pairs of a batched-MFMA bucket
#define SLX_MFMA_PWAVES 8  // producer waves of the batched-MFMA workgroup (half prepare, half write)

namespace slx {

struct MfmaBucket {
    int32_t M, K;               // rows (n_tf * n_dof) and K = n_pairs * kq
    int32_t kq;                 // K entries per derivative pair (n_gp, padded to 8 when compact)
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
    float a_unscale;            // 2^-kA of the scaled Phi
    const void* afrag;          // [mtiles][ksteps][2 (hi, lo')][64 lanes] x 8 halfs
    const int32_t* csrc;        // [n_mslices][n_pairs] coefficient slot of (slice, pair), or -1
    int32_t n_groups, max_gld;  // segment groups of the slices, longest group segment
    int32_t stg_words;          // LDS staging words per element of all groups: sum (n_tf * ld + 4)
    const int64_t* gbase;       // [n_groups] Ke base of each group
    const int32_t* gld;         // [n_groups] segment length of each group
    const int32_t* sgrp;        // [n_mslices] group of each slice
    const int32_t* scol;        // [n_mslices] column offset of each slice in its group
    int32_t slice_idx[SLX_MAX_SLICES];
};

size_t mfma_lds_bytes(const MfmaBucket& mb);
// Waves per workgroup and M tiles per wave: one tile per wave up to 8 waves,
// or (prefer_mtw == 2) two tiles per wave so each B fragment feeds 6 MFMAs.
bool mfma_variant_fits(int mtiles, int ksteps, int* nw, int* mtw, int* maxk, int prefer_mtw);
hipError_t launch_mfma_f16split(const DevBucket& b, const MfmaBucket& mb, const float* ctab, float* ke, hipStream_t st);
hipError_t launch_mfma_probe(float* out, hipStream_t st);

}  // namespace slx
