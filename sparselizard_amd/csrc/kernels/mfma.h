// Host/device description of the MFMA contraction of one bucket.
#pragma once
#include <hip/hip_runtime.h>
#include "device.h"

namespace slx {

struct MfmaBucket {
    int32_t M, K;               // rows (n_tf * n_dof) and K = n_pairs * n_gp
    int32_t n_tf, n_dof;
    int32_t mtiles, ksteps;     // ceil(M / 16), ceil(K / 32)
    int32_t n_pairs, nq;        // derivative pairs; 1 when the coefficients do not vary over Gauss points
    int32_t ne_tile;            // elements per workgroup tile
    int32_t n_mslices;          // matrix slices handled
    int32_t nw, mtw, maxk;      // waves per workgroup, M tiles per wave, K steps (template)
    int32_t compact;            // compact (constant-coefficient) column layout
    int32_t cstride, estride;   // LDS words per column / per element (all slices)
    int32_t max_blocks;         // persistent grid size
    float a_unscale;            // 2^-kA of the scaled Phi
    const void* afrag;          // [mtiles][ksteps][2 (hi, lo')][64 lanes] x 8 halfs
    const int32_t* csrc;        // [n_mslices][n_pairs] coefficient slot of (slice, pair), or -1
    const int64_t* kebase;      // [n_mslices] Ke base of each slice
    const int64_t* keld;        // [n_mslices] Ke segment length of each slice
    int32_t slice_idx[SLX_MAX_SLICES];
};

size_t mfma_lds_bytes(const MfmaBucket& mb);
bool mfma_variant_fits(int mtiles, int ksteps, int* nw, int* mtw, int* maxk);
hipError_t launch_mfma_f16split(const DevBucket& b, const MfmaBucket& mb, const float* ctab, float* ke, hipStream_t st);
hipError_t launch_mfma_probe(float* out, hipStream_t st);

}  // namespace slx
