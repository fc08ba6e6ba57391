// One wave per element holding its 16x16 MFMA tiles (up to 4 x 4) for the
// smaller slices with Gauss-point-varying coefficients (kernels/tile.h).
//
// A workgroup of SLX_TK_WAVES waves loops over a contiguous range of the
// bucket's elements, one element per wave at a time.  Up to 2 x 2 tiles, the
// weighted dof table (w_q dof_b[j][q], [q][j][4 derivatives]) and the P
// fragments (the shared tf table in the launch's arithmetic, scaled on the
// host) go to LDS once per workgroup and stay there for the whole launch; the
// 3 x 3 and 4 x 4 variants read both through the caches.  Per element: the
// lanes of the Gauss points sum
// the element's term coefficients into the 4x4 coefficient matrices
// D[q][a][b] (LDS, one region per wave) and the wave takes the element's
// power-of-two scale; then per k-chunk every lane builds its 8 entries of Q_e
//   Q_e[k][j] = sc * sum_b D[q][a][b] * w_q dof_b[j][q]
// splits them into f16 hi/lo and issues three v_mfma_f32_16x16x32_f16
// (hi*hi + hi*lo + lo*hi) into the element's FP32 accumulator, or, in the BF16
// split, into bf16 hi/mid/lo with six v_mfma_f32_16x16x32_bf16 products, or,
// in FP32, eight v_mfma_f32_16x16x4_f32 steps on the entries as they are.
#include "tile.h"

namespace slx {

typedef _Float16 half8t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8t __attribute__((ext_vector_type(8)));
typedef float float4t __attribute__((ext_vector_type(4)));
typedef unsigned int uint4t __attribute__((ext_vector_type(4)));

// arithmetic AR: 0 FP16 split, 1 BF16 split, 2 FP32 (v_mfma_f32_16x16x4_f32,
// lane group g's entry v at step v of the chunk's eight K = 4 steps)
template <int AR, int NA, int MTT, int NTT, bool GT>
__global__ void __launch_bounds__(64 * SLX_TK_WAVES) tile_qv_kernel(int n_elem, TileSlice t, const float* __restrict__ ctab, int qfast, float* __restrict__ ke) {
    constexpr int QH = tile_qh(NA);
    constexpr int NC = 16 * NTT;    // table columns
    constexpr bool BF = AR == 1, F32 = AR == 2;
    constexpr int NPL = BF ? 3 : 2;  // 16-byte operand words per lane: FP16 hi/lo, BF16 hi/mid/lo, FP32 entries 0..3 / 4..7
    static_assert(QH >= 1, "a lane group holds at least one Gauss point");
    // dynamic LDS: [ngpp][NC] float4 weighted dof table and [nkc][MTT][NPL][64] P
    // fragment words (both read from global memory instead when GT: the larger
    // tiles' tables would not fit), then [SLX_TK_WAVES][ngpp][4] float4
    // coefficient rows
    extern __shared__ __attribute__((aligned(16))) unsigned char tile_lds[];
    const int nkc = t.nkc, ngpp = t.ngpp, ngp = t.ngp, nq = t.nq, n_terms = t.n_terms;
    const float4t* tq4 = reinterpret_cast<const float4t*>(t.tq);
    const uint4t* af = reinterpret_cast<const uint4t*>(t.afrag);
    float4t* Tl = reinterpret_cast<float4t*>(tile_lds);
    uint4t* Pl = reinterpret_cast<uint4t*>(Tl + (GT ? 0 : ngpp * NC));
    float4t* Dall = reinterpret_cast<float4t*>(Pl + (GT ? 0 : nkc * MTT * NPL * 64));
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, g = lane >> 4, jl = lane & 15;
    typedef __attribute__((address_space(4))) const int32_t cint32;
    const cint32* t_ai = (const cint32*)(t.tinfo);
    const cint32* t_b = t_ai + n_terms;
    const cint32* t_slot = t_ai + 2 * n_terms;
    if constexpr (!GT) {
        for (int x = tid; x < ngpp * NC; x += blockDim.x) Tl[x] = tq4[x];
        for (int x = tid; x < nkc * MTT * NPL * 64; x += blockDim.x) Pl[x] = af[x];
    }
    const float4t* Tsrc = GT ? tq4 : Tl;
    const uint4t* Psrc = GT ? af : Pl;
    float4t* Dl = Dall + w * ngpp * 4;  // this wave's [q][a]
    float* dw = reinterpret_cast<float*>(Dl);
    const int64_t stride = (int64_t)nq * n_elem;
    // a contiguous range of elements per workgroup, SLX_TK_WAVES at a time: with
    // the point-major coefficient layout ([slot][point][element]) consecutive
    // iterations read neighbouring words of the same cache lines
    const int per = (int)(((int64_t)n_elem + gridDim.x - 1) / gridDim.x + SLX_TK_WAVES - 1) / SLX_TK_WAVES * SLX_TK_WAVES;
    const int64_t e_hi = (int64_t)per * (blockIdx.x + 1);
    const int e_end = (int)(e_hi < n_elem ? e_hi : (int64_t)n_elem);
    // the term coefficients of Gauss point q for element e; entry (slot, element,
    // point) at slot * stride + element * es + point * ps.  Lane q's are loaded
    // before any is used (and those of the next element while the current one's
    // MFMAs run); points q >= 64 (a second pass of the lanes) load on use.
    const int64_t es = qfast ? nq : 1, ps = nq == 1 ? 0 : (qfast ? 1 : (int64_t)n_elem);
    auto load_c = [&](int e, int q, float (&cv)[SLX_TK_MAXT]) {
        const bool on = e < e_end && q < ngp;
        const float* cq = ctab + (int64_t)e * es + (int64_t)q * ps;
#pragma unroll
        for (int tt = 0; tt < SLX_TK_MAXT; tt++) {
            float x = 0.f;
            if (tt < n_terms && on) x = cq[(int64_t)t_slot[tt] * stride];
            cv[tt] = x;
        }
    };
    float cv[SLX_TK_MAXT];
    load_c(per * blockIdx.x + w, lane, cv);
    for (int e0 = per * blockIdx.x; e0 < e_end; e0 += SLX_TK_WAVES) {
        const int e = e0 + w;
        const bool live = e < e_end;
        __syncthreads();  // the tables are in LDS; the previous element's coefficient rows are read
        // coefficient matrices of the element: lane q sums the terms of its Gauss point
        float m = 0.f;
        for (int q = lane; q < ngpp; q += 64) {
            float cw[SLX_TK_MAXT];
            if (q == lane) {
#pragma unroll
                for (int tt = 0; tt < SLX_TK_MAXT; tt++) cw[tt] = cv[tt];
            } else {
                load_c(e, q, cw);
            }
            float* mq = dw + q * 16;
#pragma unroll
            for (int x = 0; x < 16; x++) mq[x] = 0.f;
#pragma unroll
            for (int tt = 0; tt < SLX_TK_MAXT; tt++)
                if (tt < n_terms) mq[t_ai[tt] * 4 + t_b[tt]] += cw[tt];
#pragma unroll
            for (int a = 0; a < 4; a++) m = fmaxf(m, fabsf(mq[4 * a]) + fabsf(mq[4 * a + 1]) + fabsf(mq[4 * a + 2]) + fabsf(mq[4 * a + 3]));
        }
        for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
        // power-of-two scale: |Q| <= max_q,a sum_b |D| * max |tq|  ->  [128, 256)
        const float bound = m * t.tqmax;
        float sc = 1.f, unsc = t.a_unscale;
        if (bound > 0.f) {
            int ex = 0;
            (void)frexpf(bound, &ex);
            sc = ldexpf(1.f, 8 - ex);
            unsc = ldexpf(t.a_unscale, ex - 8);
        }
        __syncthreads();  // the coefficient rows are in LDS
        load_c(e + SLX_TK_WAVES, lane, cv);  // the next element's, in flight during the MFMAs
        float4t acc[MTT][NTT];
#pragma unroll
        for (int mt = 0; mt < MTT; mt++)
#pragma unroll
            for (int jt = 0; jt < NTT; jt++) acc[mt][jt] = float4t{0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
        for (int c = 0; c < nkc; c++) {
            const int q0 = (4 * c + g) * QH;
            // the lane's eight Q_e entries of each column tile, split into the planes
            uint4t qp[NTT][NPL];
#pragma unroll
            for (int v = 0; v < 8; v++) {
                const int qi = v / NA, a = v % NA;
                float4t d = {0.f, 0.f, 0.f, 0.f};
                if (qi < QH) d = Dl[(q0 + qi) * 4 + a];
#pragma unroll
                for (int jt = 0; jt < NTT; jt++) {
                    float val = 0.f;
                    if (qi < QH) {
                        const float4t h = Tsrc[(q0 + qi) * NC + jt * 16 + jl];
                        val = sc * (d[0] * h[0] + d[1] * h[1] + d[2] * h[2] + d[3] * h[3]);
                    }
                    if constexpr (F32) {
                        float4t x = __builtin_bit_cast(float4t, qp[jt][v >> 2]);
                        x[v & 3] = val;
                        qp[jt][v >> 2] = __builtin_bit_cast(uint4t, x);
                    } else if constexpr (BF) {
                        bf16x8t hv = __builtin_bit_cast(bf16x8t, qp[jt][0]), mv = __builtin_bit_cast(bf16x8t, qp[jt][1]),
                                lv = __builtin_bit_cast(bf16x8t, qp[jt][2]);
                        const __bf16 hi = (__bf16)val;
                        const float r1 = val - (float)hi;
                        const __bf16 mid = (__bf16)r1;
                        hv[v] = hi;
                        mv[v] = mid;
                        lv[v] = (__bf16)(r1 - (float)mid);
                        qp[jt][0] = __builtin_bit_cast(uint4t, hv);
                        qp[jt][1] = __builtin_bit_cast(uint4t, mv);
                        qp[jt][2] = __builtin_bit_cast(uint4t, lv);
                    } else {
                        half8t hv = __builtin_bit_cast(half8t, qp[jt][0]), lv = __builtin_bit_cast(half8t, qp[jt][1]);
                        const _Float16 hi = (_Float16)val;
                        hv[v] = hi;
                        lv[v] = (_Float16)(val - (float)hi);
                        qp[jt][0] = __builtin_bit_cast(uint4t, hv);
                        qp[jt][1] = __builtin_bit_cast(uint4t, lv);
                    }
                }
            }
#pragma unroll
            for (int mt = 0; mt < MTT; mt++) {
                uint4t pp[NPL];
#pragma unroll
                for (int pl = 0; pl < NPL; pl++) pp[pl] = Psrc[((c * MTT + mt) * NPL + pl) * 64 + lane];
#pragma unroll
                for (int jt = 0; jt < NTT; jt++) {
                    if constexpr (F32) {
#pragma unroll
                        for (int v = 0; v < 8; v++)
                            acc[mt][jt] = __builtin_amdgcn_mfma_f32_16x16x4f32(__builtin_bit_cast(float4t, pp[v >> 2])[v & 3],
                                                                             __builtin_bit_cast(float4t, qp[jt][v >> 2])[v & 3], acc[mt][jt], 0, 0, 0);
                    } else if constexpr (BF) {
                        const bf16x8t ah = __builtin_bit_cast(bf16x8t, pp[0]), am = __builtin_bit_cast(bf16x8t, pp[1]),
                                      al = __builtin_bit_cast(bf16x8t, pp[2]);
                        const bf16x8t xh = __builtin_bit_cast(bf16x8t, qp[jt][0]), xm = __builtin_bit_cast(bf16x8t, qp[jt][1]),
                                      xl = __builtin_bit_cast(bf16x8t, qp[jt][2]);
                        acc[mt][jt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, xh, acc[mt][jt], 0, 0, 0);
                        acc[mt][jt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, xm, acc[mt][jt], 0, 0, 0);
                        acc[mt][jt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, xh, acc[mt][jt], 0, 0, 0);
                        acc[mt][jt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, xl, acc[mt][jt], 0, 0, 0);
                        acc[mt][jt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, xm, acc[mt][jt], 0, 0, 0);
                        acc[mt][jt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, xh, acc[mt][jt], 0, 0, 0);
                    } else {
                        const half8t ph = __builtin_bit_cast(half8t, pp[0]), pl = __builtin_bit_cast(half8t, pp[1]);
                        const half8t hv = __builtin_bit_cast(half8t, qp[jt][0]), lv = __builtin_bit_cast(half8t, qp[jt][1]);
                        acc[mt][jt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ph, hv, acc[mt][jt], 0, 0, 0);
                        acc[mt][jt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ph, lv, acc[mt][jt], 0, 0, 0);
                        acc[mt][jt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(pl, hv, acc[mt][jt], 0, 0, 0);
                    }
                }
            }
        }
        // rows i = 16 mt + 4 g + r, column j = 16 jt + lane % 16
        if (live) {
            float* out = ke + t.ke_base + (int64_t)e * t.n_tf * t.ke_ld;
#pragma unroll
            for (int jt = 0; jt < NTT; jt++) {
                const int j = jt * 16 + jl;
                if (j >= t.n_dof) continue;
#pragma unroll
                for (int mt = 0; mt < MTT; mt++)
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int i = mt * 16 + 4 * g + r;
                        if (i < t.n_tf) __builtin_nontemporal_store(acc[mt][jt][r] * unsc, out + (int64_t)i * t.ke_ld + j);
                    }
            }
        }
    }
}

// P and the weighted dof table in LDS for up to 2 x 2 tiles, read through the
// caches for 3 x 3 and 4 x 4 (the square tile count of the larger side)
static bool tile_global_tables(const TileSlice& t) { return t.mtt > 2 || t.ntt > 2; }

size_t tile_lds_bytes(const TileSlice& t, int planes) {
    const size_t rows = (size_t)SLX_TK_WAVES * t.ngpp * 4 * 16;
    if (tile_global_tables(t)) return rows;
    return (size_t)t.ngpp * 16 * t.ntt * 16 + (size_t)t.nkc * t.mtt * planes * 64 * 16 + rows;
}

hipError_t launch_tile(const DevBucket& b, const TileSlice& t0, int arith, const float* ctab, int qfast, float* ke, int n_cu, hipStream_t st) {
    if (arith < 0 || arith > 2) return hipErrorInvalidValue;
    TileSlice t = t0;
    t.afrag = t0.afrag_x[arith];
    const int planes = arith == 1 ? 3 : 2;
    if (b.n_elem == 0) return hipSuccess;
    if (t.nkc < 1 || t.nkc > SLX_TK_MAXC || t.ngpp > SLX_TK_MAXQ || t.ngpp != t.nkc * 4 * tile_qh(t.na) || t.mtt < 1 || t.mtt > 4 ||
        t.ntt < 1 || t.ntt > 4 || (tile_global_tables(t) && t.mtt != t.ntt) || t.n_tf > 16 * t.mtt || t.n_dof > 16 * t.ntt ||
        !t.afrag || tile_lds_bytes(t, planes) > 64 * 1024)
        return hipErrorInvalidValue;
    const size_t lds = tile_lds_bytes(t, planes);
    const int64_t groups = ((int64_t)b.n_elem + SLX_TK_WAVES - 1) / SLX_TK_WAVES;
    const unsigned grid = (unsigned)std::min<int64_t>(groups, (int64_t)n_cu * 16);
    const dim3 block(64 * SLX_TK_WAVES);
#define SLX_TILE_LAUNCH(NA_, M_, N_, G_)                                                                                        \
    if (arith == 1) hipLaunchKernelGGL((tile_qv_kernel<1, NA_, M_, N_, G_>), dim3(grid), block, lds, st, b.n_elem, t, ctab, qfast, ke); \
    else if (arith == 2) hipLaunchKernelGGL((tile_qv_kernel<2, NA_, M_, N_, G_>), dim3(grid), block, lds, st, b.n_elem, t, ctab, qfast, ke); \
    else hipLaunchKernelGGL((tile_qv_kernel<0, NA_, M_, N_, G_>), dim3(grid), block, lds, st, b.n_elem, t, ctab, qfast, ke)
#define SLX_TILE_SHAPES(NA_)                                            \
    if (t.mtt == 1 && t.ntt == 1) { SLX_TILE_LAUNCH(NA_, 1, 1, false); }        \
    else if (t.mtt == 1 && t.ntt == 2) { SLX_TILE_LAUNCH(NA_, 1, 2, false); }   \
    else if (t.mtt == 2 && t.ntt == 1) { SLX_TILE_LAUNCH(NA_, 2, 1, false); }   \
    else if (t.mtt == 2 && t.ntt == 2) { SLX_TILE_LAUNCH(NA_, 2, 2, false); }   \
    else if (t.mtt == 3) { SLX_TILE_LAUNCH(NA_, 3, 3, true); }                  \
    else { SLX_TILE_LAUNCH(NA_, 4, 4, true); }
    switch (t.na) {
        case 1: SLX_TILE_SHAPES(1); break;
        case 2: SLX_TILE_SHAPES(2); break;
        case 3: SLX_TILE_SHAPES(3); break;
        case 4: SLX_TILE_SHAPES(4); break;
        default: return hipErrorInvalidValue;
    }
#undef SLX_TILE_SHAPES
#undef SLX_TILE_LAUNCH
    return hipGetLastError();
}

}  // namespace slx
