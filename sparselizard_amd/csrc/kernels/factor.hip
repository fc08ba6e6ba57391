// Factorized element contraction on MFMA with FP32 accumulation.
//
// One workgroup per element at a time, persistent over the bucket: 8 consumer
// waves (a 2 x 4 grid over the 16x16 output tiles) run the MFMAs while 4
// producer waves stage the next k-step.
// The K dimension runs over (Gauss point, tf derivative) pairs in k-steps of
// 32: a k-step holds qb Gauss points times the na tf derivatives, grouped in
// segments of whole Gauss points (factor_segw), so the dof table values of a
// Gauss point are read once per element and pass.  Per k-step:
//   A chunk (shared P fragments, precomputed on the host) -> LDS
//   B chunk (Q_e for the element, built from the element's coefficients in
//            LDS and the dof table) -> LDS
// both double buffered, and every wave runs its block of 16x16 output tiles
// in the arithmetic of the launch (kernels/mfma.h):
//   AR_F16   acc += Ahi*Bhi + Ahi*Blo + Alo*Bhi           (v_mfma_f32_16x16x32_f16)
//   AR_BF16  acc += Ah*Bh + Ah*Bm + Am*Bh + Ah*Bl + Am*Bm + Al*Bh
//                                                      (v_mfma_f32_16x16x32_bf16)
//   AR_F32   acc += A*B, eight 16x16x4 steps per k-step    (v_mfma_f32_16x16x4_f32)
// P is scaled on the host and Q_e per element by powers of two so that the
// largest magnitudes sit in [128, 256): in the FP16 split the lo parts of
// every entry above 2^-10 of the maximum stay normal f16 numbers.  The
// accumulators are unscaled once when the element matrix is stored.
#include "factor.h"
#include "mfma.h"

namespace slx {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float float4v __attribute__((ext_vector_type(4)));
typedef unsigned int uint4v __attribute__((ext_vector_type(4)));

size_t factor_lds_bytes(const FactorSlice& f, int arith) {
    const size_t npl = (size_t)mfma_planes(arith);
    return (size_t)2 * f.mt * npl * 64 * 16 + (size_t)2 * (4 * f.bn) * npl * 64 * 16 + (size_t)f.nks * f.qb * 16 * 4 + 16;
}

// tile blocks per wave: rows over 2 waves, columns over 4 waves; the column
// tiles of a wider matrix are taken in passes of 4 * BN
static const int FK_BM[3] = {2, 4, 7};
static const int FK_BN[3] = {1, 2, 2};

bool factor_variant_fits(int mt, int nt, int* bm, int* bn, int variant) {
    if (variant >= 0 && variant < 3 && 2 * FK_BM[variant] >= mt) {
        *bm = FK_BM[variant];
        *bn = FK_BN[variant];
        return true;
    }
    for (int v = 0; v < 3; v++)
        if (2 * FK_BM[v] >= mt && (4 * FK_BN[v] >= nt || v == 2)) {
            *bm = FK_BM[v];
            *bn = FK_BN[v];
            return true;
        }
    return false;
}

// Producer waves of factor_element: one k-step ahead of the consumers, stage
// the shared P chunk and build the Q_e chunk of the element.  Thread bt takes
// column bc = bt % CW of the pass and segment sg = bt / CW of the k-step
// (factor_segw): entries k = SEGW sg + r, r = NA qi + ai, are Gauss point
// ks QB + sg QH + qi with tf derivative ai.  The segment's QH dof-table values
// (a float4 of the four derivatives per Gauss point) are loaded one k-step
// ahead; each is combined with the NA coefficient rows Ms[q][ai] of its Gauss
// point, and the entries are packed eight at a time into 16-byte fragment
// words (F16 / BF16: lane 16 (k / 8) + j % 16 of column tile j / 16, element
// k % 8) or written as floats (F32: Q[k][column] per tile).  The table and Ms
// are zero past the last Gauss point and the table past the last dof column,
// so the entry positions are compile-time constants and no entry is checked.
template <int AR, int BM, int BN, int NA>
__device__ __forceinline__ void factor_producer(const FactorSlice& f, uint4v* As, uint4v* Bs, const float* Ms, int bt) {
    constexpr int NPL = AR == AR_BF16 ? 3 : 2;
    constexpr int NTB = 4 * BN, BCH = NTB * NPL * 64, CW = 16 * NTB;
    constexpr int SEGW = factor_segw(BN), QH = SEGW / NA, QB = factor_qb(NA, BN), NGS = SEGW / 8;
    static_assert(CW * (32 / SEGW) == 64 * SLX_FK_BWAVES, "one producer thread per (column, segment)");
    static_assert(QH >= 1, "a segment holds at least one Gauss point");
    const int nks = f.nks, ntq = f.ntq;
    const int achunk = f.mt * NPL * 64;  // a multiple of 64: a wave's words are all in or all out
    const int npass = (f.nt + NTB - 1) / NTB;
    const int wp = __builtin_amdgcn_readfirstlane(bt >> 6), lane = bt & 63;
    const int sg = __builtin_amdgcn_readfirstlane(bt / CW), bc = bt % CW;
    // global (not flat) loads: a flat load also counts against the LDS counter,
    // which every barrier drains, so the one-k-step-ahead loads would be waited
    // for at the next barrier instead of at their use
    typedef __attribute__((address_space(1))) const float4v gfloat4v;
    const gfloat4v* tq4 = (const gfloat4v*)(f.tq);
    const uint4v* afrag = reinterpret_cast<const uint4v*>(f.afrag);
    float4v hq[QH];
    auto load_h = [&](int ks, int cb) {
        const gfloat4v* row = tq4 + ((ks * QB + sg * QH) * ntq + cb * 16 + bc);
#pragma unroll
        for (int qi = 0; qi < QH; qi++) hq[qi] = row[qi * ntq];
    };
    auto produce = [&](int ks, int buf, int cb) {
        // the shared P chunk goes straight from global memory to LDS
        // (global_load_lds: lane i of a wave lands at the wave's base + 16 i),
        // in flight while Q_e is built
#pragma unroll
        for (int u = 0; u < SLX_FK_AREG; u++) {
            const int x0 = 64 * wp + u * 64 * SLX_FK_BWAVES;  // wave-uniform first entry
            if (x0 < achunk) {
                const uint4v* src = afrag + ((size_t)ks * achunk + x0);
                __builtin_amdgcn_global_load_lds((const void*)(src + lane), (__attribute__((address_space(3))) void*)(As + buf * achunk + x0),
                                                 16, 0, 0);
            }
        }
        const float* mrow = Ms + (ks * QB + sg * QH) * 16;  // the segment's coefficient rows (LDS broadcast)
        float vals[SEGW];
#pragma unroll
        for (int r = 0; r < SEGW; r++) {
            const int qi = r / NA, ai = r % NA;
            if (qi < QH) {
                const float4v m = *reinterpret_cast<const float4v*>(mrow + (qi * 4 + ai) * 4);
                const float4v h = hq[qi];
                vals[r] = m[0] * h[0] + m[1] * h[1] + m[2] * h[2] + m[3] * h[3];
            } else {
                vals[r] = 0.f;
            }
        }
        uint4v* colw = Bs + (size_t)buf * BCH + (size_t)(bc >> 4) * NPL * 64 + (bc & 15);
        if constexpr (AR == AR_F16) {
#pragma unroll
            for (int g = 0; g < NGS; g++) {
                half8 hv, lv;
#pragma unroll
                for (int v = 0; v < 8; v++) {
                    const _Float16 hi = (_Float16)vals[8 * g + v];
                    hv[v] = hi;
                    lv[v] = (_Float16)(vals[8 * g + v] - (float)hi);
                }
                const int kg = sg * NGS + g;
                colw[16 * kg] = __builtin_bit_cast(uint4v, hv);
                colw[64 + 16 * kg] = __builtin_bit_cast(uint4v, lv);
            }
        } else if constexpr (AR == AR_BF16) {
#pragma unroll
            for (int g = 0; g < NGS; g++) {
                bf16x8 hv, mv, lv;
#pragma unroll
                for (int v = 0; v < 8; v++) {
                    const float x = vals[8 * g + v];
                    const __bf16 hi = (__bf16)x;
                    const float r1 = x - (float)hi;
                    const __bf16 mid = (__bf16)r1;
                    hv[v] = hi;
                    mv[v] = mid;
                    lv[v] = (__bf16)(r1 - (float)mid);
                }
                const int kg = sg * NGS + g;
                colw[16 * kg] = __builtin_bit_cast(uint4v, hv);
                colw[64 + 16 * kg] = __builtin_bit_cast(uint4v, mv);
                colw[128 + 16 * kg] = __builtin_bit_cast(uint4v, lv);
            }
        } else {
            float* colf = reinterpret_cast<float*>(Bs + (size_t)buf * BCH) + (size_t)(bc >> 4) * 512 + (bc & 15);
#pragma unroll
            for (int r = 0; r < SEGW; r++) colf[(SEGW * sg + r) * 16] = vals[r];
        }
        // the P chunk is in LDS before this wave reaches the barrier; the
        // next k-step's table values are requested after that wait
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        if (ks + 1 < nks) load_h(ks + 1, cb);
    };
    for (int pass = 0; pass < npass; pass++) {
        load_h(0, pass * NTB);
        produce(0, 0, pass * NTB);
        __syncthreads();  // k-step 0 staged
        for (int ks = 0; ks < nks; ks++) {
            if (ks + 1 < nks) produce(ks + 1, (ks & 1) ^ 1, pass * NTB);
            __syncthreads();  // k-step ks consumed, k-step ks + 1 staged
        }
    }
}

// One k-step of a consumer wave: its BM row tiles wm + 2 mi (a tile past the
// last row tile reads the last one; its accumulator is never stored) times its
// first NB column tiles wn + 4 nj, as straight-line code: the operands are read
// first and the products of the split are issued product by product over the
// tiles, so no accumulator is updated by two MFMAs back to back.
template <int AR, int BM, int BN, int NB>
__device__ __forceinline__ void factor_consume(const uint4v* __restrict__ Ac, const uint4v* __restrict__ Bc, float4v (&acc)[BM][BN], int wm,
                                               int wn, int MT, int lane) {
    constexpr int NPL = AR == AR_BF16 ? 3 : 2;
    static_assert(NB >= 1 && NB <= BN, "column tiles of a k-step");
    int mts[BM];
#pragma unroll
    for (int mi = 0; mi < BM; mi++) mts[mi] = min(wm + 2 * mi, MT - 1);
    if constexpr (AR == AR_F32) {
        // step j (k = 4 j + lane / 16): B lane value Q[4 j + lane / 16][lane % 16]
        const float* Bf = reinterpret_cast<const float*>(Bc);
        float bq[NB][8];
#pragma unroll
        for (int nj = 0; nj < NB; nj++)
#pragma unroll
            for (int j = 0; j < 8; j++) bq[nj][j] = Bf[(wn + 4 * nj) * 512 + 64 * j + lane];
        float4v a[BM][2];
#pragma unroll
        for (int mi = 0; mi < BM; mi++) {
            a[mi][0] = __builtin_bit_cast(float4v, Ac[(mts[mi] * NPL) * 64 + lane]);
            a[mi][1] = __builtin_bit_cast(float4v, Ac[(mts[mi] * NPL + 1) * 64 + lane]);
        }
#pragma unroll
        for (int j = 0; j < 8; j++)
#pragma unroll
            for (int mi = 0; mi < BM; mi++)
#pragma unroll
                for (int nj = 0; nj < NB; nj++)
                    acc[mi][nj] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[mi][j >> 2][j & 3], bq[nj][j], acc[mi][nj], 0, 0, 0);
    } else if constexpr (AR == AR_F16) {
        half8 bh[NB], bl[NB], ah[BM], al[BM];
#pragma unroll
        for (int nj = 0; nj < NB; nj++) {
            bh[nj] = __builtin_bit_cast(half8, Bc[((wn + 4 * nj) * NPL) * 64 + lane]);
            bl[nj] = __builtin_bit_cast(half8, Bc[((wn + 4 * nj) * NPL + 1) * 64 + lane]);
        }
#pragma unroll
        for (int mi = 0; mi < BM; mi++) {
            ah[mi] = __builtin_bit_cast(half8, Ac[(mts[mi] * NPL) * 64 + lane]);
            al[mi] = __builtin_bit_cast(half8, Ac[(mts[mi] * NPL + 1) * 64 + lane]);
        }
#pragma unroll
        for (int mi = 0; mi < BM; mi++)
#pragma unroll
            for (int nj = 0; nj < NB; nj++) acc[mi][nj] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[mi], bh[nj], acc[mi][nj], 0, 0, 0);
#pragma unroll
        for (int mi = 0; mi < BM; mi++)
#pragma unroll
            for (int nj = 0; nj < NB; nj++) acc[mi][nj] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[mi], bl[nj], acc[mi][nj], 0, 0, 0);
#pragma unroll
        for (int mi = 0; mi < BM; mi++)
#pragma unroll
            for (int nj = 0; nj < NB; nj++) acc[mi][nj] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[mi], bh[nj], acc[mi][nj], 0, 0, 0);
    } else {
        // BF16: one row tile's three planes at a time (all of them would not fit
        // the register budget); the six products alternate over the column tiles
        bf16x8 xh[NB], xm[NB], xl[NB];
#pragma unroll
        for (int nj = 0; nj < NB; nj++) {
            xh[nj] = __builtin_bit_cast(bf16x8, Bc[((wn + 4 * nj) * NPL) * 64 + lane]);
            xm[nj] = __builtin_bit_cast(bf16x8, Bc[((wn + 4 * nj) * NPL + 1) * 64 + lane]);
            xl[nj] = __builtin_bit_cast(bf16x8, Bc[((wn + 4 * nj) * NPL + 2) * 64 + lane]);
        }
#pragma unroll
        for (int mi = 0; mi < BM; mi++) {
            const bf16x8 ah = __builtin_bit_cast(bf16x8, Ac[(mts[mi] * NPL) * 64 + lane]);
            const bf16x8 am = __builtin_bit_cast(bf16x8, Ac[(mts[mi] * NPL + 1) * 64 + lane]);
            const bf16x8 al = __builtin_bit_cast(bf16x8, Ac[(mts[mi] * NPL + 2) * 64 + lane]);
#pragma unroll
            for (int nj = 0; nj < NB; nj++) acc[mi][nj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, xh[nj], acc[mi][nj], 0, 0, 0);
#pragma unroll
            for (int nj = 0; nj < NB; nj++) acc[mi][nj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, xm[nj], acc[mi][nj], 0, 0, 0);
#pragma unroll
            for (int nj = 0; nj < NB; nj++) acc[mi][nj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, xh[nj], acc[mi][nj], 0, 0, 0);
#pragma unroll
            for (int nj = 0; nj < NB; nj++) acc[mi][nj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, xl[nj], acc[mi][nj], 0, 0, 0);
#pragma unroll
            for (int nj = 0; nj < NB; nj++) acc[mi][nj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, xm[nj], acc[mi][nj], 0, 0, 0);
#pragma unroll
            for (int nj = 0; nj < NB; nj++) acc[mi][nj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, xh[nj], acc[mi][nj], 0, 0, 0);
        }
    }
}

// LDS per k-step and buffer: A chunk [MT][NPL][64] 16-byte words; B chunk
// [NTB][NPL][64] words (F16 / BF16: lane l of column tile t holds the eight
// k = 8 (l / 16) + 0..7 of column l % 16, one plane per piece; F32: the tile
// as floats Q[k][column], [32][16]).
// One element of a slice (the whole workgroup; every thread must call it).
template <int AR, int BM, int BN>
__device__ __forceinline__ void factor_element(const FactorSlice& f, int n_elem, int e, const float* __restrict__ ctab,
                                               float* __restrict__ ke, unsigned char* smem_raw) {
    constexpr int NPL = AR == AR_BF16 ? 3 : 2;                  // 16-byte planes per lane
    const int MT = f.mt, NT = f.nt, ngp = f.ngp, nq = f.nq, na = f.na, nqp = f.nks * f.qb;
    constexpr int NTB = 4 * BN;                                 // column tiles per pass
    constexpr int BCH = NTB * NPL * 64;                         // B chunk words
    static_assert(BN == 1 || BN == 2, "a partial pass leaves a wave one column tile or none");
    uint4v* As = reinterpret_cast<uint4v*>(smem_raw);           // [2][MT][NPL][64]
    uint4v* Bs = As + 2 * MT * NPL * 64;                        // [2][NTB][NPL][64]
    float* Ms = reinterpret_cast<float*>(Bs + 2 * BCH);         // [nqp][4 tf derivative][4 dof derivative]
    unsigned* smax = reinterpret_cast<unsigned*>(Ms + nqp * 16);
    const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const bool mfma_wave = w < SLX_FK_WAVES;
    // term table through the constant address space (scalar loads), coefficients
    // through the global one (no flat loads, which the barriers would drain)
    typedef __attribute__((address_space(4))) const int32_t cint32;
    typedef __attribute__((address_space(1))) const float gfloat;
    const cint32* t_ai = (const cint32*)(f.tinfo);
    const cint32* t_b = t_ai + f.n_terms;
    const cint32* t_slot = t_ai + 2 * f.n_terms;
    const gfloat* ctg = (const gfloat*)(ctab);
    const int64_t stride = (int64_t)nq * n_elem;
    const int achunk = MT * NPL * 64;
    const int npass = (NT + NTB - 1) / NTB;

    {
        if (tid == 0) *smax = 0u;
        __syncthreads();  // the previous element is done with Ms, As, Bs
        // the element's coefficient matrices per Gauss point:
        //   Ms[q][ai][b] = sum of C_t(e, q) over terms with derivatives (alist[ai], b),
        // zero for the padding points ngp <= q < nqp of the last k-step
        float m = 0.f;
        for (int q = tid; q < nqp; q += blockDim.x) {
            float* mq = Ms + q * 16;  // owned by this thread
#pragma unroll
            for (int x = 0; x < 16; x++) mq[x] = 0.f;
            if (q >= ngp) continue;
            const int qq = nq == 1 ? 0 : q;
            float cv[SLX_FK_MAXT];
#pragma unroll
            for (int t = 0; t < SLX_FK_MAXT; t++)
                cv[t] = t < f.n_terms ? ctg[(int64_t)t_slot[t] * stride + (int64_t)e * nq + qq] : 0.f;
#pragma unroll
            for (int t = 0; t < SLX_FK_MAXT; t++)
                if (t < f.n_terms) mq[t_ai[t] * 4 + t_b[t]] += cv[t];
#pragma unroll
            for (int ai = 0; ai < 4; ai++) m = fmaxf(m, fabsf(mq[4 * ai]) + fabsf(mq[4 * ai + 1]) + fabsf(mq[4 * ai + 2]) + fabsf(mq[4 * ai + 3]));
        }
        for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
        if (lane == 0) atomicMax(smax, __float_as_uint(m));
        __syncthreads();
        // power-of-two scale: |Q| <= max_q,a sum_b |Ms| * max |tq|  ->  [128, 256)
        const float bound = __uint_as_float(*smax) * f.tqmax;
        float sc = 1.f, unsc = f.a_unscale;
        if (bound > 0.f) {
            int ex = 0;
            (void)frexpf(bound, &ex);  // bound in [2^(ex-1), 2^ex)
            sc = ldexpf(1.f, 8 - ex);
            unsc = ldexpf(f.a_unscale, ex - 8);
        }
        // the element scale folded into the coefficient matrices (exact: a power of two)
        for (int q = tid; q < ngp; q += blockDim.x) {
            float* mq = Ms + q * 16;
#pragma unroll
            for (int x = 0; x < 16; x++) mq[x] *= sc;
        }
        __syncthreads();
        if (mfma_wave) {
            // consumer waves: a 2 x 4 grid over the 16x16 output tiles of a pass
            const int wm = w >> 2, wn = w & 3;
            for (int pass = 0; pass < npass; pass++) {
                const int cb = pass * NTB;
                float4v acc[BM][BN];
#pragma unroll
                for (int mi = 0; mi < BM; mi++)
#pragma unroll
                    for (int nj = 0; nj < BN; nj++) acc[mi][nj] = float4v{0.f, 0.f, 0.f, 0.f};
                __syncthreads();  // k-step 0 staged
                // column tiles of this wave in the pass (a prefix of its BN tiles)
                int nbv = 0;
#pragma unroll
                for (int nj = 0; nj < BN; nj++)
                    if (cb + wn + 4 * nj < NT) nbv = nj + 1;
                for (int ks = 0; ks < f.nks; ks++) {
                    const int cur = ks & 1;
                    const uint4v* Ac = As + cur * achunk;
                    const uint4v* Bc = Bs + cur * BCH;
                    if (nbv == BN) factor_consume<AR, BM, BN, BN>(Ac, Bc, acc, wm, wn, MT, lane);
                    else if (nbv > 0) factor_consume<AR, BM, BN, 1>(Ac, Bc, acc, wm, wn, MT, lane);
                    __syncthreads();  // k-step ks consumed, k-step ks + 1 staged
                }
                // element matrix rows i = mt*16 + 4*(lane/16) + r, column j = jt*16 + lane%16:
                // a uniform tile base plus one 32-bit lane offset per r, so that no
                // per-entry address is kept live across the k-steps
                float* eb = ke + f.ke_base + (int64_t)e * f.n_tf * f.ke_ld;
                const int ld = (int)f.ke_ld, lrow = 4 * (lane >> 4), lcol = lane & 15;
#pragma unroll
                for (int mi = 0; mi < BM; mi++) {
                    const int mt = wm + 2 * mi;
#pragma unroll
                    for (int nj = 0; nj < BN; nj++) {
                        const int jt = cb + wn + 4 * nj;
                        if (mt >= MT || jt >= NT) continue;
                        const bool jok = jt * 16 + lcol < f.n_dof;
                        float* tb = eb + (mt * 16 * ld + jt * 16);
#pragma unroll
                        for (int r = 0; r < 4; r++) {
                            // streaming store: element matrices must not evict the shared
                            // P fragments and dof table from L2
                            if (jok && mt * 16 + lrow + r < f.n_tf)
                                __builtin_nontemporal_store(acc[mi][nj][r] * unsc, tb + ((lrow + r) * ld + lcol));
                        }
                    }
                }
            }
        } else {
            // producer waves (the element routine is specialised on the number
            // of tf derivatives, which fixes the K order of a k-step)
            const int bt = tid - 64 * SLX_FK_WAVES;
            switch (na) {
                case 1: factor_producer<AR, BM, BN, 1>(f, As, Bs, Ms, bt); break;
                case 2: factor_producer<AR, BM, BN, 2>(f, As, Bs, Ms, bt); break;
                case 3: factor_producer<AR, BM, BN, 3>(f, As, Bs, Ms, bt); break;
                default: factor_producer<AR, BM, BN, 4>(f, As, Bs, Ms, bt); break;
            }
        }
    }
}

template <int AR, int BM, int BN>
__global__ void __launch_bounds__(64 * (SLX_FK_WAVES + SLX_FK_BWAVES))
    factor_contract_kernel(int n_elem, FactorSlice f, const float* __restrict__ ctab, float* __restrict__ ke) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    for (int e = blockIdx.x; e < n_elem; e += gridDim.x) factor_element<AR, BM, BN>(f, n_elem, e, ctab, ke, smem_raw);
}

template <int AR, int BM, int BN>
static hipError_t launch_fk(const DevBucket& b, const FactorSlice& f, const float* ctab, float* ke, int n_cu, hipStream_t st) {
    const size_t lds = factor_lds_bytes(f, AR);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    int per_cu = (int)std::max<size_t>(1, std::min<size_t>(2, (160 * 1024) / lds));
    int64_t grid = std::min<int64_t>(b.n_elem, (int64_t)n_cu * per_cu);
    hipLaunchKernelGGL((factor_contract_kernel<AR, BM, BN>), dim3((unsigned)grid), dim3(64 * (SLX_FK_WAVES + SLX_FK_BWAVES)), lds, st, b.n_elem,
                       f, ctab, ke);
    return hipGetLastError();
}

// ------------------------------------------------------------------ FP64
//
// One workgroup (8 waves, a 2 x 4 grid over the 16x16 output tiles of a pass)
// per element at a time.  Per k-step of 16 (Gauss point, tf derivative) pairs
// every wave builds its share of the next Q_e chunk in LDS (FP64, the element's
// coefficient matrices times w_q dof_b[j][q]) while the MFMAs of the current
// chunk run; the P fragments come straight from global memory (shared by the
// bucket's elements, L2-resident).  No scaling: FP64 needs none.
typedef double double4v __attribute__((ext_vector_type(4)));

size_t factor64_lds_bytes(const FactorSlice& f) {
    return (size_t)f.ngp * 16 * 8 + (size_t)2 * 16 * (4 * f.bn * 16) * 8;
}

template <int BM, int BN>
__device__ __forceinline__ void factor64_element(const FactorSlice& f, int n_elem, int e, const double* __restrict__ ctab,
                                                 double* __restrict__ ke, unsigned char* smem_raw) {
    constexpr int NTB = 4 * BN;   // column tiles per pass
    constexpr int PC = 16 * NTB;  // columns per pass
    const int MT = f.mt, NT = f.nt, ngp = f.ngp, nq = f.nq, na = f.na, qb = f.qb64, ntp = f.ntp, nks = f.nks64;
    double* Ms = reinterpret_cast<double*>(smem_raw);  // [ngp][4 tf derivative][4 dof derivative]
    double* Bq = Ms + (size_t)ngp * 16;                // [2][16][PC]
    const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = w >> 2, wn = w & 3;
    const int32_t* t_ai = f.tinfo;
    const int32_t* t_b = f.tinfo + f.n_terms;
    const int32_t* t_slot = f.tinfo + 2 * f.n_terms;
    const int64_t stride = (int64_t)nq * n_elem;
    const double4v* tq4 = reinterpret_cast<const double4v*>(f.tq64);
    const int npass = (NT + NTB - 1) / NTB;
    // Q_e chunk of k-step ks for the pass's columns: entry (kl, c) = k slot kl
    // (Gauss point ks*qb + kl/na, tf derivative kl%na), column cb*16 + c
    auto build = [&](int ks, int buf, int cb) {
        double* dst = Bq + (size_t)buf * 16 * PC;
        for (int x = tid; x < 16 * PC; x += 512) {
            const int kl = x / PC, c = x - kl * PC;
            const int qi = kl / na, ai = kl - qi * na, q = ks * qb + qi, j = cb * 16 + c;
            double v = 0.0;
            if (qi < qb && q < ngp && j < ntp) {
                const double4v t = tq4[(size_t)q * ntp + j];
                const double* m = Ms + (q * 4 + ai) * 4;
                v = m[0] * t[0] + m[1] * t[1] + m[2] * t[2] + m[3] * t[3];
            }
            dst[x] = v;
        }
    };
    {
        __syncthreads();  // the previous element is done with Ms and Bq
        for (int q = tid; q < ngp; q += 512) {
            double* mq = Ms + q * 16;  // owned by this thread
#pragma unroll
            for (int x = 0; x < 16; x++) mq[x] = 0.0;
            const int qq = nq == 1 ? 0 : q;
#pragma unroll
            for (int t = 0; t < SLX_FK_MAXT; t++)
                if (t < f.n_terms) mq[t_ai[t] * 4 + t_b[t]] += ctab[(int64_t)t_slot[t] * stride + (int64_t)e * nq + qq];
        }
        __syncthreads();
        for (int pass = 0; pass < npass; pass++) {
            const int cb = pass * NTB;
            double4v acc[BM][BN];
#pragma unroll
            for (int mi = 0; mi < BM; mi++)
#pragma unroll
                for (int nj = 0; nj < BN; nj++) acc[mi][nj] = double4v{0.0, 0.0, 0.0, 0.0};
            build(0, 0, cb);
            __syncthreads();
            for (int ks = 0; ks < nks; ks++) {
                if (ks + 1 < nks) build(ks + 1, (ks + 1) & 1, cb);
                const double* Bc = Bq + (size_t)(ks & 1) * 16 * PC;
                double bq[BN][4];
#pragma unroll
                for (int nj = 0; nj < BN; nj++)
#pragma unroll
                    for (int j = 0; j < 4; j++) bq[nj][j] = Bc[(4 * j + (lane >> 4)) * PC + (wn + 4 * nj) * 16 + (lane & 15)];
#pragma unroll
                for (int mi = 0; mi < BM; mi++) {
                    const int mt = wm + 2 * mi;
                    if (mt < MT) {
                        const double* af = f.afrag64 + ((size_t)(ks * MT + mt) * 4) * 64 + lane;
                        double a[4];
#pragma unroll
                        for (int j = 0; j < 4; j++) a[j] = af[64 * j];
#pragma unroll
                        for (int nj = 0; nj < BN; nj++)
                            if (cb + wn + 4 * nj < NT)
#pragma unroll
                                for (int j = 0; j < 4; j++) acc[mi][nj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[j], bq[nj][j], acc[mi][nj], 0, 0, 0);
                    }
                }
                __syncthreads();  // chunk ks consumed, chunk ks + 1 built
            }
            // lane l holds rows mt*16 + l/16 + 4v, column jt*16 + l%16
            double* eb = ke + f.ke_base + (int64_t)e * f.n_tf * f.ke_ld;
            const int ld = (int)f.ke_ld;
#pragma unroll
            for (int mi = 0; mi < BM; mi++) {
                const int mt = wm + 2 * mi;
#pragma unroll
                for (int nj = 0; nj < BN; nj++) {
                    const int jt = cb + wn + 4 * nj;
                    const int j = jt * 16 + (lane & 15);
                    if (mt >= MT || jt >= NT || j >= f.n_dof) continue;
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        const int i = mt * 16 + (lane >> 4) + 4 * v;
                        if (i < f.n_tf) __builtin_nontemporal_store(acc[mi][nj][v], eb + i * ld + j);
                    }
                }
            }
        }
    }
}

template <int BM, int BN>
__global__ void __launch_bounds__(512) factor64_kernel(int n_elem, FactorSlice f, const double* __restrict__ ctab, double* __restrict__ ke) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    for (int e = blockIdx.x; e < n_elem; e += gridDim.x) factor64_element<BM, BN>(f, n_elem, e, ctab, ke, smem_raw);
}

template <int BM, int BN>
static hipError_t launch_f64(const DevBucket& b, const FactorSlice& f, const double* ctab, double* ke, int n_cu, hipStream_t st) {
    const size_t lds = factor64_lds_bytes(f);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    const int64_t grid = std::min<int64_t>(b.n_elem, (int64_t)n_cu);
    hipLaunchKernelGGL((factor64_kernel<BM, BN>), dim3((unsigned)grid), dim3(512), lds, st, b.n_elem, f, ctab, ke);
    return hipGetLastError();
}

hipError_t launch_factor64(const DevBucket& b, const FactorSlice& f, const double* ctab, double* ke, int n_cu, hipStream_t st) {
    if (b.n_elem == 0) return hipSuccess;
    if (!f.f64_ok || !f.afrag64 || !f.tq64) return hipErrorInvalidValue;
    if (f.bm == 2 && f.bn == 1) return launch_f64<2, 1>(b, f, ctab, ke, n_cu, st);
    if (f.bm == 4 && f.bn == 2) return launch_f64<4, 2>(b, f, ctab, ke, n_cu, st);
    if (f.bm == 7 && f.bn == 2) return launch_f64<7, 2>(b, f, ctab, ke, n_cu, st);
    return hipErrorInvalidValue;
}

// ------------------------------------------------------------------ fused
//
// One persistent launch over the factorized slices of every bucket: each
// workgroup takes the next (slice, element) item from a shared counter, so the
// expensive elements (highest order, numbered first) start first and the
// cheap ones fill the tail.  Every item is computed by the same per-element
// code as the per-slice launches, so the element matrices are identical.
int factor_variant_index(int bm, int bn) {
    for (int v = 0; v < 3; v++)
        if (FK_BM[v] == bm && FK_BN[v] == bn) return v;
    return -1;
}

__global__ void factor_fused_clear(unsigned long long* counter) { *counter = 0ull; }

// the item's slice: the last slice whose first item is <= it
__device__ __forceinline__ int fused_slice_of(const FusedSlice* fs, int n_fs, int64_t it) {
    int lo = 0, hi = n_fs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (fs[mid].first <= it) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <int AR>
__global__ void __launch_bounds__(64 * (SLX_FK_WAVES + SLX_FK_BWAVES))
    factor_fused_kernel(const FusedSlice* __restrict__ fs, int n_fs, int64_t total, unsigned long long* __restrict__ counter,
                        float* __restrict__ ke) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    __shared__ int64_t s_item;
    for (;;) {
        if (threadIdx.x == 0) s_item = (int64_t)atomicAdd(counter, 1ull);
        __syncthreads();
        const int64_t it = s_item;
        __syncthreads();  // every thread has the item before thread 0 takes the next
        if (it >= total) break;
        const int s = fused_slice_of(fs, n_fs, it);
        const FusedSlice& F = fs[s];
        const int e = (int)(it - F.first);
        const float* ctab = static_cast<const float*>(F.ctab);
        if (F.variant == 0) factor_element<AR, 2, 1>(F.f, F.n_elem, e, ctab, ke, smem_raw);
        else if (F.variant == 1) factor_element<AR, 4, 2>(F.f, F.n_elem, e, ctab, ke, smem_raw);
        else factor_element<AR, 7, 2>(F.f, F.n_elem, e, ctab, ke, smem_raw);
    }
}

__global__ void __launch_bounds__(512)
    factor64_fused_kernel(const FusedSlice* __restrict__ fs, int n_fs, int64_t total, unsigned long long* __restrict__ counter,
                          double* __restrict__ ke) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    __shared__ int64_t s_item;
    for (;;) {
        if (threadIdx.x == 0) s_item = (int64_t)atomicAdd(counter, 1ull);
        __syncthreads();
        const int64_t it = s_item;
        __syncthreads();
        if (it >= total) break;
        const int s = fused_slice_of(fs, n_fs, it);
        const FusedSlice& F = fs[s];
        const int e = (int)(it - F.first);
        const double* ctab = static_cast<const double*>(F.ctab);
        if (F.variant == 0) factor64_element<2, 1>(F.f, F.n_elem, e, ctab, ke, smem_raw);
        else if (F.variant == 1) factor64_element<4, 2>(F.f, F.n_elem, e, ctab, ke, smem_raw);
        else factor64_element<7, 2>(F.f, F.n_elem, e, ctab, ke, smem_raw);
    }
}

hipError_t launch_factor_fused(const FusedSlice* fs, int n_fs, int64_t total, unsigned long long* counter, int arith, size_t lds,
                               void* ke, int n_cu, hipStream_t st) {
    if (total <= 0 || n_fs <= 0) return hipSuccess;
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL(factor_fused_clear, dim3(1), dim3(1), 0, st, counter);
    const int per_cu = (int)std::max<size_t>(1, std::min<size_t>(2, (160 * 1024) / std::max<size_t>(lds, 1)));
    const unsigned grid = (unsigned)std::min<int64_t>(total, (int64_t)n_cu * per_cu);
    if (arith == 3)
        hipLaunchKernelGGL(factor64_fused_kernel, dim3(grid), dim3(512), lds, st, fs, n_fs, total, counter, (double*)ke);
    else if (arith == AR_BF16)
        hipLaunchKernelGGL(factor_fused_kernel<AR_BF16>, dim3(grid), dim3(64 * (SLX_FK_WAVES + SLX_FK_BWAVES)), lds, st, fs, n_fs, total, counter,
                           (float*)ke);
    else if (arith == AR_F32)
        hipLaunchKernelGGL(factor_fused_kernel<AR_F32>, dim3(grid), dim3(64 * (SLX_FK_WAVES + SLX_FK_BWAVES)), lds, st, fs, n_fs, total, counter,
                           (float*)ke);
    else if (arith == AR_F16)
        hipLaunchKernelGGL(factor_fused_kernel<AR_F16>, dim3(grid), dim3(64 * (SLX_FK_WAVES + SLX_FK_BWAVES)), lds, st, fs, n_fs, total, counter,
                           (float*)ke);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

template <int AR>
static hipError_t launch_fk_ar(const DevBucket& b, const FactorSlice& f, const float* ctab, float* ke, int n_cu, hipStream_t st) {
    if (f.bm == 2 && f.bn == 1) return launch_fk<AR, 2, 1>(b, f, ctab, ke, n_cu, st);
    if (f.bm == 4 && f.bn == 2) return launch_fk<AR, 4, 2>(b, f, ctab, ke, n_cu, st);
    if (f.bm == 7 && f.bn == 2) return launch_fk<AR, 7, 2>(b, f, ctab, ke, n_cu, st);
    return hipErrorInvalidValue;
}

hipError_t launch_factor(const DevBucket& b, const FactorSlice& f0, int arith, const float* ctab, float* ke, int n_cu, hipStream_t st) {
    if (b.n_elem == 0) return hipSuccess;
    if (arith < 0 || arith >= AR_COUNT || !((f0.arith_ok >> arith) & 1u) || !f0.afrag_x[arith]) return hipErrorInvalidValue;
    FactorSlice f = f0;
    f.arith = arith;
    f.afrag = f0.afrag_x[arith];
    if (arith == AR_BF16) return launch_fk_ar<AR_BF16>(b, f, ctab, ke, n_cu, st);
    if (arith == AR_F32) return launch_fk_ar<AR_F32>(b, f, ctab, ke, n_cu, st);
    return launch_fk_ar<AR_F16>(b, f, ctab, ke, n_cu, st);
}

}  // namespace slx
