// Symbolic phase on the device (rawmat::process, rawmat.cpp:127-298: the
// element rows of every free row in accumulation order, the sorted unique
// columns of every row split into free (A) and constrained (D) columns with
// local numbering, gentools::findtruefalse, gentools.cpp:2040-2068), plus the
// reduction maps of this library: a slot per element-matrix entry, rows with
// identical slot blocks sharing one copy, and the work order of the
// reduction.  Every array equals the one the host phase in plan.cpp builds,
// bit for bit: each ordering the host takes from a sequential loop or a stable
// sort is rebuilt here by a stable radix sort on keys that carry the same tie
// order.
//
// Steps (all on one stream):
//   1. free / local numbering of the DOFs (a scan of the constraint mask);
//   2. element-row instances (group, tf function, element) keyed by their free
//      row, enumerated in accumulation order and sorted stably by row: each
//      row's segments in the order of the host's counting sort;
//   3. per row, its columns gathered into LDS, bitonic-sorted with their
//      entry index as tie-break, split into A and D columns and ranked; a
//      count pass gives the row lengths, a write pass the columns and slots;
//   4. shared slot blocks: rows sorted by a hash of their block, each row's
//      representative the lowest row with an identical block;
//   5. work order: rows sorted stably by decreasing length; descriptors,
//      segments and representative slot blocks written in that order.
#include <hip/hip_runtime.h>
#include <cstring>
#include <map>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_reduce.hpp>
#include <rocprim/device/device_scan.hpp>
#include "symbolic.h"

namespace slx {
namespace {

constexpr uint64_t kOffMask = (uint64_t(1) << 48) - 1;

enum Flag {
    F_BAD_TF = 0,
    F_BAD_DOF,
    F_DUP,
    F_WIDE,       // a row with 65536 or more segments or columns
    F_REP,        // a representative after its row in the work order (internal)
    F_MAXLEN,     // longest segment of any row
    F_MAXROW,     // longest row (A + D columns)
    F_N1,         // rows of the middle tier
    F_N2,         // rows of the large tier
    F_FALLBACK,   // a row beyond the large tier
    F_COUNT
};

// Row tiers of the column kernel: (threads, entries, segments).
constexpr int T0_NT = 64, T0_CAP = 512, T0_SEG = 64;
constexpr int T1_NT = 256, T1_CAP = 2048, T1_SEG = 256;
constexpr int T2_NT = 1024, T2_CAP = 8192, T2_SEG = 1024;

// Slice tables transposed to element-major order (device).
struct SliceT {
    const int32_t* dof;   // [n_elem][n_dof]
    const int8_t* dsg;    // [n_elem][n_dof] or null
    const int8_t* tsg;    // [n_elem][n_tf] or null
    int32_t n_dof, col_off, n_tf, pad;
};

struct RowArgs {
    const int64_t* ptr;       // [nf + 1] segment range of each row
    const uint64_t* segs;     // segment descriptors in row order
    const int64_t* coo;       // [nf] entries of each row (skipped columns included)
    const int64_t* slot_ptr;  // [nf + 1]
    const SymGroup* groups;
    int n_groups;
    const SliceT* slices;
    const uint8_t* cons;      // null: nothing constrained
    const int32_t* localidx;
    int64_t n_dofs;
    int64_t* lenA;
    int64_t* lenD;
    const int64_t* Arow;
    const int64_t* Drow;
    int32_t* Acol;
    int32_t* Dcol;
    void* slots;              // row-order slot blocks (narrow type)
    int* flags;
};

// ---------------------------------------------------------------- step 1

__global__ void k_notc(int64_t nd, const uint8_t* __restrict__ cons, int64_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nd) out[i] = (cons && cons[i]) ? 0 : 1;
    else if (i == nd) out[i] = 0;
}

// nfb[i]: free DOFs before i (exclusive scan of the free mask)
__global__ void k_index(int64_t nd, const uint8_t* __restrict__ cons, const int64_t* __restrict__ nfb, int32_t* __restrict__ freeidx,
                        int32_t* __restrict__ localidx, int64_t* __restrict__ free_dofs, int64_t* __restrict__ cons_dofs) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nd) return;
    const int64_t f = nfb[i];
    if (cons && cons[i]) {
        freeidx[i] = -1;
        localidx[i] = (int32_t)(i - f);
        cons_dofs[i - f] = i;
    } else {
        freeidx[i] = (int32_t)f;
        localidx[i] = (int32_t)f;
        free_dofs[f] = i;
    }
}

__global__ void k_range_check(int64_t n, const int32_t* __restrict__ t, int64_t nd, int* __restrict__ flags, int flag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && t[i] >= nd) flags[flag] = 1;
}

template <typename T>
__global__ void k_transpose(int64_t nrow, int64_t ncol, const T* __restrict__ in, T* __restrict__ out) {
    // in [nrow][ncol] -> out [ncol][nrow]
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrow * ncol) return;
    const int64_t r = t / ncol, c = t - r * ncol;
    out[c * nrow + r] = in[t];
}

// ---------------------------------------------------------------- step 2

// Instances of one group: t = i * n_elem + e (the host's enumeration order);
// key = free row, or `sentinel` when the tf address is skipped or constrained.
__global__ void k_instances(SymGroup G, const int32_t* __restrict__ tf, const int8_t* __restrict__ tsg, const int32_t* __restrict__ freeidx,
                            int64_t nd, uint32_t sentinel, uint32_t* __restrict__ keys, uint64_t* __restrict__ vals) {
    const int64_t n = (int64_t)G.n_tf * G.n_elem;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int32_t a = tf[t];
    uint32_t key = sentinel;
    if (a >= 0 && a < nd && freeidx[a] >= 0) key = (uint32_t)freeidx[a];
    const int64_t i = t / G.n_elem, e = t - i * G.n_elem;
    uint64_t v = ((uint64_t)(uint32_t)G.len << 48) | (uint64_t)(G.base + (e * G.n_tf + i) * G.ld);
    if (G.rhs && tsg && tsg[t] < 0) v |= uint64_t(1) << 63;
    keys[t] = key;
    vals[t] = v;
}

// ptr[r] = first sorted position with key >= r, r = 0..nrows
__global__ void k_lower_bound(int64_t nrows, const uint32_t* __restrict__ sorted, int64_t n, int64_t* __restrict__ ptr) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r > nrows) return;
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if ((int64_t)sorted[mid] < r) lo = mid + 1;
        else hi = mid;
    }
    ptr[r] = lo;
}

// Slot block size (every segment padded to 4 entries), entries and tier of each row.
__global__ void k_rowsizes(int64_t nf, const int64_t* __restrict__ ptr, const uint64_t* __restrict__ segs, int64_t* __restrict__ slotsz,
                           int64_t* __restrict__ coo, int32_t* __restrict__ list1, int32_t* __restrict__ list2, int* __restrict__ flags) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r > nf) return;
    if (r == nf) {
        slotsz[nf] = 0;
        coo[nf] = 0;
        return;
    }
    int64_t s = 0, c = 0;
    int mx = 1;
    for (int64_t k = ptr[r]; k < ptr[r + 1]; k++) {
        const int len = (int)((segs[k] >> 48) & 0xFFFF);
        s += (len + 3) / 4 * 4;
        c += len;
        mx = max(mx, len);
    }
    slotsz[r] = s;
    coo[r] = c;
    atomicMax(&flags[F_MAXLEN], mx);
    const int64_t nseg = ptr[r + 1] - ptr[r];
    if (nseg >= 65536) flags[F_WIDE] = 1;
    if (c <= T0_CAP && nseg <= T0_SEG) return;
    if (c <= T1_CAP && nseg <= T1_SEG) list1[atomicAdd(&flags[F_N1], 1)] = (int32_t)r;
    else if (c <= T2_CAP && nseg <= T2_SEG) list2[atomicAdd(&flags[F_N2], 1)] = (int32_t)r;
    else flags[F_FALLBACK] = 1;
}

// ---------------------------------------------------------------- step 3

template <int NT>
__device__ int block_exclusive_scan(int v, int* wsum, int& total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(x, o, 64);
        if (lane >= o) x += t;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int base = 0;
    total = 0;
    for (int u = 0; u < NT / 64; u++) {
        const int ws = wsum[u];
        if (u < w) base += ws;
        total += ws;
    }
    __syncthreads();
    return base + x - v;
}

// Columns of rows: one workgroup per row (grid-stride).  WRITE = false: row
// lengths (and checks); WRITE = true: columns and slots.  `list` null: every
// row, skipping those of the larger tiers.
template <int NT, int CAP, int MAXSEG, bool WRITE, typename SlotT>
__global__ __launch_bounds__(NT) void k_rows(RowArgs a, const int32_t* __restrict__ list, int64_t n_list) {
    __shared__ uint64_t key[CAP];
    __shared__ uint32_t info[CAP];   // slot position in the row's block | sign (bit 31)
    __shared__ uint16_t segof[CAP];  // segment of the entry
    __shared__ int32_t s_g[MAXSEG], s_e[MAXSEG], s_i[MAXSEG], s_len[MAXSEG], s_start[MAXSEG + 1], s_slot[MAXSEG];
    __shared__ int32_t s_cnt;
    __shared__ int32_t wsum[NT / 64];
    const int tid = threadIdx.x;
    for (int64_t li = blockIdx.x; li < n_list; li += gridDim.x) {
        const int64_t r = list ? (int64_t)list[li] : li;
        const int64_t k0 = a.ptr[r];
        const int64_t nseg64 = a.ptr[r + 1] - k0, ntot64 = a.coo[r];
        if (ntot64 > CAP || nseg64 > MAXSEG) continue;  // a larger tier (uniform over the block)
        const int nseg = (int)nseg64, ntot = (int)ntot64;
        // segment table: group, element and tf function of every segment
        for (int k = tid; k < nseg; k += NT) {
            const uint64_t v = a.segs[k0 + k];
            const int64_t off = (int64_t)(v & kOffMask);
            int lo = 0, hi = a.n_groups - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (a.groups[mid].base <= off) lo = mid;
                else hi = mid - 1;
            }
            const SymGroup& G = a.groups[lo];
            const int64_t rel = (off - G.base) / G.ld;
            s_g[k] = lo;
            s_e[k] = (int)(rel / G.n_tf);
            s_i[k] = (int)(rel % G.n_tf);
            s_len[k] = (int)((v >> 48) & 0xFFFF);
        }
        if (tid == 0) s_cnt = 0;
        __syncthreads();
        // entry and slot starts of every segment (slot blocks padded to 4)
        if (tid < 64) {
            int ce = 0, cs = 0;
            for (int b0 = 0; b0 < nseg; b0 += 64) {
                const int k = b0 + tid;
                const int len = k < nseg ? s_len[k] : 0, l4 = (len + 3) / 4 * 4;
                int ie = len, is = l4;
                for (int o = 1; o < 64; o <<= 1) {
                    const int t1 = __shfl_up(ie, o, 64), t2 = __shfl_up(is, o, 64);
                    if (tid >= o) {
                        ie += t1;
                        is += t2;
                    }
                }
                if (k < nseg) {
                    s_start[k] = ce + ie - len;
                    s_slot[k] = cs + is - l4;
                }
                ce += __shfl(ie, 63, 64);
                cs += __shfl(is, 63, 64);
            }
            if (tid == 0) s_start[nseg] = ce;
        }
        __syncthreads();
        // gather the columns
        int nvl = 0;
        for (int idx = tid; idx < ntot; idx += NT) {
            int lo = 0, hi = nseg - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (s_start[mid] <= idx) lo = mid;
                else hi = mid - 1;
            }
            const int k = lo, jj = idx - s_start[k];
            const SymGroup& G = a.groups[s_g[k]];
            int s = G.slice0;
            while (s < G.slice0 + G.n_slices - 1 && jj >= a.slices[s].col_off + a.slices[s].n_dof) s++;
            const SliceT& S = a.slices[s];
            const int j = jj - S.col_off;
            const int64_t e = s_e[k];
            const int32_t c = S.dof[e * S.n_dof + j];
            uint64_t kv = ~0ull;
            uint32_t neg = 0;
            if (c >= a.n_dofs) {
                a.flags[F_BAD_DOF] = 1;
            } else if (c >= 0) {
                kv = ((uint64_t)(uint32_t)c << 32) | (uint32_t)idx;
                int sg = 1;
                if (S.tsg) sg *= S.tsg[e * S.n_tf + s_i[k]];
                if (S.dsg) sg *= S.dsg[e * S.n_dof + j];
                neg = sg < 0 ? 1u : 0u;
                nvl++;
            }
            key[idx] = kv;
            info[idx] = (uint32_t)(s_slot[k] + jj) | (neg << 31);
            segof[idx] = (uint16_t)k;
        }
        int N = 2;
        while (N < ntot) N <<= 1;
        for (int idx = ntot + tid; idx < N; idx += NT) key[idx] = ~0ull;
        if (nvl) atomicAdd(&s_cnt, nvl);
        __syncthreads();
        const int nv = s_cnt;
        // bitonic sort of key[0, N)
        for (int kk = 2; kk <= N; kk <<= 1)
            for (int j = kk >> 1; j > 0; j >>= 1) {
                for (int p = tid; p < (N >> 1); p += NT) {
                    const int lo = ((p & ~(j - 1)) << 1) | (p & (j - 1));
                    const int hi = lo + j;
                    const uint64_t x = key[lo], y = key[hi];
                    const bool asc = (lo & kk) == 0;
                    if ((x > y) == asc) {
                        key[lo] = y;
                        key[hi] = x;
                    }
                }
                __syncthreads();
            }
        // sweep: each thread a contiguous chunk of the sorted valid entries
        const int C = (nv + NT - 1) / NT;
        const int q0 = min(nv, tid * C), q1 = min(nv, q0 + C);
        int cA = 0, cD = 0;
        bool dup = false;
        for (int q = q0; q < q1; q++) {
            const uint32_t col = (uint32_t)(key[q] >> 32);
            const bool head = q == 0 || (uint32_t)(key[q - 1] >> 32) != col;
            if (head) {
                if (a.cons && a.cons[col]) cD++;
                else cA++;
            } else if (segof[(uint32_t)key[q]] == segof[(uint32_t)key[q - 1]]) {
                dup = true;
            }
        }
        if (dup) a.flags[F_DUP] = 1;
        int la, ld;
        const int exA = block_exclusive_scan<NT>(cA, wsum, la);
        const int exD = block_exclusive_scan<NT>(cD, wsum, ld);
        if (!WRITE) {
            if (tid == 0) {
                a.lenA[r] = la;
                a.lenD[r] = ld;
                atomicMax(&a.flags[F_MAXROW], la + ld);
                if (la >= 65536 || ld >= 65536) a.flags[F_WIDE] = 1;
            }
        } else {
            constexpr uint32_t top = uint32_t(1) << (8 * sizeof(SlotT) - 1);
            const int64_t arow = a.Arow[r], drow = a.Drow[r], sp0 = a.slot_ptr[r];
            SlotT* slots = (SlotT*)a.slots + sp0;
            int rA = exA, rD = exD, cur = 0;
            for (int q = q0; q < q1; q++) {
                const uint32_t col = (uint32_t)(key[q] >> 32);
                const bool head = q == 0 || (uint32_t)(key[q - 1] >> 32) != col;
                const bool isc = a.cons && a.cons[col];
                if (head) {
                    if (isc) {
                        cur = la + rD;
                        a.Dcol[drow + rD] = a.localidx[col];
                        rD++;
                    } else {
                        cur = rA;
                        a.Acol[arow + rA] = a.localidx[col];
                        rA++;
                    }
                } else if (q == q0) {
                    cur = isc ? la + rD - 1 : rA - 1;  // the column's head is in an earlier chunk
                }
                const uint32_t inf = info[(uint32_t)key[q]];
                slots[inf & 0x7FFFFFFFu] = (SlotT)((uint32_t)cur | ((inf >> 31) ? top : 0u));
            }
        }
        __syncthreads();  // LDS reused by the next row
    }
}

// ---------------------------------------------------------------- step 4

// Hash of each row's slot block (32-bit words: blocks are whole multiples of 4 entries).
__global__ void k_hash(int64_t nf, const int64_t* __restrict__ slot_ptr, const uint32_t* __restrict__ w, int sb, uint64_t* __restrict__ h,
                       int32_t* __restrict__ rows) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nf) return;
    const int64_t w0 = slot_ptr[r] * sb / 4, w1 = slot_ptr[r + 1] * sb / 4;
    uint64_t x = 1469598103934665603ull ^ (uint64_t)(w1 - w0);
    for (int64_t i = w0; i < w1; i++) x = (x ^ w[i]) * 1099511628211ull;
    h[r] = x;
    rows[r] = (int32_t)r;
}

__global__ void k_run_head(int64_t n, const uint64_t* __restrict__ h, int32_t* __restrict__ head) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) head[i] = (i == 0 || h[i] != h[i - 1]) ? (int32_t)i : 0;
}

__device__ bool same_block(const int64_t* slot_ptr, const uint32_t* w, int sb, int64_t x, int64_t y) {
    const int64_t xa = slot_ptr[x] * sb / 4, xb = slot_ptr[x + 1] * sb / 4;
    const int64_t ya = slot_ptr[y] * sb / 4, yb = slot_ptr[y + 1] * sb / 4;
    if (xb - xa != yb - ya) return false;
    for (int64_t i = 0; i < xb - xa; i++)
        if (w[xa + i] != w[ya + i]) return false;
    return true;
}

// rep[r] = the lowest row whose slot block equals r's (rows of one hash run ascend)
__global__ void k_rep(int64_t n, const int32_t* __restrict__ rows, const int32_t* __restrict__ run, const int64_t* __restrict__ slot_ptr,
                      const uint32_t* __restrict__ w, int sb, int32_t* __restrict__ rep) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int32_t r = rows[p];
    int64_t b = run[p];
    while (b < p && !same_block(slot_ptr, w, sb, rows[b], r)) b++;
    rep[r] = rows[b];
}

// ---------------------------------------------------------------- step 5

__global__ void k_order_keys(int64_t nf, const int64_t* __restrict__ lenA, const int64_t* __restrict__ lenD, uint32_t* __restrict__ k,
                             int32_t* __restrict__ rows) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nf) return;
    const int64_t len = min<int64_t>(lenA[r] + lenD[r], 65535);
    k[r] = (uint32_t)(65535 - len);
    rows[r] = (int32_t)r;
}

__global__ void k_work1(int64_t nf, const int32_t* __restrict__ order, const int64_t* __restrict__ ptr, const int64_t* __restrict__ slot_ptr,
                        const int32_t* __restrict__ rep, int32_t* __restrict__ pos_of, int64_t* __restrict__ nk, int64_t* __restrict__ own) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p > nf) return;
    if (p == nf) {
        nk[nf] = 0;
        own[nf] = 0;
        return;
    }
    const int32_t r = order[p];
    pos_of[r] = (int32_t)p;
    nk[p] = ptr[r + 1] - ptr[r];
    own[p] = rep[r] == r ? slot_ptr[r + 1] - slot_ptr[r] : 0;
}

__global__ void k_work2(int64_t nf, const int32_t* __restrict__ order, const int32_t* __restrict__ pos_of, const int32_t* __restrict__ rep,
                        const int64_t* __restrict__ ptr, const int64_t* __restrict__ kw, const int64_t* __restrict__ sw,
                        const int64_t* __restrict__ lenA, const int64_t* __restrict__ lenD, const int64_t* __restrict__ Arow,
                        const int64_t* __restrict__ Drow, const uint64_t* __restrict__ segs, const int64_t* __restrict__ slot_ptr,
                        const uint32_t* __restrict__ w_row, int sb, uint64_t* __restrict__ rdesc, uint64_t* __restrict__ erows_w,
                        uint32_t* __restrict__ w_work, int* __restrict__ flags) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nf) return;
    const int32_t r = order[p];
    const int64_t nkr = ptr[r + 1] - ptr[r];
    int64_t sp;
    if (rep[r] == r) {
        sp = sw[p];
        const int64_t src = slot_ptr[r] * sb / 4, dst = sw[p] * sb / 4, nw = (slot_ptr[r + 1] - slot_ptr[r]) * sb / 4;
        for (int64_t i = 0; i < nw; i++) w_work[dst + i] = w_row[src + i];
    } else {
        const int32_t q = pos_of[rep[r]];
        if (q >= p) flags[F_REP] = 1;
        sp = sw[q];
    }
    rdesc[4 * p + 0] = (uint64_t)Arow[r] | ((uint64_t)lenA[r] << 48);
    rdesc[4 * p + 1] = (uint64_t)Drow[r] | ((uint64_t)lenD[r] << 48);
    rdesc[4 * p + 2] = (uint64_t)kw[p] | ((uint64_t)nkr << 48);
    rdesc[4 * p + 3] = (uint64_t)sp;
    for (int64_t t = 0; t < nkr; t++) erows_w[kw[p] + t] = segs[ptr[r] + t];
}

// ---------------------------------------------------------------- host side

inline unsigned blocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

struct Scope {
    std::vector<void*> tmp;
    ~Scope() {
        for (void* p : tmp) (void)hipFree(p);
    }
};

#define SYMCHECK(x)                                   \
    do {                                              \
        const hipError_t e_ = (x);                    \
        if (e_ != hipSuccess) {                       \
            err = std::string(#x) + ": " + hipGetErrorString(e_); \
            return e_;                                \
        }                                             \
    } while (0)

}  // namespace

hipError_t symbolic_device(const SymInput& in, SymOutput& out, std::string& err, hipStream_t st) {
    err.clear();
    out = SymOutput{};
    Scope tmp;
    std::vector<SymBuf*> owned;
    // temporaries (freed on return) and outputs (freed on failure)
    auto talloc = [&](size_t bytes, void** p) -> hipError_t {
        *p = nullptr;
        const hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 16));
        if (e == hipSuccess) tmp.tmp.push_back(*p);
        return e;
    };
    struct OutGuard {
        SymOutput& o;
        bool ok = false;
        ~OutGuard() {
            if (ok) return;
            for (SymBuf* b : {&o.Arow, &o.Drow, &o.Acol, &o.Dcol, &o.rdesc, &o.erows_w, &o.slots_w, &o.rhs_ptr, &o.rhs_erows, &o.free_dofs, &o.cons_dofs})
                if (b->p) {
                    (void)hipFree(b->p);
                    b->p = nullptr;
                }
        }
    } guard{out};
    auto oalloc = [&](SymBuf& b, size_t bytes) -> hipError_t {
        b.bytes = bytes;
        return hipMalloc(&b.p, std::max<size_t>(bytes, 16));
    };
    // rocPRIM temporary storage, grown as needed
    void* rtmp = nullptr;
    size_t rtmp_bytes = 0;
    auto rgrow = [&](size_t need) -> hipError_t {
        if (need <= rtmp_bytes) return hipSuccess;
        rtmp_bytes = need;
        return talloc(need, &rtmp);
    };

    const int64_t nd = in.n_dofs;
    SYMCHECK(nd < (int64_t(1) << 31) ? hipSuccess : hipErrorInvalidValue);
    int* flags = nullptr;
    SYMCHECK(talloc(F_COUNT * sizeof(int), (void**)&flags));
    SYMCHECK(hipMemsetAsync(flags, 0, F_COUNT * sizeof(int), st));
    {
        const int one = 1;
        SYMCHECK(hipMemcpyAsync(flags + F_MAXLEN, &one, sizeof(int), hipMemcpyHostToDevice, st));
        SYMCHECK(hipMemcpyAsync(flags + F_MAXROW, &one, sizeof(int), hipMemcpyHostToDevice, st));
    }

    // ---- step 1: numbering
    uint8_t* cons = nullptr;
    if (in.constrained && nd > 0) {
        SYMCHECK(talloc(nd, (void**)&cons));
        SYMCHECK(hipMemcpyAsync(cons, in.constrained, nd, hipMemcpyHostToDevice, st));
    }
    int64_t *nfb = nullptr;
    int32_t *freeidx = nullptr, *localidx = nullptr;
    SYMCHECK(talloc((nd + 1) * sizeof(int64_t), (void**)&nfb));
    SYMCHECK(talloc(nd * sizeof(int32_t), (void**)&freeidx));
    SYMCHECK(talloc(nd * sizeof(int32_t), (void**)&localidx));
    {
        int64_t* mask = nullptr;
        SYMCHECK(talloc((nd + 1) * sizeof(int64_t), (void**)&mask));
        k_notc<<<blocks(nd + 1, 256), 256, 0, st>>>(nd, cons, mask);
        size_t need = 0;
        SYMCHECK(rocprim::exclusive_scan(nullptr, need, mask, nfb, int64_t(0), (size_t)(nd + 1), rocprim::plus<int64_t>(), st));
        SYMCHECK(rgrow(need));
        SYMCHECK(rocprim::exclusive_scan(rtmp, need, mask, nfb, int64_t(0), (size_t)(nd + 1), rocprim::plus<int64_t>(), st));
        SYMCHECK(hipMemcpyAsync(&out.nf, nfb + nd, sizeof(int64_t), hipMemcpyDeviceToHost, st));
        SYMCHECK(hipStreamSynchronize(st));
    }
    const int64_t nf = out.nf;
    out.nc = nd - nf;
    SYMCHECK(oalloc(out.free_dofs, nf * sizeof(int64_t)));
    SYMCHECK(oalloc(out.cons_dofs, out.nc * sizeof(int64_t)));
    if (nd > 0)
        k_index<<<blocks(nd, 256), 256, 0, st>>>(nd, cons, nfb, freeidx, localidx, (int64_t*)out.free_dofs.p, (int64_t*)out.cons_dofs.p);

    // ---- upload every distinct table once; transposed copies of what rows gather
    std::map<const void*, void*> dev, devT;
    auto up = [&](const void* h, size_t bytes, void** d) -> hipError_t {
        auto it = dev.find(h);
        if (it != dev.end()) {
            *d = it->second;
            return hipSuccess;
        }
        hipError_t e = talloc(bytes, d);
        if (e == hipSuccess) e = hipMemcpyAsync(*d, h, bytes, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) dev[h] = *d;
        return e;
    };
    auto upT = [&](const void* h, int64_t nrow, int64_t ncol, size_t esz, void** d) -> hipError_t {
        auto it = devT.find(h);
        if (it != devT.end()) {
            *d = it->second;
            return hipSuccess;
        }
        void* src = nullptr;
        hipError_t e = up(h, (size_t)(nrow * ncol) * esz, &src);
        if (e == hipSuccess) e = talloc((size_t)(nrow * ncol) * esz, d);
        if (e != hipSuccess) return e;
        if (nrow * ncol > 0) {
            if (esz == 4) k_transpose<int32_t><<<blocks(nrow * ncol, 256), 256, 0, st>>>(nrow, ncol, (const int32_t*)src, (int32_t*)*d);
            else k_transpose<int8_t><<<blocks(nrow * ncol, 256), 256, 0, st>>>(nrow, ncol, (const int8_t*)src, (int8_t*)*d);
        }
        devT[h] = *d;
        return hipSuccess;
    };
    std::vector<SymGroup> g_dev = in.groups;  // device pointers
    for (size_t g = 0; g < in.groups.size(); g++) {
        const SymGroup& G = in.groups[g];
        const int64_t n = (int64_t)G.n_tf * G.n_elem;
        void* d = nullptr;
        SYMCHECK(up(G.tf_addr, n * sizeof(int32_t), &d));
        g_dev[g].tf_addr = (const int32_t*)d;
        if (n > 0) k_range_check<<<blocks(n, 256), 256, 0, st>>>(n, (const int32_t*)d, nd, flags, F_BAD_TF);
        if (G.tf_sign) {
            SYMCHECK(up(G.tf_sign, n, &d));
            g_dev[g].tf_sign = (const int8_t*)d;
        }
    }
    std::vector<SliceT> s_dev(in.slices.size());
    for (size_t s = 0; s < in.slices.size(); s++) {
        const SymSlice& S = in.slices[s];
        SliceT& T = s_dev[s];
        T.n_dof = S.n_dof;
        T.col_off = S.col_off;
        T.n_tf = S.n_tf;
        T.pad = 0;
        void* d = nullptr;
        const bool fresh = dev.find(S.dof_addr) == dev.end();
        SYMCHECK(upT(S.dof_addr, S.n_dof, S.n_elem, 4, &d));
        T.dof = (const int32_t*)d;
        const int64_t n = (int64_t)S.n_dof * S.n_elem;
        if (fresh && n > 0) k_range_check<<<blocks(n, 256), 256, 0, st>>>(n, (const int32_t*)dev[S.dof_addr], nd, flags, F_BAD_DOF);
        T.dsg = nullptr;
        T.tsg = nullptr;
        if (S.dof_sign) {
            SYMCHECK(upT(S.dof_sign, S.n_dof, S.n_elem, 1, &d));
            T.dsg = (const int8_t*)d;
        }
        if (S.tf_sign) {
            SYMCHECK(upT(S.tf_sign, S.n_tf, S.n_elem, 1, &d));
            T.tsg = (const int8_t*)d;
        }
    }
    SymGroup* d_groups = nullptr;
    SliceT* d_slices = nullptr;
    SYMCHECK(talloc(g_dev.size() * sizeof(SymGroup), (void**)&d_groups));
    SYMCHECK(talloc(s_dev.size() * sizeof(SliceT), (void**)&d_slices));
    if (!g_dev.empty()) SYMCHECK(hipMemcpyAsync(d_groups, g_dev.data(), g_dev.size() * sizeof(SymGroup), hipMemcpyHostToDevice, st));
    if (!s_dev.empty()) SYMCHECK(hipMemcpyAsync(d_slices, s_dev.data(), s_dev.size() * sizeof(SliceT), hipMemcpyHostToDevice, st));

    // ---- step 2: instances of the matrix and the rhs groups, sorted by row
    const uint32_t sentinel = (uint32_t)nf;
    int end_bit = 1;
    while (end_bit < 32 && (uint64_t(1) << end_bit) <= sentinel) end_bit++;
    // returns the row pointers [nf + 1] and the sorted descriptors (valid ones first)
    auto collect = [&](const std::vector<int>& refs, int64_t** ptr_out, uint64_t** segs_out, int64_t* n_valid) -> hipError_t {
        int64_t n = 0;
        for (int g : refs) n += (int64_t)in.groups[g].n_tf * in.groups[g].n_elem;
        uint32_t *k0 = nullptr, *k1 = nullptr;
        uint64_t *v0 = nullptr, *v1 = nullptr;
        hipError_t e;
        if ((e = talloc(n * 4, (void**)&k0)) || (e = talloc(n * 4, (void**)&k1)) || (e = talloc(n * 8, (void**)&v0)) ||
            (e = talloc(n * 8, (void**)&v1)) || (e = talloc((nf + 1) * 8, (void**)ptr_out)))
            return e;
        int64_t t = 0;
        for (int g : refs) {
            const SymGroup& G = g_dev[g];
            const int64_t m = (int64_t)G.n_tf * G.n_elem;
            if (m > 0)
                k_instances<<<blocks(m, 256), 256, 0, st>>>(G, G.tf_addr, G.tf_sign, freeidx, nd, sentinel, k0 + t, v0 + t);
            t += m;
        }
        if (n > 0) {
            size_t need = 0;
            if ((e = rocprim::radix_sort_pairs(nullptr, need, k0, k1, v0, v1, (size_t)n, 0, end_bit, st))) return e;
            if ((e = rgrow(need))) return e;
            if ((e = rocprim::radix_sort_pairs(rtmp, need, k0, k1, v0, v1, (size_t)n, 0, end_bit, st))) return e;
        }
        k_lower_bound<<<blocks(nf + 1, 256), 256, 0, st>>>(nf, k1, n, *ptr_out);
        *segs_out = v1;
        if ((e = hipMemcpyAsync(n_valid, *ptr_out + nf, sizeof(int64_t), hipMemcpyDeviceToHost, st))) return e;
        return hipStreamSynchronize(st);
    };
    int64_t *ptr = nullptr, *rptr = nullptr;
    uint64_t *segs = nullptr, *rsegs = nullptr;
    SYMCHECK(collect(in.mat, &ptr, &segs, &out.n_erows));
    SYMCHECK(collect(in.rhs, &rptr, &rsegs, &out.n_rhs_erows));
    SYMCHECK(oalloc(out.rhs_ptr, (nf + 1) * sizeof(int64_t)));
    SYMCHECK(hipMemcpyAsync(out.rhs_ptr.p, rptr, (nf + 1) * sizeof(int64_t), hipMemcpyDeviceToDevice, st));
    SYMCHECK(oalloc(out.rhs_erows, out.n_rhs_erows * sizeof(uint64_t)));
    if (out.n_rhs_erows)
        SYMCHECK(hipMemcpyAsync(out.rhs_erows.p, rsegs, out.n_rhs_erows * sizeof(uint64_t), hipMemcpyDeviceToDevice, st));

    // row sizes and tiers
    int64_t *slotsz = nullptr, *coo = nullptr, *slot_ptr = nullptr;
    int32_t *list1 = nullptr, *list2 = nullptr;
    SYMCHECK(talloc((nf + 1) * 8, (void**)&slotsz));
    SYMCHECK(talloc((nf + 1) * 8, (void**)&coo));
    SYMCHECK(talloc((nf + 1) * 8, (void**)&slot_ptr));
    SYMCHECK(talloc(nf * 4, (void**)&list1));
    SYMCHECK(talloc(nf * 4, (void**)&list2));
    k_rowsizes<<<blocks(nf + 1, 256), 256, 0, st>>>(nf, ptr, segs, slotsz, coo, list1, list2, flags);
    {
        size_t need = 0;
        SYMCHECK(rocprim::exclusive_scan(nullptr, need, slotsz, slot_ptr, int64_t(0), (size_t)(nf + 1), rocprim::plus<int64_t>(), st));
        SYMCHECK(rgrow(need));
        SYMCHECK(rocprim::exclusive_scan(rtmp, need, slotsz, slot_ptr, int64_t(0), (size_t)(nf + 1), rocprim::plus<int64_t>(), st));
        need = 0;
        int64_t* coo_total = nullptr;
        SYMCHECK(talloc(8, (void**)&coo_total));
        SYMCHECK(rocprim::reduce(nullptr, need, coo, coo_total, int64_t(0), (size_t)nf, rocprim::plus<int64_t>(), st));
        SYMCHECK(rgrow(need));
        SYMCHECK(rocprim::reduce(rtmp, need, coo, coo_total, int64_t(0), (size_t)nf, rocprim::plus<int64_t>(), st));
        SYMCHECK(hipMemcpyAsync(&out.n_coo, coo_total, 8, hipMemcpyDeviceToHost, st));
    }
    int h_flags[F_COUNT];
    SYMCHECK(hipMemcpyAsync(h_flags, flags, sizeof(h_flags), hipMemcpyDeviceToHost, st));
    SYMCHECK(hipStreamSynchronize(st));
    auto fail_on_flags = [&]() -> bool {
        if (h_flags[F_BAD_TF]) err = "tf address out of range";
        else if (h_flags[F_BAD_DOF]) err = "dof address out of range";
        else if (h_flags[F_DUP]) err = "an element row maps two local functions to the same DOF (non-conforming addresses are not supported)";
        else if (h_flags[F_WIDE]) err = "a row has more than 65535 segments or columns";
        else if (h_flags[F_REP]) err = "internal: shared slot block used before its representative";
        return !err.empty();
    };
    if (fail_on_flags()) return hipErrorInvalidValue;
    if (h_flags[F_FALLBACK]) {
        out.fallback = true;
        return hipSuccess;
    }
    out.max_erow_len = h_flags[F_MAXLEN];
    const int64_t n1 = h_flags[F_N1], n2 = h_flags[F_N2];
    int64_t n_slots = 0;
    SYMCHECK(hipMemcpyAsync(&n_slots, slot_ptr + nf, 8, hipMemcpyDeviceToHost, st));
    SYMCHECK(hipStreamSynchronize(st));

    // ---- step 3: row columns (count pass, then write pass)
    int64_t *lenA = nullptr, *lenD = nullptr;
    SYMCHECK(talloc((nf + 1) * 8, (void**)&lenA));
    SYMCHECK(talloc((nf + 1) * 8, (void**)&lenD));
    SYMCHECK(hipMemsetAsync(lenA, 0, (nf + 1) * 8, st));
    SYMCHECK(hipMemsetAsync(lenD, 0, (nf + 1) * 8, st));
    RowArgs ra{};
    ra.ptr = ptr;
    ra.segs = segs;
    ra.coo = coo;
    ra.slot_ptr = slot_ptr;
    ra.groups = d_groups;
    ra.n_groups = (int)g_dev.size();
    ra.slices = d_slices;
    ra.cons = cons;
    ra.localidx = localidx;
    ra.n_dofs = nd;
    ra.lenA = lenA;
    ra.lenD = lenD;
    ra.flags = flags;
    const unsigned g0 = (unsigned)std::max<int64_t>(1, std::min<int64_t>(nf, (int64_t)in.n_cu * 32));
    auto grid = [&](int64_t n, int per_cu) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(n, (int64_t)in.n_cu * per_cu)); };
    auto rows_pass = [&](auto write_tag, auto slot_tag) {
        constexpr bool W = decltype(write_tag)::value;
        using T = decltype(slot_tag);
        if (nf > 0) k_rows<T0_NT, T0_CAP, T0_SEG, W, T><<<g0, T0_NT, 0, st>>>(ra, nullptr, nf);
        if (n1 > 0) k_rows<T1_NT, T1_CAP, T1_SEG, W, T><<<grid(n1, 8), T1_NT, 0, st>>>(ra, list1, n1);
        if (n2 > 0) k_rows<T2_NT, T2_CAP, T2_SEG, W, T><<<grid(n2, 1), T2_NT, 0, st>>>(ra, list2, n2);
    };
    rows_pass(std::false_type{}, uint8_t{});
    SYMCHECK(hipGetLastError());
    SYMCHECK(hipMemcpyAsync(h_flags, flags, sizeof(h_flags), hipMemcpyDeviceToHost, st));
    SYMCHECK(hipStreamSynchronize(st));
    if (fail_on_flags()) return hipErrorInvalidValue;
    out.maxrow = h_flags[F_MAXROW];
    {
        const int mr = out.maxrow;
        const bool sg = in.signed_slots;
        out.slot_bytes = mr < (sg ? 127 : 255) ? 1 : mr < (sg ? 32767 : 65535) ? 2 : 4;
    }
    const int sb = out.slot_bytes;
    SYMCHECK(oalloc(out.Arow, (nf + 1) * 8));
    SYMCHECK(oalloc(out.Drow, (nf + 1) * 8));
    {
        size_t need = 0;
        SYMCHECK(rocprim::exclusive_scan(nullptr, need, lenA, (int64_t*)out.Arow.p, int64_t(0), (size_t)(nf + 1), rocprim::plus<int64_t>(), st));
        SYMCHECK(rgrow(need));
        SYMCHECK(rocprim::exclusive_scan(rtmp, need, lenA, (int64_t*)out.Arow.p, int64_t(0), (size_t)(nf + 1), rocprim::plus<int64_t>(), st));
        SYMCHECK(rocprim::exclusive_scan(rtmp, need, lenD, (int64_t*)out.Drow.p, int64_t(0), (size_t)(nf + 1), rocprim::plus<int64_t>(), st));
        out.h_Arow.resize(nf + 1);
        out.h_Drow.resize(nf + 1);
        SYMCHECK(hipMemcpyAsync(out.h_Arow.data(), out.Arow.p, (nf + 1) * 8, hipMemcpyDeviceToHost, st));
        SYMCHECK(hipMemcpyAsync(out.h_Drow.data(), out.Drow.p, (nf + 1) * 8, hipMemcpyDeviceToHost, st));
        SYMCHECK(hipStreamSynchronize(st));
    }
    out.nnzA = out.h_Arow[nf];
    out.nnzD = out.h_Drow[nf];
    if (out.nnzA >= (int64_t(1) << 48) || out.nnzD >= (int64_t(1) << 48)) {
        err = "matrix exceeds the 48-bit index range";
        return hipErrorInvalidValue;
    }
    SYMCHECK(oalloc(out.Acol, out.nnzA * 4));
    SYMCHECK(oalloc(out.Dcol, out.nnzD * 4));
    void* slots_row = nullptr;
    SYMCHECK(talloc((size_t)n_slots * sb, &slots_row));
    SYMCHECK(hipMemsetAsync(slots_row, 0xFF, (size_t)n_slots * sb, st));
    ra.Arow = (const int64_t*)out.Arow.p;
    ra.Drow = (const int64_t*)out.Drow.p;
    ra.Acol = (int32_t*)out.Acol.p;
    ra.Dcol = (int32_t*)out.Dcol.p;
    ra.slots = slots_row;
    if (sb == 1) rows_pass(std::true_type{}, uint8_t{});
    else if (sb == 2) rows_pass(std::true_type{}, uint16_t{});
    else rows_pass(std::true_type{}, uint32_t{});
    SYMCHECK(hipGetLastError());

    // ---- step 4: representatives of identical slot blocks
    int32_t* rep = nullptr;
    SYMCHECK(talloc(nf * 4, (void**)&rep));
    if (nf > 0) {
        uint64_t *h0 = nullptr, *h1 = nullptr;
        int32_t *r0 = nullptr, *r1 = nullptr, *head = nullptr, *run = nullptr;
        SYMCHECK(talloc(nf * 8, (void**)&h0));
        SYMCHECK(talloc(nf * 8, (void**)&h1));
        SYMCHECK(talloc(nf * 4, (void**)&r0));
        SYMCHECK(talloc(nf * 4, (void**)&r1));
        SYMCHECK(talloc(nf * 4, (void**)&head));
        SYMCHECK(talloc(nf * 4, (void**)&run));
        k_hash<<<blocks(nf, 256), 256, 0, st>>>(nf, slot_ptr, (const uint32_t*)slots_row, sb, h0, r0);
        size_t need = 0;
        SYMCHECK(rocprim::radix_sort_pairs(nullptr, need, h0, h1, r0, r1, (size_t)nf, 0, 64, st));
        SYMCHECK(rgrow(need));
        SYMCHECK(rocprim::radix_sort_pairs(rtmp, need, h0, h1, r0, r1, (size_t)nf, 0, 64, st));
        k_run_head<<<blocks(nf, 256), 256, 0, st>>>(nf, h1, head);
        need = 0;
        SYMCHECK(rocprim::inclusive_scan(nullptr, need, head, run, (size_t)nf, rocprim::maximum<int32_t>(), st));
        SYMCHECK(rgrow(need));
        SYMCHECK(rocprim::inclusive_scan(rtmp, need, head, run, (size_t)nf, rocprim::maximum<int32_t>(), st));
        k_rep<<<blocks(nf, 256), 256, 0, st>>>(nf, r1, run, slot_ptr, (const uint32_t*)slots_row, sb, rep);
    }

    // ---- step 5: work order, descriptors, segments and slot blocks
    SYMCHECK(oalloc(out.rdesc, nf * 4 * 8));
    SYMCHECK(oalloc(out.erows_w, out.n_erows * 8));
    if (nf > 0) {
        uint32_t *k0 = nullptr, *k1 = nullptr;
        int32_t *o0 = nullptr, *order = nullptr, *pos_of = nullptr;
        int64_t *nk = nullptr, *own = nullptr, *kw = nullptr, *sw = nullptr;
        SYMCHECK(talloc(nf * 4, (void**)&k0));
        SYMCHECK(talloc(nf * 4, (void**)&k1));
        SYMCHECK(talloc(nf * 4, (void**)&o0));
        SYMCHECK(talloc(nf * 4, (void**)&order));
        SYMCHECK(talloc(nf * 4, (void**)&pos_of));
        SYMCHECK(talloc((nf + 1) * 8, (void**)&nk));
        SYMCHECK(talloc((nf + 1) * 8, (void**)&own));
        SYMCHECK(talloc((nf + 1) * 8, (void**)&kw));
        SYMCHECK(talloc((nf + 1) * 8, (void**)&sw));
        k_order_keys<<<blocks(nf, 256), 256, 0, st>>>(nf, lenA, lenD, k0, o0);
        size_t need = 0;
        SYMCHECK(rocprim::radix_sort_pairs(nullptr, need, k0, k1, o0, order, (size_t)nf, 0, 16, st));
        SYMCHECK(rgrow(need));
        SYMCHECK(rocprim::radix_sort_pairs(rtmp, need, k0, k1, o0, order, (size_t)nf, 0, 16, st));
        k_work1<<<blocks(nf + 1, 256), 256, 0, st>>>(nf, order, ptr, slot_ptr, rep, pos_of, nk, own);
        need = 0;
        SYMCHECK(rocprim::exclusive_scan(nullptr, need, nk, kw, int64_t(0), (size_t)(nf + 1), rocprim::plus<int64_t>(), st));
        SYMCHECK(rgrow(need));
        SYMCHECK(rocprim::exclusive_scan(rtmp, need, nk, kw, int64_t(0), (size_t)(nf + 1), rocprim::plus<int64_t>(), st));
        SYMCHECK(rocprim::exclusive_scan(rtmp, need, own, sw, int64_t(0), (size_t)(nf + 1), rocprim::plus<int64_t>(), st));
        SYMCHECK(hipMemcpyAsync(&out.n_slots_w, sw + nf, 8, hipMemcpyDeviceToHost, st));
        SYMCHECK(hipStreamSynchronize(st));
        SYMCHECK(oalloc(out.slots_w, (size_t)std::max<int64_t>(out.n_slots_w, 1) * sb));
        if (out.n_slots_w == 0) SYMCHECK(hipMemsetAsync(out.slots_w.p, 0, sb, st));
        k_work2<<<blocks(nf, 256), 256, 0, st>>>(nf, order, pos_of, rep, ptr, kw, sw, lenA, lenD, (const int64_t*)out.Arow.p,
                                                  (const int64_t*)out.Drow.p, segs, slot_ptr, (const uint32_t*)slots_row, sb,
                                                  (uint64_t*)out.rdesc.p, (uint64_t*)out.erows_w.p, (uint32_t*)out.slots_w.p, flags);
        // row lengths in work order (reduction launch classes)
        std::vector<uint32_t> hk(nf);
        SYMCHECK(hipMemcpyAsync(hk.data(), k1, nf * 4, hipMemcpyDeviceToHost, st));
        SYMCHECK(hipMemcpyAsync(h_flags, flags, sizeof(h_flags), hipMemcpyDeviceToHost, st));
        SYMCHECK(hipStreamSynchronize(st));
        out.h_work_len.resize(nf);
        for (int64_t p = 0; p < nf; p++) out.h_work_len[p] = 65535 - (int32_t)hk[p];
        if (fail_on_flags()) return hipErrorInvalidValue;
    } else {
        SYMCHECK(oalloc(out.slots_w, sb));
        SYMCHECK(hipMemsetAsync(out.slots_w.p, 0, sb, st));
    }
    SYMCHECK(hipGetLastError());
    SYMCHECK(hipStreamSynchronize(st));
    guard.ok = true;
    return hipSuccess;
}

}  // namespace slx
