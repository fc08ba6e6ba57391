// Assembly plan: symbolic phase (CSR pattern and reduction maps) and the
// device buffers of the numeric phase, plus the C ABI of include/slx.h.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <atomic>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <thread>
#include <vector>
#include "../../include/slx.h"
#include "host/basis.h"
#include "host/gauss.h"
#include "host/problems.h"
#include "kernels/device.h"
#include "kernels/mfma.h"
#include "kernels/factor.h"
#include "kernels/tile.h"
#include "kernels/solve.h"
#include "kernels/symbolic.h"
#include "coefjit.h"
#include <array>
#include <mutex>
#include <cmath>
#include <cstdio>
#include <dlfcn.h>

namespace slx {
hipError_t launch_coef(const DevBucket& b, const DevInputs& in, bool fp64, void* ctab, hipStream_t st, int qfast);
hipError_t launch_elem(const DevBucket& b, const void* ctab, int precision, void* ke, hipStream_t st, uint32_t slice_mask, int qfast);
hipError_t scratch_init(int device);
hipError_t launch_reduce(const DevReduce& rd, int max_erow_len, bool ke_fp64, const void* ke, double* Aval, double* Dval,
                         double* rhs, bool has_matrix, bool has_rhs, hipStream_t st, const int64_t* class_begin,
                         const int64_t* class_end, const int* class_len, int n_classes);
}  // namespace slx

using namespace slx;

namespace {
thread_local std::string g_last_error;

#define HIPCHECK(x)                                                                                  \
    do {                                                                                             \
        hipError_t _e = (x);                                                                         \
        if (_e != hipSuccess) throw DeviceError(std::string(#x) + ": " + hipGetErrorString(_e));   \
    } while (0)

struct DeviceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct UnsupportedError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

template <typename F>
slx_status guarded(F&& f) {
    try {
        f();
        return SLX_OK;
    } catch (const DeviceError& e) {
        g_last_error = e.what();
        return SLX_ERR_DEVICE;
    } catch (const UnsupportedError& e) {
        g_last_error = e.what();
        return SLX_ERR_UNSUPPORTED;
    } catch (const std::bad_alloc& e) {
        g_last_error = "allocation failed";
        return SLX_ERR_NOMEM;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return SLX_ERR_INVALID;
    }
}

// Device allocation owned by a plan.
struct DBuf {
    void* p = nullptr;
    size_t n = 0;
    ~DBuf() {
        if (p) (void)hipFree(p);
    }
    void alloc(size_t bytes) {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = bytes;
        if (bytes) HIPCHECK(hipMalloc(&p, bytes));
    }
    // take over a hipMalloc'd allocation
    void adopt(void* ptr, size_t bytes) {
        if (p) (void)hipFree(p);
        p = ptr;
        n = bytes;
    }
    template <typename T>
    T* as() const { return (T*)p; }
};

// Synchronous upload: the host staging vectors of the symbolic phase are
// temporaries, so the copy completes before returning.
template <typename T>
void upload(DBuf& d, const T* src, size_t count, hipStream_t st) {
    (void)st;
    d.alloc(count * sizeof(T));
    if (count && src) HIPCHECK(hipMemcpy(d.p, src, count * sizeof(T), hipMemcpyHostToDevice));
    else if (count) HIPCHECK(hipMemset(d.p, 0, count * sizeof(T)));
}

// Host slots: bit 30 marks a negative entry (folded orientations); -1 = skip.
#define SLX_SLOT_NEG 0x40000000

// Narrowest device slot type for rows of up to maxrow columns (all ones = skip;
// signed plans give up the top bit).
int slot_bytes_for(int maxrow, bool sgn) {
    if (maxrow < (sgn ? 127 : 255)) return 1;
    if (maxrow < (sgn ? 32767 : 65535)) return 2;
    return 4;
}

}  // namespace

struct slx_ctx_s {
    int device = 0;
    hipStream_t stream = nullptr;
    // the context and every plan created on it hold one reference: the stream
    // lives until the last of slx_destroy and slx_plan_destroy
    std::atomic<int> refs{1};
};

namespace {
void ctx_release(slx_ctx c) {
    if (c && c->refs.fetch_sub(1) == 1) {
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
    }
}
}  // namespace

struct RedClass {
    int64_t begin, end;  // positions in the row order
    int max_len;         // longest row of the class (A + D columns)
};

struct slx_plan_s {
    slx_ctx ctx = nullptr;
    hipStream_t stream = nullptr;
    const slx_plan_desc* src = nullptr;
    int64_t n_dofs = 0, nfree = 0, ncons = 0, nnzA = 0, nnzD = 0;
    int64_t n_ke = 0, n_erows = 0, n_coo = 0;
    int mesh_dim = 0, n_fields = 0;
    bool has_matrix = false, has_rhs = false;
    int max_erow_len = 1;
    // host CSR pattern
    std::vector<int64_t> Arow, Drow;
    std::vector<int32_t> Acol, Dcol;
    // pattern-only plans keep the reduction maps and the Ke layout on the host
    std::vector<int64_t> h_erow_ptr, h_slot_ptr, h_rhs_ptr, h_ke_base, h_ke_ld;
    std::vector<uint64_t> h_erows, h_rhs_erows;
    std::vector<int32_t> h_slots;
    // device: per-bucket data
    std::vector<std::unique_ptr<DBuf>> bufs;
    std::vector<DevBucket> dbuckets;
    DBuf consts;
    DBuf d_Arow, d_Drow, d_Acol, d_Dcol, d_rdesc, d_erows, d_slots, d_rhs_ptr, d_rhs_erows, d_free, d_cons;
    DBuf d_ablk;          // row blocks of the streaming FP32 product (solve)
    int64_t n_ablk = 0;
    std::vector<RedClass> red_classes;  // reduction launch classes (rows of similar length)
    DBuf d_Aval, d_Dval, d_rhs, d_ke, d_ctab;
    size_t ke_capacity_bytes = 0;
    DevReduce rd{};
    // numeric inputs staged by slx_assemble
    DBuf d_xyz, d_params;
    std::vector<std::unique_ptr<DBuf>> d_fields;
    int last_precision = -1;
    int max_param = -1;            // highest parameter index read by the programs
    int64_t max_elem_id = 0;       // parameter arrays must hold this many entries
    bool uses_fields = false;
    bool host_only = false;        // symbolic phase only (no device): pattern inspection
    bool symbolic_on_device = false;  // the pattern and maps were built by the device phase
    int64_t n_slots_w = 0, n_rhs_erows = 0;  // work-order slot entries, rhs segments
    double symbolic_seconds = 0.0;
    bool signed_slots = false;     // folded orientations: signed segment entries

    hipGraphExec_t graph_exec = nullptr;  // captured numeric phase (slx_plan_graph_capture)
    // right-hand-side-only buckets run, with the rhs reduction, on a side
    // stream forked from and joined back into the caller's stream; they keep
    // their coefficients in their own table
    std::vector<uint8_t> rhs_only;     // per bucket: every slice is a rhs slice
    bool side_ok = false;              // every rhs slice lives in a rhs-only bucket
    // compiled coefficient programs per bucket ([0] FP32, [1] FP64; null: interpreter)
    std::vector<std::array<hipFunction_t, 2>> coef_fn;
    int coef_mode = 0;                 // 0 auto, 1 interpreter only, 2 compiled everywhere
    std::string coef_jit_error;        // last compiler message (diagnostics)
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    DBuf d_ctab_side;
    std::vector<MfmaBucket> mfma;      // per bucket; M == 0 when not eligible
    std::vector<uint32_t> mfma_mask;   // matrix slices handled by the MFMA kernel
    std::vector<std::vector<std::pair<int, FactorSlice>>> factor;  // per bucket: (slice, factorized MFMA data)
    std::vector<uint32_t> factor_mask; // matrix slices handled by the factorized MFMA kernel
    std::vector<std::vector<std::pair<int, TileSlice>>> tile;  // per bucket: (slice, one-tile-per-element data), MFMA arithmetics
    std::vector<uint32_t> tile_mask;   // matrix slices handled by the one-tile kernel in the MFMA arithmetics
    int last_qfast = 0;                // coefficient table layout of the last bucket assembled
    bool factor64 = true;              // FP64 mode: factorized slices on v_mfma_f64 (SLX_TUNE_FP64_FACTOR=0: VALU)
    // fused launch: the factorized slices of every bucket in one persistent
    // launch (per arithmetic: F16, BF16, F32, and 3 = FP64); each fused bucket
    // keeps its coefficients in its own region of d_ctab_fused
    bool fused = false;
    uint32_t fused_arith = 0;              // bit a: arithmetic a (3 = FP64) uses the fused launch
    std::vector<uint8_t> fuse_bucket;      // per bucket: its factorized slices run in the fused launch
    std::vector<size_t> fuse_off;          // per bucket: byte offset of its coefficient region
    DBuf d_ctab_fused, d_fused_counter;
    DBuf d_fused[4];
    int n_fused[4] = {0, 0, 0, 0};
    int64_t fused_total[4] = {0, 0, 0, 0};
    size_t fused_lds[4] = {0, 0, 0, 0};
    bool last_fused = false;
    const void* last_ctab = nullptr;   // coefficient table of the last bucket assembled
    int n_cu = 256;

    DBuf* newbuf() {
        bufs.emplace_back(new DBuf());
        return bufs.back().get();
    }
    template <typename T>
    const T* dev(const T* src, size_t count) {
        DBuf* b = newbuf();
        upload(*b, src, count, stream);
        return b->as<T>();
    }
};

namespace {

// Element-row segment groups.  The matrix slices of a bucket that share the
// tf function set (identical tf address table) are stored side by side: the
// Ke entry (slice s, element e, tf function i, dof function j) lives at
//   ke_base[s] + (e * n_tf + i) * ke_ld[s] + j,
//   ke_base[s] = group base + column offset of s,  ke_ld[s] = total columns,
// so one CSR row reads one contiguous segment per (element, tf function)
// whatever the number of dof fields.  Matrix segments are strided by a
// multiple of 4 entries (16 when 16 or longer) from a 16-aligned base: the
// reduction reads them as aligned 4-entry vectors, and a 32-entry FP32
// segment is exactly one 128-byte line.  The padding is never read into a sum.  Slices whose dof address tables are
// identical are kept in separate groups (their columns would coincide).
// Right-hand-side slices get one value per (element, tf function).
struct SegGroup {
    int bucket;
    std::vector<int> slices;
    int n_tf;
    int64_t len;      // segment length (sum of the slices' n_dof, 1 for rhs)
    int64_t ld = 0;       // segment stride: len, rounded up to a multiple of 16 when len >= 16
    int64_t base;     // Ke offset of the group (a multiple of 16)
    int rank;         // smallest fragment rank of the group's slices
    bool rhs;
};

struct KeLayout {
    std::vector<SegGroup> groups;
    std::vector<std::vector<int64_t>> ke_base, ke_ld;  // [bucket][slice]
    std::vector<std::vector<int>> group_of;            // [bucket][slice]
    int64_t n_ke = 0;
};

uint64_t hash_ints(const int32_t* a, size_t n) {
    uint64_t h = 1469598103934665603ull;
    for (size_t i = 0; i < n; i++) h = (h ^ (uint32_t)a[i]) * 1099511628211ull;
    return h;
}

KeLayout compute_layout(const slx_plan_desc& d) {
    KeLayout L;
    L.ke_base.resize(d.n_buckets);
    L.ke_ld.resize(d.n_buckets);
    L.group_of.resize(d.n_buckets);
    for (int b = 0; b < d.n_buckets; b++) {
        const slx_bucket_desc& bd = d.buckets[b];
        L.ke_base[b].assign(bd.n_slices, 0);
        L.ke_ld[b].assign(bd.n_slices, 0);
        L.group_of[b].assign(bd.n_slices, -1);
        std::vector<uint64_t> htf(bd.n_slices, 0), hdof(bd.n_slices, 0);
        for (int s = 0; s < bd.n_slices; s++) {
            const slx_slice_desc& sd = bd.slices[s];
            htf[s] = hash_ints(sd.tf_addr, (size_t)sd.n_tf * bd.n_elem);
            if (!sd.is_rhs) hdof[s] = hash_ints(sd.dof_addr, (size_t)sd.n_dof * bd.n_elem);
        }
        auto same = [&](const int32_t* x, const int32_t* y, size_t n) { return x == y || std::memcmp(x, y, n * sizeof(int32_t)) == 0; };
        size_t first = L.groups.size();
        for (int s = 0; s < bd.n_slices; s++) {
            const slx_slice_desc& sd = bd.slices[s];
            int join = -1;
            if (!sd.is_rhs)
                for (size_t g = first; g < L.groups.size() && join < 0; g++) {
                    const SegGroup& G = L.groups[g];
                    if (G.rhs || G.n_tf != sd.n_tf) continue;
                    const int s0 = G.slices[0];
                    if (htf[s0] != htf[s] || !same(bd.slices[s0].tf_addr, sd.tf_addr, (size_t)sd.n_tf * bd.n_elem)) continue;
                    bool clash = false;
                    for (int t : G.slices)
                        if (bd.slices[t].n_dof == sd.n_dof && hdof[t] == hdof[s] &&
                            same(bd.slices[t].dof_addr, sd.dof_addr, (size_t)sd.n_dof * bd.n_elem))
                            clash = true;
                    if (!clash) join = (int)g;
                }
            if (join < 0) {
                SegGroup G;
                G.bucket = b;
                G.n_tf = sd.n_tf;
                G.len = 0;
                G.base = 0;
                G.rank = sd.frag_rank;
                G.rhs = sd.is_rhs != 0;
                L.groups.push_back(G);
                join = (int)L.groups.size() - 1;
            }
            SegGroup& G = L.groups[join];
            L.group_of[b][s] = join;
            L.ke_base[b][s] = G.len;  // column offset for now
            G.slices.push_back(s);
            G.len += sd.is_rhs ? 1 : sd.n_dof;
            G.rank = std::min(G.rank, sd.frag_rank);
        }
        for (size_t g = first; g < L.groups.size(); g++) {
            SegGroup& G = L.groups[g];
            SLX_CHECK(G.len < 65536, "element-row segment longer than 65535 entries");
            G.ld = G.rhs ? G.len : (G.len >= 16 ? (G.len + 15) / 16 * 16 : (G.len + 3) / 4 * 4);
            L.n_ke = (L.n_ke + 15) / 16 * 16;
            G.base = L.n_ke;
            for (int s : G.slices) {
                L.ke_base[b][s] += G.base;
                L.ke_ld[b][s] = G.ld;
            }
            L.n_ke += (int64_t)bd.n_elem * G.n_tf * G.ld;
        }
    }
    return L;
}

// Reduction launch classes over the work order (rows by decreasing length):
// each class's longest row is at most twice its shortest row (LDS sized per
// class).  `len` = A + D columns of the row at each work position.
void reduction_classes(slx_plan_s& P, const std::vector<int32_t>& len) {
    const int64_t nf = (int64_t)len.size();
    P.red_classes.clear();
    for (int64_t i = 0; i < nf;) {
        const int top = std::max(1, len[i]);
        int64_t j = i;
        while (j < nf && 2 * len[j] > top) j++;
        if (j == i) j = i + 1;
        P.red_classes.push_back({i, j, top});
        i = j;
    }
}

// Device buffers every plan needs once its pattern and reduction maps are on
// the device: values, the rhs with its Dirichlet entries, the reduction
// descriptor, the solve's row blocks and the element-matrix buffer.
void finish_plan(slx_plan_s& P, const slx_plan_desc& d, int slot_bytes, int maxrow) {
    hipStream_t st = P.stream;
    const int64_t nd = d.n_dofs, nf = P.nfree;
    P.d_Aval.alloc(std::max<int64_t>(P.nnzA, 1) * sizeof(double));
    P.d_Dval.alloc(std::max<int64_t>(P.nnzD, 1) * sizeof(double));
    // rhs: constrained entries hold the Dirichlet values (formulation::b(),
    // vec.cpp:43-61); the free entries are rewritten by every assembly
    {
        std::vector<double> b0(std::max<int64_t>(nd, 1), 0.0);
        if (d.dirichlet && d.constrained)
            for (int64_t i = 0; i < nd; i++)
                if (d.constrained[i]) b0[i] = d.dirichlet[i];
        upload(P.d_rhs, b0.data(), b0.size(), st);
    }
    P.rd.nrows = nf;
    P.rd.rdesc = P.d_rdesc.as<uint64_t>();
    P.rd.erows = P.d_erows.as<uint64_t>();
    P.rd.slots = P.d_slots.p;
    P.rd.slot_bytes = slot_bytes;
    P.rd.slot_signed = P.signed_slots ? 1 : 0;
    P.rd.n_cu = P.n_cu;
    P.rd.max_row_len = maxrow;
    P.rd.rhs_erow_ptr = P.d_rhs_ptr.as<int64_t>();
    P.rd.rhs_erows = P.d_rhs_erows.as<uint64_t>();
    P.rd.free_dofs = P.d_free.as<int64_t>();
    // row blocks of the solve's streaming product (kept with the pattern, so
    // that no solve allocates beyond its workspace)
    if (nf > 0) {
        std::vector<int64_t> blk;
        stream_row_blocks(P.Arow.data(), nf, blk);
        P.n_ablk = (int64_t)blk.size() - 1;
        upload(P.d_ablk, blk.data(), blk.size(), st);
    }
    // Ke buffer (FP64 capacity serves every precision)
    P.ke_capacity_bytes = std::max<int64_t>(P.n_ke, 1) * sizeof(double);
    P.d_ke.alloc(P.ke_capacity_bytes);
}

// Symbolic phase on the host: element-row lists in the reference
// accumulation order, per-row sorted unique columns, A/D split with local
// renumbering (gentools::findtruefalse, gentools.cpp:2040-2068), slots.
void build_symbolic(slx_plan_s& P, const slx_plan_desc& d, const KeLayout& L) {
    const int64_t nd = d.n_dofs;
    std::vector<int64_t> freeidx(nd), localidx(nd);
    int64_t nf = 0, nc = 0;
    for (int64_t i = 0; i < nd; i++) {
        if (d.constrained && d.constrained[i]) localidx[i] = nc++;
        else { localidx[i] = nf; freeidx[i] = nf++; }
        if (d.constrained && d.constrained[i]) freeidx[i] = -1;
    }
    P.nfree = nf;
    P.ncons = nc;
    std::vector<int64_t> free_dofs(nf);
    for (int64_t i = 0; i < nd; i++)
        if (freeidx[i] >= 0) free_dofs[freeidx[i]] = i;
    P.n_ke = L.n_ke;
    // segment groups in fragment (accumulation) order
    std::vector<int> mat, rhs;
    for (size_t g = 0; g < L.groups.size(); g++) (L.groups[g].rhs ? rhs : mat).push_back((int)g);
    auto grank = [&](int g) { return L.groups[g].rank; };
    std::stable_sort(mat.begin(), mat.end(), [&](int a, int b) { return grank(a) < grank(b); });
    std::stable_sort(rhs.begin(), rhs.end(), [&](int a, int b) { return grank(a) < grank(b); });

    // element-row segments per free row: counting sort keeps (group, i, element) order
    auto collect = [&](const std::vector<int>& refs, std::vector<int64_t>& ptr, std::vector<uint64_t>& erows,
                       std::vector<int32_t>* ref_idx, std::vector<int32_t>* elem_idx, std::vector<int32_t>* row_idx) {
        ptr.assign(nf + 1, 0);
        for (int g : refs) {
            const SegGroup& G = L.groups[g];
            const slx_bucket_desc& bd = d.buckets[G.bucket];
            const slx_slice_desc& sd = bd.slices[G.slices[0]];
            for (int i = 0; i < G.n_tf; i++)
                for (int e = 0; e < bd.n_elem; e++) {
                    int32_t a = sd.tf_addr[(size_t)i * bd.n_elem + e];
                    if (a < 0 || (d.constrained && d.constrained[a])) continue;
                    ptr[freeidx[a] + 1]++;
                }
        }
        for (int64_t i = 0; i < nf; i++) ptr[i + 1] += ptr[i];
        erows.resize(ptr[nf]);
        if (ref_idx) ref_idx->resize(ptr[nf]);
        if (elem_idx) elem_idx->resize(ptr[nf]);
        if (row_idx) row_idx->resize(ptr[nf]);
        std::vector<int64_t> fill(ptr.begin(), ptr.end() - 1);
        for (size_t ri = 0; ri < refs.size(); ri++) {
            const SegGroup& G = L.groups[refs[ri]];
            const slx_bucket_desc& bd = d.buckets[G.bucket];
            const slx_slice_desc& sd = bd.slices[G.slices[0]];
            for (int i = 0; i < G.n_tf; i++)
                for (int e = 0; e < bd.n_elem; e++) {
                    int32_t a = sd.tf_addr[(size_t)i * bd.n_elem + e];
                    if (a < 0 || (d.constrained && d.constrained[a])) continue;
                    int64_t pos = fill[freeidx[a]]++;
                    int64_t off = G.base + ((int64_t)e * G.n_tf + i) * G.ld;
                    erows[pos] = ((uint64_t)G.len << 48) | (uint64_t)off;
                    // rhs segments (length 1): bit 63 = negative tf sign (folded orientations)
                    if (G.rhs && sd.tf_sign && sd.tf_sign[(size_t)i * bd.n_elem + e] < 0) erows[pos] |= uint64_t(1) << 63;
                    if (ref_idx) (*ref_idx)[pos] = (int32_t)ri;
                    if (elem_idx) (*elem_idx)[pos] = e;
                    if (row_idx) (*row_idx)[pos] = i;
                }
        }
    };
    SLX_CHECK(L.n_ke < (int64_t(1) << 48), "element matrices exceed the 48-bit offset range");
    std::vector<int64_t> erow_ptr, rhs_ptr;
    std::vector<uint64_t> erows, rhs_erows;
    std::vector<int32_t> eref, eelem, erowi;
    collect(mat, erow_ptr, erows, &eref, &eelem, &erowi);
    collect(rhs, rhs_ptr, rhs_erows, nullptr, nullptr, nullptr);
    // folded orientations: a sign per segment entry (tf sign of the row x dof sign of the column)
    bool signed_slots = false;
    for (const SegGroup& G : L.groups)
        for (int s : G.slices) {
            const slx_slice_desc& sd = d.buckets[G.bucket].slices[s];
            if (!G.rhs && (sd.tf_sign || sd.dof_sign)) signed_slots = true;
        }
    P.signed_slots = signed_slots;
    P.n_erows = (int64_t)erows.size();
    // slot pointers: every segment's slot block padded to 4 entries (aligned vector loads)
    std::vector<int64_t> slot_ptr(nf + 1, 0);
    int64_t n_coo = 0;
    for (int64_t r = 0; r < nf; r++) {
        int64_t s = 0;
        for (int64_t k = erow_ptr[r]; k < erow_ptr[r + 1]; k++) {
            const int64_t len = (int64_t)(erows[k] >> 48);
            s += (len + 3) / 4 * 4;
            n_coo += len;
        }
        slot_ptr[r + 1] = slot_ptr[r] + s;
    }
    P.n_coo = n_coo;
    std::vector<int32_t> slots(slot_ptr[nf], -1);
    std::vector<int32_t> lenA(nf), lenD(nf);
    int nthreads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (nf < 20000) nthreads = 1;
    std::vector<std::vector<int32_t>> tA(nthreads), tD(nthreads);
    std::vector<int> tmax(nthreads, 1), tmaxrow(nthreads, 0);
    std::atomic<bool> dup_error{false};
    // the columns of segment k of a row: dof addresses of every slice of its group
    auto for_cols = [&](int64_t k, auto&& fn) {
        const SegGroup& G = L.groups[mat[eref[k]]];
        const slx_bucket_desc& bd = d.buckets[G.bucket];
        int e = eelem[k], i = erowi[k];
        int pos = 0;
        for (int s : G.slices) {
            const slx_slice_desc& sd = bd.slices[s];
            const int rs = sd.tf_sign ? sd.tf_sign[(size_t)i * bd.n_elem + e] : 1;
            for (int j = 0; j < sd.n_dof; j++) {
                const int cs = sd.dof_sign ? sd.dof_sign[(size_t)j * bd.n_elem + e] : 1;
                fn(pos++, sd.dof_addr[(size_t)j * bd.n_elem + e], rs * cs < 0);
            }
        }
        return (int)G.len;
    };
    auto work = [&](int t) {
        int64_t r0 = nf * t / nthreads, r1 = nf * (t + 1) / nthreads;
        std::vector<int64_t> cols;
        std::vector<int32_t> rankv;
        for (int64_t r = r0; r < r1; r++) {
            cols.clear();
            for (int64_t k = erow_ptr[r]; k < erow_ptr[r + 1]; k++) {
                int len = for_cols(k, [&](int, int32_t c, bool) {
                    if (c >= 0) cols.push_back(c);
                });
                tmax[t] = std::max(tmax[t], len);
            }
            std::sort(cols.begin(), cols.end());
            cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
            // A (free) columns first, then D (constrained), both ascending
            int la = 0, ldd = 0;
            for (int64_t c : cols) (d.constrained && d.constrained[c]) ? ldd++ : la++;
            lenA[r] = la;
            lenD[r] = ldd;
            tmaxrow[t] = std::max(tmaxrow[t], la + ldd);
            for (int64_t c : cols)
                if (!(d.constrained && d.constrained[c])) tA[t].push_back((int32_t)localidx[c]);
            for (int64_t c : cols)
                if (d.constrained && d.constrained[c]) tD[t].push_back((int32_t)localidx[c]);
            // rank of every column among the A or the D columns of the row
            rankv.resize(cols.size());
            {
                int na = 0, ndd = 0;
                for (size_t q = 0; q < cols.size(); q++) {
                    bool qD = d.constrained && d.constrained[cols[q]];
                    rankv[q] = qD ? la + ndd++ : na++;
                }
            }
            int64_t sp = slot_ptr[r];
            for (int64_t k = erow_ptr[r]; k < erow_ptr[r + 1]; k++) {
                int len = for_cols(k, [&](int j, int32_t c, bool neg) {
                    int32_t slot = -1;
                    if (c >= 0) slot = rankv[std::lower_bound(cols.begin(), cols.end(), (int64_t)c) - cols.begin()] | (neg ? SLX_SLOT_NEG : 0);
                    slots[sp + j] = slot;
                });
                // slots of one segment must be distinct for the lockstep reduction
                for (int j = 0; j < len; j++)
                    for (int j2 = j + 1; j2 < len; j2++)
                        if (slots[sp + j] >= 0 && (slots[sp + j] & ~SLX_SLOT_NEG) == (slots[sp + j2] & ~SLX_SLOT_NEG)) dup_error = true;
                sp += (len + 3) / 4 * 4;
            }
        }
    };
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++) th.emplace_back(work, t);
    for (auto& x : th) x.join();
    if (dup_error) fail("an element row maps two local functions to the same DOF (non-conforming addresses are not supported)");
    P.Arow.assign(nf + 1, 0);
    P.Drow.assign(nf + 1, 0);
    for (int64_t r = 0; r < nf; r++) {
        P.Arow[r + 1] = P.Arow[r] + lenA[r];
        P.Drow[r + 1] = P.Drow[r] + lenD[r];
    }
    P.Acol.clear();
    P.Dcol.clear();
    P.Acol.reserve(P.Arow[nf]);
    P.Dcol.reserve(P.Drow[nf]);
    for (int t = 0; t < nthreads; t++) {
        P.Acol.insert(P.Acol.end(), tA[t].begin(), tA[t].end());
        P.Dcol.insert(P.Dcol.end(), tD[t].begin(), tD[t].end());
    }
    P.nnzA = P.Arow[nf];
    P.nnzD = P.Drow[nf];
    int maxrow = 1;
    for (int t = 0; t < nthreads; t++) {
        P.max_erow_len = std::max(P.max_erow_len, tmax[t]);
        maxrow = std::max(maxrow, tmaxrow[t]);
    }
    for (int b = 0; b < d.n_buckets; b++)
        for (int s = 0; s < d.buckets[b].n_slices; s++) {
            P.h_ke_base.push_back(L.ke_base[b][s]);
            P.h_ke_ld.push_back(L.ke_ld[b][s]);
        }
    // rows whose slot blocks are identical share one copy: the components of a
    // vector field at one node have the same columns and the same segments, and
    // on meshes with a repeating topology most rows of one kind have the same
    // block (a jittered 16^3 tet grid at P2: 141 distinct blocks for 104,544
    // rows); the few distinct blocks then stay in cache.
    // rep[r] = lowest row with the same block
    std::vector<int64_t> rep(nf);
    {
        std::vector<uint64_t> h(nf);
        auto hwork = [&](int t) {
            for (int64_t r = nf * t / nthreads; r < nf * (t + 1) / nthreads; r++) {
                uint64_t x = 1469598103934665603ull ^ (uint64_t)(slot_ptr[r + 1] - slot_ptr[r]);
                for (int64_t i = slot_ptr[r]; i < slot_ptr[r + 1]; i++) x = (x ^ (uint32_t)slots[i]) * 1099511628211ull;
                h[r] = x;
            }
        };
        std::vector<std::thread> hth;
        for (int t = 0; t < nthreads; t++) hth.emplace_back(hwork, t);
        for (auto& x : hth) x.join();
        std::vector<int64_t> byh(nf);
        for (int64_t r = 0; r < nf; r++) byh[r] = r;
        std::sort(byh.begin(), byh.end(), [&](int64_t a, int64_t b) { return h[a] != h[b] ? h[a] < h[b] : a < b; });
        auto same = [&](int64_t a, int64_t b) {
            const int64_t la = slot_ptr[a + 1] - slot_ptr[a];
            return la == slot_ptr[b + 1] - slot_ptr[b] &&
                   std::equal(slots.begin() + slot_ptr[a], slots.begin() + slot_ptr[a + 1], slots.begin() + slot_ptr[b]);
        };
        for (int64_t i = 0; i < nf;) {
            int64_t j = i;
            while (j < nf && h[byh[j]] == h[byh[i]]) j++;
            // rows i..j-1 share a hash (ascending row order): match each to the first equal one
            for (int64_t a = i; a < j; a++) {
                rep[byh[a]] = byh[a];
                for (int64_t b = i; b < a; b++)
                    if (rep[byh[b]] == byh[b] && same(byh[b], byh[a])) {
                        rep[byh[a]] = byh[b];
                        break;
                    }
            }
            i = j;
        }
    }
    // reduction work order: rows by decreasing length (stable: a row's slot-block
    // representative, an earlier row of the same length, comes first), in
    // classes whose longest row is at most twice the shortest class bound (LDS
    // sized per class).  Row order is kept within a length so that rows sharing
    // elements, whose segments sit next to each other in the element matrices,
    // are reduced close together in time.
    std::vector<int32_t> order(nf);
    for (int64_t r = 0; r < nf; r++) order[r] = (int32_t)r;
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return lenA[a] + lenD[a] > lenA[b] + lenD[b]; });
    {
        std::vector<int32_t> wlen(nf);
        for (int64_t i = 0; i < nf; i++) wlen[i] = lenA[order[i]] + lenD[order[i]];
        reduction_classes(P, wlen);
    }
    if (P.host_only) {
        P.h_erow_ptr = std::move(erow_ptr);
        P.h_erows = std::move(erows);
        P.h_slot_ptr = std::move(slot_ptr);
        P.h_slots = std::move(slots);
        P.h_rhs_ptr = std::move(rhs_ptr);
        P.h_rhs_erows = std::move(rhs_erows);
        return;
    }
    // device copies
    hipStream_t st = P.stream;
    upload(P.d_Arow, P.Arow.data(), P.Arow.size(), st);
    upload(P.d_Drow, P.Drow.data(), P.Drow.size(), st);
    upload(P.d_Acol, P.Acol.data(), P.Acol.size(), st);
    upload(P.d_Dcol, P.Dcol.data(), P.Dcol.size(), st);
    // reduction streams in work order: a 4-word descriptor per position, and the
    // segments and slot blocks of consecutive positions stored consecutively, so
    // the rows one block reduces read contiguous metadata
    {
        std::vector<uint64_t> rdesc((size_t)nf * 4), erows_w(erows.size());
        std::vector<int64_t> slot_w(nf, -1);  // work-order start of each position's slots (-1: shares its rep's block)
        std::vector<int64_t> rep_sw(nf, -1);  // slot block start of a representative row
        int64_t kw = 0, sw = 0;
        for (int64_t p = 0; p < nf; p++) {
            const int64_t r = order[p];
            const int64_t nk = erow_ptr[r + 1] - erow_ptr[r];
            SLX_CHECK(nk < 65536 && lenA[r] < 65536 && lenD[r] < 65536, "a row has more than 65535 segments or columns");
            int64_t sp;
            if (rep[r] == r) {
                sp = sw;
                rep_sw[r] = sw;
                slot_w[p] = sw;
                sw += slot_ptr[r + 1] - slot_ptr[r];
            } else {
                sp = rep_sw[rep[r]];  // the representative comes first in the work order
                SLX_CHECK(sp >= 0, "internal: shared slot block used before its representative");
            }
            rdesc[4 * p + 0] = (uint64_t)P.Arow[r] | ((uint64_t)lenA[r] << 48);
            rdesc[4 * p + 1] = (uint64_t)P.Drow[r] | ((uint64_t)lenD[r] << 48);
            rdesc[4 * p + 2] = (uint64_t)kw | ((uint64_t)nk << 48);
            rdesc[4 * p + 3] = (uint64_t)sp;
            std::copy(erows.begin() + erow_ptr[r], erows.begin() + erow_ptr[r + 1], erows_w.begin() + kw);
            kw += nk;
        }
        SLX_CHECK(P.Arow[nf] < (int64_t(1) << 48) && P.Drow[nf] < (int64_t(1) << 48), "matrix exceeds the 48-bit index range");
        upload(P.d_rdesc, rdesc.data(), rdesc.size(), st);
        upload(P.d_erows, erows_w.data(), erows_w.size(), st);
        P.n_slots_w = sw;
        // slots in the narrowest type that holds every row length (all ones = skip)
        // signed plans: the top bit of a slot is the entry's sign
        auto narrow = [&](auto tag) {
            using T = decltype(tag);
            const T top = (T)(T(1) << (8 * sizeof(T) - 1));
            std::vector<T> v(std::max<int64_t>(sw, 1));
            for (int64_t p = 0; p < nf; p++) {
                if (slot_w[p] < 0) continue;
                const int64_t r = order[p];
                T* dst = v.data() + slot_w[p];
                for (int64_t i = slot_ptr[r]; i < slot_ptr[r + 1]; i++) {
                    const int32_t sl = slots[i];
                    *dst++ = sl < 0 ? (T)~(T)0 : (T)((sl & ~SLX_SLOT_NEG) | ((sl & SLX_SLOT_NEG) ? top : 0));
                }
            }
            upload(P.d_slots, v.data(), v.size(), st);
        };
        const int sb = slot_bytes_for(maxrow, signed_slots);
        if (sb == 1) narrow(uint8_t{});
        else if (sb == 2) narrow(uint16_t{});
        else narrow(uint32_t{});
    }
    int slot_bytes = slot_bytes_for(maxrow, signed_slots);
    upload(P.d_rhs_ptr, rhs_ptr.data(), rhs_ptr.size(), st);
    upload(P.d_rhs_erows, rhs_erows.data(), rhs_erows.size(), st);
    P.n_rhs_erows = (int64_t)rhs_erows.size();
    upload(P.d_free, free_dofs.data(), free_dofs.size(), st);
    {
        std::vector<int64_t> cons_dofs;
        cons_dofs.reserve(nc);
        for (int64_t i = 0; i < nd; i++)
            if (d.constrained && d.constrained[i]) cons_dofs.push_back(i);
        upload(P.d_cons, cons_dofs.data(), cons_dofs.size(), st);
    }
    finish_plan(P, d, slot_bytes, maxrow);
    // the column arrays live on the device from here on
    std::vector<int32_t>().swap(P.Acol);
    std::vector<int32_t>().swap(P.Dcol);
}

// The same phase on the device (kernels/symbolic.hip).  Returns false, having
// changed nothing, when a row exceeds the device kernels' capacity (more than
// 8192 entries or 1024 element rows): the host phase then runs.
bool build_symbolic_device(slx_plan_s& P, const slx_plan_desc& d, const KeLayout& L) {
    SymInput in;
    in.n_dofs = d.n_dofs;
    in.constrained = d.constrained;
    in.n_cu = P.n_cu;
    for (size_t g = 0; g < L.groups.size(); g++) {
        const SegGroup& G = L.groups[g];
        const slx_bucket_desc& bd = d.buckets[G.bucket];
        const slx_slice_desc& sd0 = bd.slices[G.slices[0]];
        SymGroup sg{};
        sg.tf_addr = sd0.tf_addr;
        sg.tf_sign = sd0.tf_sign;
        sg.base = G.base;
        sg.ld = (int32_t)G.ld;
        sg.len = (int32_t)G.len;
        sg.n_tf = G.n_tf;
        sg.n_elem = bd.n_elem;
        sg.slice0 = (int32_t)in.slices.size();
        sg.n_slices = G.rhs ? 0 : (int32_t)G.slices.size();
        sg.rhs = G.rhs ? 1 : 0;
        if (!G.rhs)
            for (int s : G.slices) {
                const slx_slice_desc& sd = bd.slices[s];
                SymSlice ss{};
                ss.dof_addr = sd.dof_addr;
                ss.tf_sign = sd.tf_sign;
                ss.dof_sign = sd.dof_sign;
                ss.n_dof = sd.n_dof;
                ss.col_off = (int32_t)(L.ke_base[G.bucket][s] - G.base);
                ss.n_tf = sd.n_tf;
                ss.n_elem = bd.n_elem;
                in.slices.push_back(ss);
                if (sd.tf_sign || sd.dof_sign) in.signed_slots = true;
            }
        in.groups.push_back(sg);
    }
    for (size_t g = 0; g < L.groups.size(); g++) (L.groups[g].rhs ? in.rhs : in.mat).push_back((int)g);
    auto grank = [&](int g) { return L.groups[g].rank; };
    std::stable_sort(in.mat.begin(), in.mat.end(), [&](int a, int b) { return grank(a) < grank(b); });
    std::stable_sort(in.rhs.begin(), in.rhs.end(), [&](int a, int b) { return grank(a) < grank(b); });
    SLX_CHECK(L.n_ke < (int64_t(1) << 48), "element matrices exceed the 48-bit offset range");

    SymOutput o;
    std::string err;
    const hipError_t e = symbolic_device(in, o, err, P.stream);
    if (e != hipSuccess) {
        if (e == hipErrorInvalidValue && !err.empty()) fail(err);
        throw DeviceError("device symbolic phase: " + (err.empty() ? std::string(hipGetErrorString(e)) : err));
    }
    if (o.fallback) return false;
    P.nfree = o.nf;
    P.ncons = o.nc;
    P.n_ke = L.n_ke;
    P.signed_slots = in.signed_slots;
    P.n_erows = o.n_erows;
    P.n_coo = o.n_coo;
    P.max_erow_len = std::max(1, o.max_erow_len);
    P.nnzA = o.nnzA;
    P.nnzD = o.nnzD;
    P.Arow = std::move(o.h_Arow);
    P.Drow = std::move(o.h_Drow);
    P.Acol.clear();
    P.Dcol.clear();
    for (int b = 0; b < d.n_buckets; b++)
        for (int s = 0; s < d.buckets[b].n_slices; s++) {
            P.h_ke_base.push_back(L.ke_base[b][s]);
            P.h_ke_ld.push_back(L.ke_ld[b][s]);
        }
    reduction_classes(P, o.h_work_len);
    P.d_Arow.adopt(o.Arow.p, o.Arow.bytes);
    P.d_Drow.adopt(o.Drow.p, o.Drow.bytes);
    P.d_Acol.adopt(o.Acol.p, o.Acol.bytes);
    P.d_Dcol.adopt(o.Dcol.p, o.Dcol.bytes);
    P.d_rdesc.adopt(o.rdesc.p, o.rdesc.bytes);
    P.d_erows.adopt(o.erows_w.p, o.erows_w.bytes);
    P.d_slots.adopt(o.slots_w.p, o.slots_w.bytes);
    P.d_rhs_ptr.adopt(o.rhs_ptr.p, o.rhs_ptr.bytes);
    P.d_rhs_erows.adopt(o.rhs_erows.p, o.rhs_erows.bytes);
    P.d_free.adopt(o.free_dofs.p, o.free_dofs.bytes);
    P.d_cons.adopt(o.cons_dofs.p, o.cons_dofs.bytes);
    P.symbolic_on_device = true;
    P.n_slots_w = o.n_slots_w;
    P.n_rhs_erows = o.n_rhs_erows;
    finish_plan(P, d, o.slot_bytes, o.maxrow);
    return true;
}

// Address range check of the description (the device phase checks on the device).
void check_addresses(const slx_plan_desc* desc) {
    for (int b = 0; b < desc->n_buckets; b++) {
        const slx_bucket_desc& bd = desc->buckets[b];
        for (int s = 0; s < bd.n_slices; s++) {
            const slx_slice_desc& sd = bd.slices[s];
            size_t ntf = (size_t)sd.n_tf * bd.n_elem, ndf = sd.is_rhs ? 0 : (size_t)sd.n_dof * bd.n_elem;
            for (size_t k = 0; k < ntf; k++) SLX_CHECK(sd.tf_addr[k] < desc->n_dofs, "tf address out of range");
            for (size_t k = 0; k < ndf; k++) SLX_CHECK(sd.dof_addr[k] < desc->n_dofs, "dof address out of range");
        }
    }
}

// IEEE binary16 round-to-nearest-even conversion (host side of the split).
uint16_t f32_to_f16(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    int32_t exp = (int32_t)((x >> 23) & 0xff) - 127 + 15;
    uint32_t mant = x & 0x7fffffu;
    if (((x >> 23) & 0xff) == 0xff) return (uint16_t)(sign | 0x7c00u | (mant ? 0x200u : 0));
    if (exp >= 31) return (uint16_t)(sign | 0x7c00u);
    if (exp <= 0) {
        if (exp < -10) return (uint16_t)sign;
        mant |= 0x800000u;
        int shift = 14 - exp;
        uint32_t h = mant >> shift, rem = mant & ((1u << shift) - 1), half = 1u << (shift - 1);
        if (rem > half || (rem == half && (h & 1u))) h++;
        return (uint16_t)(sign | h);
    }
    uint32_t h = ((uint32_t)exp << 10) | (mant >> 13);
    uint32_t rem = mant & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
    return (uint16_t)(sign | h);
}

float f16_to_f32(uint16_t h) {
    uint32_t sign = (uint32_t)(h & 0x8000u) << 16, exp = (h >> 10) & 0x1f, mant = h & 0x3ffu, x;
    if (exp == 0) {
        if (mant == 0) x = sign;
        else {
            float v = std::ldexp((float)mant, -24);
            std::memcpy(&x, &v, 4);
            x |= sign;
        }
    } else if (exp == 31) x = sign | 0x7f800000u | (mant << 13);
    else x = sign | ((exp - 15 + 127) << 23) | (mant << 13);
    float f;
    std::memcpy(&f, &x, 4);
    return f;
}

// BF16 round-to-nearest-even (host side of the BF16 split; finite inputs).
uint16_t f32_to_bf16(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t r = x + 0x7fffu + ((x >> 16) & 1u);
    return (uint16_t)(r >> 16);
}

float bf16_to_f32(uint16_t h) {
    const uint32_t x = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &x, 4);
    return f;
}

// The 16-byte fragment plane entries of one A operand value x in every
// arithmetic: F16 (hi, lo' = fp16((x - hi) 2^11)), BF16 (hi, mid, lo).
void split_f16(float x, uint16_t& hi, uint16_t& lo) {
    hi = f32_to_f16(x);
    lo = f32_to_f16((x - f16_to_f32(hi)) * 2048.f);
}
void split_bf16(float x, uint16_t& hi, uint16_t& mid, uint16_t& lo) {
    hi = f32_to_bf16(x);
    const float r1 = x - bf16_to_f32(hi);
    mid = f32_to_bf16(r1);
    lo = f32_to_bf16(r1 - bf16_to_f32(mid));
}

// Builds the MFMA description of a bucket's matrix slices (all sharing one
// Phi table) or leaves M = 0 when the bucket is not eligible.
// Kernel choice per bucket class, from the table the autotuner writes
// (sparselizard_amd/tune/kernel_table.txt, found next to the library's
// directory, or the file SLX_TUNE_TABLE names; SLX_TUNE_TABLE=none disables
// it).  A line "elem_type n_tf n_dof q_invariant kernel mtw" picks the kernel
// of the matrix slices of buckets of that class: mfma (batched, mtw M tiles
// per wave), valu (FP32 element kernel) or factor (factorized MFMA).
// Classes the table does not list use the built-in rules below.
enum { TK_NONE = 0, TK_MFMA = 1, TK_VALU = 2, TK_FACTOR = 3, TK_TILE = 4 };
// kernel_table.txt rows: elem_type n_tf n_dof q_invariant kernel mtw [variant]
// (variant: tile blocks per wave of the factorized kernel, 0 = (2, 1),
// 1 = (4, 2), 2 = (7, 2); absent or -1: the smallest that covers the matrix)
struct TuneEntry {
    int type, n_tf, n_dof, q_inv, kernel, mtw, variant;
};

// The table is read again when SLX_TUNE_TABLE changes; callers copy what they
// need under the lock (tune_lookup).
std::mutex g_tune_mu;
const std::vector<TuneEntry>& tune_table() {
    static std::string loaded_key = "\x01unset";
    static std::vector<TuneEntry> table;
    const char* env = std::getenv("SLX_TUNE_TABLE");
    const std::string key = env ? env : "";
    if (key == loaded_key) return table;
    loaded_key = key;
    table = [&] {
        std::vector<TuneEntry> t;
        std::string path;
        if (env) {
            path = env;
            if (path == "none") return t;
        } else {
            Dl_info info;
            if (dladdr((void*)&tune_table, &info) && info.dli_fname) {
                std::string lib = info.dli_fname;
                const size_t k = lib.rfind('/');
                path = (k == std::string::npos ? std::string(".") : lib.substr(0, k)) + "/../tune/kernel_table.txt";
            }
        }
        FILE* f = path.empty() ? nullptr : std::fopen(path.c_str(), "r");
        if (!f) return t;
        char line[512];
        while (std::fgets(line, sizeof(line), f)) {
            if (line[0] == '#') continue;
            TuneEntry x{};
            x.variant = -1;
            char kern[32] = {0};
            if (std::sscanf(line, "%d %d %d %d %31s %d %d", &x.type, &x.n_tf, &x.n_dof, &x.q_inv, kern, &x.mtw, &x.variant) < 6) continue;
            const std::string k = kern;
            x.kernel = k == "mfma" ? TK_MFMA : k == "valu" ? TK_VALU : k == "factor" ? TK_FACTOR : k == "tile" ? TK_TILE : TK_NONE;
            if (x.kernel != TK_NONE) t.push_back(x);
        }
        std::fclose(f);
        return t;
    }();
    return table;
}

int tune_lookup(int type, int n_tf, int n_dof, int q_inv, int* mtw, int* variant = nullptr) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (const TuneEntry& x : tune_table())
        if (x.type == type && x.n_tf == n_tf && x.n_dof == n_dof && x.q_inv == q_inv) {
            if (mtw) *mtw = x.mtw;
            if (variant) *variant = x.variant;
            return x.kernel;
        }
    return TK_NONE;
}

// Largest element matrix (n_tf * n_dof entries) of the batched MFMA kernel
// without a table entry; larger slices go to the factorized kernel.
// SLX_TUNE_BATCHED_MAXM overrides it (tuning knob).
int batched_max_m() {
    const char* t = std::getenv("SLX_TUNE_BATCHED_MAXM");
    return t ? std::atoi(t) : 256;
}

void build_mfma(slx_plan_s& P, const slx_bucket_desc& bd, int b, const KeLayout& L) {
    MfmaBucket mb;
    std::memset(&mb, 0, sizeof(mb));
    uint32_t mask = 0;
    int tf_table = -1, dof_table = -1, n_tf = 0, n_dof = 0;
    std::vector<std::pair<int, int>> pairs;
    std::vector<int> ms;
    for (int s = 0; s < bd.n_slices; s++) {
        const slx_slice_desc& sd = bd.slices[s];
        if (sd.is_rhs) continue;
        if (tf_table < 0) {
            tf_table = sd.tf_table;
            dof_table = sd.dof_table;
            n_tf = sd.n_tf;
            n_dof = sd.n_dof;
        }
        if (sd.tf_table != tf_table || sd.dof_table != dof_table) return;  // different Phi per slice
        ms.push_back(s);
        for (int t = 0; t < sd.n_terms; t++) {
            std::pair<int, int> pr(sd.terms[t].tf_deriv, sd.terms[t].dof_deriv);
            if (std::find(pairs.begin(), pairs.end(), pr) == pairs.end()) pairs.push_back(pr);
        }
    }
    if (ms.empty() || (int)ms.size() > SLX_MAX_SLICES) return;
    int tuned_mtw = 1;
    const int tuned = tune_lookup(bd.elem_type, n_tf, n_dof, bd.q_invariant ? 1 : 0, &tuned_mtw);
    if (tuned == TK_VALU || tuned == TK_FACTOR || tuned == TK_TILE) return;
    if (tuned == TK_NONE && n_tf * n_dof > batched_max_m()) return;  // the factorized kernel takes larger slices
    mb.M = n_tf * n_dof;
    mb.n_tf = n_tf;
    mb.n_dof = n_dof;
    mb.n_pairs = (int)pairs.size();
    if (mb.n_pairs > SLX_MFMA_MAXP) return;
    mb.nq = bd.q_invariant ? 1 : bd.n_gp;
    // coefficients that vary over the Gauss points: the per-point columns make
    // the batched kernel producer-bound; the FP32 VALU element kernel is faster
    // there (measured: 2D P2 Newton 6.9 vs 2.1 ms, 512 order-3 quad problems
    // 78 vs 33 ms) at the same 1e-7 error.  SLX_TUNE_MFMA_QVARYING=1 keeps them
    // on MFMA (tuning knob; the tests use it to cover that path).
    if (mb.nq != 1 && tuned != TK_MFMA) {
        const char* t = std::getenv("SLX_TUNE_MFMA_QVARYING");
        if (!(t && std::atoi(t) != 0)) return;
    }
    // compact columns (coefficients constant over the Gauss points): the
    // Gauss-point sum of every derivative pair is taken here, in FP64
    // (sum_q c w_q tf_i dof_j = c sum_q w_q tf_i dof_j), so a pair needs one K
    // row; it keeps a block of 8 (the kernel broadcasts one pair's coefficient
    // over the eight k of a lane) whose first row holds the sum
    mb.compact = mb.nq == 1 ? 1 : 0;
    // SLX_MFMA_GAUSS_SUMS=0 keeps one K row per Gauss point in compact buckets
    // (the per-Gauss-point contraction of the FP64 path, padded to a multiple
    // of 8 per pair): a like-for-like comparison knob
    bool gauss_sums = true;
    if (const char* t = std::getenv("SLX_MFMA_GAUSS_SUMS")) gauss_sums = std::atoi(t) != 0;
    mb.kq = mb.compact ? (gauss_sums ? 8 : ((bd.n_gp + 7) / 8) * 8) : bd.n_gp;
    mb.K = mb.n_pairs * mb.kq;
    mb.mtiles = (mb.M + 15) / 16;
    mb.ksteps = (mb.K + 31) / 32;
    mb.n_mslices = (int)ms.size();
    mb.sgroup = 1;
    // one slice at a time (the software-pipelined consumer loop); SLX_TUNE_MFMA_SGROUP
    // = 2 or 3 takes compact slices in groups of that many (tuning knob)
    mb.sgroup = 1;
    if (const char* t = std::getenv("SLX_TUNE_MFMA_SGROUP")) {
        const int g = std::atoi(t);
        if (mb.compact && (g == 2 || g == 3) && mb.n_mslices % g == 0) mb.sgroup = g;
    }
    // tile split: SLX_TUNE_MFMA_MTW=2 selects two M tiles per wave (tuning knob)
    int prefer_mtw = tuned == TK_MFMA ? tuned_mtw : 1;
    if (const char* t = std::getenv("SLX_TUNE_MFMA_MTW")) prefer_mtw = std::atoi(t);
    if (!mfma_variant_fits(mb.mtiles, mb.ksteps, &mb.nw, &mb.mtw, &mb.maxk, prefer_mtw)) return;
    if (mb.mtw != 1) mb.sgroup = 1;
    mb.cstride = mb.compact ? mb.n_pairs + 1 : 32 * mb.ksteps;
    int es = mb.n_mslices * mb.cstride;
    if (mb.compact) es |= 1;  // odd: conflict-free 32-bit reads across elements
    else es = ((es + 59) / 64) * 64 + 4;  // = 4 (mod 64): 16-byte reads spread over the banks
    mb.estride = es;
    // segment groups of the slices (output blocks)
    std::vector<int> gl;  // layout group index per kernel group
    std::vector<int32_t> sgrp(ms.size()), scol(ms.size());
    std::vector<int64_t> gb;
    std::vector<int32_t> gd;
    for (size_t si = 0; si < ms.size(); si++) {
        const int g = L.group_of[b][ms[si]];
        int k = (int)(std::find(gl.begin(), gl.end(), g) - gl.begin());
        if (k == (int)gl.size()) {
            gl.push_back(g);
            gb.push_back(L.groups[g].base);
            gd.push_back((int32_t)L.groups[g].ld);
        }
        sgrp[si] = k;
        scol[si] = (int32_t)(L.ke_base[b][ms[si]] - L.groups[g].base);
    }
    // every slice of a group must be computed here (the block is written whole)
    for (int g : gl)
        for (int s2 : L.groups[g].slices)
            if (std::find(ms.begin(), ms.end(), s2) == ms.end()) return;
    mb.n_groups = (int)gl.size();
    mb.max_gld = *std::max_element(gd.begin(), gd.end());
    mb.stg_words = 0;
    for (int32_t l : gd) mb.stg_words += n_tf * l + 4;
    // per slice: staging offset (group block + column), segment length, staging row
    std::vector<int32_t> sprm(3 * ms.size());
    for (size_t si = 0; si < ms.size(); si++) {
        int off = 0;
        for (int g2 = 0; g2 < sgrp[si]; g2++) off += 16 * (n_tf * gd[g2] + 4);
        sprm[3 * si] = off + scol[si];
        sprm[3 * si + 1] = gd[sgrp[si]];
        sprm[3 * si + 2] = n_tf * gd[sgrp[si]] + 4;
    }
    // 16-element tiles; double-buffered columns and staging of one workgroup per CU
    mb.ne_tile = 16;
    if (mfma_lds_bytes(mb) > 160 * 1024) return;
    // coefficient slot of every (slice, pair)
    std::vector<int32_t> csrc((size_t)mb.n_mslices * mb.n_pairs, -1);
    for (size_t si = 0; si < ms.size(); si++) {
        const slx_slice_desc& sd = bd.slices[ms[si]];
        mb.slice_idx[si] = ms[si];
        mask |= 1u << ms[si];
        for (int t = 0; t < sd.n_terms; t++) {
            std::pair<int, int> pr(sd.terms[t].tf_deriv, sd.terms[t].dof_deriv);
            int p = (int)(std::find(pairs.begin(), pairs.end(), pr) - pairs.begin());
            // two terms of one slice with the same derivative pair: not a single coefficient
            if (csrc[si * mb.n_pairs + p] >= 0) return;
            csrc[si * mb.n_pairs + p] = sd.terms[t].coef;
        }
    }
    // Phi[m = i*n_dof + j][k = p*kq + q] = tf_r[i][q] w_q dof_s[j][q]; with
    // constant coefficients the Gauss-point sum of a pair sits in row p*kq
    // (kq = 8 compact, kq = 1 expanded) and the pair's other rows are zero
    const double* tft = bd.tables[tf_table];
    const double* dft = bd.tables[dof_table];
    const int ng = bd.n_gp;
    const bool presum = mb.nq == 1 && gauss_sums;
    std::vector<double> psum;  // [m][p] Gauss-point sums, or [m][p][q] products
    const int qcols = presum ? 1 : ng;
    psum.assign((size_t)mb.M * mb.n_pairs * qcols, 0.0);
    for (int i = 0; i < n_tf; i++)
        for (int j = 0; j < n_dof; j++)
            for (int p = 0; p < mb.n_pairs; p++)
                for (int q = 0; q < ng; q++) {
                    double v = tft[((size_t)pairs[p].first * n_tf + i) * ng + q] * bd.gp_weights[q] *
                               dft[((size_t)pairs[p].second * n_dof + j) * ng + q];
                    psum[(((size_t)i * n_dof + j) * mb.n_pairs + p) * qcols + (presum ? 0 : q)] += v;
                }
    double mx = 0.0;
    for (double v : psum) mx = std::max(mx, std::fabs(v));
    int ex = 0;
    if (mx > 0) std::frexp(mx, &ex);
    const double sc = std::ldexp(1.0, -ex);
    mb.a_unscale = (float)std::ldexp(1.0, ex);
    // fragments [mt][ks][plane][lane] x 16 bytes of the K layout (kq, K,
    // ksteps).  16x16x32 (F16, BF16): lane l holds k = 8 (l / 16) + v of row
    // l % 16; 16x16x4 (F32): plane pl, entry v is step j = 4 pl + v, k = 4 j + l / 16
    auto make_frags = [&](int kq, int K, int ksteps, bool f16, bool others) {
        auto phix = [&](int m, int k) -> float {
            if (m >= mb.M || k >= K) return 0.f;
            const int p = k / kq, q = k % kq;
            if (presum) return q == 0 ? (float)(psum[(size_t)m * mb.n_pairs + p] * sc) : 0.f;
            if (q >= qcols) return 0.f;  // Gauss points padded to the pair's K block
            return (float)(psum[((size_t)m * mb.n_pairs + p) * qcols + q] * sc);
        };
        const size_t nfr = (size_t)mb.mtiles * ksteps * 64;
        std::vector<uint16_t> frag(f16 ? nfr * 2 * 8 : 0, 0), fragb(others ? nfr * 3 * 8 : 0, 0);
        std::vector<float> fragf(others ? nfr * 2 * 4 : 0, 0.f);
        for (int mt = 0; mt < mb.mtiles; mt++)
            for (int ks = 0; ks < ksteps; ks++)
                for (int lane = 0; lane < 64; lane++) {
                    const int m = mt * 16 + lane % 16;
                    const size_t w = (size_t)(mt * ksteps + ks);
                    for (int v = 0; v < 8; v++) {
                        const float x = phix(m, ks * 32 + 8 * (lane / 16) + v);
                        uint16_t hi, mid, lo;
                        if (f16) {
                            split_f16(x, hi, lo);
                            frag[((w * 2 + 0) * 64 + lane) * 8 + v] = hi;
                            frag[((w * 2 + 1) * 64 + lane) * 8 + v] = lo;
                        }
                        if (others) {
                            split_bf16(x, hi, mid, lo);
                            fragb[((w * 3 + 0) * 64 + lane) * 8 + v] = hi;
                            fragb[((w * 3 + 1) * 64 + lane) * 8 + v] = mid;
                            fragb[((w * 3 + 2) * 64 + lane) * 8 + v] = lo;
                        }
                    }
                    if (others)
                        for (int pl = 0; pl < 2; pl++)
                            for (int v = 0; v < 4; v++)
                                fragf[((w * 2 + pl) * 64 + lane) * 4 + v] = phix(m, ks * 32 + 4 * (4 * pl + v) + lane / 16);
                }
        if (f16) mb.afrag_x[AR_F16] = P.dev(frag.data(), frag.size());
        if (others) {
            mb.afrag_x[AR_BF16] = P.dev(fragb.data(), fragb.size());
            mb.afrag_x[AR_F32] = P.dev(fragf.data(), fragf.size());
        }
    };
    make_frags(mb.kq, mb.K, mb.ksteps, true, true);
    mb.ck_kq = mb.kq;
    mb.ck_K = mb.K;
    mb.ck_ksteps = mb.ksteps;
    // F16 with constant coefficients: presummed columns with one K row per
    // pair (kq = 1; a lane reads the words of eight pairs, so up to 16 pairs
    // take one K step instead of one per four pairs), when that layout fits
    // the LDS; BF16 and F32 keep the compact layout above.
    // SLX_TUNE_MFMA_PRESUM=0 keeps F16 on the compact layout (tuning knob).
    bool try_expanded = presum && mb.sgroup == 1;
    if (const char* t = std::getenv("SLX_TUNE_MFMA_PRESUM")) try_expanded = try_expanded && std::atoi(t) != 0;
    mb.presum = 0;
    if (try_expanded) {
        MfmaBucket e = mb;
        e.presum = 1;
        e.kq = 1;
        e.K = e.n_pairs;
        e.ksteps = 1;  // at most 16 pairs
        // the software-pipelined consumer (one M tile per wave, one slice at a time)
        if (mfma_variant_fits(e.mtiles, e.ksteps, &e.nw, &e.mtw, &e.maxk, prefer_mtw) && e.mtw == 1) {
            // 16 words per slice; the lanes of k 16..31 read the next slice's
            // words (or the 16 zero words after the last slice), against zero
            // rows of Phi
            e.cstride = 16;
            e.estride = ((e.n_mslices * e.cstride + 12 + 63) / 64) * 64 + 4;  // = 4 (mod 64), >= ns * 16 + 16
            if (mfma_lds_bytes(e) <= 160 * 1024) {
                mb = e;
                make_frags(mb.kq, mb.K, mb.ksteps, true, false);
            }
        }
    }
    mb.afrag = mb.afrag_x[AR_F16];
    mb.arith = AR_F16;
    // arithmetics with a kernel variant: F16 always; BF16 and F32 take compact
    // columns with one M tile per wave, within the LDS of one CU
    mb.arith_ok = 1u << AR_F16;
    if (mb.compact && mb.mtiles <= 8)
        for (int a : {AR_BF16, AR_F32})
            if (mfma_lds_bytes(mfma_for(mb, a)) <= 160 * 1024) mb.arith_ok |= 1u << a;
    mb.csrc = P.dev(csrc.data(), csrc.size());
    mb.gbase = P.dev(gb.data(), gb.size());
    mb.gld = P.dev(gd.data(), gd.size());
    mb.sgrp = P.dev(sgrp.data(), sgrp.size());
    mb.scol = P.dev(scol.data(), scol.size());
    mb.sprm = P.dev(sprm.data(), sprm.size());
    mb.max_blocks = P.n_cu;
    P.mfma[b] = mb;
    P.mfma_mask[b] = mask;
}

// Gauss-point-major copy of a shape table, rows padded with zeros so that a
// tile of up to 16 functions past the end stays inside the allocation:
// out[(q * 4 + d) * pad + f] = tab[(d * nff + f) * ngp + q].
int table_pad(int nff) { return ((nff + 7) / 8) * 8 + 16; }

std::vector<double> table_qmajor(const double* tab, int nff, int ngp) {
    const int pad = table_pad(nff);
    std::vector<double> out((size_t)ngp * 4 * pad, 0.0);
    for (int q = 0; q < ngp; q++)
        for (int d = 0; d < 4; d++)
            for (int f = 0; f < nff; f++) out[((size_t)q * 4 + d) * pad + f] = tab[((size_t)d * nff + f) * ngp + q];
    return out;
}

// Factorized MFMA data for the large matrix slices of a bucket (those the
// batched MFMA kernel does not take): Ke = P^T Q_e with K over (Gauss point,
// tf derivative); see kernels/factor.h.
void build_factor(slx_plan_s& P, const slx_bucket_desc& bd, int b, const KeLayout& L) {
    for (int s = 0; s < bd.n_slices; s++) {
        const slx_slice_desc& sd = bd.slices[s];
        if (sd.is_rhs || ((P.mfma_mask[b] >> s) & 1u)) continue;
        int tuned_variant = -1;
        const int tuned = tune_lookup(bd.elem_type, sd.n_tf, sd.n_dof, bd.q_invariant ? 1 : 0, nullptr, &tuned_variant);
        if (tuned == TK_VALU || sd.n_terms > SLX_FK_MAXT) continue;
        if (tuned != TK_FACTOR && (int64_t)sd.n_tf * sd.n_dof <= batched_max_m()) continue;
        FactorSlice f;
        std::memset(&f, 0, sizeof(f));
        f.n_tf = sd.n_tf;
        f.n_dof = sd.n_dof;
        f.ngp = bd.n_gp;
        f.nq = bd.q_invariant ? 1 : bd.n_gp;
        f.mt = (sd.n_tf + 15) / 16;
        f.nt = (sd.n_dof + 15) / 16;
        f.ntp = f.nt * 16;
        f.n_terms = sd.n_terms;
        if (!factor_variant_fits(f.mt, f.nt, &f.bm, &f.bn, tuned == TK_FACTOR ? tuned_variant : -1)) continue;
        // tf derivatives in use; every term's (derivative index, dof derivative, slot)
        std::vector<int> alist;
        for (int t = 0; t < sd.n_terms; t++)
            if (std::find(alist.begin(), alist.end(), sd.terms[t].tf_deriv) == alist.end()) alist.push_back(sd.terms[t].tf_deriv);
        std::sort(alist.begin(), alist.end());
        f.na = (int)alist.size();
        std::vector<int32_t> tinfo(3 * sd.n_terms);
        for (int t = 0; t < sd.n_terms; t++) {
            tinfo[t] = (int32_t)(std::find(alist.begin(), alist.end(), sd.terms[t].tf_deriv) - alist.begin());
            tinfo[sd.n_terms + t] = sd.terms[t].dof_deriv;
            tinfo[2 * sd.n_terms + t] = sd.terms[t].coef;
            f.bmask |= 1 << sd.terms[t].dof_deriv;
        }
        // Gauss points per k-step: segments of whole Gauss points (factor_segw)
        if (f.na < 1 || f.na > 4) continue;
        f.qb = factor_qb(f.na, f.bn);
        f.nks = (f.ngp + f.qb - 1) / f.qb;
        f.ntq = (f.ntp + 127) / 128 * 128;
        // arithmetics whose LDS and P chunk staging fit; the FP16 split must
        f.arith_ok = 0;
        for (int a = 0; a < AR_COUNT; a++)
            if (factor_lds_bytes(f, a) <= 160 * 1024 && f.mt * mfma_planes(a) * 64 <= SLX_FK_AREG * 64 * SLX_FK_BWAVES)
                f.arith_ok |= 1u << a;
        if (!(f.arith_ok & (1u << AR_F16))) continue;
        // the MFMA arithmetics of a slice the one-tile kernel takes run there
        if ((P.tile_mask[b] >> s) & 1u) f.arith_ok &= ~((1u << AR_F16) | (1u << AR_BF16) | (1u << AR_F32));
        // P fragments: lane l of tile (ks, mt) holds P^T[i = mt*16 + l%16][k = 8*(l/16) + v],
        // k -> (Gauss point, tf derivative) by segments (factor_segw); power-of-two scale to [128, 256)
        const double* tft = bd.tables[sd.tf_table];
        const int nff = bd.table_nff[sd.tf_table], ngp = bd.n_gp;
        double amax = 0.0;
        for (int a : alist)
            for (int i = 0; i < sd.n_tf; i++)
                for (int q = 0; q < ngp; q++) amax = std::max(amax, std::fabs(tft[((size_t)a * nff + i) * ngp + q]));
        int aex = 0;
        if (amax > 0) (void)std::frexp(amax, &aex);
        const double ascale = std::ldexp(1.0, 8 - aex);
        f.a_unscale = (float)std::ldexp(1.0, aex - 8);
        // P^T[i][kl] of k-step ks (scaled), kl = local k: segment kl / segw, entry r = kl % segw of
        // the segment is Gauss point ks*qb + (kl / segw)*qh + r / na, tf derivative r % na (r < qh*na)
        const int segw = factor_segw(f.bn), qh = segw / f.na;
        auto px = [&](int ks, int i, int kl) {
            const int r = kl % segw, qi = r / f.na, ai = r % f.na, q = ks * f.qb + (kl / segw) * qh + qi;
            if (i < sd.n_tf && qi < qh && q < ngp) return (float)(tft[((size_t)alist[ai] * nff + i) * ngp + q] * ascale);
            return 0.f;
        };
        // fragments [ks][mt][plane][lane] x 16 bytes: F16 (hi, lo), BF16 (hi, mid, lo)
        // with k = 8 (l / 16) + v; F32 plane pl entry v: step j = 4 pl + v, k = 4 j + l / 16
        // (F16 lo unscaled here: P sits in [128, 256))
        const size_t nfr = (size_t)f.nks * f.mt * 64;
        std::vector<uint16_t> frag(nfr * 2 * 8, 0), fragb(nfr * 3 * 8, 0);
        std::vector<float> fragf(nfr * 2 * 4, 0.f);
        for (int ks = 0; ks < f.nks; ks++)
            for (int mt = 0; mt < f.mt; mt++)
                for (int l = 0; l < 64; l++) {
                    const int i = mt * 16 + l % 16;
                    const size_t w = (size_t)ks * f.mt + mt;
                    for (int v = 0; v < 8; v++) {
                        const float x = px(ks, i, 8 * (l / 16) + v);
                        const uint16_t hi = f32_to_f16(x);
                        frag[((w * 2 + 0) * 64 + l) * 8 + v] = hi;
                        frag[((w * 2 + 1) * 64 + l) * 8 + v] = f32_to_f16(x - f16_to_f32(hi));
                        uint16_t bh, bm, bl;
                        split_bf16(x, bh, bm, bl);
                        fragb[((w * 3 + 0) * 64 + l) * 8 + v] = bh;
                        fragb[((w * 3 + 1) * 64 + l) * 8 + v] = bm;
                        fragb[((w * 3 + 2) * 64 + l) * 8 + v] = bl;
                    }
                    for (int pl = 0; pl < 2; pl++)
                        for (int v = 0; v < 4; v++) fragf[((w * 2 + pl) * 64 + l) * 4 + v] = px(ks, i, 4 * (4 * pl + v) + l / 16);
                }
        // dof table times the weights, Gauss point major; zero rows up to nks * qb and
        // zero columns up to ntq, so the kernel reads every k-step's rows unchecked
        const double* dft = bd.tables[sd.dof_table];
        const int nfd = bd.table_nff[sd.dof_table];
        std::vector<float> tq((size_t)f.nks * f.qb * 4 * f.ntq, 0.f);
        double tqmax = 0.0;
        for (int q = 0; q < ngp; q++)
            for (int d = 0; d < 4; d++)
                for (int j = 0; j < sd.n_dof; j++) {
                    const double x = bd.gp_weights[q] * dft[((size_t)d * nfd + j) * ngp + q];
                    tq[((size_t)q * f.ntq + j) * 4 + d] = (float)x;
                    tqmax = std::max(tqmax, std::fabs(x));
                }
        f.tqmax = (float)(tqmax * 1.0001);
        // FP64 variant: k-steps of 16, unscaled P, FP64 w_q dof_b[j][q]
        f.qb64 = std::min(16 / f.na, SLX_FK_RU);
        f.nks64 = (ngp + f.qb64 - 1) / f.qb64;
        f.f64_ok = f.qb64 >= 1 && factor64_lds_bytes(f) <= 160 * 1024;
        if (f.f64_ok) {
            std::vector<double> a64((size_t)f.nks64 * f.mt * 4 * 64, 0.0), t64((size_t)ngp * 4 * f.ntp, 0.0);
            for (int ks = 0; ks < f.nks64; ks++)
                for (int mt = 0; mt < f.mt; mt++)
                    for (int j = 0; j < 4; j++)
                        for (int l = 0; l < 64; l++) {
                            const int i = mt * 16 + l % 16, kl = 4 * j + l / 16;
                            const int qi = kl / f.na, ai = kl % f.na, q = ks * f.qb64 + qi;
                            if (i < sd.n_tf && qi < f.qb64 && q < ngp)
                                a64[(((size_t)ks * f.mt + mt) * 4 + j) * 64 + l] = tft[((size_t)alist[ai] * nff + i) * ngp + q];
                        }
            for (int q = 0; q < ngp; q++)
                for (int dd = 0; dd < 4; dd++)
                    for (int j = 0; j < sd.n_dof; j++)
                        t64[((size_t)q * f.ntp + j) * 4 + dd] = bd.gp_weights[q] * dft[((size_t)dd * nfd + j) * ngp + q];
            f.afrag64 = P.dev(a64.data(), a64.size());
            f.tq64 = P.dev(t64.data(), t64.size());
        }
        f.tinfo = P.dev(tinfo.data(), tinfo.size());
        f.afrag_x[AR_F16] = reinterpret_cast<const uint32_t*>(P.dev(frag.data(), frag.size()));
        f.afrag_x[AR_BF16] = reinterpret_cast<const uint32_t*>(P.dev(fragb.data(), fragb.size()));
        f.afrag_x[AR_F32] = reinterpret_cast<const uint32_t*>(P.dev(fragf.data(), fragf.size()));
        f.afrag = f.afrag_x[AR_F16];
        f.arith = AR_F16;
        f.tq = P.dev(tq.data(), tq.size());
        f.ke_base = L.ke_base[b][s];
        f.ke_ld = L.ke_ld[b][s];
        P.factor[b].push_back({s, f});
        P.factor_mask[b] |= 1u << s;
    }
}

// One-tile-per-element data (kernels/tile.h) for the small matrix slices of a
// bucket whose coefficients vary over the Gauss points and that the batched
// MFMA kernel does not take (the three MFMA arithmetics; the FP32 VALU and
// FP64 modes keep their kernels there).
void build_tile(slx_plan_s& P, const slx_bucket_desc& bd, int b, const KeLayout& L) {
    if (bd.q_invariant) return;
    if (const char* t = std::getenv("SLX_TUNE_TILE"))
        if (std::atoi(t) == 0) return;
    // smaller slices waste most of the tile's Q_e build: the VALU element kernel
    // measured as fast or faster there (6 x 6: P2 triangles, order-1 prisms)
    int min_m = 100;
    if (const char* t = std::getenv("SLX_TUNE_TILE_MIN")) min_m = std::atoi(t);
    for (int s = 0; s < bd.n_slices; s++) {
        const slx_slice_desc& sd = bd.slices[s];
        if (sd.is_rhs || ((P.mfma_mask[b] >> s) & 1u)) continue;
        if (sd.n_tf > 64 || sd.n_dof > 64 || sd.n_tf * sd.n_dof < min_m || sd.n_terms < 1 || sd.n_terms > SLX_TK_MAXT) continue;
        // a kernel-table entry for the class decides; without one the slice takes the kernel
        const int tuned = tune_lookup(bd.elem_type, sd.n_tf, sd.n_dof, 0, nullptr);
        if (tuned != TK_NONE && tuned != TK_TILE) continue;
        std::vector<int> alist;
        for (int t = 0; t < sd.n_terms; t++)
            if (std::find(alist.begin(), alist.end(), sd.terms[t].tf_deriv) == alist.end()) alist.push_back(sd.terms[t].tf_deriv);
        std::sort(alist.begin(), alist.end());
        const int na = (int)alist.size(), ngp = bd.n_gp;
        if (na < 1 || na > 4) continue;
        const int qh = tile_qh(na), nkc = (ngp + 4 * qh - 1) / (4 * qh), ngpp = nkc * 4 * qh;
        if (nkc > SLX_TK_MAXC || ngpp > SLX_TK_MAXQ) continue;
        TileSlice t;
        std::memset(&t, 0, sizeof(t));
        t.mtt = (sd.n_tf + 15) / 16;
        t.ntt = (sd.n_dof + 15) / 16;
        if (t.mtt > 2 || t.ntt > 2) t.mtt = t.ntt = std::max(t.mtt, t.ntt);  // 3 x 3 or 4 x 4
        t.n_tf = sd.n_tf;
        t.n_dof = sd.n_dof;
        t.ngp = ngp;
        t.nq = bd.n_gp;
        t.na = na;
        t.nkc = nkc;
        t.ngpp = ngpp;
        t.n_terms = sd.n_terms;
        if (tile_lds_bytes(t, 3) > 64 * 1024) continue;
        std::vector<int32_t> tinfo(3 * sd.n_terms);
        for (int k = 0; k < sd.n_terms; k++) {
            tinfo[k] = (int32_t)(std::find(alist.begin(), alist.end(), sd.terms[k].tf_deriv) - alist.begin());
            tinfo[sd.n_terms + k] = sd.terms[k].dof_deriv;
            tinfo[2 * sd.n_terms + k] = sd.terms[k].coef;
        }
        // P fragments: lane l of chunk c and row tile mt holds P^T[i = 16 mt + l % 16][k = 32 c + 8 (l / 16) + v],
        // k -> Gauss point (4 c + l / 16) qh + v / na, tf derivative alist[v % na]; power-of-two scale to [128, 256)
        const double* tft = bd.tables[sd.tf_table];
        const int nff = bd.table_nff[sd.tf_table];
        double amax = 0.0;
        for (int a : alist)
            for (int i = 0; i < sd.n_tf; i++)
                for (int q = 0; q < ngp; q++) amax = std::max(amax, std::fabs(tft[((size_t)a * nff + i) * ngp + q]));
        int aex = 0;
        if (amax > 0) (void)std::frexp(amax, &aex);
        const double ascale = std::ldexp(1.0, 8 - aex);
        t.a_unscale = (float)std::ldexp(1.0, aex - 8);
        // FP16 split: planes hi, lo; BF16 split: planes hi, mid, lo; FP32: the entries
        std::vector<uint16_t> frag((size_t)nkc * t.mtt * 2 * 64 * 8, 0), fragb((size_t)nkc * t.mtt * 3 * 64 * 8, 0);
        std::vector<float> fragf((size_t)nkc * t.mtt * 2 * 64 * 4, 0.f);
        for (int c = 0; c < nkc; c++)
            for (int mt = 0; mt < t.mtt; mt++)
                for (int l = 0; l < 64; l++)
                    for (int v = 0; v < 8; v++) {
                        const int i = 16 * mt + l % 16, qi = v / na, ai = v % na, q = (4 * c + l / 16) * qh + qi;
                        float x = 0.f;
                        if (i < sd.n_tf && qi < qh && q < ngp) x = (float)(tft[((size_t)alist[ai] * nff + i) * ngp + q] * ascale);
                        const uint16_t hi = f32_to_f16(x);
                        const size_t w = ((size_t)c * t.mtt + mt) * 2;
                        frag[((w + 0) * 64 + l) * 8 + v] = hi;
                        frag[((w + 1) * 64 + l) * 8 + v] = f32_to_f16(x - f16_to_f32(hi));
                        uint16_t bh, bm, bl;
                        split_bf16(x, bh, bm, bl);
                        const size_t wb = ((size_t)c * t.mtt + mt) * 3;
                        fragb[((wb + 0) * 64 + l) * 8 + v] = bh;
                        fragb[((wb + 1) * 64 + l) * 8 + v] = bm;
                        fragb[((wb + 2) * 64 + l) * 8 + v] = bl;
                        fragf[((((size_t)c * t.mtt + mt) * 2 + v / 4) * 64 + l) * 4 + v % 4] = x;
                    }
        // weighted dof table [q][j][4], zero past the last Gauss point and dof
        const double* dft = bd.tables[sd.dof_table];
        const int nfd = bd.table_nff[sd.dof_table];
        const int nc = 16 * t.ntt;
        std::vector<float> tq((size_t)ngpp * nc * 4, 0.f);
        double tqmax = 0.0;
        for (int q = 0; q < ngp; q++)
            for (int j = 0; j < sd.n_dof; j++)
                for (int dd = 0; dd < 4; dd++) {
                    const double x = bd.gp_weights[q] * dft[((size_t)dd * nfd + j) * ngp + q];
                    tq[((size_t)q * nc + j) * 4 + dd] = (float)x;
                    tqmax = std::max(tqmax, std::fabs(x));
                }
        t.tqmax = (float)(tqmax * 1.0001);
        t.tinfo = P.dev(tinfo.data(), tinfo.size());
        t.afrag_x[0] = reinterpret_cast<const uint32_t*>(P.dev(frag.data(), frag.size()));
        t.afrag_x[1] = reinterpret_cast<const uint32_t*>(P.dev(fragb.data(), fragb.size()));
        t.afrag_x[2] = reinterpret_cast<const uint32_t*>(P.dev(fragf.data(), fragf.size()));
        t.afrag = t.afrag_x[0];
        t.tq = P.dev(tq.data(), tq.size());
        t.ke_base = L.ke_base[b][s];
        t.ke_ld = L.ke_ld[b][s];
        P.tile[b].push_back({s, t});
        P.tile_mask[b] |= 1u << s;
    }
}

void compile_coef_programs(slx_plan_s& P, const slx_plan_desc& d);

// The fused launch (kernels/factor.hip, launch_factor_fused): when two or more
// buckets have factorized slices, their coefficient tables get regions of
// their own and the slice tables of the launch are built for every
// arithmetic, slices in decreasing per-element cost.  By default the FP16 and
// BF16 splits use it (measured faster than per-slice launches at 48^3 hp
// Poisson: 13.1 vs 13.7 ms and 15.6 vs 17.1 ms); FP32 MFMA and FP64 keep
// per-slice launches (25.0 vs 24.7 ms, 50.9 vs 47.7 ms: the fused kernel
// takes the register budget of the largest tile variant for every slice).
// SLX_TUNE_FUSED=0 turns it off, 1 uses it in every arithmetic (also for a
// single bucket).
void build_fused(slx_plan_s& P, const slx_plan_desc& d) {
    int mode = -1;
    if (const char* t = std::getenv("SLX_TUNE_FUSED")) mode = std::atoi(t);
    P.fused_arith = mode == 1 ? 0xFu : ((1u << AR_F16) | (1u << AR_BF16));
    P.fuse_bucket.assign(d.n_buckets, 0);
    P.fuse_off.assign(d.n_buckets, 0);
    int nb = 0;
    for (int b = 0; b < d.n_buckets; b++)
        if (!P.factor[b].empty() && !P.rhs_only[b]) nb++;
    P.fused = mode != 0 && (nb >= 2 || (mode == 1 && nb >= 1));
    if (!P.fused) return;
    size_t bytes = 0;
    for (int b = 0; b < d.n_buckets; b++) {
        if (P.factor[b].empty() || P.rhs_only[b]) continue;
        const slx_bucket_desc& bd = d.buckets[b];
        P.fuse_bucket[b] = 1;
        P.fuse_off[b] = bytes;
        const size_t words = (size_t)std::max(bd.n_coefs, 1) * (bd.q_invariant ? 1 : bd.n_gp) * std::max(bd.n_elem, 1);
        bytes += (words * sizeof(double) + 255) / 256 * 256;
    }
    P.d_ctab_fused.alloc(std::max<size_t>(bytes, 256));
    P.d_fused_counter.alloc(sizeof(unsigned long long));
    for (int a = 0; a < 4; a++) {
        struct Entry {
            FusedSlice fs;
            double cost;
        };
        std::vector<Entry> v;
        size_t lds = 0;
        for (int b = 0; b < d.n_buckets; b++) {
            if (!P.fuse_bucket[b]) continue;
            for (auto& sf : P.factor[b]) {
                const FactorSlice& f0 = sf.second;
                if (a < 3 ? !((f0.arith_ok >> a) & 1u) : !f0.f64_ok) continue;
                Entry en;
                std::memset(&en, 0, sizeof(en));
                en.fs.f = f0;
                if (a < 3) {
                    en.fs.f.arith = a;
                    en.fs.f.afrag = f0.afrag_x[a];
                }
                en.fs.ctab = (const char*)P.d_ctab_fused.p + P.fuse_off[b];
                en.fs.n_elem = d.buckets[b].n_elem;
                en.fs.variant = factor_variant_index(f0.bm, f0.bn);
                SLX_CHECK(en.fs.variant >= 0, "internal: factorized slice without a kernel variant");
                en.cost = (double)f0.mt * f0.nt * (a < 3 ? f0.nks : f0.nks64);
                lds = std::max(lds, a < 3 ? factor_lds_bytes(f0, a) : factor64_lds_bytes(f0));
                v.push_back(en);
            }
        }
        std::stable_sort(v.begin(), v.end(), [](const Entry& x, const Entry& y) { return x.cost > y.cost; });
        std::vector<FusedSlice> tab;
        int64_t first = 0;
        for (Entry& en : v) {
            en.fs.first = first;
            first += en.fs.n_elem;
            tab.push_back(en.fs);
        }
        P.n_fused[a] = (int)tab.size();
        P.fused_total[a] = first;
        P.fused_lds[a] = lds;
        if (!tab.empty()) upload(P.d_fused[a], tab.data(), tab.size(), P.stream);
    }
}

void upload_buckets(slx_plan_s& P, const slx_plan_desc& d, const KeLayout& L) {
    P.dbuckets.resize(d.n_buckets);
    P.mfma.assign(d.n_buckets, MfmaBucket());
    for (auto& m : P.mfma) m.M = 0;
    P.mfma_mask.assign(d.n_buckets, 0);
    P.factor.assign(d.n_buckets, {});
    P.factor_mask.assign(d.n_buckets, 0);
    P.tile.assign(d.n_buckets, {});
    P.tile_mask.assign(d.n_buckets, 0);
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, P.ctx->device) == hipSuccess && prop.multiProcessorCount > 0) P.n_cu = prop.multiProcessorCount;
    }
    size_t ctab_words = 1, ctab_side_words = 1;
    P.rhs_only.assign(d.n_buckets, 0);
    bool rhs_in_mixed = false, any_rhs_only = false;
    for (int b = 0; b < d.n_buckets; b++) {
        const slx_bucket_desc& bd = d.buckets[b];
        DevBucket& db = P.dbuckets[b];
        {
            int nr = 0;
            for (int s = 0; s < bd.n_slices; s++) nr += bd.slices[s].is_rhs ? 1 : 0;
            P.rhs_only[b] = (bd.n_slices > 0 && nr == bd.n_slices) ? 1 : 0;
            if (nr > 0 && !P.rhs_only[b]) rhs_in_mixed = true;
            any_rhs_only = any_rhs_only || P.rhs_only[b];
        }
        std::memset(&db, 0, sizeof(db));
        SLX_CHECK(bd.n_slices <= SLX_MAX_SLICES, "bucket has " << bd.n_slices << " slices (max " << SLX_MAX_SLICES << ")");
        SLX_CHECK(bd.n_regs <= SLX_MAX_PROG_REGS, "coefficient program needs " << bd.n_regs << " registers (max " << SLX_MAX_PROG_REGS << ")");
        SLX_CHECK(bd.n_geo <= SLX_MAX_GEO_CURVED, "too many geometry nodes");
        SLX_CHECK(d.n_fields <= SLX_MAX_FIELDS, "too many fields in coefficient programs");
        int edim = bd.elem_type == 1 ? 1 : (bd.elem_type <= 3 ? 2 : 3);
        SLX_CHECK(edim == d.mesh_dim, "only integrals over elements of the mesh dimension are provided");
        db.type = bd.elem_type;
        db.n_elem = bd.n_elem;
        db.n_geo = bd.n_geo;
        db.n_gp = bd.n_gp;
        db.elemdim = edim;
        db.meshdim = d.mesh_dim;
        db.n_instr = bd.n_instr;
        db.n_regs = bd.n_regs;
        db.n_coefs = bd.n_coefs;
        db.q_inv = bd.q_invariant;
        db.n_slices = bd.n_slices;
        db.n_fields = d.n_fields;
        db.geo_nodes = P.dev(bd.geo_nodes, (size_t)bd.n_elem * bd.n_geo);
        db.elem_ids = P.dev(bd.elem_ids, (size_t)bd.n_elem);
        for (int e = 0; e < bd.n_elem; e++) {
            SLX_CHECK(bd.elem_ids[e] >= 0, "negative element id");
            P.max_elem_id = std::max<int64_t>(P.max_elem_id, (int64_t)bd.elem_ids[e] + 1);
        }
        db.gp_w = P.dev(bd.gp_weights, (size_t)bd.n_gp);
        std::vector<double> pts(bd.gp_coords, bd.gp_coords + 3 * bd.n_gp), geo;
        geometry_tables(bd.elem_type, bd.n_geo, pts, geo);  // order 1, or order 2 (curved)
        db.geo_tab = P.dev(geo.data(), geo.size());
        db.prog = P.dev(bd.prog, (size_t)bd.n_instr);
        std::vector<const double*> dtabs(bd.n_tables);
        std::vector<const float*> dtabs32(bd.n_tables);
        for (int t = 0; t < bd.n_tables; t++) {
            std::vector<double> tq = table_qmajor(bd.tables[t], bd.table_nff[t], bd.n_gp);
            dtabs[t] = P.dev(tq.data(), tq.size());
            // FP32 copy for the FP32 element kernels (rounded once here, not per use)
            std::vector<float> tq32(tq.begin(), tq.end());
            dtabs32[t] = P.dev(tq32.data(), tq32.size());
        }
        const size_t words = (size_t)std::max(bd.n_coefs, 1) * (bd.q_invariant ? 1 : bd.n_gp) * std::max(bd.n_elem, 1);
        if (P.rhs_only[b]) ctab_side_words = std::max(ctab_side_words, words);
        else ctab_words = std::max(ctab_words, words);
        for (int s = 0; s < bd.n_slices; s++) {
            const slx_slice_desc& sd = bd.slices[s];
            DevSlice& ds = db.slices[s];
            ds.n_tf = sd.n_tf;
            ds.n_dof = sd.is_rhs ? 1 : sd.n_dof;
            ds.is_rhs = sd.is_rhs;
            ds.n_terms = sd.n_terms;
            ds.terms = P.dev(sd.terms, (size_t)sd.n_terms);
            SLX_CHECK(sd.tf_table >= 0 && sd.tf_table < bd.n_tables && bd.table_nff[sd.tf_table] == sd.n_tf,
                      "slice tf table does not match n_tf");
            SLX_CHECK(sd.is_rhs || (sd.dof_table >= 0 && sd.dof_table < bd.n_tables && bd.table_nff[sd.dof_table] == sd.n_dof),
                      "slice dof table does not match n_dof");
            ds.tf_tabq = dtabs[sd.tf_table];
            ds.dof_tabq = sd.is_rhs ? nullptr : dtabs[sd.dof_table];
            ds.tf_tabq32 = dtabs32[sd.tf_table];
            ds.dof_tabq32 = sd.is_rhs ? nullptr : dtabs32[sd.dof_table];
            ds.tf_pad = table_pad(sd.n_tf);
            ds.dof_pad = sd.is_rhs ? 0 : table_pad(sd.n_dof);
            ds.ke_base = L.ke_base[b][s];
            ds.ke_ld = L.ke_ld[b][s];
            for (int t = 0; t < sd.n_terms; t++) {
                SLX_CHECK(sd.terms[t].coef >= 0 && sd.terms[t].coef < bd.n_coefs, "term coefficient slot out of range");
                SLX_CHECK(sd.terms[t].tf_deriv >= 0 && sd.terms[t].tf_deriv <= 3 && sd.terms[t].dof_deriv >= 0 && sd.terms[t].dof_deriv <= 3,
                          "term derivative out of range");
            }
        }
        db.dslices = P.dev(db.slices, (size_t)bd.n_slices);
        for (int f = 0; f < d.n_fields; f++) {
            size_t k = (size_t)f * d.n_buckets + b;
            db.fields[f].nff = d.field_nff[k];
            db.fields[f].addr = P.dev(d.field_addr[k], (size_t)d.field_nff[k] * bd.n_elem);
            db.fields[f].tab = P.dev(d.field_tables[k], (size_t)4 * d.field_nff[k] * bd.n_gp);
        }
        // program sanity: registers and opcodes in range
        for (int i = 0; i < bd.n_instr; i++) {
            const slx_instr& in = bd.prog[i];
            SLX_CHECK(in.a >= 0 && in.b >= 0 && in.a < SLX_MAX_PROG_REGS && in.b < SLX_MAX_PROG_REGS, "bad operand register");
            if (in.op == SLX_OP_PARAM) {
                SLX_CHECK(in.imm >= 0, "bad parameter index");
                P.max_param = std::max(P.max_param, in.imm);
            }
            if (in.op == SLX_OP_FIELD) P.uses_fields = true;
            SLX_CHECK(in.op >= 0 && in.op < SLX_OP_COUNT, "bad opcode");
            SLX_CHECK(in.dst >= 0 && in.dst < bd.n_regs, "bad destination register");
            if (in.op == SLX_OP_OUT) SLX_CHECK(in.imm >= 0 && in.imm < bd.n_coefs, "bad coefficient slot");
            if (in.op == SLX_OP_CONST) SLX_CHECK(in.imm >= 0 && in.imm < d.n_consts, "bad constant index");
            if (in.op == SLX_OP_FIELD) SLX_CHECK(in.imm >= 0 && in.imm < d.n_fields && in.a >= 0 && in.a <= 3, "bad field reference");
            if (in.op == SLX_OP_JAC || in.op == SLX_OP_INVJAC) SLX_CHECK(in.a >= 0 && in.a < 3 && in.b >= 0 && in.b < 3, "bad Jacobian entry");
            if (in.op == SLX_OP_COORD) SLX_CHECK(in.a >= 0 && in.a < 3, "bad coordinate");
        }
        // element node numbers must be in range of the node array
        for (size_t k = 0; k < (size_t)bd.n_elem * bd.n_geo; k++)
            SLX_CHECK(bd.geo_nodes[k] >= 0 && bd.geo_nodes[k] < d.n_nodes, "element node out of range");
        build_mfma(P, bd, b, L);
        build_tile(P, bd, b, L);
        build_factor(P, bd, b, L);
    }
    upload(P.consts, d.consts, (size_t)std::max(d.n_consts, 1), P.stream);
    P.d_ctab.alloc(ctab_words * sizeof(double));
    P.d_ctab_side.alloc(ctab_side_words * sizeof(double));
    build_fused(P, d);
    compile_coef_programs(P, d);
    P.side_ok = d.has_matrix && d.has_rhs && any_rhs_only && !rhs_in_mixed;
    if (P.side_ok) {
        HIPCHECK(hipStreamCreateWithFlags(&P.side, hipStreamNonBlocking));
        HIPCHECK(hipEventCreateWithFlags(&P.ev_fork, hipEventDisableTiming));
        HIPCHECK(hipEventCreateWithFlags(&P.ev_join, hipEventDisableTiming));
    }
}

// Compiled coefficient programs for the buckets the mode selects (auto: buckets
// of at least SLX_COEF_JIT_MIN elements); unique programs compiled in parallel.
#define SLX_COEF_JIT_MIN 4096
void compile_coef_programs(slx_plan_s& P, const slx_plan_desc& d) {
    P.coef_fn.assign(d.n_buckets, {nullptr, nullptr});
    if (P.coef_mode == 1) return;
    std::vector<std::pair<int, int>> todo;  // (bucket, fp64)
    for (int b = 0; b < d.n_buckets; b++) {
        if (d.buckets[b].n_coefs == 0 || d.buckets[b].n_elem == 0) continue;
        // programs beyond the interpreter's register file are always compiled
        if (P.coef_mode == 0 && d.buckets[b].n_elem < SLX_COEF_JIT_MIN && d.buckets[b].n_regs <= SLX_MAX_REGS) continue;
        todo.push_back({b, 0});
        todo.push_back({b, 1});
    }
    std::mutex mu;
    std::vector<std::thread> th;
    const int device = P.ctx->device;
    const size_t nth = std::min<size_t>(todo.size(), 8);
    for (size_t t = 0; t < nth; t++)
        th.emplace_back([&, t]() {
            (void)hipSetDevice(device);
            for (size_t i = t; i < todo.size(); i += nth) {
                const slx_bucket_desc& bd = d.buckets[todo[i].first];
                std::vector<int> fnff(d.n_fields);
                for (int f = 0; f < d.n_fields; f++) fnff[f] = d.field_nff[(size_t)f * d.n_buckets + todo[i].first];
                hipFunction_t fn = nullptr;
                std::string err;
                if (coef_compile(bd.prog, bd.n_instr, todo[i].second != 0, bd.q_invariant == 0, fnff, &fn, &err)) {
                    P.coef_fn[todo[i].first][todo[i].second] = fn;
                } else {
                    std::lock_guard<std::mutex> lk(mu);
                    P.coef_jit_error = err;
                }
            }
        });
    for (auto& x : th) x.join();
}

// Per-launch HIP events by kernel class (profiling only).
enum { T_COEF = 0, T_MFMA = 1, T_VALU = 2, T_REDUCE = 3, T_RHS = 4, T_COUNT = 5 };
struct LaunchTimer {
    hipStream_t st;
    std::vector<hipEvent_t> pool;
    std::vector<std::pair<int, size_t>> marks;  // (class, index of the start event)
    size_t used = 0;
    hipEvent_t next() {
        if (used == pool.size()) {
            hipEvent_t e;
            HIPCHECK(hipEventCreate(&e));
            pool.push_back(e);
        }
        return pool[used++];
    }
    void begin(int cls) {
        marks.push_back({cls, used});
        HIPCHECK(hipEventRecord(next(), st));
    }
    void end() { HIPCHECK(hipEventRecord(next(), st)); }
    void collect(double* ms) {
        HIPCHECK(hipEventSynchronize(pool[used - 1]));
        for (auto& m : marks) {
            float t = 0;
            HIPCHECK(hipEventElapsedTime(&t, pool[m.second], pool[m.second + 1]));
            ms[m.first] += t;
        }
        marks.clear();
        used = 0;
    }
    ~LaunchTimer() {
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

void numeric(slx_plan_s& P, const double* xyz, const double* params, int64_t plen, const double* const* fields, int precision,
             hipStream_t st, LaunchTimer* tm = nullptr) {
    SLX_CHECK(precision >= 0 && precision <= 4, "unknown precision " << precision);
    if (P.max_param >= 0) SLX_CHECK(params && plen >= P.max_elem_id, "per-element parameters missing or shorter than the element count");
    if (P.uses_fields) {
        SLX_CHECK(fields, "field data required by the coefficient programs");
        for (int f = 0; f < P.n_fields; f++) SLX_CHECK(fields[f], "field data pointer is null");
    }
    DevInputs in{};
    in.xyz = xyz;
    in.params = params;
    in.param_len = plen;
    in.consts = P.consts.as<double>();
    for (int f = 0; f < P.n_fields; f++) in.field_data[f] = fields ? fields[f] : nullptr;
    bool ke64 = precision == SLX_PREC_FP64;
    // MFMA arithmetic of the mode (-1: VALU element kernels only)
    const int arith = precision == SLX_PREC_F16_SPLIT ? AR_F16
                      : precision == SLX_PREC_BF16_SPLIT ? AR_BF16
                      : precision == SLX_PREC_FP32_MFMA ? AR_F32 : -1;
    // fork: rhs-only buckets and the rhs reduction overlap the matrix work
    // (serial when every launch is timed)
    const bool overlap = P.side_ok && !tm;
    if (overlap) {
        HIPCHECK(hipEventRecord(P.ev_fork, st));
        HIPCHECK(hipStreamWaitEvent(P.side, P.ev_fork, 0));
    }
    // fused launch: coefficients of every fused bucket into its region, then one
    // persistent launch over all their factorized slices
    // slices of bucket bi on the factorized kernel in this arithmetic (or in FP64)
    auto factor_mask = [&](size_t bi) {
        uint32_t m = 0;
        if (arith >= 0) {
            for (auto& fs : P.factor[bi])
                if ((fs.second.arith_ok >> arith) & 1u) m |= 1u << fs.first;
        } else if (ke64 && P.factor64) {
            for (auto& fs : P.factor[bi])
                if (fs.second.f64_ok) m |= 1u << fs.first;
        }
        return m;
    };
    const int fa = arith >= 0 ? arith : ((ke64 && P.factor64) ? 3 : -1);
    const bool fused = P.fused && fa >= 0 && ((P.fused_arith >> fa) & 1u) && P.n_fused[fa] > 0;
    P.last_fused = fused;
    if (fused) {
        for (size_t bi = 0; bi < P.dbuckets.size(); bi++) {
            if (!P.fuse_bucket[bi] || factor_mask(bi) == 0) continue;
            const DevBucket& b = P.dbuckets[bi];
            void* ctab = (char*)P.d_ctab_fused.p + P.fuse_off[bi];
            if (tm) tm->begin(T_COEF);
            const hipFunction_t cfn = bi < P.coef_fn.size() ? P.coef_fn[bi][ke64 ? 1 : 0] : nullptr;
            if (cfn) HIPCHECK(coef_launch(cfn, b, in, b.q_inv ? 1 : b.n_gp, 1, ctab, st));
            else {
                SLX_CHECK(b.n_regs <= SLX_MAX_REGS, "coefficient program of " << b.n_regs << " registers has no compiled kernel and the interpreter holds "
                                                   << SLX_MAX_REGS << " (" << P.coef_jit_error << ")");
                HIPCHECK(launch_coef(b, in, ke64, ctab, st, 1));
            }
            if (tm) tm->end();
        }
        if (tm) tm->begin(T_MFMA);
        HIPCHECK(launch_factor_fused(P.d_fused[fa].as<FusedSlice>(), P.n_fused[fa], P.fused_total[fa],
                                     P.d_fused_counter.as<unsigned long long>(), fa, P.fused_lds[fa], P.d_ke.p, P.n_cu, st));
        if (tm) tm->end();
    }
    for (size_t bi = 0; bi < P.dbuckets.size(); bi++) {
        const DevBucket& b = P.dbuckets[bi];
        const hipStream_t bst = (overlap && P.rhs_only[bi]) ? P.side : st;
        const uint32_t fmask = factor_mask(bi);  // slices on the factorized kernel
        const uint32_t tmask = arith >= 0 ? P.tile_mask[bi] & ~fmask : 0u;  // slices on the one-tile kernel
        const bool in_fused = fused && P.fuse_bucket[bi] && fmask != 0;
        void* ctab = in_fused ? (void*)((char*)P.d_ctab_fused.p + P.fuse_off[bi]) : (P.rhs_only[bi] ? P.d_ctab_side.p : P.d_ctab.p);
        uint32_t all = (b.n_slices >= 32) ? 0xffffffffu : ((1u << b.n_slices) - 1u);
        uint32_t done = 0;
        // the factorized kernel reads the coefficients of one element contiguously
        const int qfast = fmask != 0 ? 1 : 0;
        P.last_qfast = qfast;
        P.last_ctab = ctab;
        if (!in_fused) {
            if (tm) tm->begin(T_COEF);
            const hipFunction_t cfn = bi < P.coef_fn.size() ? P.coef_fn[bi][ke64 ? 1 : 0] : nullptr;
            if (cfn) HIPCHECK(coef_launch(cfn, b, in, b.q_inv ? 1 : b.n_gp, qfast, ctab, bst));
            else {
                SLX_CHECK(b.n_regs <= SLX_MAX_REGS, "coefficient program of " << b.n_regs << " registers has no compiled kernel and the interpreter holds "
                                                   << SLX_MAX_REGS << " (" << P.coef_jit_error << ")");
                HIPCHECK(launch_coef(b, in, ke64, ctab, bst, qfast));
            }
            if (tm) tm->end();
        }
        if (arith >= 0 && P.mfma[bi].M > 0 && ((P.mfma[bi].arith_ok >> arith) & 1u)) {
            if (tm) tm->begin(T_MFMA);
            HIPCHECK(launch_mfma(b, mfma_for(P.mfma[bi], arith), (const float*)ctab, P.d_ke.as<float>(), bst));
            if (tm) tm->end();
            done = P.mfma_mask[bi];
        }
        if (fmask != 0 && in_fused) {
            done |= fmask;  // computed by the fused launch
        } else if (fmask != 0) {
            if (tm) tm->begin(T_MFMA);
            for (auto& fs : P.factor[bi])
                if ((fmask >> fs.first) & 1u) {
                    if (ke64) HIPCHECK(launch_factor64(b, fs.second, (const double*)ctab, P.d_ke.as<double>(), P.n_cu, bst));
                    else HIPCHECK(launch_factor(b, fs.second, arith, (const float*)ctab, P.d_ke.as<float>(), P.n_cu, bst));
                }
            if (tm) tm->end();
            done |= fmask;
        }
        if (tmask != 0) {
            if (tm) tm->begin(T_MFMA);
            for (auto& ts : P.tile[bi])
                if ((tmask >> ts.first) & 1u) HIPCHECK(launch_tile(b, ts.second, arith == AR_BF16 ? 1 : arith == AR_F32 ? 2 : 0, (const float*)ctab, qfast, P.d_ke.as<float>(),
                                                                P.n_cu, bst));
            if (tm) tm->end();
            done |= tmask;
        }
        int p = (precision == SLX_PREC_FP64) ? SLX_PREC_FP64 : SLX_PREC_FP32;
        if ((all & ~done) != 0) {
            if (tm) tm->begin(T_VALU);
            HIPCHECK(launch_elem(b, ctab, p, P.d_ke.p, bst, all & ~done, qfast));
            if (tm) tm->end();
        }
    }
    if (P.has_matrix) {
        if (tm) tm->begin(T_REDUCE);
        std::vector<int64_t> cb, ce;
        std::vector<int> cl;
        for (const RedClass& c : P.red_classes) {
            cb.push_back(c.begin);
            ce.push_back(c.end);
            cl.push_back(c.max_len);
        }
        HIPCHECK(launch_reduce(P.rd, P.max_erow_len, ke64, P.d_ke.p, P.d_Aval.as<double>(), P.d_Dval.as<double>(), P.d_rhs.as<double>(),
                               true, false, st, cb.data(), ce.data(), cl.data(), (int)cl.size()));
        if (tm) tm->end();
    }
    if (P.has_rhs) {
        if (tm) tm->begin(T_RHS);
        HIPCHECK(launch_reduce(P.rd, P.max_erow_len, ke64, P.d_ke.p, P.d_Aval.as<double>(), P.d_Dval.as<double>(), P.d_rhs.as<double>(),
                               false, true, overlap ? P.side : st, nullptr, nullptr, nullptr, 0));
        if (tm) tm->end();
    }
    if (overlap) {
        // join
        HIPCHECK(hipEventRecord(P.ev_join, P.side));
        HIPCHECK(hipStreamWaitEvent(st, P.ev_join, 0));
    }
    P.last_precision = precision;
}

}  // namespace

extern "C" {

const char* slx_last_error(void) { return g_last_error.c_str(); }

slx_status slx_create(int device, slx_ctx* out) {
    return guarded([&] {
        SLX_CHECK(out, "null output pointer");
        int n = 0;
        HIPCHECK(hipGetDeviceCount(&n));
        SLX_CHECK(device >= 0 && device < n, "device " << device << " not available (" << n << " devices)");
        HIPCHECK(hipSetDevice(device));
        // device address of the no-allocation solve's workspace array (once per device)
        HIPCHECK(scratch_init(device));
        auto c = new slx_ctx_s();
        c->device = device;
        if (hipStreamCreate(&c->stream) != hipSuccess) {
            delete c;
            throw DeviceError("hipStreamCreate failed");
        }
        *out = c;
    });
}

slx_status slx_destroy(slx_ctx ctx) {
    return guarded([&] { ctx_release(ctx); });
}

slx_status slx_plan_create(slx_ctx ctx, const slx_plan_desc* desc, slx_plan* out) {
    return slx_plan_create_ex(ctx, desc, 0, out);
}

slx_status slx_plan_create_ex(slx_ctx ctx, const slx_plan_desc* desc, int32_t flags, slx_plan* out) {
    return guarded([&] {
        SLX_CHECK(desc && out, "null argument");
        SLX_CHECK((flags & ~SLX_PLAN_HOST_SYMBOLIC) == 0, "unknown plan flags");
        auto P = std::make_unique<slx_plan_s>();
        P->host_only = (ctx == nullptr);
        if (const char* t = std::getenv("SLX_TUNE_FP64_FACTOR")) P->factor64 = std::atoi(t) != 0;
        if (!P->host_only) HIPCHECK(hipSetDevice(ctx->device));
        P->ctx = ctx;
        P->stream = ctx ? ctx->stream : nullptr;
        P->src = desc;
        P->n_dofs = desc->n_dofs;
        P->mesh_dim = desc->mesh_dim;
        P->n_fields = desc->n_fields;
        P->has_matrix = desc->has_matrix != 0;
        P->has_rhs = desc->has_rhs != 0;
        SLX_CHECK(desc->n_dofs >= 0 && desc->n_nodes >= 0 && desc->n_buckets >= 0, "negative sizes in plan description");
        const auto t0 = std::chrono::steady_clock::now();
        const KeLayout L = compute_layout(*desc);
        if (!P->host_only) upload_buckets(*P, *desc, L);
        bool done = false;
        if (!P->host_only && !(flags & SLX_PLAN_HOST_SYMBOLIC)) done = build_symbolic_device(*P, *desc, L);
        if (!done) {
            check_addresses(desc);
            build_symbolic(*P, *desc, L);
        }
        P->symbolic_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (!P->host_only) HIPCHECK(hipStreamSynchronize(P->stream));
        if (ctx) ctx->refs.fetch_add(1);
        *out = P.release();
    });
}

slx_status slx_plan_symbolic_info(slx_plan p, int32_t* on_device, double* seconds) {
    return guarded([&] {
        SLX_CHECK(p, "null plan");
        if (on_device) *on_device = p->symbolic_on_device ? 1 : 0;
        if (seconds) *seconds = p->symbolic_seconds;
    });
}

slx_status slx_plan_symbolic_array(slx_plan p, int32_t which, void* out, int64_t cap_bytes, int64_t* n_bytes) {
    return guarded([&] {
        SLX_CHECK(p && n_bytes, "null argument");
        SLX_CHECK(!p->host_only, "pattern-only plans keep their maps on the host (slx_plan_reduction_maps)");
        const int64_t nf = p->nfree;
        const void* src = nullptr;
        int64_t nb = 0;
        std::vector<int64_t> host;
        switch (which) {
            case 0: src = p->d_rdesc.p; nb = nf * 4 * 8; break;
            case 1: src = p->d_erows.p; nb = p->n_erows * 8; break;
            case 2: src = p->d_slots.p; nb = p->n_slots_w * p->rd.slot_bytes; break;
            case 3: src = p->d_rhs_ptr.p; nb = (nf + 1) * 8; break;
            case 4: src = p->d_rhs_erows.p; nb = p->n_rhs_erows * 8; break;
            case 5: src = p->d_free.p; nb = nf * 8; break;
            case 6: src = p->d_cons.p; nb = p->ncons * 8; break;
            case 7: src = p->d_ablk.p; nb = nf > 0 ? (p->n_ablk + 1) * 8 : 0; break;
            case 8:
                for (const RedClass& c : p->red_classes) {
                    host.push_back(c.begin);
                    host.push_back(c.end);
                    host.push_back(c.max_len);
                }
                nb = (int64_t)host.size() * 8;
                break;
            case 9:
                host = {p->rd.slot_bytes, p->rd.max_row_len, p->max_erow_len, p->n_erows, p->n_coo};
                nb = (int64_t)host.size() * 8;
                break;
            default: fail("unknown symbolic array");
        }
        *n_bytes = nb;
        if (!out) return;
        SLX_CHECK(cap_bytes >= nb, "output buffer too small");
        if (which >= 8) {
            std::memcpy(out, host.data(), nb);
        } else if (nb > 0) {
            HIPCHECK(hipStreamSynchronize(p->stream));
            HIPCHECK(hipMemcpy(out, src, nb, hipMemcpyDeviceToHost));
        }
    });
}

slx_status slx_plan_shape(slx_plan p, int64_t* nfree, int64_t* ncons, int64_t* nnzA, int64_t* nnzD) {
    return guarded([&] {
        SLX_CHECK(p, "null plan");
        if (nfree) *nfree = p->nfree;
        if (ncons) *ncons = p->ncons;
        if (nnzA) *nnzA = p->nnzA;
        if (nnzD) *nnzD = p->nnzD;
    });
}

slx_status slx_plan_stats(slx_plan p, int64_t* n_elem_rows, int64_t* n_ke, int64_t* n_coo) {
    return guarded([&] {
        SLX_CHECK(p, "null plan");
        if (n_elem_rows) *n_elem_rows = p->n_erows;
        if (n_ke) *n_ke = p->n_ke;
        if (n_coo) *n_coo = p->n_coo;
    });
}

namespace {
void assemble_host_inputs(slx_plan p, const slx_numeric_desc* in, hipStream_t st);
}

slx_status slx_assemble(slx_plan p, const slx_numeric_desc* in) {
    return guarded([&] {
        SLX_CHECK(p && in, "null argument");
        assemble_host_inputs(p, in, p->stream);
    });
}

slx_status slx_assemble_async(slx_plan p, const slx_numeric_desc* in, void* hip_stream) {
    return guarded([&] {
        SLX_CHECK(p && in, "null argument");
        assemble_host_inputs(p, in, hip_stream ? (hipStream_t)hip_stream : p->stream);
    });
}

slx_status slx_plan_dof_maps(slx_plan p, int64_t* ainds, int64_t* dinds) {
    return guarded([&] {
        SLX_CHECK(p && p->src, "null plan");
        const slx_plan_desc& d = *p->src;
        int64_t na = 0, nc = 0;
        for (int64_t i = 0; i < d.n_dofs; i++) {
            const bool c = d.constrained && d.constrained[i];
            if (c) {
                if (dinds) dinds[nc] = i;
                nc++;
            } else {
                if (ainds) ainds[na] = i;
                na++;
            }
        }
    });
}

namespace {
void assemble_host_inputs(slx_plan p, const slx_numeric_desc* in, hipStream_t st) {
    {
        SLX_CHECK(!p->host_only, "plan was created without a device context (pattern only)");
        // the input buffers are replaced below: earlier work that reads them must be done
        HIPCHECK(hipStreamSynchronize(st));
        if (st != p->stream) HIPCHECK(hipStreamSynchronize(p->stream));
        const slx_plan_desc& d = *p->src;
        upload(p->d_xyz, in->node_xyz, (size_t)3 * d.n_nodes, p->stream);
        SLX_CHECK(in->n_params > p->max_param, "numeric description has " << in->n_params << " parameters, programs read index " << p->max_param);
        size_t np = (size_t)in->n_params * in->param_len;
        upload(p->d_params, in->params, std::max<size_t>(np, 1), p->stream);
        for (int b = 0; b < d.n_buckets; b++)
            for (int e = 0; e < d.buckets[b].n_elem && in->n_params > 0; e++)
                SLX_CHECK(d.buckets[b].elem_ids[e] >= 0 && d.buckets[b].elem_ids[e] < in->param_len, "element id beyond parameter length");
        std::vector<const double*> fptr(p->n_fields);
        p->d_fields.resize(p->n_fields);
        for (int f = 0; f < p->n_fields; f++) {
            if (!p->d_fields[f]) p->d_fields[f].reset(new DBuf());
            SLX_CHECK(in->field_data && in->field_len, "field data required by the coefficient program");
            upload(*p->d_fields[f], in->field_data[f], (size_t)std::max<int64_t>(in->field_len[f], 1), p->stream);
            fptr[f] = p->d_fields[f]->as<double>();
        }
        numeric(*p, p->d_xyz.as<double>(), p->d_params.as<double>(), in->param_len, fptr.data(), in->precision, st);
    }
}
}  // namespace

slx_status slx_assemble_device(slx_plan p, const double* d_node_xyz, const double* d_params, int64_t param_len,
                               const double* const* d_field_data, int32_t precision, void* hip_stream) {
    return guarded([&] {
        SLX_CHECK(p && d_node_xyz, "null argument");
        SLX_CHECK(!p->host_only, "plan was created without a device context (pattern only)");
        hipStream_t st = hip_stream ? (hipStream_t)hip_stream : p->stream;
        numeric(*p, d_node_xyz, d_params, param_len, d_field_data, precision, st);
    });
}

slx_status slx_plan_profile(slx_plan p, const double* d_node_xyz, const double* d_params, int64_t param_len,
                            const double* const* d_field_data, int32_t precision, int32_t iters, double* ms) {
    return guarded([&] {
        SLX_CHECK(p && d_node_xyz && ms && iters > 0, "bad argument");
        SLX_CHECK(!p->host_only, "plan was created without a device context (pattern only)");
        LaunchTimer tm;
        tm.st = p->stream;
        double acc[T_COUNT] = {0, 0, 0, 0, 0};
        for (int it = 0; it < iters; it++) {
            numeric(*p, d_node_xyz, d_params, param_len, d_field_data, precision, p->stream, &tm);
            tm.collect(acc);
        }
        for (int k = 0; k < T_COUNT; k++) ms[k] = acc[k] / iters;
    });
}

slx_status slx_plan_sync(slx_plan p) {
    return guarded([&] {
        SLX_CHECK(p, "null plan");
        if (p->host_only) return;
        HIPCHECK(hipStreamSynchronize(p->stream));
        HIPCHECK(hipDeviceSynchronize());
    });
}

slx_status slx_csr_device(slx_plan p, const int64_t** Arow, const int32_t** Acol, const double** Aval, const int64_t** Drow,
                          const int32_t** Dcol, const double** Dval, const double** rhs) {
    return guarded([&] {
        SLX_CHECK(p, "null plan");
        if (Arow) *Arow = p->d_Arow.as<int64_t>();
        if (Acol) *Acol = p->d_Acol.as<int32_t>();
        if (Aval) *Aval = p->d_Aval.as<double>();
        if (Drow) *Drow = p->d_Drow.as<int64_t>();
        if (Dcol) *Dcol = p->d_Dcol.as<int32_t>();
        if (Dval) *Dval = p->d_Dval.as<double>();
        if (rhs) *rhs = p->d_rhs.as<double>();
    });
}

slx_status slx_csr_copyout(slx_plan p, int64_t* Arow, int32_t* Acol, double* Aval, int64_t* Drow, int32_t* Dcol, double* Dval,
                           double* rhs) {
    return guarded([&] {
        SLX_CHECK(p, "null plan");
        if (!p->host_only) HIPCHECK(hipStreamSynchronize(p->stream));
        SLX_CHECK(!p->host_only || (!Aval && !Dval && !rhs), "plan was created without a device context: only the pattern is available");
        if (Arow) std::copy(p->Arow.begin(), p->Arow.end(), Arow);
        if (Drow) std::copy(p->Drow.begin(), p->Drow.end(), Drow);
        if (p->host_only) {
            if (Acol) std::copy(p->Acol.begin(), p->Acol.end(), Acol);
            if (Dcol) std::copy(p->Dcol.begin(), p->Dcol.end(), Dcol);
        } else {
            // device plans keep the columns on the device only
            if (Acol && p->nnzA) HIPCHECK(hipMemcpy(Acol, p->d_Acol.p, p->nnzA * sizeof(int32_t), hipMemcpyDeviceToHost));
            if (Dcol && p->nnzD) HIPCHECK(hipMemcpy(Dcol, p->d_Dcol.p, p->nnzD * sizeof(int32_t), hipMemcpyDeviceToHost));
        }
        if (Aval && p->nnzA) HIPCHECK(hipMemcpy(Aval, p->d_Aval.p, p->nnzA * sizeof(double), hipMemcpyDeviceToHost));
        if (Dval && p->nnzD) HIPCHECK(hipMemcpy(Dval, p->d_Dval.p, p->nnzD * sizeof(double), hipMemcpyDeviceToHost));
        if (rhs && p->n_dofs) HIPCHECK(hipMemcpy(rhs, p->d_rhs.p, p->n_dofs * sizeof(double), hipMemcpyDeviceToHost));
    });
}

namespace {
// device view of the assembled system for the solve kernels (the row blocks
// of the streaming product are built with the pattern)
SolveSystem solve_view(slx_plan p) {
    return SolveSystem{p->n_dofs, p->nfree, p->ncons, p->nnzA, p->nnzD, p->d_Arow.as<int64_t>(), p->d_Drow.as<int64_t>(),
                       p->d_Acol.as<int32_t>(), p->d_Dcol.as<int32_t>(), p->d_Aval.as<double>(), p->d_Dval.as<double>(),
                       p->d_rhs.as<double>(), p->d_free.as<int64_t>(), p->d_cons.as<int64_t>(), p->d_ablk.as<int64_t>(),
                       p->n_ablk};
}

void check_assembled(slx_plan p) {
    SLX_CHECK(p, "null plan");
    SLX_CHECK(!p->host_only, "plan was created without a device context");
    SLX_CHECK(p->has_matrix, "the formulation has no matrix terms");
    SLX_CHECK(p->last_precision >= 0, "no assembly has run on this plan");
}
}  // namespace

slx_status slx_plan_eliminate(slx_plan p, const double* b, double* out) {
    return guarded([&] {
        check_assembled(p);
        SLX_CHECK(out || p->nfree == 0, "null output");
        HIPCHECK(hipSetDevice(p->ctx->device));
        const SolveSystem sys = solve_view(p);
        DBuf ws;
        ws.alloc((size_t)(p->n_dofs + p->nfree + p->ncons + 1) * sizeof(double));
        double* db = ws.as<double>();
        double* dout = db + p->n_dofs;
        double* dxd = dout + p->nfree;
        if (b) HIPCHECK(hipMemcpyAsync(db, b, p->n_dofs * sizeof(double), hipMemcpyHostToDevice, p->stream));
        else HIPCHECK(hipMemcpyAsync(db, p->d_rhs.p, p->n_dofs * sizeof(double), hipMemcpyDeviceToDevice, p->stream));
        HIPCHECK(eliminate_system(sys, db, dout, dxd, p->stream));
        if (p->nfree) HIPCHECK(hipMemcpyAsync(out, dout, p->nfree * sizeof(double), hipMemcpyDeviceToHost, p->stream));
        HIPCHECK(hipStreamSynchronize(p->stream));
    });
}

slx_status slx_plan_matvec(slx_plan p, const double* x, double* y) {
    return guarded([&] {
        check_assembled(p);
        SLX_CHECK((x && y) || p->n_dofs == 0, "null argument");
        HIPCHECK(hipSetDevice(p->ctx->device));
        const SolveSystem sys = solve_view(p);
        DBuf ws;
        ws.alloc((size_t)(2 * p->n_dofs + 2 * p->nfree + p->ncons + 1) * sizeof(double));
        double* dx = ws.as<double>();
        double* dy = dx + p->n_dofs;
        double* xa = dy + p->n_dofs;
        double* ya = xa + p->nfree;
        double* xd = ya + p->nfree;
        if (p->n_dofs) HIPCHECK(hipMemcpyAsync(dx, x, p->n_dofs * sizeof(double), hipMemcpyHostToDevice, p->stream));
        HIPCHECK(matvec_system(sys, dx, dy, xa, xd, ya, p->stream));
        if (p->n_dofs) HIPCHECK(hipMemcpyAsync(y, dy, p->n_dofs * sizeof(double), hipMemcpyDeviceToHost, p->stream));
        HIPCHECK(hipStreamSynchronize(p->stream));
    });
}

namespace {
// Device workspace of one solve, in bytes: a 256-byte header (the symmetry
// check's flag), then the doubles, then the floats (16-byte aligned).
struct SolveLayout {
    int64_t nD, nF, nV;
    size_t bytes;
};
SolveLayout solve_layout(slx_plan p, bool gm, int restart) {
    const int64_t nf = p->nfree, nc = p->ncons, nd = p->n_dofs;
    SolveLayout L;
    const int64_t nDg = gm ? (int64_t)(restart + 1) * (SLX_SOLVE_BLOCKS + 1) : 0;
    L.nD = 3 * nf + nc + nd + 2 * SLX_SOLVE_BLOCKS + 8 + nDg;
    L.nF = 6 * nf + p->nnzA + 4;
    L.nV = gm ? (int64_t)(restart + 1) * nf : 0;
    L.bytes = 256 + (size_t)L.nD * sizeof(double) + (size_t)(L.nF + L.nV) * sizeof(float) + 64;
    return L;
}

void check_solve_args(slx_plan p, int32_t method, double rel_tol, int32_t max_outer, int32_t max_inner, double* x_out, double* d_x_out) {
    check_assembled(p);
    SLX_CHECK(method == SLX_SOLVE_AUTO || method == SLX_SOLVE_CG || method == SLX_SOLVE_GMRES, "solve: unknown method " << method);
    SLX_CHECK(rel_tol > 0 && max_outer > 0 && max_inner > 0, "solve: rel_tol, max_outer and max_inner must be positive");
    SLX_CHECK((x_out != nullptr) != (d_x_out != nullptr), "solve: give exactly one of x_out and d_x_out");
}

// The solve on caller workspace `ws` (device memory of `ws_bytes`, 256-byte
// aligned); ws == nullptr: the workspace is allocated here for this call.
// Allocates no device memory when ws is given.
void solve_on(slx_plan p, int32_t method, double rel_tol, int32_t max_outer, int32_t max_inner, double* x_out, double* d_x_out,
              int32_t* outer_its, int32_t* inner_its, double* rel_residual, int32_t* method_used, void* ws, size_t ws_bytes) {
    HIPCHECK(hipSetDevice(p->ctx->device));
    const int64_t nf = p->nfree, nd = p->n_dofs;
    if (nf == 0) {
        // every DOF constrained: the solution is the Dirichlet values, which
        // the assembled rhs holds at the constrained entries
        if (x_out && nd) HIPCHECK(hipMemcpyAsync(x_out, p->d_rhs.as<double>(), nd * sizeof(double), hipMemcpyDeviceToHost, p->stream));
        if (d_x_out && nd)
            HIPCHECK(hipMemcpyAsync(d_x_out, p->d_rhs.as<double>(), nd * sizeof(double), hipMemcpyDeviceToDevice, p->stream));
        HIPCHECK(hipStreamSynchronize(p->stream));
        if (outer_its) *outer_its = 0;
        if (inner_its) *inner_its = 0;
        if (rel_residual) *rel_residual = 0.0;
        if (method_used) *method_used = method == SLX_SOLVE_AUTO ? SLX_SOLVE_CG : method;
        return;
    }
    const SolveSystem sys = solve_view(p);
    DBuf own;
    const int restart = std::min<int>(max_inner, SLX_GMRES_MAX);
    if (method == SLX_SOLVE_AUTO) {
        // the check's flag: the first word of the workspace (or of a small
        // allocation of its own when the solve allocates)
        int* flag = nullptr;
        if (ws) {
            SLX_CHECK(ws_bytes >= 256, "solve: workspace smaller than its header");
            flag = static_cast<int*>(ws);
        } else {
            own.alloc(256);
            flag = own.as<int>();
        }
        int nonsym = 0;
        HIPCHECK(check_symmetric(sys, flag, &nonsym, p->stream));
        method = nonsym ? SLX_SOLVE_GMRES : SLX_SOLVE_CG;
    }
    const bool gm = method == SLX_SOLVE_GMRES;
    const SolveLayout L = solve_layout(p, gm, restart);
    unsigned char* base = static_cast<unsigned char*>(ws);
    if (ws) {
        SLX_CHECK(L.bytes <= ws_bytes, "solve: the workspace of this system (" << L.bytes << " bytes: " << nf << " free rows, "
                                          << p->nnzA << " nonzeros" << (gm ? ", GMRES basis" : "") << ") exceeds the "
                                          << ws_bytes << " bytes available");
    } else {
        own.alloc(L.bytes);
        base = own.as<unsigned char>();
    }
    SolveWork w;
    double* dp = reinterpret_cast<double*>(base + 256);
    w.f = dp; dp += nf;
    w.r = dp; dp += nf;
    w.x = dp; dp += nf;
    w.xd = dp; dp += p->ncons;
    w.xfull = dp; dp += nd;
    w.part = dp; dp += 2 * SLX_SOLVE_BLOCKS;
    w.sc = dp; dp += 8;
    w.mpart = gm ? dp : nullptr; dp += gm ? (int64_t)(restart + 1) * SLX_SOLVE_BLOCKS : 0;
    w.hv = gm ? dp : nullptr; dp += gm ? restart + 1 : 0;
    float* fp = (float*)dp;
    w.r32 = fp; fp += nf;
    w.z = fp; fp += nf;
    w.p = fp; fp += nf;
    w.q = fp; fp += nf;
    w.d = fp; fp += nf;
    w.dinv = fp; fp += nf;
    w.val32 = (float*)(((uintptr_t)fp + 15) & ~(uintptr_t)15);
    w.V = gm ? w.val32 + p->nnzA + 4 : nullptr;
    SolveParams prm{rel_tol, 1e-3, max_outer, max_inner, gm ? 1 : 0, restart};
    SolveStats stats;
    HIPCHECK(solve_system(sys, prm, w, stats, p->stream));
    if (x_out && nd) HIPCHECK(hipMemcpyAsync(x_out, w.xfull, nd * sizeof(double), hipMemcpyDeviceToHost, p->stream));
    if (d_x_out && nd) HIPCHECK(hipMemcpyAsync(d_x_out, w.xfull, nd * sizeof(double), hipMemcpyDeviceToDevice, p->stream));
    HIPCHECK(hipStreamSynchronize(p->stream));
    if (outer_its) *outer_its = stats.outer;
    if (inner_its) *inner_its = stats.inner;
    if (rel_residual) *rel_residual = stats.rel_residual;
    if (method_used) *method_used = method;
}
}  // namespace

slx_status slx_plan_solve(slx_plan p, int32_t method, double rel_tol, int32_t max_outer, int32_t max_inner, double* x_out,
                          double* d_x_out, int32_t* outer_its, int32_t* inner_its, double* rel_residual, int32_t* method_used) {
    return guarded([&] {
        check_solve_args(p, method, rel_tol, max_outer, max_inner, x_out, d_x_out);
        solve_on(p, method, rel_tol, max_outer, max_inner, x_out, d_x_out, outer_its, inner_its, rel_residual, method_used,
                 nullptr, 0);
    });
}

slx_status slx_plan_kernels(slx_plan p, int32_t* n_mfma, int32_t* n_buckets) {
    return guarded([&] {
        SLX_CHECK(p, "null plan");
        int n = 0;
        for (auto& m : p->mfma) n += m.M > 0 ? 1 : 0;
        if (n_mfma) *n_mfma = n;
        if (n_buckets) *n_buckets = (int32_t)p->dbuckets.size();
    });
}

slx_status slx_plan_fused_info(slx_plan p, int32_t precision, int32_t* n_slices, int64_t* n_items) {
    return guarded([&] {
        SLX_CHECK(p, "null plan");
        SLX_CHECK(precision >= 0 && precision <= 4, "unknown precision " << precision);
        const int fa = precision == SLX_PREC_F16_SPLIT ? AR_F16
                       : precision == SLX_PREC_BF16_SPLIT ? AR_BF16
                       : precision == SLX_PREC_FP32_MFMA ? AR_F32
                       : (precision == SLX_PREC_FP64 && p->factor64) ? 3 : -1;
        const bool on = p->fused && fa >= 0 && ((p->fused_arith >> fa) & 1u);
        if (n_slices) *n_slices = on ? p->n_fused[fa] : 0;
        if (n_items) *n_items = on ? p->fused_total[fa] : 0;
    });
}

slx_status slx_plan_kernel_counts(slx_plan p, int32_t precision, int32_t* n_batched, int32_t* n_factorized) {
    return guarded([&] {
        SLX_CHECK(p, "null plan");
        SLX_CHECK(precision >= 0 && precision <= 4, "unknown precision " << precision);
        const int arith = precision == SLX_PREC_F16_SPLIT ? AR_F16
                          : precision == SLX_PREC_BF16_SPLIT ? AR_BF16
                          : precision == SLX_PREC_FP32_MFMA ? AR_F32 : -1;
        int nb = 0, nf = 0;
        if (arith >= 0) {
            for (auto& m : p->mfma) nb += (m.M > 0 && ((m.arith_ok >> arith) & 1u)) ? 1 : 0;
            for (auto& v : p->factor)
                for (auto& fs : v) nf += ((fs.second.arith_ok >> arith) & 1u) ? 1 : 0;
        } else if (precision == SLX_PREC_FP64 && p->factor64) {
            for (auto& v : p->factor)
                for (auto& fs : v) nf += fs.second.f64_ok ? 1 : 0;
        }
        if (n_batched) *n_batched = nb;
        if (n_factorized) *n_factorized = nf;
    });
}

slx_status slx_plan_tile_count(slx_plan p, int32_t precision, int32_t* n_tile) {
    return guarded([&] {
        SLX_CHECK(p && n_tile, "null argument");
        SLX_CHECK(precision >= 0 && precision <= 4, "unknown precision " << precision);
        int n = 0;
        if (precision == SLX_PREC_F16_SPLIT || precision == SLX_PREC_BF16_SPLIT || precision == SLX_PREC_FP32_MFMA)
            for (auto& v : p->tile) n += (int)v.size();
        *n_tile = n;
    });
}

slx_status slx_plan_reduction_maps(slx_plan p, int64_t* sizes, int64_t* erow_ptr, uint64_t* erows, int64_t* slot_ptr,
                                   int32_t* slots, int64_t* rhs_ptr, uint64_t* rhs_erows, int64_t* ke_base, int64_t* ke_ld) {
    return guarded([&] {
        SLX_CHECK(p && sizes, "null argument");
        SLX_CHECK(p->host_only || (!erow_ptr && !erows && !slot_ptr && !slots && !rhs_ptr && !rhs_erows),
                  "reduction maps are kept only by pattern-only plans (ctx == NULL); the layout is always available");
        sizes[0] = (int64_t)p->h_erows.size();
        sizes[1] = (int64_t)p->h_slots.size();
        sizes[2] = (int64_t)p->h_rhs_erows.size();
        sizes[3] = (int64_t)p->h_ke_base.size();
        auto cp = [](auto* dst, const auto& v) {
            if (dst) std::copy(v.begin(), v.end(), dst);
        };
        cp(erow_ptr, p->h_erow_ptr);
        cp(erows, p->h_erows);
        cp(slot_ptr, p->h_slot_ptr);
        cp(slots, p->h_slots);
        cp(rhs_ptr, p->h_rhs_ptr);
        cp(rhs_erows, p->h_rhs_erows);
        cp(ke_base, p->h_ke_base);
        cp(ke_ld, p->h_ke_ld);
    });
}

slx_status slx_plan_graph_capture(slx_plan p, const double* d_node_xyz, const double* d_params, int64_t param_len,
                                  const double* const* d_field_data, int32_t precision) {
    return guarded([&] {
        SLX_CHECK(p && d_node_xyz, "null argument");
        SLX_CHECK(!p->host_only, "plan was created without a device context (pattern only)");
        HIPCHECK(hipSetDevice(p->ctx->device));
        HIPCHECK(hipStreamSynchronize(p->stream));
        if (p->graph_exec) {
            (void)hipGraphExecDestroy(p->graph_exec);
            p->graph_exec = nullptr;
        }
        hipGraph_t g = nullptr;
        HIPCHECK(hipStreamBeginCapture(p->stream, hipStreamCaptureModeRelaxed));
        try {
            numeric(*p, d_node_xyz, d_params, param_len, d_field_data, precision, p->stream);
        } catch (...) {
            hipGraph_t dropped = nullptr;
            (void)hipStreamEndCapture(p->stream, &dropped);
            if (dropped) (void)hipGraphDestroy(dropped);
            throw;
        }
        HIPCHECK(hipStreamEndCapture(p->stream, &g));
        hipError_t e = hipGraphInstantiate(&p->graph_exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCHECK(e);
    });
}

slx_status slx_plan_graph_launch(slx_plan p, void* hip_stream) {
    return guarded([&] {
        SLX_CHECK(p && p->graph_exec, "no captured graph on this plan");
        HIPCHECK(hipGraphLaunch(p->graph_exec, hip_stream ? (hipStream_t)hip_stream : p->stream));
    });
}

slx_status slx_plan_coefficients(slx_plan p, double* out, int64_t cap, int64_t* n_out) {
    return guarded([&] {
        SLX_CHECK(p && n_out, "null argument");
        SLX_CHECK(!p->host_only, "plan was created without a device context (pattern only)");
        SLX_CHECK(p->last_precision >= 0 && !p->dbuckets.empty(), "no assembly has run on this plan");
        const DevBucket& b = p->dbuckets.back();
        const void* ctab = p->last_ctab;
        const int64_t n = (int64_t)b.n_coefs * (b.q_inv ? 1 : b.n_gp) * b.n_elem;
        *n_out = n;
        if (!out) return;
        SLX_CHECK(cap >= n, "output holds " << cap << " entries, the table has " << n);
        HIPCHECK(hipStreamSynchronize(p->stream));
        std::vector<double> tmp(n);
        if (p->last_precision == SLX_PREC_FP64) {
            HIPCHECK(hipMemcpy(tmp.data(), ctab, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
        } else {
            std::vector<float> t32(n);
            HIPCHECK(hipMemcpy(t32.data(), ctab, (size_t)n * sizeof(float), hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < n; i++) tmp[i] = t32[i];
        }
        const int64_t nq = b.q_inv ? 1 : b.n_gp, ne = b.n_elem;
        if (p->last_qfast) {
            // the factorized kernels' layout (slot, element, q) -> (slot, q, element)
            for (int64_t i = 0; i < n; i++) {
                const int64_t c = i / (nq * ne), r = i % (nq * ne), e = r / nq, q = r % nq;
                out[(c * nq + q) * ne + e] = tmp[i];
            }
        } else {
            for (int64_t i = 0; i < n; i++) out[i] = tmp[i];
        }
    });
}

slx_status slx_plan_element_matrices(slx_plan p, double* out, int64_t cap) {
    return guarded([&] {
        SLX_CHECK(p && out, "null argument");
        SLX_CHECK(!p->host_only, "plan was created without a device context (pattern only)");
        SLX_CHECK(p->last_precision >= 0, "no assembly has run on this plan");
        SLX_CHECK(cap >= p->n_ke, "output holds " << cap << " entries, the element matrices have " << p->n_ke);
        HIPCHECK(hipStreamSynchronize(p->stream));
        if (p->last_precision == SLX_PREC_FP64) {
            HIPCHECK(hipMemcpy(out, p->d_ke.p, (size_t)p->n_ke * sizeof(double), hipMemcpyDeviceToHost));
        } else {
            std::vector<float> tmp(p->n_ke);
            HIPCHECK(hipMemcpy(tmp.data(), p->d_ke.p, (size_t)p->n_ke * sizeof(float), hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < p->n_ke; i++) out[i] = tmp[i];
        }
    });
}

slx_status slx_plan_coef_mode(slx_plan p, int32_t mode, int32_t* n_compiled) {
    return guarded([&] {
        SLX_CHECK(p, "null plan");
        SLX_CHECK(mode >= 0 && mode <= 2, "coefficient mode must be 0 (auto), 1 (interpreter) or 2 (compiled)");
        SLX_CHECK(!p->host_only, "plan was created without a device context");
        for (int b = 0; mode == 1 && b < p->src->n_buckets; b++)
            SLX_CHECK(p->src->buckets[b].n_regs <= SLX_MAX_REGS, "coefficient program of bucket " << b << " needs " << p->src->buckets[b].n_regs
                                                              << " registers; the interpreter holds " << SLX_MAX_REGS);
        HIPCHECK(hipSetDevice(p->ctx->device));
        if (mode != p->coef_mode) {
            HIPCHECK(hipStreamSynchronize(p->stream));
            p->coef_mode = mode;
            compile_coef_programs(*p, *p->src);
        }
        int n = 0;
        for (auto& f : p->coef_fn) n += (f[0] ? 1 : 0) + (f[1] ? 1 : 0);
        if (n_compiled) *n_compiled = n;
        if (mode == 2) {
            int want = 0;
            for (int b = 0; b < p->src->n_buckets; b++)
                if (p->src->buckets[b].n_coefs > 0 && p->src->buckets[b].n_elem > 0) want += 2;
            SLX_CHECK(n == want, "coefficient program compilation failed: " << p->coef_jit_error);
        }
    });
}

slx_status slx_plan_destroy(slx_plan p) {
    return guarded([&] {
        if (!p) return;
        if (!p->host_only) (void)hipStreamSynchronize(p->stream);
        if (p->side) (void)hipStreamSynchronize(p->side);
        if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
        if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
        if (p->ev_join) (void)hipEventDestroy(p->ev_join);
        if (p->side) (void)hipStreamDestroy(p->side);
        slx_ctx c = p->ctx;
        delete p;
        ctx_release(c);
    });
}

slx_status slx_selftest_mfma_layout(slx_ctx ctx, float* out256) {
    return guarded([&] {
        SLX_CHECK(ctx && out256, "null argument");
        HIPCHECK(hipSetDevice(ctx->device));
        DBuf d;
        d.alloc(256 * sizeof(float));
        HIPCHECK(launch_mfma_probe(0, d.as<float>(), ctx->stream));
        HIPCHECK(hipStreamSynchronize(ctx->stream));
        HIPCHECK(hipMemcpy(out256, d.p, 256 * sizeof(float), hipMemcpyDeviceToHost));
    });
}

slx_status slx_selftest_mfma_layout_kind(slx_ctx ctx, int32_t kind, float* out256) {
    return guarded([&] {
        SLX_CHECK(ctx && out256, "null argument");
        SLX_CHECK(kind >= 0 && kind <= 3, "probe kind must be 0..3");
        HIPCHECK(hipSetDevice(ctx->device));
        DBuf d;
        d.alloc(256 * sizeof(float));
        HIPCHECK(launch_mfma_probe(kind, d.as<float>(), ctx->stream));
        HIPCHECK(hipStreamSynchronize(ctx->stream));
        HIPCHECK(hipMemcpy(out256, d.p, 256 * sizeof(float), hipMemcpyDeviceToHost));
    });
}

slx_status slx_selftest_spmv_stream(slx_ctx ctx, int64_t nrows, const int64_t* rowp, const int32_t* col, const double* val,
                                    const double* x, double* y, double* xy) {
    return guarded([&] {
        SLX_CHECK(ctx && rowp && y && xy && nrows > 0, "null argument");
        const int64_t nnz = rowp[nrows];
        SLX_CHECK(rowp[0] == 0 && nnz >= 0, "row pointer must start at 0");
        for (int64_t r = 0; r < nrows; r++) SLX_CHECK(rowp[r + 1] >= rowp[r], "row pointer decreases at row " << r);
        for (int64_t k = 0; k < nnz; k++) SLX_CHECK(col[k] >= 0 && col[k] < nrows, "column index out of range at " << k);
        HIPCHECK(hipSetDevice(ctx->device));
        std::vector<int64_t> blk;
        stream_row_blocks(rowp, nrows, blk);
        std::vector<float> v32(val, val + nnz), x32(x, x + nrows);
        DBuf drow, dcol, dval, dblk, dx, dy, dpart;
        upload(drow, rowp, nrows + 1, ctx->stream);
        upload(dcol, nnz ? col : nullptr, std::max<int64_t>(nnz, 1), ctx->stream);
        upload(dval, nnz ? v32.data() : nullptr, std::max<int64_t>(nnz, 1), ctx->stream);
        upload(dblk, blk.data(), blk.size(), ctx->stream);
        upload(dx, x32.data(), nrows, ctx->stream);
        dy.alloc(nrows * sizeof(float));
        dpart.alloc((SLX_SOLVE_BLOCKS + 8) * sizeof(double));
        HIPCHECK(spmv_stream_probe(nrows, drow.as<int64_t>(), dcol.as<int32_t>(), dval.as<float>(), nnz, dblk.as<int64_t>(),
                                   (int64_t)blk.size() - 1, dx.as<float>(), dy.as<float>(), dpart.as<double>(),
                                   dpart.as<double>() + SLX_SOLVE_BLOCKS, ctx->stream));
        HIPCHECK(hipStreamSynchronize(ctx->stream));
        std::vector<float> y32(nrows);
        HIPCHECK(hipMemcpy(y32.data(), dy.p, nrows * sizeof(float), hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(xy, dpart.as<double>() + SLX_SOLVE_BLOCKS + 1, sizeof(double), hipMemcpyDeviceToHost));
        std::copy(y32.begin(), y32.end(), y);
    });
}

// ---- reference tables ----

slx_status slx_gauss_rule(int32_t elem_type, int32_t order, int32_t cap, int32_t* npts, double* coords, double* weights) {
    return guarded([&] {
        SLX_CHECK(npts, "null argument");
        GaussRule r = gauss_rule(elem_type, order);
        *npts = r.count();
        SLX_CHECK(r.count() <= cap || !coords, "output capacity " << cap << " below " << r.count() << " points");
        if (coords) std::copy(r.coords.begin(), r.coords.end(), coords);
        if (weights) std::copy(r.weights.begin(), r.weights.end(), weights);
    });
}

slx_status slx_h1_count(int32_t elem_type, int32_t order, int32_t* nff) {
    return guarded([&] {
        SLX_CHECK(nff, "null argument");
        *nff = h1_count(elem_type, order);
    });
}

slx_status slx_h1_tables(int32_t elem_type, int32_t order, int32_t total_orientation, int32_t npts, const double* pts, double* out) {
    return guarded([&] {
        SLX_CHECK(pts && out && npts >= 0, "bad argument");
        std::vector<double> p(pts, pts + 3 * (size_t)npts), o;
        h1_tables(elem_type, order, total_orientation, p, o);
        std::copy(o.begin(), o.end(), out);
    });
}

slx_status slx_hcurl_count(int32_t elem_type, int32_t order, int32_t* nff) {
    return guarded([&] {
        SLX_CHECK(nff && order >= 0, "bad argument");
        *nff = hcurl_count(elem_type, order);
    });
}

slx_status slx_hcurl_tables(int32_t elem_type, int32_t order, int32_t total_orientation, int32_t family, int32_t npts,
                            const double* pts, double* out) {
    return guarded([&] {
        SLX_CHECK(pts && out && npts >= 0 && order >= 0, "bad argument");
        std::vector<double> p(pts, pts + 3 * (size_t)npts), o;
        hcurl_tables(elem_type, order, total_orientation, family, p, o);
        std::copy(o.begin(), o.end(), out);
    });
}

slx_status slx_geometry_tables(int32_t elem_type, int32_t n_geo, int32_t npts, const double* pts, double* out) {
    return guarded([&] {
        SLX_CHECK(pts && out && npts >= 0 && elem_type >= 1 && elem_type <= 7, "bad argument");
        std::vector<double> p(pts, pts + 3 * (size_t)npts), o;
        geometry_tables(elem_type, n_geo, p, o);
        std::copy(o.begin(), o.end(), out);
    });
}

// ---- formulation layer ----

struct slx_problem_s {
    slx_ctx ctx = nullptr;
    std::shared_ptr<Problem> prob;
    slx_plan plan = nullptr;
};

slx_status slx_problem_create(slx_ctx ctx, const char* kind, const char* spec, slx_problem* out) {
    return guarded([&] {
        SLX_CHECK(kind && out, "null argument");
        auto p = new slx_problem_s();
        p->ctx = ctx;
        try {
            if (ctx) HIPCHECK(hipSetDevice(ctx->device));
            p->prob = make_problem(kind, spec ? spec : "", ctx ? (void*)ctx->stream : nullptr);
        } catch (...) {
            delete p;
            throw;
        }
        *out = p;
    });
}

slx_status slx_problem_destroy(slx_problem p) {
    return guarded([&] {
        if (!p) return;
        if (p->plan) slx_plan_destroy(p->plan);
        delete p;
    });
}

slx_status slx_problem_plan(slx_problem p, slx_plan* plan_out) {
    return guarded([&] {
        SLX_CHECK(p, "null problem");
        if (!p->plan) {
            slx_plan pl = nullptr;
            if (slx_plan_create(p->ctx, &p->prob->plan->desc, &pl) != SLX_OK) throw std::runtime_error(g_last_error);
            p->plan = pl;
        }
        if (plan_out) *plan_out = p->plan;
    });
}

slx_status slx_problem_numeric(slx_problem p, slx_numeric_desc* out) {
    return guarded([&] {
        SLX_CHECK(p && out, "null argument");
        PlanStorage& ps = *p->prob->plan;
        std::memset(out, 0, sizeof(*out));
        out->node_xyz = p->prob->mesh->coords().data();
        out->n_params = ps.n_params;
        out->param_len = ps.param_len;
        out->params = ps.params.data();
        ps.field_data_ptrs.clear();
        for (auto& f : ps.plan_fields) ps.field_data_ptrs.push_back(f->data.data());
        out->field_data = ps.field_data_ptrs.data();
        out->field_len = ps.field_len.data();
        out->precision = SLX_PREC_FP64;
    });
}

slx_status slx_problem_generate(slx_problem p, int32_t precision) {
    return guarded([&] {
        slx_plan pl = nullptr;
        if (slx_problem_plan(p, &pl) != SLX_OK) throw std::runtime_error(g_last_error);
        slx_numeric_desc nd;
        if (slx_problem_numeric(p, &nd) != SLX_OK) throw std::runtime_error(g_last_error);
        nd.precision = precision;
        if (slx_assemble(pl, &nd) != SLX_OK) throw std::runtime_error(g_last_error);
        if (slx_plan_sync(pl) != SLX_OK) throw std::runtime_error(g_last_error);
    });
}

slx_status slx_problem_info(slx_problem p, int64_t* ndofs, int64_t* nfree, int64_t* nelem, int64_t* nbuckets) {
    return guarded([&] {
        SLX_CHECK(p, "null problem");
        PlanStorage& ps = *p->prob->plan;
        if (ndofs) *ndofs = ps.desc.n_dofs;
        if (nfree) {
            int64_t n = 0;
            for (uint8_t c : ps.constrained) n += c ? 0 : 1;
            *nfree = n;
        }
        if (nelem) *nelem = ps.n_elem_total;
        if (nbuckets) *nbuckets = ps.desc.n_buckets;
    });
}

slx_status slx_problem_desc(slx_problem p, const slx_plan_desc** desc) {
    return guarded([&] {
        SLX_CHECK(p && desc, "null argument");
        *desc = &p->prob->plan->desc;
    });
}

slx_status slx_problem_setdata_vec(slx_problem p, const double* x, int64_t n) {
    return guarded([&] {
        SLX_CHECK(p && x, "null argument");
        p->prob->form->dofmanager().setdata(x, n);
    });
}

slx_status slx_problem_setdata(slx_problem p, int32_t field, const double* values, int64_t n) {
    return guarded([&] {
        SLX_CHECK(p && values, "null argument");
        PlanStorage& ps = *p->prob->plan;
        SLX_CHECK(field >= 0 && field < (int)ps.plan_fields.size(), "field index out of range");
        auto& f = ps.plan_fields[field];
        SLX_CHECK(n == f->data_size(), "field data has " << f->data_size() << " entries, got " << n);
        std::copy(values, values + n, f->data.begin());
    });
}

}  // extern "C"

namespace slx {
// Entry of the no-allocation solve (kernels/noalloc.hip): the same solve on
// the caller's device workspace.
slx_status solve_on_workspace(slx_plan p, int32_t method, double rel_tol, int32_t max_outer, int32_t max_inner, double* x_out,
                              double* d_x_out, int32_t* outer_its, int32_t* inner_its, double* rel_residual, int32_t* method_used,
                              void* ws, size_t ws_bytes) {
    return guarded([&] {
        check_solve_args(p, method, rel_tol, max_outer, max_inner, x_out, d_x_out);
        SLX_CHECK(ws && ((uintptr_t)ws & 255u) == 0, "solve: workspace must be 256-byte aligned device memory");
        solve_on(p, method, rel_tol, max_outer, max_inner, x_out, d_x_out, outer_its, inner_its, rel_residual, method_used, ws,
                 ws_bytes);
    });
}
int plan_device(slx_plan p) { return p && p->ctx ? p->ctx->device : -1; }
void set_last_error(const char* msg) { g_last_error = msg; }
}  // namespace slx
