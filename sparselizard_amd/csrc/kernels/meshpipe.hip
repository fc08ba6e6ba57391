// Mesh loading pipeline on the device (rawmesh::load, rawmesh.cpp:254-325):
// elements::explode (elements.cpp:1115-1208), removeduplicates of every
// element type (rawmesh.cpp:160-182, gentools.cpp:136-246, elements.cpp:
// 965-998), definedisjointregions (elements.cpp:1257-1342),
// reorderbydisjointregions (elements.cpp:1344-1372) and orient
// (elements.cpp:1412-1441).  It leaves the Mesh exactly as Mesh::finalize does
// on the host: every numbering the sequential loops fix (first occurrence,
// last duplicate's data, stable region order) is rebuilt by stable radix
// sorts and scans.  Duplicate nodes (a tolerance test, gentools.cpp:89-134)
// are merged on the host before, which commutes with the explosion (both
// only copy or rename node numbers).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <climits>
#include <cstring>
#include <numeric>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include "../host/mesh.h"

namespace slx {
namespace {

struct DevTopo {
    int nn[8], ne[8], ntf[8], nqf[8];
    int cnt[8][8];                    // sub-entities of type st in type t (st 0: corner nodes, st == t: the element)
    int en_off[8], fn_off[8], fe_off[8];
    int tab[448];                     // edge nodes, face nodes, face edges of every type
};

inline unsigned nblk(int64_t n, int bs = 256) { return (unsigned)std::max<int64_t>(1, (n + bs - 1) / bs); }

// ---------------------------------------------------------------- explode

__global__ void k_explode(int t, int ne_el, DevTopo T, const int* __restrict__ nodes, int line_base, int tri_base, int quad_base,
                          int* __restrict__ line_nodes, int* __restrict__ t_lines, int* __restrict__ tri_nodes, int* __restrict__ t_tris,
                          int* __restrict__ tri_lines, int* __restrict__ quad_nodes, int* __restrict__ t_quads,
                          int* __restrict__ quad_lines) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne_el) return;
    const int nn = T.nn[t], ne = T.ne[t], ntf = T.ntf[t], nqf = T.nqf[t];
    const int* en = T.tab + T.en_off[t];
    const int* fn = T.tab + T.fn_off[t];
    const int* fe = T.tab + T.fe_off[t];
    const int* nd = nodes + (size_t)e * nn;
    const int firstline = line_base + e * ne;
    for (int l = 0; l < ne; l++) {
        line_nodes[2 * (size_t)(firstline + l)] = nd[en[2 * l]];
        line_nodes[2 * (size_t)(firstline + l) + 1] = nd[en[2 * l + 1]];
        t_lines[(size_t)e * ne + l] = firstline + l;
    }
    if (!tri_nodes) return;  // 2D: lines only
    for (int f = 0; f < ntf; f++) {
        const int idx = tri_base + e * ntf + f;
        for (int k = 0; k < 3; k++) {
            tri_nodes[3 * (size_t)idx + k] = nd[fn[3 * f + k]];
            tri_lines[3 * (size_t)idx + k] = firstline + abs(fe[3 * f + k]) - 1;
        }
        t_tris[(size_t)e * ntf + f] = idx;
    }
    for (int f = 0; f < nqf; f++) {
        const int idx = quad_base + e * nqf + f;
        for (int k = 0; k < 4; k++) {
            quad_nodes[4 * (size_t)idx + k] = nd[fn[3 * ntf + 4 * f + k]];
            quad_lines[4 * (size_t)idx + k] = firstline + abs(fe[3 * ntf + 4 * f + k]) - 1;
        }
        t_quads[(size_t)e * nqf + f] = idx;
    }
}

// ---------------------------------------------------------------- duplicates

__global__ void k_sorted_keys(int n, int nn, const int* __restrict__ nodes, int* __restrict__ key, int* __restrict__ perm) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    int v[8];
    for (int p = 0; p < nn; p++) v[p] = nodes[(size_t)e * nn + p];
    for (int i = 1; i < nn; i++)
        for (int j = i; j > 0 && v[j - 1] > v[j]; j--) {
            const int x = v[j];
            v[j] = v[j - 1];
            v[j - 1] = x;
        }
    for (int p = 0; p < nn; p++) key[(size_t)e * nn + p] = v[p];
    perm[e] = e;
}

__global__ void k_gather_word(int n, int nn, int p, const int* __restrict__ key, const int* __restrict__ perm, unsigned* __restrict__ w) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) w[j] = (unsigned)key[(size_t)perm[j] * nn + p];
}

__global__ void k_heads(int n, int nn, const int* __restrict__ key, const int* __restrict__ perm, int* __restrict__ head) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int h = j == 0;
    if (!h)
        for (int p = 0; p < nn; p++)
            if (key[(size_t)perm[j] * nn + p] != key[(size_t)perm[j - 1] * nn + p]) h = 1;
    head[j] = h;
}

// gid = inclusive scan of head, minus one
__global__ void k_groups(int n, const int* __restrict__ perm, const int* __restrict__ head, const int* __restrict__ gidp1,
                         int* __restrict__ first, int* __restrict__ last, int* __restrict__ g_of, int* __restrict__ isfirst) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int g = gidp1[j] - 1, e = perm[j];
    if (head[j]) first[g] = e;  // the stable sort puts the first occurrence first
    if (j == n - 1 || head[j + 1]) last[g] = e;
    g_of[e] = g;
    isfirst[e] = head[j];
}

__global__ void k_renum(int n, const int* __restrict__ g_of, const int* __restrict__ first, const int* __restrict__ num, int* __restrict__ renum) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) renum[e] = num[first[g_of[e]]];
}

// row u of the deduplicated table = the group's last duplicate (elements.cpp:965-998)
__global__ void k_copy_rows(int ng, int per, const int* __restrict__ first, const int* __restrict__ last, const int* __restrict__ num,
                            const int* __restrict__ src, int* __restrict__ dst) {
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= (int64_t)ng * per) return;
    const int g = (int)(x / per), c = (int)(x - (int64_t)g * per);
    dst[(size_t)num[first[g]] * per + c] = src[(size_t)last[g] * per + c];
}

__global__ void k_map(int64_t n, int* __restrict__ a, const int* __restrict__ renum) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = renum[a[i]];
}

// ---------------------------------------------------------------- disjoint regions

__global__ void k_mark(int m, const int* __restrict__ elems, int t, int st, int nsub, const int* __restrict__ sub, unsigned bit,
                       unsigned* __restrict__ mask) {
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= (int64_t)m * nsub) return;
    const int i = (int)(x / nsub), s = (int)(x - (int64_t)i * nsub);
    const int elem = elems[i];
    const int se = st == t ? elem : sub[(size_t)elem * nsub + s];
    atomicOr(mask + se, bit);
}

__global__ void k_first_of_mask(int n, const unsigned* __restrict__ mask, int* __restrict__ first) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) atomicMin(first + mask[e], e);
}

__global__ void k_dr_of(int n, const unsigned* __restrict__ mask, const int* __restrict__ table, unsigned* __restrict__ dr) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) dr[e] = (unsigned)table[mask[e]];
}

// ---------------------------------------------------------------- reorder, orientation

__global__ void k_iota(int n, int* __restrict__ a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = i;
}

__global__ void k_invert(int n, const int* __restrict__ ro, int* __restrict__ renum) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) renum[ro[i]] = i;
}

template <typename T>
__global__ void k_gather_rows(int n, int per, const int* __restrict__ ro, const T* __restrict__ src, T* __restrict__ dst) {
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= (int64_t)n * per) return;
    const int i = (int)(x / per), c = (int)(x - (int64_t)i * per);
    dst[x] = src[(size_t)ro[i] * per + c];
}

// orientation::gettotalorientation: edge bits, then one base-8 digit per face
// (elements.cpp:1412-1441, orientation.cpp)
__global__ void k_orient(int n, int t, DevTopo T, const int* __restrict__ nodes, int* __restrict__ out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int nn = T.nn[t], ne = T.ne[t], ntf = T.ntf[t], nqf = T.nqf[t];
    const int* en = T.tab + T.en_off[t];
    const int* fn = T.tab + T.fn_off[t];
    const int* nd = nodes + (size_t)e * nn;
    int total = 0, factor = 1;
    for (int i = 0; i < ne; i++) {
        total += (nd[en[2 * i]] > nd[en[2 * i + 1]] ? 0 : 1) * factor;
        factor *= 2;
    }
    for (int i = 0; i < ntf + nqf; i++) {
        const int k = i < ntf ? 3 : 4;
        const int* f = fn + (i < ntf ? 3 * i : 3 * ntf + 4 * (i - ntf));
        int maxpos = 0;
        for (int q = 1; q < k; q++)
            if (nd[f[q]] > nd[f[maxpos]]) maxpos = q;
        const int a = nd[f[(maxpos + 1) % k]], b = nd[f[(maxpos + k - 1) % k]];
        total += (a > b ? 2 * maxpos : 2 * maxpos + 1) * factor;
        factor *= 8;
    }
    out[e] = total;
}

struct Pool {
    std::vector<void*> v;
    ~Pool() {
        for (void* p : v) (void)hipFree(p);
    }
    template <typename T>
    T* get(int64_t n) {
        void* p = nullptr;
        if (hipMalloc(&p, (size_t)std::max<int64_t>(n, 1) * sizeof(T)) != hipSuccess) throw std::bad_alloc();
        v.push_back(p);
        return (T*)p;
    }
};

#define MPCHECK(x)                                                                                         \
    do {                                                                                                   \
        const hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) fail(std::string("device mesh pipeline: ") + #x + ": " + hipGetErrorString(e_)); \
    } while (0)

int bits_for(int64_t n) {
    int b = 1;
    while (b < 32 && (int64_t(1) << b) <= n) b++;
    return b;
}

}  // namespace

// Runs on `stream`; every array of the mesh is downloaded at the end.
void finalize_on_device(Mesh& m, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    Pool pool;
    void* rtmp = nullptr;
    size_t rbytes = 0;
    auto rgrow = [&](size_t need) {
        if (need > rbytes) {
            rbytes = need;
            rtmp = pool.get<char>((int64_t)need);
        }
    };
    // topology tables
    DevTopo T;
    std::memset(&T, 0, sizeof(T));
    int off = 0;
    for (int t = 0; t < 8; t++) {
        T.nn[t] = topo::count_nodes(t);
        if (t >= 1) {
            T.ne[t] = topo::count_edges(t);
            T.ntf[t] = topo::count_tri_faces(t);
            T.nqf[t] = topo::count_quad_faces(t);
        }
        for (int s = 0; s < 8; s++) T.cnt[t][s] = s == 0 ? topo::count_nodes(t) : topo::count_type(t, s);
        auto put = [&](const std::vector<int>& v) {
            const int o = off;
            SLX_CHECK(off + (int)v.size() <= (int)(sizeof(T.tab) / sizeof(int)), "internal: topology table overflow");
            for (int x : v) T.tab[off++] = x;
            return o;
        };
        if (t >= 1) {
            T.en_off[t] = put(topo::edge_nodes(t));
            T.fn_off[t] = put(topo::face_nodes(t));
            T.fe_off[t] = put(topo::face_edges(t));
        }
    }
    // device tables sub[t][st] (the mesh's layout), with their row widths
    int* sub[8][4] = {};
    int64_t rows[8] = {};  // element count of every type (t >= 1)
    auto per_of = [&](int t, int s) { return s == 0 ? T.nn[t] : T.cnt[t][s]; };
    for (int t = 1; t < 8; t++) {
        rows[t] = (int64_t)m.sub_[t][0].size() / T.nn[t];
        sub[t][0] = pool.get<int>(std::max<int64_t>(rows[t] * T.nn[t], 1));
        if (rows[t]) MPCHECK(hipMemcpyAsync(sub[t][0], m.sub_[t][0].data(), rows[t] * T.nn[t] * sizeof(int), hipMemcpyHostToDevice, st));
    }
    const int64_t n_nodes = m.count(0);

    // ---- explode: new lines, triangles, quadrangles appended in type order
    {
        int64_t nl = rows[LINE], nt = rows[TRI], nq = rows[QUAD];
        int64_t addl = 0, addt = 0, addq = 0;
        for (int t = 2; t < 8; t++) {
            addl += rows[t] * T.ne[t];
            if (topo::dimension(t) == 3) {
                addt += rows[t] * T.ntf[t];
                addq += rows[t] * T.nqf[t];
            }
        }
        auto grow = [&](int t, int s, int64_t old_rows, int64_t new_rows) {
            const int per = per_of(t, s);
            int* p = pool.get<int>(std::max<int64_t>(new_rows * per, 1));
            if (sub[t][s] && old_rows) MPCHECK(hipMemcpyAsync(p, sub[t][s], old_rows * per * sizeof(int), hipMemcpyDeviceToDevice, st));
            sub[t][s] = p;
        };
        grow(LINE, 0, nl, nl + addl);
        grow(TRI, 0, nt, nt + addt);
        grow(QUAD, 0, nq, nq + addq);
        // line references of triangles and quadrangles: [TRI][LINE] exists for every triangle once exploded
        int* tri_lines = pool.get<int>(std::max<int64_t>((nt + addt) * 3, 1));
        int* quad_lines = pool.get<int>(std::max<int64_t>((nq + addq) * 4, 1));
        int64_t lb = nl, tb = nt, qb = nq;
        for (int t = 2; t < 8; t++) {
            if (!rows[t]) continue;
            sub[t][LINE] = pool.get<int>(rows[t] * T.ne[t]);
            const bool is3d = topo::dimension(t) == 3;
            if (is3d && T.ntf[t]) sub[t][TRI] = pool.get<int>(rows[t] * T.ntf[t]);
            if (is3d && T.nqf[t]) sub[t][QUAD] = pool.get<int>(rows[t] * T.nqf[t]);
            // a 2D type's own lines: the triangle / quadrangle line tables of its rows
            int* own = t == TRI ? tri_lines : t == QUAD ? quad_lines : sub[t][LINE];
            hipLaunchKernelGGL(k_explode, dim3(nblk(rows[t])), dim3(256), 0, st, t, (int)rows[t], T, sub[t][0], (int)lb, (int)tb, (int)qb,
                               sub[LINE][0], own, is3d ? sub[TRI][0] : nullptr, sub[t][TRI], tri_lines, sub[QUAD][0], sub[t][QUAD], quad_lines);
            if (t == TRI || t == QUAD) sub[t][LINE] = own;
            lb += rows[t] * T.ne[t];
            if (is3d) {
                tb += rows[t] * T.ntf[t];
                qb += rows[t] * T.nqf[t];
            }
        }
        rows[LINE] = lb;
        rows[TRI] = tb;
        rows[QUAD] = qb;
        if (rows[TRI]) sub[TRI][LINE] = tri_lines;
        if (rows[QUAD]) sub[QUAD][LINE] = quad_lines;
        SLX_CHECK(lb < INT_MAX && tb < INT_MAX && qb < INT_MAX, "mesh too large for 32-bit element numbers");
        MPCHECK(hipGetLastError());
    }
    auto renumber_refs = [&](int type, const int* renum) {
        // elements::renumber: node, line, triangle and quadrangle references only
        if (type > 3) return;
        for (int t = 1; t < 8; t++)
            if (sub[t][type] && rows[t] && t != type) {
                const int64_t n = rows[t] * per_of(t, type);
                hipLaunchKernelGGL(k_map, dim3(nblk(n)), dim3(256), 0, st, n, sub[t][type], renum);
            }
    };
    std::vector<int> h_renum;
    auto phys_renum = [&](int type, const int* d_renum, int64_t n) {
        h_renum.resize(n);
        if (n) MPCHECK(hipMemcpyAsync(h_renum.data(), d_renum, n * sizeof(int), hipMemcpyDeviceToHost, st));
        MPCHECK(hipStreamSynchronize(st));
        for (auto& pe : m.phys_elems_)
            for (int& e : pe[type]) e = h_renum[e];
    };

    // ---- duplicates of every element type, in type order
    for (int t = 1; t < 8; t++) {
        const int n = (int)rows[t];
        if (n == 0) continue;
        const int nn = T.nn[t];
        int* key = pool.get<int>((int64_t)n * nn);
        int* perm = pool.get<int>(n);
        int* perm2 = pool.get<int>(n);
        unsigned* w = pool.get<unsigned>(n);
        unsigned* w2 = pool.get<unsigned>(n);
        hipLaunchKernelGGL(k_sorted_keys, dim3(nblk(n)), dim3(256), 0, st, n, nn, sub[t][0], key, perm);
        const int eb = bits_for(n_nodes);
        for (int p = nn - 1; p >= 0; p--) {
            hipLaunchKernelGGL(k_gather_word, dim3(nblk(n)), dim3(256), 0, st, n, nn, p, key, perm, w);
            size_t need = 0;
            MPCHECK(rocprim::radix_sort_pairs(nullptr, need, w, w2, perm, perm2, (size_t)n, 0, eb, st));
            rgrow(need);
            MPCHECK(rocprim::radix_sort_pairs(rtmp, need, w, w2, perm, perm2, (size_t)n, 0, eb, st));
            std::swap(perm, perm2);
        }
        int* head = pool.get<int>(n);
        int* gidp1 = pool.get<int>(n);
        hipLaunchKernelGGL(k_heads, dim3(nblk(n)), dim3(256), 0, st, n, nn, key, perm, head);
        size_t need = 0;
        MPCHECK(rocprim::inclusive_scan(nullptr, need, head, gidp1, (size_t)n, rocprim::plus<int>(), st));
        rgrow(need);
        MPCHECK(rocprim::inclusive_scan(rtmp, need, head, gidp1, (size_t)n, rocprim::plus<int>(), st));
        int ng = 0;
        MPCHECK(hipMemcpyAsync(&ng, gidp1 + n - 1, sizeof(int), hipMemcpyDeviceToHost, st));
        int* first = pool.get<int>(n);
        int* last = pool.get<int>(n);
        int* g_of = pool.get<int>(n);
        int* isfirst = pool.get<int>(n);
        int* num = pool.get<int>(n);
        int* renum = pool.get<int>(n);
        hipLaunchKernelGGL(k_groups, dim3(nblk(n)), dim3(256), 0, st, n, perm, head, gidp1, first, last, g_of, isfirst);
        need = 0;
        MPCHECK(rocprim::exclusive_scan(nullptr, need, isfirst, num, 0, (size_t)n, rocprim::plus<int>(), st));
        rgrow(need);
        MPCHECK(rocprim::exclusive_scan(rtmp, need, isfirst, num, 0, (size_t)n, rocprim::plus<int>(), st));
        hipLaunchKernelGGL(k_renum, dim3(nblk(n)), dim3(256), 0, st, n, g_of, first, num, renum);
        MPCHECK(hipStreamSynchronize(st));
        for (int s = 0; s < 4; s++) {
            if (!sub[t][s] || (s != 0 && s == t)) continue;
            const int per = per_of(t, s);
            if (per == 0) continue;
            int* dst = pool.get<int>((int64_t)ng * per);
            hipLaunchKernelGGL(k_copy_rows, dim3(nblk((int64_t)ng * per)), dim3(256), 0, st, ng, per, first, last, num, sub[t][s], dst);
            sub[t][s] = dst;
        }
        rows[t] = ng;
        renumber_refs(t, renum);
        phys_renum(t, renum, n);
        MPCHECK(hipGetLastError());
    }
    for (auto& pe : m.phys_elems_)
        for (int t = 0; t <= 7; t++) {
            std::sort(pe[t].begin(), pe[t].end());
            pe[t].erase(std::unique(pe[t].begin(), pe[t].end()), pe[t].end());
        }

    // ---- disjoint regions: membership masks, numbered by first occurrence per type
    const int np = (int)m.phys_number_.size();
    SLX_CHECK(np <= 16, "internal: device mesh pipeline takes at most 16 physical regions");
    auto count_of = [&](int t) { return t == 0 ? n_nodes : rows[t]; };
    unsigned* mask[8] = {};
    for (int s = 0; s < 8; s++) {
        mask[s] = pool.get<unsigned>(count_of(s));
        MPCHECK(hipMemsetAsync(mask[s], 0, std::max<int64_t>(count_of(s), 1) * sizeof(unsigned), st));
    }
    for (int pi = 0; pi < np; pi++)
        for (int t = 0; t <= 7; t++) {
            const std::vector<int>& el = m.phys_elems_[pi][t];
            if (el.empty()) continue;
            int* d_el = pool.get<int>((int64_t)el.size());
            MPCHECK(hipMemcpyAsync(d_el, el.data(), el.size() * sizeof(int), hipMemcpyHostToDevice, st));
            for (int s = 0; s <= 7; s++) {
                const int nsub = T.cnt[t][s];
                if (nsub <= 0) continue;
                if (s != t && (s > 3 || !sub[t][s])) continue;
                const int64_t x = (int64_t)el.size() * nsub;
                hipLaunchKernelGGL(k_mark, dim3(nblk(x)), dim3(256), 0, st, (int)el.size(), d_el, t, s, nsub, s == t ? nullptr : sub[t][s],
                                   1u << pi, mask[s]);
            }
            MPCHECK(hipStreamSynchronize(st));  // d_el is host-staged per list
        }
    const int nm = 1 << np;
    std::vector<int> table((size_t)8 * nm, -1);
    m.dr_type_.clear();
    m.dr_def_.clear();
    unsigned* dr[8] = {};
    {
        int* first = pool.get<int>((int64_t)8 * nm);
        std::vector<int> big((size_t)8 * nm, INT_MAX), hfirst((size_t)8 * nm);
        MPCHECK(hipMemcpyAsync(first, big.data(), big.size() * sizeof(int), hipMemcpyHostToDevice, st));
        for (int t = 0; t <= 7; t++)
            if (count_of(t)) hipLaunchKernelGGL(k_first_of_mask, dim3(nblk(count_of(t))), dim3(256), 0, st, (int)count_of(t), mask[t], first + (size_t)t * nm);
        MPCHECK(hipMemcpyAsync(hfirst.data(), first, hfirst.size() * sizeof(int), hipMemcpyDeviceToHost, st));
        MPCHECK(hipStreamSynchronize(st));
        for (int t = 0; t <= 7; t++) {
            std::vector<std::pair<int, int>> seen;  // (first element, mask)
            for (int k = 0; k < nm; k++)
                if (hfirst[(size_t)t * nm + k] != INT_MAX) seen.push_back({hfirst[(size_t)t * nm + k], k});
            std::sort(seen.begin(), seen.end());
            for (auto& sm : seen) {
                table[(size_t)t * nm + sm.second] = (int)m.dr_type_.size();
                m.dr_type_.push_back(t);
                std::vector<char> key(np);
                for (int i = 0; i < np; i++) key[i] = (sm.second >> i) & 1;
                m.dr_def_.push_back(key);
            }
        }
        int* d_table = pool.get<int>((int64_t)8 * nm);
        MPCHECK(hipMemcpyAsync(d_table, table.data(), table.size() * sizeof(int), hipMemcpyHostToDevice, st));
        for (int t = 0; t <= 7; t++) {
            dr[t] = pool.get<unsigned>(count_of(t));
            if (count_of(t)) hipLaunchKernelGGL(k_dr_of, dim3(nblk(count_of(t))), dim3(256), 0, st, (int)count_of(t), mask[t], d_table + (size_t)t * nm, dr[t]);
        }
    }

    // ---- reorder every type by disjoint region (stable), renumber references
    double* coords = pool.get<double>(3 * n_nodes);
    if (n_nodes) MPCHECK(hipMemcpyAsync(coords, m.coords_.data(), 3 * n_nodes * sizeof(double), hipMemcpyHostToDevice, st));
    const int drb = bits_for((int64_t)m.dr_type_.size());
    for (int t = 0; t <= 7; t++) {
        const int n = (int)count_of(t);
        if (n == 0) continue;
        int* io = pool.get<int>(n);
        int* ro = pool.get<int>(n);
        unsigned* drs = pool.get<unsigned>(n);
        int* renum = pool.get<int>(n);
        hipLaunchKernelGGL(k_iota, dim3(nblk(n)), dim3(256), 0, st, n, io);
        size_t need = 0;
        MPCHECK(rocprim::radix_sort_pairs(nullptr, need, dr[t], drs, io, ro, (size_t)n, 0, drb, st));
        rgrow(need);
        MPCHECK(rocprim::radix_sort_pairs(rtmp, need, dr[t], drs, io, ro, (size_t)n, 0, drb, st));
        dr[t] = drs;
        hipLaunchKernelGGL(k_invert, dim3(nblk(n)), dim3(256), 0, st, n, ro, renum);
        if (t == 0) {
            double* c2 = pool.get<double>(3 * (int64_t)n);
            hipLaunchKernelGGL(k_gather_rows<double>, dim3(nblk(3 * (int64_t)n)), dim3(256), 0, st, n, 3, ro, coords, c2);
            coords = c2;
        } else {
            for (int s = 0; s < 4; s++) {
                if (!sub[t][s] || (s != 0 && s == t)) continue;
                const int per = per_of(t, s);
                if (per == 0) continue;
                int* dst = pool.get<int>((int64_t)n * per);
                hipLaunchKernelGGL(k_gather_rows<int>, dim3(nblk((int64_t)n * per)), dim3(256), 0, st, n, per, ro, sub[t][s], dst);
                sub[t][s] = dst;
            }
        }
        renumber_refs(t, renum);
        phys_renum(t, renum, n);
        MPCHECK(hipGetLastError());
    }

    // ---- orientation
    int* orient[8] = {};
    for (int t = 1; t <= 7; t++) {
        orient[t] = pool.get<int>(rows[t]);
        if (rows[t]) hipLaunchKernelGGL(k_orient, dim3(nblk(rows[t])), dim3(256), 0, st, (int)rows[t], t, T, sub[t][0], orient[t]);
    }
    MPCHECK(hipGetLastError());

    // ---- download into the mesh
    m.coords_.resize(3 * (size_t)n_nodes);
    if (n_nodes) MPCHECK(hipMemcpyAsync(m.coords_.data(), coords, 3 * n_nodes * sizeof(double), hipMemcpyDeviceToHost, st));
    for (int t = 0; t <= 7; t++) {
        const int64_t n = count_of(t);
        m.indisjreg_[t].resize(n);
        if (n) MPCHECK(hipMemcpyAsync(m.indisjreg_[t].data(), dr[t], n * sizeof(int), hipMemcpyDeviceToHost, st));
        m.totalorient_[t].assign(n, 0);
        if (t >= 1) {
            if (n) MPCHECK(hipMemcpyAsync(m.totalorient_[t].data(), orient[t], n * sizeof(int), hipMemcpyDeviceToHost, st));
            for (int s = 0; s < 4; s++) {
                if (!sub[t][s] || (s != 0 && s == t)) {
                    if (s != 0) m.sub_[t][s].clear();
                    continue;
                }
                const int per = per_of(t, s);
                m.sub_[t][s].resize((size_t)n * per);
                if (n * per) MPCHECK(hipMemcpyAsync(m.sub_[t][s].data(), sub[t][s], n * per * sizeof(int), hipMemcpyDeviceToHost, st));
            }
        }
    }
    MPCHECK(hipStreamSynchronize(st));
}

}  // namespace slx
