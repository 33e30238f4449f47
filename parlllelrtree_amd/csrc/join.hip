// join.hip -- MBR-overlap spatial join of query rectangles against a BU or
// TD tree, producing every (q, r) with overlap(q, rect r), sorted by (q, r).
//
// Reference: the GPU query of both files only advances a per-query BFS
// frontier, one launch per level, with an atomicAdd push into an in-place
// 1000-slot queue and no hit output (query1/init, BotomUp-rtree.cu:592-630,
// TopDown-rtree.cu:355-377, 450-463).  The hit semantics are those of the CPU
// search (RtreeSearch, BotomUp-rtree.cu:439-483, TopDown-rtree.cu:56-76):
// closed-interval overlap, descend into overlapping internal entries, every
// overlapping leaf entry is a hit.
// Here: one launch traverses the whole tree depth first.  A query owns G =
// next_pow2(fanout) lanes of a wave64 (4 queries per wave at fanout 16); lane
// j tests entry j of the node on top of the query's stack (kept in LDS), a
// ballot gives the overlapping entries, internal ones are pushed in reverse
// order and leaf ones are counted (pass 1) or written at
// offset[q] + running count + rank (pass 2).  An exclusive scan of the
// counts gives the offsets; a final kernel sorts each query's hits by r.
// Pass 1 also keeps the first JN_KEEP hits of every query: pass 2 copies
// them for a query with no more hits than that instead of traversing again
// (a layer join has one or two hits per query).
#include <mutex>

#include "prt_common.hpp"
#include "prt_profile.hpp"

namespace prt {
namespace {

constexpr int JN_THREADS = 256;
constexpr int JN_WAVES = JN_THREADS / 64;
constexpr int JN_KEEP = 8;  // hits per query kept by the count pass

struct TreeView {
    int mode;            // 0 = BU, 1 = TD
    int f, G;            // fanout, lanes per query
    int64_t root;        // BU: root node id; TD: root block slot (0)
    int64_t num_leaves;  // BU only
    const int32_t *cnt;  // BU node_count
    const float4 *mbr;   // BU entry_mbr / TD slot_mbr
    const int32_t *ref;  // BU entry_ref / TD slot_child
    const uint32_t *perm;  // TD perm
};

template <bool WRITE>
__global__ __launch_bounds__(JN_THREADS) void k_join(TreeView tv, const float4 *__restrict__ qs,
                                                     int64_t nq, int stack_cap,
                                                     int64_t *__restrict__ counts,
                                                     const int64_t *__restrict__ offsets,
                                                     uint32_t *__restrict__ pairs, int64_t cap,
                                                     uint32_t *__restrict__ kept) {
    extern __shared__ int32_t stk_all[];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int G = tv.G, QW = 64 / G;
    const int g = lane / G, l = lane % G;
    const int64_t q = (static_cast<int64_t>(blockIdx.x) * JN_WAVES + wave) * QW + g;
    const bool valid = q < nq;
    int32_t *stk = stk_all + (wave * QW + g) * stack_cap;
    const uint64_t gmask = (G == 64) ? ~0ull : ((1ull << G) - 1ull);
    const uint64_t below = (l == 0) ? 0ull : ((1ull << l) - 1ull);

    float4 qb = valid ? qs[q] : empty_box();
    int top = 0;
    const int64_t base = (WRITE && valid) ? offsets[q] : 0;
    // write pass: a query whose hits the count pass kept copies them
    const int64_t cq = (WRITE && valid) ? counts[q] : 0;
    const bool copied = WRITE && valid && cq <= JN_KEEP;
    if (copied) {
        for (int i = l; i < cq; i += G) {
            const int64_t o = base + i;
            if (o < cap) {
                pairs[2 * o] = static_cast<uint32_t>(q);
                pairs[2 * o + 1] = kept[q * JN_KEEP + i];
            }
        }
    }
    if (valid && !copied) {
        if (l == 0) stk[0] = static_cast<int32_t>(tv.root);
        top = 1;
    }
    __builtin_amdgcn_wave_barrier();
    int64_t found = 0;

    while (__any(top > 0)) {
        int64_t node = -1;
        if (top > 0) {
            node = stk[top - 1];
            --top;
        }
        __builtin_amdgcn_wave_barrier();
        bool hit = false;
        int32_t ref = 0;
        if (node >= 0 && l < tv.f) {
            if (tv.mode == 0) {
                if (l < tv.cnt[node]) {
                    const int64_t e = node * tv.f + l;
                    hit = overlap(qb, tv.mbr[e]);
                    ref = tv.ref[e];
                }
            } else {
                const int64_t sl = node + l;
                ref = tv.ref[sl];
                if (ref != 0) hit = overlap(tv.mbr[sl], qb);
            }
        }
        const bool leaf = (tv.mode == 0) ? (node >= 0 && node < tv.num_leaves) : (ref < 0);
        const uint64_t inner = (__ballot(hit && !leaf) >> (g * G)) & gmask;
        const uint64_t leafm = (__ballot(hit && leaf) >> (g * G)) & gmask;
        const int ninner = __popcll(inner);
        if (hit && !leaf) {
            const int r = __popcll(inner & below);
            stk[top + (ninner - 1 - r)] = ref;  // lowest entry ends on top
        }
        top += ninner;
        if (WRITE && hit && leaf) {
            const int64_t o = base + found + __popcll(leafm & below);
            if (o < cap) {
                const uint32_t rid = (tv.mode == 0) ? static_cast<uint32_t>(ref)
                                                    : tv.perm[-static_cast<int64_t>(ref) - 1];
                pairs[2 * o] = static_cast<uint32_t>(q);
                pairs[2 * o + 1] = rid;
            }
        }
        if (!WRITE && hit && leaf) {
            const int64_t k = found + __popcll(leafm & below);
            if (k < JN_KEEP)
                kept[q * JN_KEEP + k] = (tv.mode == 0) ? static_cast<uint32_t>(ref)
                                                       : tv.perm[-static_cast<int64_t>(ref) - 1];
        }
        found += __popcll(leafm);
        __builtin_amdgcn_wave_barrier();
    }
    if (!WRITE && valid && l == 0) counts[q] = found;
}

// Sort the r's of every query segment: short segments by one thread, up to
// 1024 hits by one wave (bitonic in LDS), longer ones by one block (up to
// 8192 in LDS, beyond in place in the pair buffer).
constexpr int SS_WAVE_MAX = 1024;
constexpr int SS_BLOCK_MAX = 8192;

__device__ inline void bitonic_lds_u32(uint32_t *v, int n2, int tid, int nthreads, bool block) {
    for (int k = 2; k <= n2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int p = tid; p < n2 / 2; p += nthreads) {
                int i = ((p & ~(j - 1)) << 1) | (p & (j - 1));
                int m = i | j;
                bool up = (i & k) == 0;
                uint32_t x = v[i], y = v[m];
                if ((x > y) == up) {
                    v[i] = y;
                    v[m] = x;
                }
            }
            if (block) __syncthreads();
            else __builtin_amdgcn_wave_barrier();
        }
    }
}

constexpr int SS_THREAD_MAX = 16;  // segments sorted by one thread

// Per query: segments of up to SS_THREAD_MAX hits are sorted here by one
// thread (insertion sort, in place); longer ones are appended to the list
// of the wave sort (<= 1024) or of the block sort.
// lists[0] / lists[nq]: the two lists; lcount[0..1]: their lengths.
__global__ __launch_bounds__(JN_THREADS) void k_seg_classify(const int64_t *__restrict__ counts,
                                                             const int64_t *__restrict__ offsets,
                                                             int64_t nq, int64_t cap,
                                                             uint32_t *__restrict__ pairs,
                                                             uint32_t *__restrict__ lists,
                                                             unsigned long long *lcount) {
    const int64_t q = static_cast<int64_t>(blockIdx.x) * JN_THREADS + threadIdx.x;
    if (q >= nq) return;
    const int64_t o = offsets[q];
    int64_t len = counts[q];
    if (o + len > cap) len = cap - o;
    if (len <= 1) return;
    if (len <= SS_THREAD_MAX) {
        uint32_t *r = pairs + 2 * o + 1;  // r of hit i at r[2 * i]
        for (int i = 1; i < len; ++i) {
            const uint32_t v = r[2 * i];
            int j = i;
            while (j > 0 && r[2 * (j - 1)] > v) {
                r[2 * j] = r[2 * (j - 1)];
                --j;
            }
            r[2 * j] = v;
        }
    } else if (len <= SS_WAVE_MAX) {
        lists[atomicAdd(&lcount[0], 1ull)] = static_cast<uint32_t>(q);
    } else {
        lists[nq + atomicAdd(&lcount[1], 1ull)] = static_cast<uint32_t>(q);
    }
}

// One wave per listed query (grid-stride over the list).
__global__ __launch_bounds__(JN_THREADS) void k_seg_sort_wave(const int64_t *__restrict__ counts,
                                                              const int64_t *__restrict__ offsets,
                                                              int64_t cap,
                                                              const uint32_t *__restrict__ list,
                                                              const unsigned long long *lcount,
                                                              uint32_t *__restrict__ pairs) {
    __shared__ uint32_t buf[JN_WAVES][SS_WAVE_MAX];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t nl = static_cast<int64_t>(*lcount);
    for (int64_t x = static_cast<int64_t>(blockIdx.x) * JN_WAVES + wave; x < nl;
         x += static_cast<int64_t>(gridDim.x) * JN_WAVES) {
        const int64_t q = list[x];
        const int64_t o = offsets[q];
        int64_t len = counts[q];
        if (o + len > cap) len = cap - o;
        int n2 = 1;
        while (n2 < len) n2 <<= 1;
        uint32_t *v = buf[wave];
        for (int i = lane; i < n2; i += 64) v[i] = i < len ? pairs[2 * (o + i) + 1] : 0xFFFFFFFFu;
        __builtin_amdgcn_wave_barrier();
        bitonic_lds_u32(v, n2, lane, 64, false);
        for (int i = lane; i < len; i += 64) pairs[2 * (o + i) + 1] = v[i];
        __builtin_amdgcn_wave_barrier();
    }
}

// A segment of more than SS_BLOCK_MAX hits, sorted in place by one block:
// r[2 * i] for i < len, positions from len to the next power of two are
// +inf and never stored.  Bitonic sort in its ascending form (the first
// compare of each merge pairs i with its mirror i ^ (k - 1), then
// half-cleaners i, i ^ j): every compare puts the smaller value at the
// lower position, so a virtual +inf never moves.  Chunks of SS_BLOCK_MAX
// are sorted in LDS, merge steps with j >= SS_BLOCK_MAX run over the pair
// buffer, the rest of each merge chunk by chunk in LDS.
__device__ void sort_big_segment(uint32_t *r, int64_t len, uint32_t *v) {
    constexpr int C = SS_BLOCK_MAX;
    int64_t n2 = C;
    while (n2 < len) n2 <<= 1;
    const int tid = threadIdx.x, nt = blockDim.x;
    auto load = [&](int64_t c0) {
        for (int i = tid; i < C; i += nt) v[i] = c0 + i < len ? r[2 * (c0 + i)] : 0xFFFFFFFFu;
        __syncthreads();
    };
    auto store = [&](int64_t c0) {
        for (int i = tid; i < C && c0 + i < len; i += nt) r[2 * (c0 + i)] = v[i];
        __syncthreads();
    };
    for (int64_t c0 = 0; c0 < len; c0 += C) {  // every chunk ascending
        load(c0);
        bitonic_lds_u32(v, C, tid, nt, true);
        store(c0);
    }
    for (int64_t k = 2 * C; k <= n2; k <<= 1) {
        for (int64_t j = k >> 1; j >= C; j >>= 1) {  // over the pair buffer
            for (int64_t p = tid; p < n2 / 2; p += nt) {
                const int64_t i = ((p & ~(j - 1)) << 1) | (p & (j - 1));
                const int64_t m = j == (k >> 1) ? (i ^ (k - 1)) : (i | j);
                if (i >= len || m >= len) continue;  // m >= len: +inf stays above
                const uint32_t a = r[2 * i], b = r[2 * m];
                if (a > b) {
                    r[2 * i] = b;
                    r[2 * m] = a;
                }
            }
            __syncthreads();
        }
        for (int64_t c0 = 0; c0 < len; c0 += C) {  // half-cleaners j < C
            load(c0);
            for (int j = C >> 1; j > 0; j >>= 1) {
                for (int p = tid; p < C / 2; p += nt) {
                    const int i = ((p & ~(j - 1)) << 1) | (p & (j - 1));
                    const int m = i | j;
                    const uint32_t a = v[i], b = v[m];
                    if (a > b) {
                        v[i] = b;
                        v[m] = a;
                    }
                }
                __syncthreads();
            }
            store(c0);
        }
    }
}

// One block per listed query (grid-stride over the list).
__global__ __launch_bounds__(1024) void k_seg_sort_block(const int64_t *__restrict__ counts,
                                                         const int64_t *__restrict__ offsets,
                                                         int64_t cap,
                                                         const uint32_t *__restrict__ list,
                                                         const unsigned long long *lcount,
                                                         uint32_t *__restrict__ pairs) {
    __shared__ uint32_t v[SS_BLOCK_MAX];
    const int64_t nl = static_cast<int64_t>(*lcount);
    for (int64_t x = blockIdx.x; x < nl; x += gridDim.x) {
        const int64_t q = list[x];
        const int64_t o = offsets[q];
        int64_t len = counts[q];
        if (o + len > cap) len = cap - o;
        if (len > SS_BLOCK_MAX) {
            sort_big_segment(pairs + 2 * o + 1, len, v);
            continue;
        }
        int n2 = 1;
        while (n2 < len) n2 <<= 1;
        for (int i = threadIdx.x; i < n2; i += blockDim.x)
            v[i] = i < len ? pairs[2 * (o + i) + 1] : 0xFFFFFFFFu;
        __syncthreads();
        bitonic_lds_u32(v, n2, threadIdx.x, blockDim.x, true);
        for (int i = threadIdx.x; i < len; i += blockDim.x) pairs[2 * (o + i) + 1] = v[i];
        __syncthreads();
    }
}

__global__ void k_zero_i64(int64_t *p, int n) {
    if (threadIdx.x < n) p[threadIdx.x] = 0;
}

int run_join(const TreeView &tv, int levels, const float *queries, int64_t nq, uint32_t *pairs,
             int64_t cap, int64_t *count_d, void *ws, size_t ws_bytes, hipStream_t s) {
    PRT_REQUIRE(nq >= 0 && nq < (int64_t(1) << 32), PRT_E_INVAL, "join: nq out of range");
    PRT_REQUIRE(count_d != nullptr && (nq == 0 || queries), PRT_E_INVAL, "join: null pointer");
    PRT_REQUIRE(cap >= 0 && (cap == 0 || pairs), PRT_E_INVAL, "join: bad pair buffer");
    PRT_REQUIRE(ws_bytes >= prt_join_workspace_size(nq), PRT_E_NOMEM, "join: workspace too small");
    Carver c(ws, ws_bytes);
    int64_t *counts = c.take<int64_t>(nq + 1);
    int64_t *offs = c.take<int64_t>(nq + 1);
    void *scan_ws = c.take<char>(scan_i64_ws_bytes(nq));
    uint32_t *lists = c.take<uint32_t>(2 * (nq > 0 ? nq : 1));
    int64_t *lcount = c.take<int64_t>(2);
    uint32_t *kept = c.take<uint32_t>(JN_KEEP * (nq > 0 ? nq : 1));
    PRT_REQUIRE(c.ok(), PRT_E_NOMEM, "join: workspace too small");
    PRT_LAUNCH(k_zero_i64, 1, 64, 0, s, count_d, 2);
    PRT_LAUNCHED();
    if (nq == 0) return PRT_OK;

    const int QW = 64 / tv.G;
    const int stack_cap = levels * tv.f + 1;
    const size_t lds = static_cast<size_t>(JN_WAVES) * QW * stack_cap * sizeof(int32_t);
    PRT_REQUIRE(lds <= 64 * 1024, PRT_E_UNSUPPORTED, "join: traversal stack too large");
    const unsigned grid = static_cast<unsigned>(ceil_div(nq, JN_WAVES * QW));
    const float4 *q4 = reinterpret_cast<const float4 *>(queries);
    PRT_LAUNCH(k_join<false>, grid, JN_THREADS, lds, s, tv, q4, nq, stack_cap, counts, nullptr,
               nullptr, 0, kept);
    PRT_LAUNCHED();
    int rc = scan_i64(counts, offs, nq, count_d, scan_ws, s);
    if (rc) return rc;
    if (cap == 0) return PRT_OK;
    PRT_LAUNCH(k_join<true>, grid, JN_THREADS, lds, s, tv, q4, nq, stack_cap, counts, offs, pairs,
               cap, kept);
    PRT_LAUNCHED();
    PRT_LAUNCH(k_zero_i64, 1, 64, 0, s, lcount, 2);
    PRT_LAUNCHED();
    unsigned long long *lc = reinterpret_cast<unsigned long long *>(lcount);
    PRT_LAUNCH(k_seg_classify, dim3(static_cast<unsigned>(ceil_div(nq, JN_THREADS))), JN_THREADS, 0,
               s, counts, offs, nq, cap, pairs, lists, lc);
    PRT_LAUNCHED();
    int64_t wgrid = ceil_div(nq, JN_WAVES);
    if (wgrid > 65536) wgrid = 65536;  // one listed query per wave
    PRT_LAUNCH(k_seg_sort_wave, dim3(static_cast<unsigned>(wgrid)), JN_THREADS, 0, s, counts, offs,
               cap, lists, lc, pairs);
    PRT_LAUNCHED();
    PRT_LAUNCH(k_seg_sort_block, 256, 1024, 0, s, counts, offs, cap, lists + nq, lc + 1, pairs);
    PRT_LAUNCHED();
    return PRT_OK;
}

int lanes_for(int f) {
    int G = 1;
    while (G < f) G <<= 1;
    return G;
}

}  // namespace
}  // namespace prt

using namespace prt;

extern "C" size_t prt_join_workspace_size(int64_t nq) {
    Sizer s;
    s.take<int64_t>(nq + 1);
    s.take<int64_t>(nq + 1);
    s.take<char>(scan_i64_ws_bytes(nq));
    s.take<uint32_t>(2 * (nq > 0 ? nq : 1));
    s.take<int64_t>(2);
    s.take<uint32_t>(JN_KEEP * (nq > 0 ? nq : 1));
    return s.bytes();
}

extern "C" int prt_join_bu(const prt_bu_tree *tree, const float *queries, int64_t nq,
                           uint32_t *pairs, int64_t cap, int64_t *count_d, void *ws,
                           size_t ws_bytes, void *stream) {
    PRT_REQUIRE(tree && tree->num_nodes > 0 && tree->node_count && tree->entry_mbr && tree->entry_ref,
                PRT_E_INVAL, "join_bu: tree not built");
    PRT_REQUIRE(tree->fanout >= 2 && tree->fanout <= PRT_MAX_FANOUT, PRT_E_INVAL, "join_bu: fanout");
    TreeView tv;
    tv.mode = 0;
    tv.f = tree->fanout;
    tv.G = lanes_for(tree->fanout);
    tv.root = tree->num_nodes - 1;
    tv.num_leaves = tree->level_count[0];
    tv.cnt = tree->node_count;
    tv.mbr = reinterpret_cast<const float4 *>(tree->entry_mbr);
    tv.ref = tree->entry_ref;
    tv.perm = nullptr;
    return run_join(tv, tree->num_levels, queries, nq, pairs, cap, count_d, ws, ws_bytes,
                    as_stream(stream));
}

extern "C" int prt_join_td(const prt_td_tree *tree, const float *queries, int64_t nq,
                           uint32_t *pairs, int64_t cap, int64_t *count_d, void *ws,
                           size_t ws_bytes, void *stream) {
    PRT_REQUIRE(tree && tree->num_slots > 0 && tree->slot_child && tree->slot_mbr && tree->perm,
                PRT_E_INVAL, "join_td: tree not built");
    PRT_REQUIRE(tree->fanout >= 2 && tree->fanout <= PRT_MAX_FANOUT, PRT_E_INVAL, "join_td: fanout");
    TreeView tv;
    tv.mode = 1;
    tv.f = tree->fanout;
    tv.G = lanes_for(tree->fanout);
    tv.root = 0;
    tv.num_leaves = 0;
    tv.cnt = nullptr;
    tv.mbr = reinterpret_cast<const float4 *>(tree->slot_mbr);
    tv.ref = tree->slot_child;
    tv.perm = tree->perm;
    return run_join(tv, tree->num_levels, queries, nq, pairs, cap, count_d, ws, ws_bytes,
                    as_stream(stream));
}

// ---- the join with no allocation: workspace in a file-scope device array --
// For callers that cannot allocate (graph capture, serving loops): the
// workspace of prt_join_bu for up to JN_NOALLOC_MAX_QUERIES queries lives in
// `scratch`, one copy per device, held until the process ends.  Host code
// never passes the array's name to a kernel (on the host it is a stand-in,
// not a device address): its address is resolved once per device with
// hipGetSymbolAddress, with that device current, and only that is used.
namespace prt {
namespace {
constexpr int64_t JN_NOALLOC_MAX_QUERIES = 262144;
constexpr size_t JN_SCRATCH_BYTES = 15u << 20;  // >= prt_join_workspace_size(262144)
constexpr int JN_MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[JN_SCRATCH_BYTES];

void *g_scratch_ptr[JN_MAX_DEVICES] = {};
std::mutex g_scratch_mu;

int scratch_for_device(void **out) {
    int dev = 0;
    PRT_HIP(hipGetDevice(&dev));
    PRT_REQUIRE(dev >= 0 && dev < JN_MAX_DEVICES, PRT_E_UNSUPPORTED,
                "join_bu_noalloc: device %d beyond %d", dev, JN_MAX_DEVICES);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        PRT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        PRT_REQUIRE(p != nullptr, PRT_E_HIP, "join_bu_noalloc: no scratch address");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return PRT_OK;
}
}  // namespace
}  // namespace prt

extern "C" int prt_join_bu_noalloc(const prt_bu_tree *tree, const float *queries, int64_t nq,
                                   uint32_t *pairs, int64_t cap, int64_t *count_d, void *stream) {
    PRT_REQUIRE(nq >= 0 && nq <= JN_NOALLOC_MAX_QUERIES, PRT_E_UNSUPPORTED,
                "join_bu_noalloc: %lld queries, at most %lld per call (the scratch holds the "
                "workspace of that many)",
                (long long)nq, (long long)JN_NOALLOC_MAX_QUERIES);
    const size_t need = prt_join_workspace_size(nq);
    PRT_REQUIRE(need <= JN_SCRATCH_BYTES, PRT_E_UNSUPPORTED,
                "join_bu_noalloc: workspace %zu > scratch %zu bytes", need, JN_SCRATCH_BYTES);
    void *ws = nullptr;
    int rc = scratch_for_device(&ws);
    if (rc) return rc;
    return prt_join_bu(tree, queries, nq, pairs, cap, count_d, ws, JN_SCRATCH_BYTES, stream);
}
