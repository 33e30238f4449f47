// bu_build.hip -- Morton keys and the bottom-up (packed) R-tree build.
//
// Reference: BotomUp-rtree.cu.  The reference computes a 64-bit 4-D key
// (CalcuateZOrderValue/Interleave, :73-123) in ONE block, sorts 112-byte AoS
// records with sixteen 4-bit LSD rounds (:841-849), then packs levels with
// one thread per node (Construction_init/Construction1, :373-426).
// Here: one fused kernel writes (key << 32 | id) items, the 8-bit radix sort
// of radix.hip sorts them in 4 passes, and each level is packed with one
// lane per entry slot (SoA outputs, coalesced), a node's union computed
// in LDS by its f lanes and handed up as the parent's entry.
#include "prt_common.hpp"
#include "prt_profile.hpp"

namespace prt {
namespace {

constexpr int PK_THREADS = 256;

__global__ void k_morton_keys(const float4 *__restrict__ rects, int64_t n,
                              uint32_t *__restrict__ keys) {
    int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) keys[i] = morton_center30(rects[i]);
}

// Reference 4-D key: input bit k (k < 16) of coordinate w (x1, y1, x2, y2)
// goes to key bit 8*(k/2) + (6 - 2w) + (k & 1)  (BotomUp-rtree.cu:98-123).
__device__ inline uint64_t spread_ref4(uint64_t a) {
    uint64_t out = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) out |= ((a >> k) & 1ull) << (8 * (k >> 1) + (k & 1));
    return out;
}

__global__ void k_morton_ref4d(const int64_t *__restrict__ rects, int64_t n, int64_t scale,
                               uint64_t *__restrict__ keys) {
    int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t k = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        uint64_t a = static_cast<uint64_t>(rects[4 * i + w] / scale);
        k |= spread_ref4(a) << (6 - 2 * w);
    }
    keys[i] = k;
}

// Leaf level: lane (node g, slot j).  Writes the sorted keys / leaf order,
// entry boxes and ids, node counts and ranges, and the node's union.
__global__ __launch_bounds__(PK_THREADS) void k_bu_leaves(
    const uint64_t *__restrict__ sorted, const float4 *__restrict__ rects, int64_t n, int f,
    int nodes_per_block, int64_t num_leaves, uint32_t *__restrict__ keys,
    uint32_t *__restrict__ perm, int32_t *__restrict__ node_count,
    int32_t *__restrict__ node_start, int32_t *__restrict__ node_end,
    float4 *__restrict__ entry_mbr, int32_t *__restrict__ entry_ref,
    float4 *__restrict__ node_mbr) {
    __shared__ float4 box[PK_THREADS];
    const int t = threadIdx.x, g = t / f, j = t % f;
    const int64_t node = static_cast<int64_t>(blockIdx.x) * nodes_per_block + g;
    const bool live = g < nodes_per_block && node < num_leaves;
    if (live) {
        const int64_t pos = node * f + j;
        float4 m = empty_box();
        int32_t ref = -1;
        if (pos < n) {
            uint64_t it = sorted[pos];
            uint32_t id = static_cast<uint32_t>(it);
            keys[pos] = static_cast<uint32_t>(it >> 32);
            perm[pos] = id;
            m = rects[id];
            ref = static_cast<int32_t>(id);
        }
        entry_mbr[pos] = m;
        entry_ref[pos] = ref;
        box[t] = m;
        if (j == 0) {
            int64_t cnt = n - node * f < f ? n - node * f : f;
            node_count[node] = static_cast<int32_t>(cnt);
            node_start[node] = static_cast<int32_t>(node * f);
            node_end[node] = static_cast<int32_t>(node * f + cnt - 1);
        }
    }
    __syncthreads();
    if (live && j == 0) {
        float4 u = box[t];
        for (int k = 1; k < f; ++k) u = box_union(u, box[t + k]);
        node_mbr[node] = u;
    }
}

// One upper level: entry j of node i is child c = first_child(i) + j, its box
// is the child's union (= the union of all the child's entries, BU:816-826).
__global__ __launch_bounds__(PK_THREADS) void k_bu_upper(
    int f, int nodes_per_block, int64_t ls_prev, int64_t lc_prev, int64_t ls, int64_t lc,
    int32_t *__restrict__ node_count, int32_t *__restrict__ node_start,
    int32_t *__restrict__ node_end, float4 *__restrict__ entry_mbr,
    int32_t *__restrict__ entry_ref, float4 *__restrict__ node_mbr) {
    __shared__ float4 box[PK_THREADS];
    const int t = threadIdx.x, g = t / f, j = t % f;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * nodes_per_block + g;
    const bool live = g < nodes_per_block && i < lc;
    const int64_t node = ls + i;
    if (live) {
        const int64_t b = ls_prev + i * f;
        const int64_t last = ls_prev + lc_prev - 1;
        const int64_t e = b + f - 1 < last ? b + f - 1 : last;
        const int64_t c = b + j;
        float4 m = empty_box();
        int32_t ref = -1;
        if (c <= e) {
            m = node_mbr[c];
            ref = static_cast<int32_t>(c);
        }
        entry_mbr[node * f + j] = m;
        entry_ref[node * f + j] = ref;
        box[t] = m;
        if (j == 0) {
            node_count[node] = static_cast<int32_t>(e - b + 1);
            node_start[node] = node_start[b];
            node_end[node] = node_end[e];
        }
    }
    __syncthreads();
    if (live && j == 0) {
        float4 u = box[t];
        for (int k = 1; k < f; ++k) u = box_union(u, box[t + k]);
        node_mbr[node] = u;
    }
}

int bu_levels(int64_t n, int f, int64_t *ls, int64_t *lc) {
    int L = 0;
    int64_t cnt = ceil_div(n, f), start = 0;
    for (;;) {
        if (L >= PRT_MAX_LEVELS) return -1;
        ls[L] = start;
        lc[L] = cnt;
        start += cnt;
        ++L;
        if (cnt <= 1) break;
        cnt = ceil_div(cnt, f);
    }
    return L;
}

}  // namespace
}  // namespace prt

using namespace prt;

extern "C" int prt_morton_center30(const float *rects, int64_t n, uint32_t *keys, void *stream) {
    PRT_REQUIRE(n >= 0 && (n == 0 || (rects && keys)), PRT_E_INVAL, "morton: bad arguments");
    if (n == 0) return PRT_OK;
    PRT_LAUNCH(k_morton_keys, dim3(static_cast<unsigned>(ceil_div(n, 256))), dim3(256), 0,
               as_stream(stream), reinterpret_cast<const float4 *>(rects), n, keys);
    PRT_LAUNCHED();
    return PRT_OK;
}

extern "C" int prt_morton_ref4d(const int64_t *rects, int64_t n, int64_t scale, uint64_t *keys,
                                void *stream) {
    PRT_REQUIRE(n >= 0 && scale > 0 && (n == 0 || (rects && keys)), PRT_E_INVAL,
                "morton_ref4d: bad arguments");
    if (n == 0) return PRT_OK;
    PRT_LAUNCH(k_morton_ref4d, dim3(static_cast<unsigned>(ceil_div(n, 256))), dim3(256), 0,
               as_stream(stream), rects, n, scale, keys);
    PRT_LAUNCHED();
    return PRT_OK;
}

extern "C" size_t prt_sort_workspace_size(int64_t n) {
    Sizer s;
    s.take<uint64_t>(n > 0 ? n : 1);
    s.take<uint64_t>(n > 0 ? n : 1);
    return s.bytes() + radix_ws_bytes(n, 0);
}

namespace {
__global__ void k_pack_pairs(const uint32_t *__restrict__ keys, int64_t n, int key_bits,
                             uint64_t *__restrict__ items) {
    int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t k = key_bits >= 32 ? keys[i] : (keys[i] & ((1u << key_bits) - 1u));
    items[i] = (static_cast<uint64_t>(k) << 32) | static_cast<uint64_t>(i);
}
__global__ void k_unpack_pairs(const uint64_t *__restrict__ items, int64_t n,
                               uint32_t *__restrict__ keys, uint32_t *__restrict__ idx) {
    int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t v = items[i];
    if (keys) keys[i] = static_cast<uint32_t>(v >> 32);
    if (idx) idx[i] = static_cast<uint32_t>(v);
}
}  // namespace

extern "C" int prt_sort_pairs_u32(const uint32_t *keys_in, int64_t n, int key_bits,
                                  uint32_t *keys_out, uint32_t *idx_out, void *ws, size_t ws_bytes,
                                  void *stream) {
    PRT_REQUIRE(n >= 0 && n < (int64_t(1) << 32), PRT_E_INVAL, "sort: n out of range");
    PRT_REQUIRE(key_bits >= 1 && key_bits <= 32, PRT_E_INVAL, "sort: key_bits must be 1..32");
    PRT_REQUIRE(ws_bytes >= prt_sort_workspace_size(n), PRT_E_NOMEM, "sort: workspace too small");
    if (n == 0) return PRT_OK;
    hipStream_t s = as_stream(stream);
    Carver c(ws, ws_bytes);
    uint64_t *a = c.take<uint64_t>(n), *b = c.take<uint64_t>(n);
    RadixWs rw;
    PRT_REQUIRE(radix_carve(c, n, 0, rw) == PRT_OK, PRT_E_NOMEM, "sort: workspace too small");
    unsigned grid = static_cast<unsigned>(ceil_div(n, 256));
    PRT_LAUNCH(k_pack_pairs, grid, 256, 0, s, keys_in, n, key_bits, a);
    PRT_LAUNCHED();
    uint64_t *res = nullptr;
    int rc = radix_sort_u64(a, a, b, n, 0, 32, 32 + key_bits, rw, s, &res);
    if (rc) return rc;
    PRT_LAUNCH(k_unpack_pairs, grid, 256, 0, s, res, n, keys_out, idx_out);
    PRT_LAUNCHED();
    return PRT_OK;
}

namespace {
// Leaves from the sorted items (key word << 32 | rect id), then every upper
// level.
int bu_pack(const uint64_t *sorted, const float4 *r4, int64_t n, prt_bu_tree *tree,
            float4 *node_mbr, hipStream_t s) {
    const int f = tree->fanout;
    const int npb = PK_THREADS / f;
    float4 *em = reinterpret_cast<float4 *>(tree->entry_mbr);
    PRT_LAUNCH(k_bu_leaves, dim3(static_cast<unsigned>(ceil_div(tree->level_count[0], npb))),
               dim3(PK_THREADS), 0, s, sorted, r4, n, f, npb, tree->level_count[0], tree->keys,
               tree->perm, tree->node_count, tree->node_start, tree->node_end, em, tree->entry_ref,
               node_mbr);
    PRT_LAUNCHED();
    for (int l = 1; l < tree->num_levels; ++l) {
        PRT_LAUNCH(k_bu_upper, dim3(static_cast<unsigned>(ceil_div(tree->level_count[l], npb))),
                   dim3(PK_THREADS), 0, s, f, npb, tree->level_start[l - 1],
                   tree->level_count[l - 1], tree->level_start[l], tree->level_count[l],
                   tree->node_count, tree->node_start, tree->node_end, em, tree->entry_ref,
                   node_mbr);
        PRT_LAUNCHED();
    }
    return PRT_OK;
}
}  // namespace

extern "C" int prt_bu_plan(int64_t n, int fanout, prt_bu_tree *tree) {
    PRT_REQUIRE(tree != nullptr, PRT_E_INVAL, "bu_plan: null tree");
    PRT_REQUIRE(n >= 1 && n < (int64_t(1) << 31), PRT_E_INVAL, "bu_plan: n=%lld out of range",
                (long long)n);
    PRT_REQUIRE(fanout >= 2 && fanout <= PRT_MAX_FANOUT, PRT_E_INVAL,
                "bu_plan: fanout %d not in [2,%d]", fanout, PRT_MAX_FANOUT);
    int64_t ls[PRT_MAX_LEVELS], lc[PRT_MAX_LEVELS];
    int L = bu_levels(n, fanout, ls, lc);
    PRT_REQUIRE(L > 0, PRT_E_UNSUPPORTED, "bu_plan: too many levels");
    tree->fanout = fanout;
    tree->num_levels = L;
    tree->num_rects = n;
    tree->num_nodes = ls[L - 1] + lc[L - 1];
    for (int l = 0; l < PRT_MAX_LEVELS; ++l) {
        tree->level_start[l] = l < L ? ls[l] : 0;
        tree->level_count[l] = l < L ? lc[l] : 0;
    }
    return PRT_OK;
}

extern "C" size_t prt_bu_workspace_size(int64_t n, int fanout) {
    prt_bu_tree t;
    if (prt_bu_plan(n, fanout, &t) != PRT_OK) return 0;
    Sizer s;
    s.take<uint64_t>(n);
    s.take<uint64_t>(n);
    s.take<float4>(t.num_nodes);
    return s.bytes() + radix_ws_bytes(n, 0);
}

extern "C" int prt_bu_build(const float *rects, int64_t n, int fanout, prt_bu_tree *tree, void *ws,
                            size_t ws_bytes, void *stream) {
    int rc = prt_bu_plan(n, fanout, tree);
    if (rc) return rc;
    PRT_REQUIRE(rects && tree->keys && tree->perm && tree->node_count && tree->node_start &&
                    tree->node_end && tree->entry_mbr && tree->entry_ref,
                PRT_E_INVAL, "bu_build: null device pointer");
    PRT_REQUIRE(ws_bytes >= prt_bu_workspace_size(n, fanout), PRT_E_NOMEM,
                "bu_build: workspace %zu < %zu bytes", ws_bytes, prt_bu_workspace_size(n, fanout));
    hipStream_t s = as_stream(stream);
    Carver c(ws, ws_bytes);
    uint64_t *a = c.take<uint64_t>(n), *b = c.take<uint64_t>(n);
    float4 *node_mbr = c.take<float4>(tree->num_nodes);
    RadixWs rw;
    PRT_REQUIRE(radix_carve(c, n, 0, rw) == PRT_OK && a && b && node_mbr, PRT_E_NOMEM,
                "bu_build: workspace too small");
    const float4 *r4 = reinterpret_cast<const float4 *>(rects);

    uint64_t *sorted = nullptr;
    rc = radix_sort_morton(r4, a, b, n, rw, s, &sorted);
    if (rc) return rc;
    return bu_pack(sorted, r4, n, tree, node_mbr, s);
}

namespace {
// Sign of f - v for an fp32 value f rounded from the int64 v, compared
// exactly (as integers: every fp32 of magnitude below 2^63 is an integer
// there once it exceeds 2^24, and the rounding of v is within a unit of it).
__device__ inline int f32_cmp(float f, int64_t v) {
    if (f >= 0x1p63f) return 1;
    if (f < -0x1p63f) return -1;
    const int64_t fi = static_cast<int64_t>(f);
    return fi > v ? 1 : (fi < v ? -1 : 0);
}
__device__ inline float f32_down(int64_t v) {
    const float f = static_cast<float>(v);
    return f32_cmp(f, v) > 0 ? nextafterf(f, -__builtin_huge_valf()) : f;
}
__device__ inline float f32_up(int64_t v) {
    const float f = static_cast<float>(v);
    return f32_cmp(f, v) < 0 ? nextafterf(f, __builtin_huge_valf()) : f;
}

// Reference keys (as k_morton_ref4d) and the fp32 boxes of the integer rects
// (rounded outward: exact for |v| <= 2^24, containing otherwise).  A
// coordinate that fp32 cannot hold raises PRT_STATUS_INEXACT: the tree's
// boxes then contain the rects without equalling them, and joins over it
// return a superset of the int64 Overlap hit set (BU:496-506).
__global__ void k_ref4d_prepare(const int64_t *__restrict__ rects, int64_t n, int64_t scale,
                                uint64_t *__restrict__ keys, float4 *__restrict__ r4,
                                uint32_t *flags) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    bool inexact = false;
    if (i < n) {
        uint64_t k = 0;
        int64_t v[4];
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            v[w] = rects[4 * i + w];
            const uint64_t a = static_cast<uint64_t>(v[w] / scale);
            k |= spread_ref4(a) << (6 - 2 * w);
        }
        keys[i] = k;
        const float4 b = make_float4(f32_down(v[0]), f32_down(v[1]), f32_up(v[2]), f32_up(v[3]));
        r4[i] = b;
        inexact = f32_cmp(b.x, v[0]) != 0 || f32_cmp(b.y, v[1]) != 0 || f32_cmp(b.z, v[2]) != 0 ||
                  f32_cmp(b.w, v[3]) != 0;
    }
    // one flag write per wave that holds an inexact coordinate
    if (__ballot(inexact) != 0ull && (threadIdx.x & 63) == 0)
        atomicOr(flags, static_cast<uint32_t>(PRT_STATUS_INEXACT));
}
}  // namespace

extern "C" size_t prt_bu_ref4d_workspace_size(int64_t n, int fanout) {
    prt_bu_tree t;
    if (prt_bu_plan(n, fanout, &t) != PRT_OK) return 0;
    Sizer s;
    for (int k = 0; k < 6; ++k) s.take<uint64_t>(n);  // keys, items, 4 sort buffers
    s.take<float4>(n);
    s.take<float4>(t.num_nodes);
    return s.bytes() + radix_ws_bytes(n, 0);
}

extern "C" int prt_bu_build_ref4d(const int64_t *rects, int64_t n, int fanout, int64_t scale,
                                  prt_bu_tree *tree, uint64_t *keys_sorted, void *ws,
                                  size_t ws_bytes, void *stream) {
    int rc = prt_bu_plan(n, fanout, tree);
    if (rc) return rc;
    PRT_REQUIRE(scale > 0, PRT_E_INVAL, "bu_build_ref4d: scale must be positive");
    PRT_REQUIRE(rects && tree->keys && tree->perm && tree->node_count && tree->node_start &&
                    tree->node_end && tree->entry_mbr && tree->entry_ref,
                PRT_E_INVAL, "bu_build_ref4d: null device pointer");
    const size_t need = prt_bu_ref4d_workspace_size(n, fanout);
    PRT_REQUIRE(ws_bytes >= need, PRT_E_NOMEM, "bu_build_ref4d: workspace %zu < %zu bytes",
                ws_bytes, need);
    hipStream_t s = as_stream(stream);
    Carver c(ws, ws_bytes);
    uint64_t *keys = c.take<uint64_t>(n), *items = c.take<uint64_t>(n);
    uint64_t *b0 = c.take<uint64_t>(n), *b1 = c.take<uint64_t>(n);
    uint64_t *b2 = c.take<uint64_t>(n), *b3 = c.take<uint64_t>(n);
    float4 *r4 = c.take<float4>(n);
    float4 *node_mbr = c.take<float4>(tree->num_nodes);
    RadixWs rw;
    PRT_REQUIRE(radix_carve(c, n, 0, rw) == PRT_OK && c.ok(), PRT_E_NOMEM,
                "bu_build_ref4d: workspace too small");
    uint32_t *flags = device_status_flags();
    PRT_REQUIRE(flags != nullptr, PRT_E_HIP, "bu_build_ref4d: no device status word");
    PRT_LAUNCH(k_ref4d_prepare, dim3(static_cast<unsigned>(ceil_div(n, 256))), dim3(256), 0, s,
               rects, n, scale, keys, r4, flags);
    PRT_LAUNCHED();
    rc = radix_sort_keys64(keys, n, b0, b1, b2, b3, rw, s, items, keys_sorted);
    if (rc) return rc;
    return bu_pack(items, r4, n, tree, node_mbr, s);
}
