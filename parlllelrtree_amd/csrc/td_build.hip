// td_build.hip -- top-down (sort-tile-recursive style) R-tree build.
//
// Reference: TopDown-rtree.cu.  Per level p the reference recomputes a key
// (x1 on even levels, y1 on odd, plus the segment number, :577-586 host /
// :465-481 GPU), sorts 48-byte records with sixteen 4-bit LSD rounds whose
// scans take ~670 launches per level (:648-669), then packs the level
// (:591-613 host / fun :309-337 GPU).
// Here, per level:
//   * segments larger than one LDS tile (4096 items) are sorted by the
//     segmented 8-bit radix sort of radix.hip on (ord32(coord) << 32 | pos);
//   * smaller segments are sorted in LDS with a bitonic network per segment,
//     the gather into the new order fused into the same kernel;
//   * the packing is done once, bottom-up, after the last level: the set of
//     rects under an entry never changes after its level is sorted (deeper
//     levels only permute inside its range), so an entry's box is the union
//     of its child block -- bit-identical to the reference's direct union.
#include "prt_common.hpp"
#include "prt_profile.hpp"

namespace prt {
namespace {

constexpr int LS_THREADS = 1024;
constexpr int LS_SLOTS = 4096;

__global__ void k_td_items(const float4 *__restrict__ rect_cur, int64_t n, int comp,
                           uint64_t *__restrict__ items) {
    int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 r = rect_cur[i];
    items[i] = (static_cast<uint64_t>(ord32(comp ? r.y : r.x)) << 32) | static_cast<uint64_t>(i);
}

__global__ void k_td_gather(const uint64_t *__restrict__ sorted, const float4 *__restrict__ rect_cur,
                            const uint32_t *__restrict__ id_cur, int64_t n,
                            float4 *__restrict__ rect_next, uint32_t *__restrict__ id_next) {
    int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t pos = static_cast<uint32_t>(sorted[i]);
    rect_next[i] = rect_cur[pos];
    id_next[i] = id_cur ? id_cur[pos] : pos;
}

// Sort every segment [s*a, min((s+1)*a, n)) by (coord, position) in LDS.
// Segments are padded to A2 = next_pow2(a) slots; a block holds 4096/A2.
__global__ __launch_bounds__(LS_THREADS) void k_td_lds_sort(
    const float4 *__restrict__ rect_cur, const uint32_t *__restrict__ id_cur, int64_t n, int64_t a,
    int A2, int comp, float4 *__restrict__ rect_next, uint32_t *__restrict__ id_next) {
    __shared__ uint64_t key[LS_SLOTS];
    const int tid = threadIdx.x;
    const int spb = LS_SLOTS / A2;
    const int64_t seg0 = static_cast<int64_t>(blockIdx.x) * spb;
    for (int i = tid; i < LS_SLOTS; i += LS_THREADS) {
        int64_t seg = seg0 + i / A2, k = i % A2;
        int64_t pos = seg * a + k;
        uint64_t v = ~0ull;
        if (k < a && pos < n) {
            float4 r = rect_cur[pos];
            v = (static_cast<uint64_t>(ord32(comp ? r.y : r.x)) << 32) | static_cast<uint64_t>(pos);
        }
        key[i] = v;
    }
    __syncthreads();
    for (int k = 2; k <= A2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int p = tid; p < LS_SLOTS / 2; p += LS_THREADS) {
                int i = ((p & ~(j - 1)) << 1) | (p & (j - 1));
                int l = i | j;
                bool up = (k == A2) || ((i & k) == 0);
                uint64_t x = key[i], y = key[l];
                if ((x > y) == up) {
                    key[i] = y;
                    key[l] = x;
                }
            }
            __syncthreads();
        }
    }
    for (int i = tid; i < LS_SLOTS; i += LS_THREADS) {
        int64_t seg = seg0 + i / A2, k = i % A2;
        int64_t dst = seg * a + k;
        if (k < a && dst < n) {
            uint32_t src = static_cast<uint32_t>(key[i]);
            rect_next[dst] = rect_cur[src];
            id_next[dst] = id_cur ? id_cur[src] : src;
        }
    }
}

// Last level: leaf slot i holds the rect at sorted position i (-i-1).
__global__ void k_td_leaves(const float4 *__restrict__ rect_final, int64_t n, int64_t off,
                            int64_t count, int32_t *__restrict__ slot_child,
                            float4 *__restrict__ slot_mbr) {
    int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    if (i < n) {
        slot_child[off + i] = static_cast<int32_t>(-i - 1);
        slot_mbr[off + i] = rect_final[i];
    } else {
        slot_child[off + i] = 0;
        slot_mbr[off + i] = empty_box();
    }
}

// Level p entry i covers sorted [i*a, i*a + a); child block off_next + i*f.
__global__ void k_td_level(int64_t n, int f, int64_t a, int64_t off, int64_t count,
                           int64_t off_next, int32_t *__restrict__ slot_child,
                           float4 *__restrict__ slot_mbr) {
    int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float4 u = empty_box();
    int32_t child = 0;
    if (i * a < n) {
        int64_t c = off_next + i * f;
        child = static_cast<int32_t>(c);
        for (int k = 0; k < f; ++k) u = box_union(u, slot_mbr[c + k]);
    }
    slot_child[off + i] = child;
    slot_mbr[off + i] = u;
}

int td_levels(int64_t n, int f, int64_t *off) {
    int64_t pushed[PRT_MAX_LEVELS];
    int L = 0;
    int64_t num = n;
    do {
        if (L >= PRT_MAX_LEVELS) return -1;
        num = ceil_div(num, f) * f;
        pushed[L++] = num;
        num /= f;
    } while (num > 1);
    off[0] = 0;
    for (int p = 0; p < L; ++p) off[p + 1] = off[p] + pushed[L - 1 - p];
    return L;
}

int64_t next_pow2(int64_t v) {
    int64_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

}  // namespace
}  // namespace prt

using namespace prt;

extern "C" int prt_td_plan(int64_t n, int fanout, prt_td_tree *tree) {
    PRT_REQUIRE(tree != nullptr, PRT_E_INVAL, "td_plan: null tree");
    PRT_REQUIRE(n >= 1 && n < (int64_t(1) << 31), PRT_E_INVAL, "td_plan: n=%lld out of range",
                (long long)n);
    PRT_REQUIRE(fanout >= 2 && fanout <= PRT_MAX_FANOUT, PRT_E_INVAL,
                "td_plan: fanout %d not in [2,%d]", fanout, PRT_MAX_FANOUT);
    int64_t off[PRT_MAX_LEVELS + 1];
    int L = td_levels(n, fanout, off);
    PRT_REQUIRE(L > 0, PRT_E_UNSUPPORTED, "td_plan: too many levels");
    PRT_REQUIRE(off[L] < (int64_t(1) << 31), PRT_E_UNSUPPORTED, "td_plan: slot count overflows int32");
    tree->fanout = fanout;
    tree->num_levels = L;
    tree->num_rects = n;
    tree->num_slots = off[L];
    for (int p = 0; p <= PRT_MAX_LEVELS; ++p) tree->level_offset[p] = p <= L ? off[p] : 0;
    return PRT_OK;
}

namespace {
// segment size used by the sort of level p: a_p = f^(L-p), clipped to n
int64_t td_seg(int64_t n, int f, int L, int p) {
    int64_t a = 1;
    for (int k = 0; k < L - p; ++k) {
        a *= f;
        if (a >= n) return n;
    }
    return a;
}
}  // namespace

extern "C" size_t prt_td_workspace_size(int64_t n, int fanout) {
    prt_td_tree t;
    if (prt_td_plan(n, fanout, &t) != PRT_OK) return 0;
    Sizer s;
    s.take<uint64_t>(n);
    s.take<uint64_t>(n);
    s.take<float4>(n);
    s.take<float4>(n);
    s.take<uint32_t>(n);
    s.take<uint32_t>(n);
    size_t rad = 0;
    for (int p = 0; p < t.num_levels; ++p) {
        int64_t a = td_seg(n, fanout, t.num_levels, p);
        if (a > LS_SLOTS) {
            size_t r = radix_ws_bytes(n, a);
            if (r > rad) rad = r;
        }
    }
    return s.bytes() + rad;
}

extern "C" int prt_td_build(const float *rects, int64_t n, int fanout, prt_td_tree *tree, void *ws,
                            size_t ws_bytes, void *stream) {
    int rc = prt_td_plan(n, fanout, tree);
    if (rc) return rc;
    PRT_REQUIRE(rects && tree->slot_child && tree->slot_mbr && tree->perm, PRT_E_INVAL,
                "td_build: null device pointer");
    size_t need = prt_td_workspace_size(n, fanout);
    PRT_REQUIRE(ws_bytes >= need, PRT_E_NOMEM, "td_build: workspace %zu < %zu bytes", ws_bytes, need);
    hipStream_t s = as_stream(stream);
    Carver c(ws, ws_bytes);
    uint64_t *ia = c.take<uint64_t>(n), *ib = c.take<uint64_t>(n);
    float4 *ra = c.take<float4>(n), *rb = c.take<float4>(n);
    uint32_t *da = c.take<uint32_t>(n), *db = c.take<uint32_t>(n);
    const int L = tree->num_levels, f = fanout;
    RadixWs rw{};
    {
        int64_t worst = 0;
        size_t rad = 0;
        for (int p = 0; p < L; ++p) {
            int64_t a = td_seg(n, f, L, p);
            if (a > LS_SLOTS && radix_ws_bytes(n, a) > rad) {
                rad = radix_ws_bytes(n, a);
                worst = a;
            }
        }
        if (worst > 0) PRT_REQUIRE(radix_carve(c, n, worst, rw) == PRT_OK, PRT_E_NOMEM, "td_build: ws");
    }
    PRT_REQUIRE(c.ok(), PRT_E_NOMEM, "td_build: workspace too small");

    const float4 *cur_r = reinterpret_cast<const float4 *>(rects);
    const uint32_t *cur_i = nullptr;  // identity at level 0
    const unsigned grid = static_cast<unsigned>(ceil_div(n, 256));
    for (int p = 0; p < L; ++p) {
        const int comp = p & 1;  // x1 on even levels, y1 on odd (TopDown-rtree.cu:574-586)
        const int64_t a = td_seg(n, f, L, p);
        float4 *nxt_r = (cur_r == ra) ? rb : ra;
        uint32_t *nxt_i = (p == L - 1) ? tree->perm : ((cur_i == da) ? db : da);
        if (a > LS_SLOTS) {
            PRT_LAUNCH(k_td_items, grid, 256, 0, s, cur_r, n, comp, ia);
            PRT_LAUNCHED();
            uint64_t *sorted = nullptr;
            rc = radix_sort_u64(ia, ia, ib, n, a, 32, 64, rw, s, &sorted);
            if (rc) return rc;
            PRT_LAUNCH(k_td_gather, grid, 256, 0, s, sorted, cur_r, cur_i, n, nxt_r, nxt_i);
            PRT_LAUNCHED();
        } else {
            const int A2 = static_cast<int>(next_pow2(a));
            const int64_t segs = ceil_div(n, a);
            const int64_t blocks = ceil_div(segs, LS_SLOTS / A2);
            PRT_LAUNCH(k_td_lds_sort, dim3(static_cast<unsigned>(blocks)), dim3(LS_THREADS), 0, s,
                       cur_r, cur_i, n, a, A2, comp, nxt_r, nxt_i);
            PRT_LAUNCHED();
        }
        cur_r = nxt_r;
        cur_i = nxt_i;
    }
    float4 *sm = reinterpret_cast<float4 *>(tree->slot_mbr);
    const int64_t *off = tree->level_offset;
    {
        int64_t cnt = off[L] - off[L - 1];
        PRT_LAUNCH(k_td_leaves, dim3(static_cast<unsigned>(ceil_div(cnt, 256))), dim3(256), 0, s,
                   cur_r, n, off[L - 1], cnt, tree->slot_child, sm);
        PRT_LAUNCHED();
    }
    int64_t a = 1;  // entry range of level L-2 is f, of level p is f^(L-1-p)
    for (int p = L - 2; p >= 0; --p) {
        a *= f;
        int64_t cnt = off[p + 1] - off[p];
        PRT_LAUNCH(k_td_level, dim3(static_cast<unsigned>(ceil_div(cnt, 256))), dim3(256), 0, s, n,
                   f, a, off[p], cnt, off[p + 1], tree->slot_child, sm);
        PRT_LAUNCHED();
    }
    return PRT_OK;
}
