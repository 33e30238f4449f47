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

__global__ void k_zero_words(uint32_t *__restrict__ p, int64_t n) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        p[i] = 0u;
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

// Fused lower levels for power-of-two fanouts.  One group = one segment of
// the first level whose segments fit LDS (a0 <= 4096 items, padded to
// A2 = next_pow2(a0); a block holds 4096 / A2 groups of one layer).  Items,
// rects and ids of all layers sit side by side (layer l at positions
// l*n ..), ids are positions in that joint order.  The kernel
//   1. gathers the group's rects in the order of the previous level's sort
//      (src = low word of its sorted items, or the identity at level 0),
//   2. sorts the group for every remaining level p0..L-1 in LDS: level p
//      sorts each aligned block of a_p positions by (ord32(x1 or y1), current
//      position) -- the stable per-segment sort of those levels -- carrying
//      the group-local index of each rect,
//   3. writes the leaf slots and perm in the final order, and
//   4. reduces the boxes bottom-up in LDS into the slots of levels L-2..p0
//      (entry of level q = union of its f children; empty slots get the
//      empty box and child 0).
// Levels above p0 are packed by k_td_level from these slots.  Same slots as
// the level-by-level path (unions are exact, so bit-identical).
struct SubtreeArgs {
    const uint64_t *items;   // sorted items of level p0 - 1, or null (p0 == 0)
    RectLayers rect_cur;     // rects in the order the items index
    const uint32_t *id_cur;  // ids in that order, or null (identity)
    int64_t n, a0;           // n: rects per layer
    int A2, logA2, L, p0, f, logf;
    int64_t seg[PRT_MAX_LEVELS];        // a_p of the sort of level p
    int64_t off[PRT_MAX_LEVELS + 1];    // slot offsets per level
    int32_t *slot_child[PRT_MAX_LAYERS];
    float4 *slot_mbr[PRT_MAX_LAYERS];
    uint32_t *perm[PRT_MAX_LAYERS];
    int64_t num_groups;                 // per layer
    int64_t blocks_per_layer;
};

// Thread t holds the keys of positions 4t..4t+3 in registers.  A bitonic
// compare-exchange at distance j pairs registers of one thread (j < 4) or
// lanes of one wave (4 <= j < 256: lane t ^ (j/4), by cross-lane shuffles),
// so blocks of up to 256 positions sort without LDS or block barriers;
// larger blocks go to block_radix.
constexpr int ST_PER = LS_SLOTS / LS_THREADS;  // 4 keys per thread

// dynamic LDS, 72 KB so that two blocks share a CU: kx | ky (ord32 of x1 and
// y1 per local index, 2 x 16 KB) | src (joint position of each local index,
// 16 KB) | ord and its radix twin (2 x 8 KB) | per-wave 16-bit digit
// counters (8 KB).  Once the sorts are done the first 64 KB hold the rects
// by local index, then the boxes of the block's upper levels.
constexpr int ST_WAVES = LS_THREADS / 64;
constexpr size_t ST_LDS = LS_SLOTS * (4 + 4 + 4 + 2 + 2) + ST_WAVES * 256 * 2;
constexpr uint16_t ST_NONE = 0xFFFFu;  // ord entry of a padding position
constexpr int ST_RADIX_MIN = 512;       // blocks of this many positions or more: block_radix
// LDS requested when there are no more blocks than CUs: above half of a CU's
// 160 KB, so that no CU takes two blocks while another idles
constexpr size_t ST_LDS_ALONE = 96 * 1024;

// Sorts every aligned block of B positions (B a power of two <= 256)
// ascending; position of register r of thread t is 4t + r.
// v of lane (this lane ^ lx) for a power of two lx below 64, every lane
// active: within a quad and across half a row (xor 8, a rotation of the row
// by 8) a DPP move, within 32 lanes a swizzle, across the wave's halves a
// permute.  Called with constant lx, the choice folds away.
__device__ inline uint32_t lane_xor(uint32_t v, int lx) {
    const int x = static_cast<int>(v);
    switch (lx) {
    case 1: return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, false));
    case 2: return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, x, 0x4E, 0xF, 0xF, false));
    case 4: return static_cast<uint32_t>(__builtin_amdgcn_ds_swizzle(x, 0x101F));
    case 8: return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, x, 0x128, 0xF, 0xF, false));
    case 16: return static_cast<uint32_t>(__builtin_amdgcn_ds_swizzle(x, 0x401F));
    default: return static_cast<uint32_t>(__shfl_xor(x, lx, 64));
    }
}
__device__ inline uint64_t lane_xor64(uint64_t v, int lx) {
    return (static_cast<uint64_t>(lane_xor(static_cast<uint32_t>(v >> 32), lx)) << 32) |
           lane_xor(static_cast<uint32_t>(v), lx);
}

// The network for a block size fixed at compile time: fully unrolled, every
// exchange distance a constant.
template <int B> __device__ inline void block_bitonic_fixed(uint64_t (&kk)[ST_PER]) {
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 2; k <= B; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j >= ST_PER; j >>= 1) {
            const int i0 = ST_PER * t;
            const bool want_max = ((i0 & j) == 0) != ((k == B) || ((i0 & k) == 0));
#pragma unroll
            for (int r = 0; r < ST_PER; ++r) {
                const uint64_t o = lane_xor64(kk[r], j / ST_PER);
                const bool take = (o < kk[r]) != want_max;
                kk[r] = take ? o : kk[r];
            }
        }
#pragma unroll
        for (int j = (k >> 1) < ST_PER ? (k >> 1) : ST_PER / 2; j > 0; j >>= 1) {
#pragma unroll
            for (int r = 0; r < ST_PER; ++r) {
                if (r & j) continue;
                const int i = ST_PER * t + r;
                const bool up = (k == B) || ((i & k) == 0);
                const uint64_t x = kk[r], y = kk[r | j];
                if ((x > y) == up) {
                    kk[r] = y;
                    kk[r | j] = x;
                }
            }
        }
    }
}

__device__ inline void block_bitonic(uint64_t (&kk)[ST_PER], int B) {
    if (B == 256) return block_bitonic_fixed<256>(kk);
    if (B == 16) return block_bitonic_fixed<16>(kk);
    const int t = threadIdx.x;
    for (int k = 2; k <= B; k <<= 1) {
        int j = k >> 1;
        // 4 <= j < 256: the partners of positions 4t..4t+3 are the positions
        // 4t'..4t'+3 of lane t' = t ^ (j/4) of the same wave; the direction
        // bits come from t alone, and each key takes its partner's when the
        // partner belongs on its side (one 64-bit compare and one select)
        for (; j >= ST_PER; j >>= 1) {
            const int i0 = ST_PER * t;
            const bool want_max = ((i0 & j) == 0) != ((k == B) || ((i0 & k) == 0));
            const int lx = j / ST_PER;
#pragma unroll
            for (int r = 0; r < ST_PER; ++r) {
                const uint64_t o = __shfl_xor(kk[r], lx, 64);
                const bool take = (o < kk[r]) != want_max;
                kk[r] = take ? o : kk[r];
            }
        }
        for (; j > 0; j >>= 1) {
#pragma unroll
            for (int r = 0; r < ST_PER; ++r) {
                if (r & j) continue;
                const int i = ST_PER * t + r;
                const bool up = (k == B) || ((i & k) == 0);
                const uint64_t x = kk[r], y = kk[r | j];
                if ((x > y) == up) {
                    kk[r] = y;
                    kk[r | j] = x;
                }
            }
        }
    }
}

// Stable LSD radix sort (four 8-bit passes) of the positions of every
// aligned segment of B positions (B a multiple of 256) by the 32-bit key
// key_of[local index] of the index at each position: ord -> ord (ord2 is
// the twin buffer).  Wave w ranks positions w*256 + r*64 + lane (r = 0..3,
// in order) with ballots over the digit bits and per-wave counters; one
// exclusive scan over the counters in (segment, digit, wave) order gives
// every (wave, digit) its first destination, since segments are
// consecutive in that order.  Padding positions carry key 0xFFFFFFFF and
// stay behind the others of their segment (they come last in position
// order).  The same order as a sort by (key, position).
__device__ void block_radix(uint16_t *ord, uint16_t *ord2, const uint32_t *key_of, uint16_t *wcnt,
                            int B) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wps = B / 256;  // waves per segment (a power of two)
    const int lwps = __builtin_ctz(wps);
    for (int i = tid; i < ST_WAVES * 256; i += LS_THREADS) wcnt[i] = 0;
    __syncthreads();
    for (int pass = 0; pass < 4; ++pass) {
        const int shift = 8 * pass;
        uint32_t dig[ST_PER], rank[ST_PER];
        uint16_t o[ST_PER];
#pragma unroll
        for (int r = 0; r < ST_PER; ++r) {
            const int pos = wave * 256 + r * 64 + lane;
            o[r] = ord[pos];
            const uint32_t key = o[r] == ST_NONE ? 0xFFFFFFFFu : key_of[o[r]];
            dig[r] = (key >> shift) & 0xFFu;
        }
#pragma unroll
        for (int r = 0; r < ST_PER; ++r) {
            const uint32_t d = dig[r];
            const uint64_t mism = digit_mismatch(d, true);
            const uint32_t base = wcnt[wave * 256 + d];
            const uint32_t before = static_cast<uint32_t>(lane) - digit_lanes_below(mism);
            __builtin_amdgcn_wave_barrier();
            if (before == 0)
                wcnt[wave * 256 + d] =
                    static_cast<uint16_t>(base + 64u - static_cast<uint32_t>(__popcll(mism)));
            __builtin_amdgcn_wave_barrier();
            rank[r] = base + before;
        }
        __syncthreads();
        // exclusive scan of the counters in (segment, digit, wave) order:
        // element e = s * 256 * wps + d * wps + wl is wave s * wps + wl's count of d
        uint32_t c[ST_PER], run = 0;
#pragma unroll
        for (int q = 0; q < ST_PER; ++q) {
            const int e = ST_PER * tid + q;
            const int seg = e >> (8 + lwps), rem = e & ((256 << lwps) - 1);
            const int d = rem >> lwps, w = (seg << lwps) + (rem & (wps - 1));
            c[q] = wcnt[w * 256 + d];
            run += c[q];
        }
        const uint32_t x = wave_incl_scan(run);
        __shared__ uint32_t wsum[ST_WAVES];
        if (lane == 63) wsum[wave] = x;
        __syncthreads();
        uint32_t before = x - run;
        for (int w = 0; w < wave; ++w) before += wsum[w];
#pragma unroll
        for (int q = 0; q < ST_PER; ++q) {
            const int e = ST_PER * tid + q;
            const int seg = e >> (8 + lwps), rem = e & ((256 << lwps) - 1);
            const int d = rem >> lwps, w = (seg << lwps) + (rem & (wps - 1));
            wcnt[w * 256 + d] = static_cast<uint16_t>(before);  // first destination of (wave, digit)
            before += c[q];
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < ST_PER; ++r) ord2[wcnt[wave * 256 + dig[r]] + rank[r]] = o[r];
        __syncthreads();
        for (int i = tid; i < ST_WAVES * 256; i += LS_THREADS) wcnt[i] = 0;
        uint16_t *to = ord;
        ord = ord2;
        ord2 = to;
        __syncthreads();
    }
}

// box of lane + x (lanes past the wave's end read their own)
__device__ inline float4 box_from_above(float4 v, int x) {
    return make_float4(__shfl_down(v.x, x, 64), __shfl_down(v.y, x, 64), __shfl_down(v.z, x, 64),
                       __shfl_down(v.w, x, 64));
}
// the same for x = CTRL - 0x100 < 16 within a row of 16 lanes, by a DPP
// row shift (lanes past the row's end read their own)
template <int CTRL> __device__ inline float dpp_shl(float v) {
    const int b = __float_as_int(v);
    return __int_as_float(__builtin_amdgcn_update_dpp(b, b, CTRL, 0xF, 0xF, false));
}
template <int CTRL> __device__ inline float4 box_from_above_row(float4 v) {
    return make_float4(dpp_shl<CTRL>(v.x), dpp_shl<CTRL>(v.y), dpp_shl<CTRL>(v.z),
                       dpp_shl<CTRL>(v.w));
}

// Slot i of level q (entries of cover = 2^lc leaves) for box x of the
// block: group g0 + x / epg, entry x % epg of that group (epg = A2 / cover).
// Powers of two throughout, so shifts and masks instead of divisions.
__device__ inline void write_entry(const SubtreeArgs &A, int32_t *slot_child, float4 *slot_mbr,
                                   int64_t g0, int q, int lc, int64_t x, float4 u) {
    const int le = A.logA2 - lc;  // log2(entries per group)
    const int64_t gg = g0 + (x >> le), e = x & ((int64_t(1) << le) - 1);
    const int64_t i = ((gg * A.a0) >> lc) + e;  // slot index in level q
    const int64_t cnt = A.off[q + 1] - A.off[q];
    if (gg < A.num_groups && i < cnt && (e << lc) < A.a0) {
        const bool full = (i << lc) < A.n;
        slot_child[A.off[q] + i] = full ? static_cast<int32_t>(A.off[q + 1] + i * A.f) : 0;
        slot_mbr[A.off[q] + i] = full ? u : empty_box();
    }
}

// slots of level q past the last group's entries are empty
__device__ inline void write_empty_tail(const SubtreeArgs &A, int32_t *slot_child, float4 *slot_mbr,
                                        int q, int lc) {
    const int64_t cnt = A.off[q + 1] - A.off[q];
    const int64_t first = (A.num_groups * A.a0 + (int64_t(1) << lc) - 1) >> lc;
    for (int64_t i = first + threadIdx.x; i < cnt; i += LS_THREADS) {
        slot_child[A.off[q] + i] = 0;
        slot_mbr[A.off[q] + i] = empty_box();
    }
}

// WPE: waves per SIMD the register budget allows -- 8 when two blocks share
// a CU (64 VGPRs), 4 when each block has a CU to itself (no spills)
template <int WPE>
__global__ __launch_bounds__(LS_THREADS) __attribute__((amdgpu_waves_per_eu(WPE, WPE)))
void k_td_subtree(SubtreeArgs A) {
    extern __shared__ __align__(16) char st_lds[];
    uint32_t *kx = reinterpret_cast<uint32_t *>(st_lds);
    uint32_t *ky = kx + LS_SLOTS;
    uint32_t *src = ky + LS_SLOTS;
    uint16_t *ord = reinterpret_cast<uint16_t *>(src + LS_SLOTS);
    uint16_t *ord2 = ord + LS_SLOTS;
    uint16_t *wcnt = ord2 + LS_SLOTS;
    float4 *box = reinterpret_cast<float4 *>(st_lds);  // after the sorts
    const int tid = threadIdx.x, lane = tid & 63;
    const int spb = LS_SLOTS / A.A2;
    // XCD-aware block order: blocks are dealt to the 8 XCDs round-robin, so
    // block b runs on XCD b % 8; XCD x gets a contiguous range of groups,
    // whose random gathers then stay within a few segments that fit its L2
    const int64_t nb = gridDim.x, xq = nb / 8, xr = nb % 8, xcd = blockIdx.x % 8;
    const int64_t blk_all = xcd * xq + (xcd < xr ? xcd : xr) + blockIdx.x / 8;
    const int layer = static_cast<int>(blk_all / A.blocks_per_layer);
    const int64_t blk = blk_all - layer * A.blocks_per_layer;
    const int64_t base = layer * A.n;  // first joint position of the layer
    int32_t *const slot_child = A.slot_child[layer];
    float4 *const slot_mbr = A.slot_mbr[layer];
    uint32_t *const perm = A.perm[layer];
    const int64_t g0 = blk * spb;  // first group of the block (in its layer)
    const bool last_block = blk == A.blocks_per_layer - 1;
    // 1. the groups' sort keys and joint positions into LDS (local index k),
    //    in the order of the previous level's sort; every index load of the
    //    thread is issued before the dependent rect loads, and the loads are
    //    unconditional (padding positions read position 0 and keep nothing)
    uint32_t sidx[ST_PER];
    bool live[ST_PER];
#pragma unroll
    for (int m = 0; m < ST_PER; ++m) {
        const int k = tid + m * LS_THREADS;
        const int64_t gg = g0 + (k >> A.logA2), w = k & (A.A2 - 1), pos = gg * A.a0 + w;
        live[m] = gg < A.num_groups && w < A.a0 && pos < A.n;
        sidx[m] = static_cast<uint32_t>(base + pos);
        if (live[m] && A.items) sidx[m] = static_cast<uint32_t>(A.items[base + pos]);
        if (!live[m]) sidx[m] = 0u;
    }
    // with the full register budget x2 and y2 stay in registers until the
    // leaves are written (x1 and y1 come back from their sort keys); with
    // 64 registers the leaves read them again from global memory instead
    constexpr bool keep_zw = WPE <= 4;
    float2 zw[ST_PER];
#pragma unroll
    for (int m = 0; m < ST_PER; ++m) {
        const int k = tid + m * LS_THREADS;
        const float4 r = A.rect_cur.at(sidx[m]);
        if (keep_zw) zw[m] = make_float2(r.z, r.w);
        kx[k] = ord32(r.x);
        ky[k] = ord32(r.y);
        src[k] = sidx[m];
        // ord[pos] = local index of the rect at position pos of the current
        // order (the previous level's), ST_NONE at padding positions
        ord[k] = live[m] ? static_cast<uint16_t>(k) : ST_NONE;
    }
    __syncthreads();
    // 2. one sort per level, each within aligned blocks of B = a_p positions
    //    by (ord32(x1 or y1), position): block_radix for blocks of 512
    //    positions or more, the bitonic network in registers below
    for (int p = A.p0; p < A.L; ++p) {
        const uint32_t *key_of = (p & 1) ? ky : kx;
        int B = 1;
        while (B < A.seg[p]) B <<= 1;
        if (B > A.A2) B = A.A2;
        if (B >= ST_RADIX_MIN) {
            block_radix(ord, ord2, key_of, wcnt, B);
            continue;
        }
        // key of position k = 4t + r: (ord32(coord), position, local index)
        uint64_t kk[ST_PER];
#pragma unroll
        for (int r = 0; r < ST_PER; ++r) {
            const int k = ST_PER * tid + r;
            const uint16_t o = ord[k];
            kk[r] = o == ST_NONE ? ~0ull
                                 : (static_cast<uint64_t>(key_of[o]) << 32) |
                                       (static_cast<uint64_t>(k) << 12) | o;
        }
        block_bitonic(kk, B);
#pragma unroll
        for (int r = 0; r < ST_PER; ++r)
            ord[ST_PER * tid + r] = kk[r] == ~0ull ? ST_NONE : static_cast<uint16_t>(kk[r] & 0xFFFu);
        __syncthreads();
    }
    // 3. leaves in the final order (position k = tid + m * 1024, coalesced).
    //    With the full register budget: x1 and y1 come back from their sort
    //    keys (ord32 is one-to-one but for -0 and +0, which share a key: a
    //    zero coordinate is read again from global memory for its sign), x2
    //    and y2 move from the registers of step 1 to the final positions
    //    through LDS (the kx and ky arrays, free once everything the leaves
    //    need from LDS is in registers).  With 64 registers the rects are
    //    read again from global memory (L2) through the joint positions.
    const int64_t offL = A.off[A.L - 1], cntL = A.off[A.L] - A.off[A.L - 1];
    uint16_t fo[ST_PER];
    uint32_t s4[ST_PER];
    float4 mine[ST_PER];
    bool has[ST_PER];
#pragma unroll
    for (int m = 0; m < ST_PER; ++m) {
        fo[m] = ord[tid + m * LS_THREADS];
        has[m] = fo[m] != ST_NONE;
        const int o = has[m] ? fo[m] : 0;
        s4[m] = has[m] ? src[o] : 0u;
        if (keep_zw) {
            mine[m].x = unord32(kx[o]);
            mine[m].y = unord32(ky[o]);
        }
    }
    if (keep_zw) {
        float *zb = reinterpret_cast<float *>(kx), *wb = reinterpret_cast<float *>(ky);
        __syncthreads();
#pragma unroll
        for (int m = 0; m < ST_PER; ++m) {
            zb[tid + m * LS_THREADS] = zw[m].x;
            wb[tid + m * LS_THREADS] = zw[m].y;
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < ST_PER; ++m) {
            const int o = has[m] ? fo[m] : 0;
            mine[m].z = zb[o];
            mine[m].w = wb[o];
        }
#pragma unroll
        for (int m = 0; m < ST_PER; ++m) {
            if (has[m] && (mine[m].x == 0.0f || mine[m].y == 0.0f)) {
                const float4 r = A.rect_cur.at(s4[m]);
                mine[m].x = r.x;
                mine[m].y = r.y;
            }
        }
    } else {
#pragma unroll
        for (int m = 0; m < ST_PER; ++m) mine[m] = A.rect_cur.at(s4[m]);
    }
    // the id loads go out after the barriers (which would wait for them)
    uint32_t idv[ST_PER];
    if (A.id_cur) {
#pragma unroll
        for (int m = 0; m < ST_PER; ++m) idv[m] = A.id_cur[s4[m]];
    } else {
#pragma unroll
        for (int m = 0; m < ST_PER; ++m) idv[m] = s4[m];
    }
#pragma unroll
    for (int m = 0; m < ST_PER; ++m) {
        const int k = tid + m * LS_THREADS;
        const int64_t gg = g0 + (k >> A.logA2), w = k & (A.A2 - 1), dst = gg * A.a0 + w;
        mine[m] = has[m] ? mine[m] : empty_box();
        if (gg < A.num_groups && w < A.a0 && dst < cntL) {
            slot_child[offL + dst] = dst < A.n ? static_cast<int32_t>(-dst - 1) : 0;
            slot_mbr[offL + dst] = mine[m];
        }
    }
#pragma unroll
    for (int m = 0; m < ST_PER; ++m) {
        const int k = tid + m * LS_THREADS;
        const int64_t gg = g0 + (k >> A.logA2), w = k & (A.A2 - 1), dst = gg * A.a0 + w;
        if (has[m]) perm[dst] = idv[m] - static_cast<uint32_t>(base);
    }
    // leaf slots past the last group are empty
    if (last_block) {
        for (int64_t i = A.num_groups * A.a0 + tid; i < cntL; i += LS_THREADS) {
            slot_child[offL + i] = 0;
            slot_mbr[offL + i] = empty_box();
        }
    }
    // 4. levels L-2 .. p0: box of an entry of level q = union of the boxes
    //    of its `cover` leaf positions.  While cover <= 64 those positions
    //    are lanes of one wave (position k = tid + m * 1024 is lane k % 64):
    //    a tree reduction toward the group's first lane (it adds the lane x
    //    above for x = 1, 2, 4, ...; within a row of 16 lanes by DPP row
    //    shifts), and that lane writes the entry's slot (the union is order
    //    free: min and max with -0 below +0).  The levels above go through
    //    LDS.
    int lc = 0;  // log2(cover)
    int q = A.L - 2;
    for (; q >= A.p0 && lc + A.logf <= 6; --q) {
        int x = 1 << lc;
        lc += A.logf;
        if (x == 1 && lc >= 4) {
#pragma unroll
            for (int m = 0; m < ST_PER; ++m) {
                mine[m] = box_union(mine[m], box_from_above_row<0x101>(mine[m]));
                mine[m] = box_union(mine[m], box_from_above_row<0x102>(mine[m]));
                mine[m] = box_union(mine[m], box_from_above_row<0x104>(mine[m]));
                mine[m] = box_union(mine[m], box_from_above_row<0x108>(mine[m]));
            }
            x = 16;
        }
        for (; x < (1 << lc); x <<= 1) {
#pragma unroll
            for (int m = 0; m < ST_PER; ++m) mine[m] = box_union(mine[m], box_from_above(mine[m], x));
        }
        if ((lane & ((1 << lc) - 1)) == 0) {
#pragma unroll
            for (int m = 0; m < ST_PER; ++m)
                write_entry(A, slot_child, slot_mbr, g0, q, lc, (tid + m * LS_THREADS) >> lc, mine[m]);
        }
        if (last_block) write_empty_tail(A, slot_child, slot_mbr, q, lc);
    }
    if (q < A.p0) return;
    __syncthreads();  // every read of the rects in LDS is done
    if ((lane & ((1 << lc) - 1)) == 0) {
#pragma unroll
        for (int m = 0; m < ST_PER; ++m) box[(tid + m * LS_THREADS) >> lc] = mine[m];
    }
    __syncthreads();
    int width = LS_SLOTS >> lc;  // boxes of level q + 1 in LDS
    for (; q >= A.p0; --q) {
        lc += A.logf;
        width >>= A.logf;
        float4 u = empty_box();
        if (tid < width)
            for (int c = 0; c < A.f; ++c) u = box_union(u, box[tid * A.f + c]);
        __syncthreads();
        if (tid < width) {
            box[tid] = u;
            write_entry(A, slot_child, slot_mbr, g0, q, lc, tid, u);
        }
        if (last_block) write_empty_tail(A, slot_child, slot_mbr, q, lc);
        __syncthreads();
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

// Levels p_top .. 0 in one block, level by level (a block barrier between
// two levels): the same slots as k_td_level launched once per level.  a =
// leaves under one entry of level p_top.
constexpr int TOP_THREADS = 1024;
struct TopArgs {
    int64_t n, a;
    int f, p_top;
    int64_t off[PRT_MAX_LEVELS + 1];
    int32_t *slot_child[PRT_MAX_LAYERS];  // block l packs layer l
    float4 *slot_mbr[PRT_MAX_LAYERS];
};
__global__ __launch_bounds__(TOP_THREADS) void k_td_top(TopArgs T) {
    int32_t *const slot_child = T.slot_child[blockIdx.x];
    float4 *const slot_mbr = T.slot_mbr[blockIdx.x];
    int64_t a = T.a;
    for (int p = T.p_top; p >= 0; --p) {
        const int64_t off = T.off[p], cnt = T.off[p + 1] - T.off[p], off_next = T.off[p + 1];
        for (int64_t i = threadIdx.x; i < cnt; i += TOP_THREADS) {
            float4 u = empty_box();
            int32_t child = 0;
            if (i * a < T.n) {
                const int64_t c = off_next + i * T.f;
                child = static_cast<int32_t>(c);
                for (int k = 0; k < T.f; ++k) u = box_union(u, slot_mbr[c + k]);
            }
            slot_child[off + i] = child;
            slot_mbr[off + i] = u;
        }
        a *= T.f;
        __syncthreads();
    }
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

int td_device_cus() {
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            cus <= 0)
            cus = 256;
    }
    return cus;
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
// Radix levels (segments larger than LS_SLOTS): look-back words of the
// largest tiling, shared by the levels in turn, and counter words of every
// level side by side, so that one launch clears all of them.  `nl` layers of
// n rects sort together.
struct TdRadixWords {
    int64_t status = 0, ghist = 0;
};
TdRadixWords td_radix_words(int64_t n, int f, int L, int nl) {
    TdRadixWords w;
    for (int p = 0; p < L; ++p) {
        const int64_t a = td_seg(n, f, L, p);
        if (a <= LS_SLOTS) continue;
        const int64_t st = radix_status_words(nl * n, a, n);
        if (st > w.status) w.status = st;
        w.ghist += radix_ghist_words(nl * n, a, n);
    }
    return w;
}

// Workspace of a build of nl layers of n rects (L levels): items, rects and
// ids twice each, look-back and counter words.
size_t td_ws_bytes(int nl, int64_t n, int f, int L) {
    const int64_t N = nl * n;
    Sizer s;
    s.take<uint64_t>(N);
    s.take<uint64_t>(N);
    s.take<float4>(N);
    s.take<float4>(N);
    s.take<uint32_t>(N);
    s.take<uint32_t>(N);
    const TdRadixWords w = td_radix_words(n, f, L, nl);
    s.take<uint64_t>(w.status);
    s.take<uint32_t>(w.ghist);
    return s.bytes();
}

// Power-of-two fanouts: the radix levels, then one fused launch for the LDS
// levels and their slots, then the levels above.  `in` holds nl layers of n
// rects (planned trees[0..nl-1], all alike); every launch covers all layers.
int td_build_pow2(int nl, const RectLayers &in, int64_t n, int f, prt_td_tree *trees, void *ws,
                  size_t ws_bytes, hipStream_t s) {
    const int64_t N = nl * n;
    const int L = trees[0].num_levels;
    Carver c(ws, ws_bytes);
    uint64_t *ia = c.take<uint64_t>(N), *ib = c.take<uint64_t>(N);
    float4 *ra = c.take<float4>(N), *rb = c.take<float4>(N);
    uint32_t *da = c.take<uint32_t>(N), *db = c.take<uint32_t>(N);
    const TdRadixWords rwords = td_radix_words(n, f, L, nl);
    uint64_t *status = c.take<uint64_t>(rwords.status);
    uint32_t *ghist = c.take<uint32_t>(rwords.ghist);
    PRT_REQUIRE(c.ok(), PRT_E_NOMEM, "td_build: workspace too small");

    RectLayers cur_r = in;
    const uint32_t *cur_i = nullptr;  // identity at level 0
    int p0 = L - 1;
    for (int p = 0; p < L; ++p)
        if (td_seg(n, f, L, p) <= LS_SLOTS) {
            p0 = p;
            break;
        }
    // the counters of every radix level are cleared by one launch.  Between
    // two radix levels the rects and ids move into the order of the finished
    // level (in the last digit pass), so the next level's sort and the
    // subtree kernel read neighbouring rects from a few segments.
    if (rwords.ghist > 0) {
        const int64_t zw = rwords.ghist;
        const unsigned zg = static_cast<unsigned>(ceil_div(zw, 1024) < 1024 ? ceil_div(zw, 1024) : 1024);
        PRT_LAUNCH(k_zero_words, zg, 1024, 0, s, ghist, zw);
        PRT_LAUNCHED();
    }
    const uint64_t *items = nullptr;
    uint64_t *made = nullptr;  // this level's items, written by the level before
    uint32_t *lg = ghist;
    for (int p = 0; p < p0; ++p) {
        const int64_t a = td_seg(n, f, L, p);
        const RadixWs lw{lg, status, radix_ghist_words(N, a, n), rwords.status};
        lg += lw.ghist_cap;
        uint64_t *sorted = nullptr;
        int rc;
        if (p + 2 < p0) {
            // two or more radix levels follow: the last digit pass writes
            // only the next level's items, and their payload stays the
            // original index (the rects are not moved: the next level that
            // moves them reads them from the input, one random read per rect,
            // instead of from a moved copy plus its id array)
            rc = radix_sort_td_coord(cur_r, p & 1, made, ia, ib, N, a, n, lw, s, &sorted, true,
                                     nullptr, nullptr, nullptr, true);
            if (rc) return rc;
            made = sorted;
        } else if (p + 1 < p0) {
            // the last digit pass also moves the rects and ids into this
            // level's order, and writes the next level's items
            float4 *nxt_r = (cur_r.p[0] == ra) ? rb : ra;
            uint32_t *nxt_i = (cur_i == da) ? db : da;
            rc = radix_sort_td_coord(cur_r, p & 1, made, ia, ib, N, a, n, lw, s, &sorted, true,
                                     cur_i, nxt_r, nxt_i);
            if (rc) return rc;
            cur_r = one_layer(nxt_r, N);
            cur_i = nxt_i;
            made = sorted;
        } else {
            rc = radix_sort_td_coord(cur_r, p & 1, made, ia, ib, N, a, n, lw, s, &sorted, true);
            if (rc) return rc;
            items = sorted;  // the subtree kernel gathers through them
        }
    }
    SubtreeArgs A{};
    A.items = items;
    A.rect_cur = cur_r;
    A.id_cur = cur_i;
    A.n = n;
    A.a0 = td_seg(n, f, L, p0);
    A.A2 = static_cast<int>(next_pow2(A.a0));
    A.logA2 = __builtin_ctz(static_cast<unsigned>(A.A2));
    A.L = L;
    A.p0 = p0;
    A.f = f;
    A.logf = 0;
    while ((1 << A.logf) < f) ++A.logf;
    for (int p = 0; p < PRT_MAX_LEVELS; ++p) A.seg[p] = p < L ? td_seg(n, f, L, p) : 0;
    const int64_t *off = trees[0].level_offset;
    for (int p = 0; p <= PRT_MAX_LEVELS; ++p) A.off[p] = off[p];
    for (int l = 0; l < nl; ++l) {
        A.slot_child[l] = trees[l].slot_child;
        A.slot_mbr[l] = reinterpret_cast<float4 *>(trees[l].slot_mbr);
        A.perm[l] = trees[l].perm;
    }
    A.num_groups = ceil_div(n, A.a0);
    A.blocks_per_layer = ceil_div(A.num_groups, LS_SLOTS / A.A2);
    const int64_t blocks = nl * A.blocks_per_layer;
    if (blocks <= td_device_cus()) {
        PRT_LAUNCH(k_td_subtree<4>, dim3(static_cast<unsigned>(blocks)), dim3(LS_THREADS),
                   ST_LDS_ALONE, s, A);
    } else {
        PRT_LAUNCH(k_td_subtree<8>, dim3(static_cast<unsigned>(blocks)), dim3(LS_THREADS), ST_LDS,
                   s, A);
    }
    PRT_LAUNCHED();
    int64_t a = 1;
    for (int k = 0; k < L - p0; ++k) a *= f;  // entries of level p0 - 1 cover f^(L-p0)
    // levels above p0: one launch per level while a level has more slots
    // than a block has threads, then one block per layer packs the rest
    int p = p0 - 1;
    for (; p >= 0 && off[p + 1] - off[p] > TOP_THREADS; --p) {
        const int64_t cnt = off[p + 1] - off[p];
        for (int l = 0; l < nl; ++l) {
            PRT_LAUNCH(k_td_level, dim3(static_cast<unsigned>(ceil_div(cnt, 256))), dim3(256), 0, s,
                       n, f, a, off[p], cnt, off[p + 1], trees[l].slot_child,
                       reinterpret_cast<float4 *>(trees[l].slot_mbr));
            PRT_LAUNCHED();
        }
        a *= f;
    }
    if (p >= 0) {
        TopArgs T{};
        T.n = n;
        T.a = a;
        T.f = f;
        T.p_top = p;
        for (int q = 0; q <= PRT_MAX_LEVELS; ++q) T.off[q] = off[q];
        for (int l = 0; l < nl; ++l) {
            T.slot_child[l] = trees[l].slot_child;
            T.slot_mbr[l] = reinterpret_cast<float4 *>(trees[l].slot_mbr);
        }
        PRT_LAUNCH(k_td_top, dim3(static_cast<unsigned>(nl)), dim3(TOP_THREADS), 0, s, T);
        PRT_LAUNCHED();
    }
    return PRT_OK;
}
}  // namespace

extern "C" size_t prt_td_workspace_size(int64_t n, int fanout) {
    prt_td_tree t;
    if (prt_td_plan(n, fanout, &t) != PRT_OK) return 0;
    return td_ws_bytes(1, n, fanout, t.num_levels);
}

extern "C" size_t prt_td_layers_workspace_size(int num_layers, int64_t n, int fanout) {
    prt_td_tree t;
    if (num_layers < 1 || num_layers > PRT_MAX_LAYERS) return 0;
    if (prt_td_plan(n, fanout, &t) != PRT_OK) return 0;
    if ((fanout & (fanout - 1)) != 0) return td_ws_bytes(1, n, fanout, t.num_levels);
    return td_ws_bytes(num_layers, n, fanout, t.num_levels);
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
    if ((fanout & (fanout - 1)) == 0)
        return td_build_pow2(1, one_layer(reinterpret_cast<const float4 *>(rects), n), n, fanout,
                             tree, ws, ws_bytes, s);
    Carver c(ws, ws_bytes);
    uint64_t *ia = c.take<uint64_t>(n), *ib = c.take<uint64_t>(n);
    float4 *ra = c.take<float4>(n), *rb = c.take<float4>(n);
    uint32_t *da = c.take<uint32_t>(n), *db = c.take<uint32_t>(n);
    const int L = tree->num_levels, f = fanout;
    const TdRadixWords rwords = td_radix_words(n, f, L, 1);
    uint64_t *status = c.take<uint64_t>(rwords.status);
    uint32_t *ghist = c.take<uint32_t>(rwords.ghist);
    PRT_REQUIRE(c.ok(), PRT_E_NOMEM, "td_build: workspace too small");
    // one sort at a time may use all counter words (the level-by-level path)
    RadixWs rw{ghist, status, rwords.ghist, rwords.status};

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

extern "C" int prt_td_build_layers(int num_layers, const float *const *rects, int64_t n, int fanout,
                                   prt_td_tree *trees, void *ws, size_t ws_bytes, void *stream) {
    PRT_REQUIRE(num_layers >= 1 && num_layers <= PRT_MAX_LAYERS, PRT_E_INVAL,
                "td_build_layers: %d layers not in [1,%d]", num_layers, PRT_MAX_LAYERS);
    PRT_REQUIRE(rects && trees, PRT_E_INVAL, "td_build_layers: null host array");
    PRT_REQUIRE(n >= 1 && n * num_layers < (int64_t(1) << 31), PRT_E_INVAL,
                "td_build_layers: %d layers of %lld rects out of range", num_layers, (long long)n);
    for (int l = 0; l < num_layers; ++l) {
        const int rc = prt_td_plan(n, fanout, &trees[l]);
        if (rc) return rc;
        PRT_REQUIRE(rects[l] && trees[l].slot_child && trees[l].slot_mbr && trees[l].perm,
                    PRT_E_INVAL, "td_build_layers: null device pointer in layer %d", l);
    }
    const size_t need = prt_td_layers_workspace_size(num_layers, n, fanout);
    PRT_REQUIRE(ws_bytes >= need, PRT_E_NOMEM, "td_build_layers: workspace %zu < %zu bytes",
                ws_bytes, need);
    if ((fanout & (fanout - 1)) != 0) {
        // the level-by-level path builds one layer after another
        for (int l = 0; l < num_layers; ++l) {
            const int rc = prt_td_build(rects[l], n, fanout, &trees[l], ws, ws_bytes, stream);
            if (rc) return rc;
        }
        return PRT_OK;
    }
    RectLayers in{};
    for (int l = 0; l < num_layers; ++l) in.p[l] = reinterpret_cast<const float4 *>(rects[l]);
    in.len = n;
    in.count = num_layers;
    return td_build_pow2(num_layers, in, n, fanout, trees, ws, ws_bytes, as_stream(stream));
}
