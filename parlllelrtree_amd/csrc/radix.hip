// radix.hip -- stable, segmented LSD radix sort of 64-bit items for gfx950,
// plus the int64 exclusive scan used by the join.
//
// Replaces the reference's 4-bit LSD sorts (LocalSort/scan1/scan2/Reorder,
// BotomUp-rtree.cu:125-266; fun/prefix*/arraysChange*/sort,
// TopDown-rtree.cu:162-261).  Differences by design:
//   * 8-bit digits; the item is (key << 32 | payload), so 30/32-bit keys need
//     4 passes instead of 16, and only 8 bytes move per item per pass
//     instead of the 112/48-byte records of the reference;
//   * ranking inside a tile uses wave64 ballots (match-any over the 8 digit
//     bits) and per-wave LDS counters, so the order of equal digits is the
//     input order (stable);
//   * the scatter is staged through LDS so that each digit's run is written
//     contiguously (coalesced);
//   * onesweep: one launch histograms every digit pass at once; each pass is
//     then a single launch whose tiles get their global offsets by decoupled
//     look-back over the earlier tiles (no per-pass histogram or scan
//     launches; the reference runs a histogram and two scan levels per
//     4-bit pass);
//   * segments: a tile never spans two segments and histograms and look-back
//     stay inside a segment, so one launch sequence sorts many independent
//     segments (the top-down build's per-level segmented sort,
//     TopDown-rtree.cu:577-588).
#include <mutex>

#include "prt_common.hpp"
#include "prt_profile.hpp"

namespace prt {
namespace {

constexpr int RS_THREADS = 1024;
constexpr int RS_WAVES = RS_THREADS / 64;
// Items per thread: 4 (tiles of 4096) for sorts of fewer than
// RS_BIG_ITEMS items, 8 (tiles of 8192) from there on.  Larger tiles halve
// the tiles a pass walks through and the look-back words each item pays for,
// and keep twice the items in flight per block; below RS_BIG_ITEMS they would
// leave resident blocks idle (measured: 16.7M-rect bottom-up build 1.064 ->
// 0.870 ms, top-down 2.929 -> 2.568 ms; two 1M-rect top-down layers
// 0.267 -> 0.281 ms).  LDS per block stays under half a CU's (two blocks of
// 16 waves per CU).
constexpr int RS_BIG_ITEMS = 1 << 22;
inline int tile_items_for(int64_t n) { return n >= RS_BIG_ITEMS ? 8 : 4; }
constexpr int RS_BINS = 256;
constexpr int RS_MAX_PASSES = 4;  // keys of up to 32 bits

// n items in n / layer layers of `layer` items; each layer in segments of
// seg items (the last one of a layer may be shorter), each segment in tiles.
struct Tiling {
    int64_t n, seg, layer;
    int64_t tiles_per_seg, segs_per_layer, num_segs, num_tiles;
    int64_t tile;  // items per tile: RS_THREADS * (4 or 8)
};

Tiling make_tiling(int64_t n, int64_t seg, int64_t layer = 0) {
    Tiling t;
    t.n = n;
    t.tile = static_cast<int64_t>(RS_THREADS) * tile_items_for(n);
    t.layer = (layer <= 0 || layer > n) ? n : layer;
    t.seg = (seg <= 0 || seg > t.layer) ? t.layer : seg;
    t.tiles_per_seg = ceil_div(t.seg, t.tile);
    t.segs_per_layer = ceil_div(t.layer, t.seg);
    t.num_segs = t.segs_per_layer * (n / t.layer);
    t.num_tiles = t.tiles_per_seg * t.num_segs;
    return t;
}

// first item and end of segment s
__device__ inline void seg_range(const Tiling &t, int64_t s, int64_t &b, int64_t &e) {
    const int64_t l = s / t.segs_per_layer, ls = s - l * t.segs_per_layer;
    b = l * t.layer + ls * t.seg;
    e = l * t.layer + ((ls + 1) * t.seg < t.layer ? (ls + 1) * t.seg : t.layer);
}

__device__ inline void tile_range(const Tiling &t, int64_t tile, int64_t &b, int64_t &e) {
    const int64_t s = tile / t.tiles_per_seg, j = tile % t.tiles_per_seg;
    int64_t sb, se;
    seg_range(t, s, sb, se);
    b = sb + j * t.tile;
    e = b + t.tile < se ? b + t.tile : se;
    if (b > e) b = e;
}

// ---- onesweep: one histogram launch for all digit passes, then one scatter
// launch per pass whose tiles find their global offsets by decoupled look-back
//
// Look-back state: one 64-bit word per (pass, tile, digit): the top two bits
// are the flag (0 = not yet published, 1 = the tile's own count of the digit,
// 2 = inclusive count over the segment's tiles up to and including this one),
// the low 62 bits the count.  Tiles take their index from an atomic ticket, so
// every tile a block waits for belongs to a block that is already running and
// publishes its own count before it looks back: the waits always end.  When
// the whole grid fits the GPU at once (every block can be resident together)
// the tile is the block index instead: no block waits for one that cannot
// start, and a block loads its items at once, without the ticket's round trip.
constexpr uint64_t LB_AGG = 1ull << 62;
constexpr uint64_t LB_INC = 2ull << 62;
constexpr uint64_t LB_VAL = (1ull << 62) - 1;
constexpr int LB_SPIN_LIMIT = 1 << 22;  // safety net only (seconds of polling)
constexpr int LB_WIN = 8;                // look-back words loaded per step

// Look-back test hook (prt_debug_lookback): polls before giving up, and a
// tile that publishes nothing.  Read on the host at every launch.
int g_spin_limit = LB_SPIN_LIMIT;
int64_t g_mute_tile = -1;

__device__ inline uint64_t lb_load(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void lb_store(uint64_t *p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_zero_u32(uint32_t *__restrict__ p, int64_t n) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        p[i] = 0u;
}

// Item producers of the histogram launch: read the items, or make them from
// the rects (and store them) in the same pass.
struct SrcItems {
    const uint64_t *in;
    __device__ uint64_t operator()(int64_t i) const { return in[i]; }
};
struct SrcTdCoord {  // (ord32(x1 or y1) << 32) | position
    RectLayers rects;
    uint64_t *out;
    int comp;
    __device__ uint64_t operator()(int64_t i) const {
        const float4 r = rects.at(i);
        const uint64_t v = (static_cast<uint64_t>(ord32(comp ? r.y : r.x)) << 32) |
                           static_cast<uint64_t>(i);
        out[i] = v;
        return v;
    }
};
struct SrcMorton {  // (morton_center30 << 32) | position
    const float4 *rects;
    uint64_t *out;
    __device__ uint64_t operator()(int64_t i) const {
        const uint64_t v = (static_cast<uint64_t>(morton_center30(rects[i])) << 32) |
                           static_cast<uint64_t>(i);
        out[i] = v;
        return v;
    }
};

struct SrcKeyLo {  // (low word of a 64-bit key << 32) | position
    const uint64_t *keys;
    uint64_t *out;
    __device__ uint64_t operator()(int64_t i) const {
        const uint64_t v = (keys[i] << 32) | static_cast<uint64_t>(i);
        out[i] = v;
        return v;
    }
};
struct SrcKeyHi {  // (high word of the key of the rect at position j of the
                   // low-word order << 32) | j
    const uint64_t *keys;
    const uint64_t *first;
    uint64_t *out;
    __device__ uint64_t operator()(int64_t j) const {
        const uint32_t id = static_cast<uint32_t>(first[j]);
        const uint64_t v = (keys[id] & 0xFFFFFFFF00000000ull) | static_cast<uint64_t>(j);
        out[j] = v;
        return v;
    }
};

// ghist[(pass * num_segs + s) * 256 + d] += items of tile blockIdx.x (in
// segment s) whose digit of `pass` is d, for every pass; also clears the
// tile's look-back words of every pass.
template <class Src, int IT>
__global__ __launch_bounds__(RS_THREADS) void k_hist_all(Src src, Tiling t, int bit_lo, int bit_hi,
                                                         int passes, uint32_t *__restrict__ ghist,
                                                         uint64_t *__restrict__ status) {
    __shared__ uint32_t cnt[RS_MAX_PASSES][RS_BINS];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t tile = blockIdx.x;
    for (int i = tid; i < RS_MAX_PASSES * RS_BINS; i += RS_THREADS) (&cnt[0][0])[i] = 0u;
    for (int i = tid; i < passes * RS_BINS; i += RS_THREADS)
        status[(static_cast<int64_t>(i / RS_BINS) * t.num_tiles + tile) * RS_BINS + i % RS_BINS] = 0ull;
    __syncthreads();
    int64_t b, e;
    tile_range(t, tile, b, e);
    const int count = static_cast<int>(e - b);
    // all loads of the thread are issued before the first use (one memory
    // latency per tile, not one per item); then plain LDS atomics: a digit
    // shared by many lanes serialises only its own additions, which costs
    // less than ranking every item by ballots.  A digit shared by every lane
    // of the wave (the high digits inside a top-down segment, whose keys span
    // a narrow range) is added once, by the first lane, instead of 64 times
    // to one counter.
    uint64_t vals[IT];
#pragma unroll
    for (int j = 0; j < IT; ++j) {
        const int local = wave * (IT * 64) + j * 64 + lane;
        vals[j] = local < count ? src(b + local) : 0ull;
    }
#pragma unroll
    for (int j = 0; j < IT; ++j) {
        const int local = wave * (IT * 64) + j * 64 + lane;
        const bool valid = local < count;
        const uint64_t vmask = __ballot(valid);
        const uint64_t v = vals[j];
        for (int p = 0; p < passes; ++p) {
            const int shift = bit_lo + 8 * p;
            const int bits = bit_hi - shift < 8 ? bit_hi - shift : 8;
            const uint32_t d = static_cast<uint32_t>(v >> shift) & ((1u << bits) - 1u);
            const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
            if (vmask == ~0ull && __ballot(d != d0) == 0ull) {
                if (lane == 0) atomicAdd(&cnt[p][d0], 64u);
            } else if (valid) {
                atomicAdd(&cnt[p][d], 1u);
            }
        }
    }
    __syncthreads();
    const int64_t s = tile / t.tiles_per_seg;
    for (int i = tid; i < passes * RS_BINS; i += RS_THREADS) {
        const uint32_t c = cnt[i / RS_BINS][i % RS_BINS];
        if (c) atomicAdd(&ghist[(static_cast<int64_t>(i / RS_BINS) * t.num_segs + s) * RS_BINS +
                                i % RS_BINS], c);
    }
}

// Optional gather of the last digit pass: with rin set, the rect (and id)
// of every item's low word is copied to its sorted position as well, so the
// rects of the next level come in sorted order without a pass of their own;
// the item stored there is the next level's (ord32 of coordinate next_comp
// of the rect << 32 | sorted position), which its sort then reads as is.
// With rout null the pass writes only the next level's items, keeping the
// low word (the rect's index in rin) as their payload.
struct PassGather {
    RectLayers rin;       // rin.count == 0: no gather
    const uint32_t *iin;  // null: the id is the low word itself
    float4 *rout;         // null: next items only
    uint32_t *iout;
    int next_comp;
};

// ---- one tile of one digit pass (k_onesweep), in phases.  Wave w ranks
// the items [w*64*IT, (w+1)*64*IT) of the tile in input order; lane l of
// round j holds item w*64*IT + j*64 + l.  Output position of an item with
// digit d = segment start + (items of the segment with a smaller digit: scan
// of ghist) + (items with digit d in earlier tiles of the segment: look-back)
// + rank in the tile.  Per-wave digit counters are 16-bit (a tile holds at
// most 8192 items).
template <int IT> struct OsShared {
    uint64_t stage[RS_THREADS * IT];
    uint16_t wcnt[RS_WAVES][RS_BINS];
    uint32_t dstart[RS_BINS];
    uint32_t goff[RS_BINS];
    uint32_t red[2][RS_WAVES];
    uint32_t s_tile;
};

// the tile's items [b, b + count) into the thread's registers; every load
// is issued before the first use (one memory latency per tile)
template <int IT>
__device__ inline void os_load(const uint64_t *__restrict__ in, int64_t b, int count,
                               uint64_t (&item)[IT]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < IT; ++j) {
        const int local = wave * (IT * 64) + j * 64 + lane;
        item[j] = local < count ? in[b + local] : 0ull;
    }
}

// stable rank of each item among the wave's earlier items of its digit
// (wcnt cleared and a barrier passed); 0xFFFFFFFF for positions past count
template <int IT>
__device__ inline void os_rank(OsShared<IT> &S, const uint64_t (&item)[IT], int count, int shift,
                               uint32_t mask, uint32_t (&rank)[IT]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < IT; ++j) {
        const int local = wave * (IT * 64) + j * 64 + lane;
        const bool valid = local < count;
        const uint64_t v = item[j];
        const uint32_t d = static_cast<uint32_t>(v >> shift) & mask;
        const uint64_t mism = digit_mismatch(d, valid);
        uint32_t base = 0;
        if (valid) base = S.wcnt[wave][d];
        const uint32_t r = static_cast<uint32_t>(lane) - digit_lanes_below(mism);
        __builtin_amdgcn_wave_barrier();
        if (valid && r == 0)
            S.wcnt[wave][d] = static_cast<uint16_t>(base + 64u - static_cast<uint32_t>(__popcll(mism)));
        __builtin_amdgcn_wave_barrier();
        rank[j] = valid ? base + r : 0xFFFFFFFFu;
    }
}

// After the ranking's barrier.  Threads 0..255 own digit tid: the tile's
// count of it (published at once: the aggregate, or the inclusive count for
// the first tile of a segment) and the prefix over waves.  Every thread takes
// part in the wave scans of the tile counts (x) and of the segment's digit
// totals (g), whose wave sums go to red.
template <class Sh>
__device__ inline void os_publish(Sh &S, int64_t tile, int64_t jt, uint32_t tot,
                                  uint64_t *status_p, int64_t mute_tile, uint32_t &run,
                                  uint32_t &x, uint32_t &g) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    run = 0;
    if (tid < RS_BINS) {
        for (int w = 0; w < RS_WAVES; ++w) {
            const uint32_t c = S.wcnt[w][tid];
            S.wcnt[w][tid] = static_cast<uint16_t>(run);
            run += c;
        }
        if (tile != mute_tile)
            lb_store(&status_p[tile * RS_BINS + tid], (jt == 0 ? LB_INC : LB_AGG) | run);
    }
    x = wave_incl_scan(run);
    g = wave_incl_scan(tot);
    if (lane == 63) {
        S.red[0][wave] = x;
        S.red[1][wave] = g;
    }
}

// After the barrier that follows os_publish, threads 0..255: the tile's
// digit starts (dstart); returns the digit's start in the segment (the scan
// of the segment's digit totals).
template <class Sh>
__device__ inline uint32_t os_digit_starts(Sh &S, uint32_t run, uint32_t x, uint32_t g,
                                           uint32_t tot) {
    const int tid = threadIdx.x, wave = tid >> 6;
    uint32_t before = 0, gbefore = 0;
    for (int w = 0; w < wave; ++w) {
        before += S.red[0][w];
        gbefore += S.red[1][w];
    }
    S.dstart[tid] = before + x - run;
    return gbefore + g - tot;
}

// Threads 0..255: the segment's output offset of digit tid (goff), from a
// look-back over the earlier tiles of the segment, LB_WIN words in flight
// per step: counts are added from the nearest tile outwards up to the first
// inclusive word; a word not yet published ends the step there (the first
// tile of a segment publishes its inclusive count at once, so every walk
// ends inside the segment).  A walk that polls an unpublished word more than
// spin_limit times gives up and raises PRT_STATUS_LOOKBACK in *flags (the
// order written is then wrong, and the flag says so); mute_tile (test hook,
// -1 = none) publishes no look-back word.
template <class Sh>
__device__ inline void os_lookback(Sh &S, const Tiling &t, int64_t tile, int64_t seg,
                                   int64_t jt, uint32_t run, uint32_t gbase, uint64_t *status_p,
                                   uint32_t *flags, int spin_limit, int64_t mute_tile) {
    const int tid = threadIdx.x;
    uint64_t excl = 0;
    if (jt > 0) {
        int64_t k = tile - 1;
        const int64_t first = tile - jt;
        int spins = 0;
        for (;;) {
            uint64_t v[LB_WIN];
#pragma unroll
            for (int w = 0; w < LB_WIN; ++w)
                v[w] = k - w >= first ? lb_load(&status_p[(k - w) * RS_BINS + tid]) : LB_INC;
            int used = 0;
            bool done = false;
#pragma unroll
            for (int w = 0; w < LB_WIN; ++w) {
                if (done || used != w || (v[w] >> 62) == 0) continue;
                excl += v[w] & LB_VAL;
                used = w + 1;
                done = (v[w] >> 62) == 2;
            }
            if (done) break;
            if (used == 0) {
                if (++spins > spin_limit) {
                    if (flags) atomicOr(flags, static_cast<uint32_t>(PRT_STATUS_LOOKBACK));
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            k -= used;
        }
        if (tile != mute_tile) lb_store(&status_p[tile * RS_BINS + tid], LB_INC | (excl + run));
    }
    int64_t sb, se;
    seg_range(t, seg, sb, se);
    S.goff[tid] = static_cast<uint32_t>(sb) + gbase + static_cast<uint32_t>(excl);
}

// After the barrier that follows os_digit_starts: the items in digit order
template <int IT>
__device__ inline void os_stage(OsShared<IT> &S, const uint64_t (&item)[IT],
                                const uint32_t (&rank)[IT], int shift, uint32_t mask) {
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < IT; ++j) {
        if (rank[j] == 0xFFFFFFFFu) continue;
        const uint32_t d = static_cast<uint32_t>(item[j] >> shift) & mask;
        S.stage[S.dstart[d] + S.wcnt[wave][d] + rank[j]] = item[j];
    }
}

// After os_stage and os_lookback and a barrier: each digit's run written to
// its place, consecutive threads on consecutive items (coalesced)
template <class Sh>
__device__ inline void os_write(Sh &S, uint64_t *__restrict__ out, int count, int shift,
                                uint32_t mask) {
    for (int i = threadIdx.x; i < count; i += RS_THREADS) {
        const uint64_t v = S.stage[i];
        const uint32_t d = static_cast<uint32_t>(v >> shift) & mask;
        out[S.goff[d] + (i - S.dstart[d])] = v;
    }
}

// One tile per block: the tile is the block index when the whole grid fits
// the GPU at once, else taken from an atomic ticket (so every tile a block
// waits for belongs to a block that is already running).  With gat set
// (the last pass of a top-down level), the pass also gathers the rects.
template <int IT>
__global__ __launch_bounds__(RS_THREADS) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_onesweep(
    const uint64_t *__restrict__ in, uint64_t *__restrict__ out, Tiling t, int shift, int bits,
    const uint32_t *__restrict__ ghist_p, uint64_t *status_p, uint32_t *__restrict__ ticket,
    PassGather gat, uint32_t *flags, int spin_limit, int64_t mute_tile) {
    __shared__ OsShared<IT> S;
    const int tid = threadIdx.x;
    const uint32_t mask = (1u << bits) - 1u;
    const bool by_block = ticket == nullptr;  // co-resident grid: tile = block
    if (tid == 0 && !by_block) S.s_tile = atomicAdd(ticket, 1u);
    for (int i = tid; i < RS_WAVES * RS_BINS; i += RS_THREADS) (&S.wcnt[0][0])[i] = 0;
    if (!by_block) __syncthreads();
    const int64_t tile = by_block ? static_cast<int64_t>(blockIdx.x)
                                  : static_cast<int64_t>(S.s_tile);
    const int64_t seg = tile / t.tiles_per_seg, jt = tile % t.tiles_per_seg;
    int64_t b, e;
    tile_range(t, tile, b, e);
    const int count = static_cast<int>(e - b);
    // the segment's digit total is loaded now and used after the ranking
    const uint32_t tot = tid < RS_BINS ? ghist_p[seg * RS_BINS + tid] : 0u;

    uint64_t item[IT];
    uint32_t rank[IT];
    os_load<IT>(in, b, count, item);
    if (by_block) __syncthreads();  // the counters are cleared
    os_rank<IT>(S, item, count, shift, mask, rank);
    __syncthreads();
    uint32_t run, x, g;
    os_publish(S, tile, jt, tot, status_p, mute_tile, run, x, g);
    __syncthreads();
    uint32_t gbase = 0;
    if (tid < RS_BINS) gbase = os_digit_starts(S, run, x, g, tot);
    __syncthreads();
    // the tile goes to LDS in digit order before the look-back, so that the
    // items and ranks hold no registers during the walk
    os_stage<IT>(S, item, rank, shift, mask);
    if (tid < RS_BINS)
        os_lookback(S, t, tile, seg, jt, run, gbase, status_p, flags, spin_limit, mute_tile);
    __syncthreads();
    if (gat.rin.count == 0) {
        os_write(S, out, count, shift, mask);
        return;
    }
    // the items' rects and ids, four per thread at a time, are read (at
    // random) before the first store, unconditionally (positions past the
    // tile's count read entry 0, which exists since n >= 1, and store
    // nothing) and into named registers: loads under a branch or into an
    // array of rects were issued one at a time and parked in scratch memory
#pragma unroll
    for (int h = 0; h < IT; h += 4) {
        uint64_t it4[4];
        uint32_t src[4], idv[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int i = tid + (h + j) * RS_THREADS;
            it4[j] = i < count ? S.stage[i] : 0ull;
            src[j] = static_cast<uint32_t>(it4[j]);
        }
        const float4 rc0 = gat.rin.at(src[0]), rc1 = gat.rin.at(src[1]),
                     rc2 = gat.rin.at(src[2]), rc3 = gat.rin.at(src[3]);
        if (gat.iin) {
#pragma unroll
            for (int j = 0; j < 4; ++j) idv[j] = gat.iin[src[j]];
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) idv[j] = src[j];
        }
        auto put = [&](int j, const float4 &rc) {
            const int i = tid + (h + j) * RS_THREADS;
            if (i < count) {
                const uint32_t d = static_cast<uint32_t>(it4[j] >> shift) & mask;
                const uint32_t pos = S.goff[d] + (i - S.dstart[d]);
                const uint64_t key = static_cast<uint64_t>(ord32(gat.next_comp ? rc.y : rc.x)) << 32;
                if (gat.rout) {
                    out[pos] = key | pos;
                    gat.rout[pos] = rc;
                    gat.iout[pos] = idv[j];
                } else {
                    out[pos] = key | src[j];
                }
            }
        };
        put(0, rc0);
        put(1, rc1);
        put(2, rc2);
        put(3, rc3);
    }
}

struct SortLayout {
    int64_t ghist_words;   // passes * segs * 256 + tickets
    int64_t status_words;  // passes * tiles * 256
};

SortLayout sort_layout(const Tiling &t) {
    SortLayout l;
    l.ghist_words = RS_MAX_PASSES * t.num_segs * RS_BINS + RS_MAX_PASSES;
    l.status_words = RS_MAX_PASSES * t.num_tiles * RS_BINS;
    return l;
}

}  // namespace

size_t radix_ws_bytes(int64_t n, int64_t seg) {
    const SortLayout l = sort_layout(make_tiling(n > 0 ? n : 1, seg));
    Sizer s;
    s.take<uint32_t>(l.ghist_words);
    s.take<uint64_t>(l.status_words);
    return s.bytes();
}

int radix_carve(Carver &c, int64_t n, int64_t seg, RadixWs &w) {
    const SortLayout l = sort_layout(make_tiling(n > 0 ? n : 1, seg));
    w.ghist = c.take<uint32_t>(l.ghist_words);
    w.status = c.take<uint64_t>(l.status_words);
    w.ghist_cap = l.ghist_words;
    w.status_cap = l.status_words;
    return c.ok() ? PRT_OK : PRT_E_NOMEM;
}

int64_t radix_ghist_words(int64_t n, int64_t seg, int64_t layer) {
    return sort_layout(make_tiling(n > 0 ? n : 1, seg, layer)).ghist_words;
}

int64_t radix_status_words(int64_t n, int64_t seg, int64_t layer) {
    return sort_layout(make_tiling(n > 0 ? n : 1, seg, layer)).status_words;
}


namespace {
// blocks of k_onesweep<IT> the GPU holds at once (CUs x blocks per CU), per
// device; 0 when unknown (then every pass takes tickets)
template <int IT> int64_t onesweep_resident_blocks() {
    static int64_t cache[64] = {};
    static std::mutex mu;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
    std::lock_guard<std::mutex> lk(mu);
    if (cache[dev] == 0) {
        int cus = 0, per_cu = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_onesweep<IT>, RS_THREADS, 0) !=
                hipSuccess)
            cache[dev] = -1;
        else
            cache[dev] = static_cast<int64_t>(cus) * per_cu;
    }
    return cache[dev] > 0 ? cache[dev] : 0;
}

template <class Src>
int radix_sort_impl(const Src &prod, const uint64_t *in, uint64_t *buf0, uint64_t *buf1, int64_t n,
                    int64_t seg, int bit_lo, int bit_hi, const RadixWs &w, hipStream_t s,
                    uint64_t **result, bool zeroed, const PassGather &gat = PassGather{},
                    int64_t layer = 0) {
    PRT_REQUIRE(n >= 0 && n < (int64_t(1) << 32), PRT_E_INVAL, "radix sort: n=%lld out of range",
                (long long)n);
    PRT_REQUIRE(bit_lo >= 0 && bit_hi <= 64 && bit_lo < bit_hi && bit_hi - bit_lo <= 8 * RS_MAX_PASSES,
                PRT_E_INVAL, "radix sort: bad bit range [%d,%d)", bit_lo, bit_hi);
    if (n == 0) {
        *result = buf0;
        return PRT_OK;
    }
    PRT_REQUIRE(layer <= 0 || n % layer == 0, PRT_E_INVAL,
                "radix sort: %lld items are not whole layers of %lld", (long long)n,
                (long long)layer);
    const Tiling t = make_tiling(n, seg, layer);
    const SortLayout l = sort_layout(t);
    PRT_REQUIRE(l.ghist_words <= w.ghist_cap && l.status_words <= w.status_cap, PRT_E_NOMEM,
                "radix sort: workspace carved for a smaller tiling");
    const int passes = static_cast<int>((bit_hi - bit_lo + 7) / 8);
    uint32_t *ghist = w.ghist;
    uint32_t *tickets = w.ghist + RS_MAX_PASSES * t.num_segs * RS_BINS;
    if (!zeroed) {
        const int64_t zero_words = l.ghist_words;
        const unsigned zg = static_cast<unsigned>(ceil_div(zero_words, 1024) < 1024 ? ceil_div(zero_words, 1024) : 1024);
        PRT_LAUNCH(k_zero_u32, zg, 1024, 0, s, ghist, zero_words);
        PRT_LAUNCHED();
    }
    const dim3 grid(static_cast<unsigned>(t.num_tiles));
    const bool big = t.tile > RS_THREADS * 4;
    if (big) {
        auto kh = k_hist_all<Src, 8>;
        PRT_LAUNCH_AS("k_hist_all<Src>", kh, grid, dim3(RS_THREADS), 0, s, prod, t, bit_lo, bit_hi,
                      passes, ghist, w.status);
    } else {
        auto kh = k_hist_all<Src, 4>;
        PRT_LAUNCH_AS("k_hist_all<Src>", kh, grid, dim3(RS_THREADS), 0, s, prod, t, bit_lo, bit_hi,
                      passes, ghist, w.status);
    }
    PRT_LAUNCHED();
    const uint64_t *src = in;
    uint64_t *dst = (in == buf0) ? buf1 : buf0;
    const bool by_block =
        t.num_tiles <= (big ? onesweep_resident_blocks<8>() : onesweep_resident_blocks<4>());
    uint32_t *flags = device_status_flags();
    PRT_REQUIRE(flags != nullptr, PRT_E_HIP, "radix sort: no device status word");
    for (int p = 0; p < passes; ++p) {
        const int shift = bit_lo + 8 * p;
        const int bits = bit_hi - shift < 8 ? bit_hi - shift : 8;
        const PassGather none{};
        const PassGather &pg = p + 1 == passes ? gat : none;
        const uint32_t *gh = ghist + static_cast<int64_t>(p) * t.num_segs * RS_BINS;
        uint64_t *st = w.status + static_cast<int64_t>(p) * t.num_tiles * RS_BINS;
        uint32_t *tk = by_block ? nullptr : tickets + p;
        if (big) {
            auto ko = k_onesweep<8>;
            PRT_LAUNCH_AS("k_onesweep", ko, grid, dim3(RS_THREADS), 0, s, src, dst, t, shift, bits,
                          gh, st, tk, pg, flags, g_spin_limit, g_mute_tile);
        } else {
            auto ko = k_onesweep<4>;
            PRT_LAUNCH_AS("k_onesweep", ko, grid, dim3(RS_THREADS), 0, s, src, dst, t, shift, bits,
                          gh, st, tk, pg, flags, g_spin_limit, g_mute_tile);
        }
        PRT_LAUNCHED();
        src = dst;
        dst = (dst == buf0) ? buf1 : buf0;
    }
    *result = const_cast<uint64_t *>(src);
    return PRT_OK;
}
}  // namespace

int radix_sort_u64(const uint64_t *in, uint64_t *buf0, uint64_t *buf1, int64_t n, int64_t seg,
                   int bit_lo, int bit_hi, const RadixWs &w, hipStream_t s, uint64_t **result) {
    return radix_sort_impl(SrcItems{in}, in, buf0, buf1, n, seg, bit_lo, bit_hi, w, s, result,
                           false);
}

int radix_sort_td_coord(const RectLayers &rects, int comp, const uint64_t *items, uint64_t *buf0,
                        uint64_t *buf1, int64_t n, int64_t seg, int64_t layer, const RadixWs &w,
                        hipStream_t s, uint64_t **result, bool zeroed, const uint32_t *ids,
                        float4 *rects_out, uint32_t *ids_out, bool next_items_only) {
    PassGather g{};
    if (rects_out) g = PassGather{rects, ids, rects_out, ids_out, comp ^ 1};
    else if (next_items_only) g = PassGather{rects, nullptr, nullptr, nullptr, comp ^ 1};
    if (items) {
        PRT_REQUIRE(items == buf0 || items == buf1, PRT_E_INVAL,
                    "radix sort: items outside the sort's buffers");
        return radix_sort_impl(SrcItems{items}, items, buf0, buf1, n, seg, 32, 64, w, s, result,
                               zeroed, g, layer);
    }
    return radix_sort_impl(SrcTdCoord{rects, buf0, comp}, buf0, buf0, buf1, n, seg, 32, 64, w, s,
                           result, zeroed, g, layer);
}

int radix_sort_morton(const float4 *rects, uint64_t *buf0, uint64_t *buf1, int64_t n,
                      const RadixWs &w, hipStream_t s, uint64_t **result) {
    return radix_sort_impl(SrcMorton{rects, buf0}, buf0, buf0, buf1, n, 0, 32, 62, w, s, result,
                           false);
}

namespace {
// items[k] = (high word of key << 32) | id of the k-th rect in (key, id)
// order; keys_sorted[k] (optional) = its full key
__global__ void k_compose64(const uint64_t *__restrict__ first, const uint64_t *__restrict__ second,
                            const uint64_t *__restrict__ keys, int64_t n,
                            uint64_t *__restrict__ items, uint64_t *__restrict__ keys_sorted) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t j = static_cast<uint32_t>(second[k]);
    const uint32_t id = static_cast<uint32_t>(first[j]);
    const uint64_t key = keys[id];
    items[k] = (key & 0xFFFFFFFF00000000ull) | id;
    if (keys_sorted) keys_sorted[k] = key;
}
}  // namespace

// Stable sort by a 64-bit key as two stable 32-bit sorts (low word, then
// high word): the order is (key, index).
int radix_sort_keys64(const uint64_t *keys, int64_t n, uint64_t *b0, uint64_t *b1, uint64_t *b2,
                      uint64_t *b3, const RadixWs &w, hipStream_t s, uint64_t *items,
                      uint64_t *keys_sorted) {
    if (n == 0) return PRT_OK;
    uint64_t *first = nullptr, *second = nullptr;
    int rc = radix_sort_impl(SrcKeyLo{keys, b0}, b0, b0, b1, n, 0, 32, 64, w, s, &first, false);
    if (rc) return rc;
    rc = radix_sort_impl(SrcKeyHi{keys, first, b2}, b2, b2, b3, n, 0, 32, 64, w, s, &second,
                         false);
    if (rc) return rc;
    PRT_LAUNCH(k_compose64, dim3(static_cast<unsigned>(ceil_div(n, 256))), 256, 0, s, first, second,
               keys, n, items, keys_sorted);
    PRT_LAUNCHED();
    return PRT_OK;
}

// ---------------------------------------------------------------------------
// int64 exclusive scan (join offsets)
// ---------------------------------------------------------------------------
namespace {
constexpr int SC_THREADS = 256;
constexpr int SC_PER = 16;
constexpr int SC_BLOCK = SC_THREADS * SC_PER;

__device__ inline int64_t block_excl_scan_i64(int64_t v, int64_t *lds, int64_t &total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    int64_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        int64_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) lds[wave] = x;
    __syncthreads();
    int64_t before = 0, tot = 0;
    for (int w = 0; w < nw; ++w) {
        if (w < wave) before += lds[w];
        tot += lds[w];
    }
    total = tot;
    __syncthreads();
    return before + x - v;
}

__global__ __launch_bounds__(SC_THREADS) void k_i64_reduce(const int64_t *__restrict__ in, int64_t n,
                                                           int64_t *__restrict__ part) {
    __shared__ int64_t lds[SC_THREADS / 64];
    int64_t b = static_cast<int64_t>(blockIdx.x) * SC_BLOCK;
    int64_t s = 0;
    for (int i = threadIdx.x; i < SC_BLOCK; i += SC_THREADS)
        if (b + i < n) s += in[b + i];
    int64_t tot;
    block_excl_scan_i64(s, lds, tot);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// single block of 1024 threads scans up to 1024*SC_PER partials
__global__ __launch_bounds__(1024) void k_i64_partials(int64_t *__restrict__ part, int64_t np,
                                                       int64_t *__restrict__ total) {
    __shared__ int64_t lds[16];
    int64_t v[SC_PER], s = 0;
    for (int j = 0; j < SC_PER; ++j) {
        int64_t k = static_cast<int64_t>(threadIdx.x) * SC_PER + j;
        v[j] = k < np ? part[k] : 0;
        s += v[j];
    }
    int64_t tot;
    int64_t run = block_excl_scan_i64(s, lds, tot);
    for (int j = 0; j < SC_PER; ++j) {
        int64_t k = static_cast<int64_t>(threadIdx.x) * SC_PER + j;
        if (k < np) part[k] = run;
        run += v[j];
    }
    if (threadIdx.x == 0) *total = tot;
}

__global__ __launch_bounds__(SC_THREADS) void k_i64_down(const int64_t *__restrict__ in, int64_t n,
                                                         const int64_t *__restrict__ part,
                                                         int64_t *__restrict__ out) {
    __shared__ int64_t lds[SC_THREADS / 64];
    int64_t b = static_cast<int64_t>(blockIdx.x) * SC_BLOCK + threadIdx.x * SC_PER;
    int64_t v[SC_PER], s = 0;
    for (int j = 0; j < SC_PER; ++j) {
        v[j] = b + j < n ? in[b + j] : 0;
        s += v[j];
    }
    int64_t tot;
    int64_t run = block_excl_scan_i64(s, lds, tot) + part[blockIdx.x];
    for (int j = 0; j < SC_PER; ++j) {
        if (b + j < n) out[b + j] = run;
        run += v[j];
    }
}
}  // namespace

size_t scan_i64_ws_bytes(int64_t n) {
    return static_cast<size_t>(ceil_div(n > 0 ? n : 1, SC_BLOCK) + 1) * sizeof(int64_t) + 256;
}

int scan_i64(const int64_t *in, int64_t *out, int64_t n, int64_t *total, void *ws, hipStream_t s) {
    int64_t np = ceil_div(n > 0 ? n : 1, SC_BLOCK);
    PRT_REQUIRE(np <= 1024 * SC_PER, PRT_E_UNSUPPORTED, "scan: n=%lld too large", (long long)n);
    int64_t *part = static_cast<int64_t *>(ws);
    // with n == 0 the single reduce block writes part[0] = 0
    PRT_LAUNCH(k_i64_reduce, dim3(static_cast<unsigned>(np)), dim3(SC_THREADS), 0, s, in, n, part);
    PRT_LAUNCHED();
    PRT_LAUNCH(k_i64_partials, 1, 1024, 0, s, part, np, total);
    PRT_LAUNCHED();
    if (n > 0) {
        PRT_LAUNCH(k_i64_down, dim3(static_cast<unsigned>(np)), dim3(SC_THREADS), 0, s, in, n, part,
                   out);
        PRT_LAUNCHED();
    }
    return PRT_OK;
}

}  // namespace prt

extern "C" int prt_debug_lookback(int32_t spin_limit, int64_t mute_tile) {
    prt::g_spin_limit = spin_limit < 0 ? prt::LB_SPIN_LIMIT : spin_limit;
    prt::g_mute_tile = mute_tile < 0 ? -1 : mute_tile;
    return PRT_OK;
}
