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
//   * segments: a tile never spans two segments and the tile histograms are
//     stored segment-major, so one linear scan gives every segment its own
//     offsets and one launch sequence sorts many independent segments (the
//     top-down build's per-level segmented sort, TopDown-rtree.cu:577-588).
#include "prt_common.hpp"
#include "prt_profile.hpp"

namespace prt {
namespace {

constexpr int RS_THREADS = 1024;
constexpr int RS_WAVES = RS_THREADS / 64;
constexpr int RS_ITEMS = 4;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;  // 4096 items per tile
constexpr int RS_BINS = 256;
constexpr int SCAN_THREADS = 1024;
constexpr int SCAN_CHUNK = 4096;  // histogram entries per scan block

struct Tiling {
    int64_t n, seg;
    int64_t tiles_per_seg, num_segs, num_tiles;
};

Tiling make_tiling(int64_t n, int64_t seg) {
    Tiling t;
    t.n = n;
    t.seg = (seg <= 0 || seg > n) ? n : seg;
    t.tiles_per_seg = ceil_div(t.seg, RS_TILE);
    t.num_segs = ceil_div(n, t.seg);
    t.num_tiles = t.tiles_per_seg * t.num_segs;
    return t;
}

__device__ inline void tile_range(const Tiling &t, int64_t tile, int64_t &b, int64_t &e) {
    int64_t s = tile / t.tiles_per_seg, j = tile % t.tiles_per_seg;
    b = s * t.seg + j * RS_TILE;
    int64_t se = (s + 1) * t.seg;
    if (se > t.n) se = t.n;
    e = b + RS_TILE < se ? b + RS_TILE : se;
    if (b > e) b = e;
}

// Histogram layout, segment-major: entry (segment s, digit d, tile j of the
// segment) lives at (s * 256 + d) * tiles_per_seg + j.  A linear exclusive
// scan of this array then yields every item's output position directly:
// the prefix before segment s is the size of the segments before it.
__device__ inline int64_t hist_index(const Tiling &t, int64_t tile, int d) {
    const int64_t s = tile / t.tiles_per_seg, j = tile % t.tiles_per_seg;
    return (s * RS_BINS + d) * t.tiles_per_seg + j;
}

// hist[hist_index(tile, d)] = number of items of the tile with digit d
__global__ __launch_bounds__(RS_THREADS) void k_hist(const uint64_t *__restrict__ in, Tiling t,
                                                     int shift, uint32_t mask,
                                                     uint32_t *__restrict__ hist) {
    __shared__ uint32_t cnt[RS_WAVES][RS_BINS];
    const int tid = threadIdx.x, wave = tid >> 6;
    for (int i = tid; i < RS_WAVES * RS_BINS; i += RS_THREADS) (&cnt[0][0])[i] = 0;
    __syncthreads();
    int64_t b, e;
    tile_range(t, blockIdx.x, b, e);
    for (int64_t i = b + tid; i < e; i += RS_THREADS) {
        uint32_t d = static_cast<uint32_t>(in[i] >> shift) & mask;
        atomicAdd(&cnt[wave][d], 1u);
    }
    __syncthreads();
    if (tid < RS_BINS) {
        uint32_t s = 0;
        for (int w = 0; w < RS_WAVES; ++w) s += cnt[w][tid];
        hist[hist_index(t, blockIdx.x, tid)] = s;
    }
}

// Block-wide exclusive scan of one value per thread (blockDim.x == 1024).
__device__ inline uint32_t block_excl_scan_1024(uint32_t v, uint32_t *lds /*[32]*/, uint32_t &total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) lds[wave] = x;
    __syncthreads();
    if (wave == 0) {
        uint32_t w = lane < 16 ? lds[lane] : 0u;
        for (int o = 1; o < 16; o <<= 1) {
            uint32_t y = __shfl_up(w, o, 64);
            if (lane >= o) w += y;
        }
        if (lane < 16) lds[16 + lane] = w;
    }
    __syncthreads();
    total = lds[16 + 15];
    uint32_t before = wave > 0 ? lds[16 + wave - 1] : 0u;
    uint32_t r = before + x - v;
    __syncthreads();
    return r;
}

// Linear exclusive scan of the histogram: phase 1 reduces chunks of
// SCAN_CHUNK entries (coalesced), phase 2 scans the partials in one block,
// phase 3 rescans each chunk and writes the offsets.
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(const uint32_t *__restrict__ hist,
                                                              int64_t len,
                                                              uint32_t *__restrict__ partials) {
    __shared__ uint32_t lds[32];
    int64_t b = static_cast<int64_t>(blockIdx.x) * SCAN_CHUNK;
    uint32_t s = 0;
    for (int i = threadIdx.x; i < SCAN_CHUNK; i += SCAN_THREADS)
        if (b + i < len) s += hist[b + i];
    uint32_t total;
    block_excl_scan_1024(s, lds, total);
    if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_partials(uint32_t *__restrict__ partials,
                                                                int64_t np) {
    __shared__ uint32_t lds[32];
    // np <= SCAN_THREADS * 4 (checked on the host)
    uint32_t v[4], s = 0;
    for (int j = 0; j < 4; ++j) {
        int64_t k = threadIdx.x * 4 + j;
        v[j] = k < np ? partials[k] : 0u;
        s += v[j];
    }
    uint32_t total;
    uint32_t run = block_excl_scan_1024(s, lds, total);
    for (int j = 0; j < 4; ++j) {
        int64_t k = threadIdx.x * 4 + j;
        if (k < np) partials[k] = run;
        run += v[j];
    }
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_down(const uint32_t *__restrict__ hist,
                                                            int64_t len,
                                                            const uint32_t *__restrict__ partials,
                                                            uint32_t *__restrict__ offs) {
    __shared__ uint32_t lds[32];
    constexpr int PER = SCAN_CHUNK / SCAN_THREADS;
    int64_t b = static_cast<int64_t>(blockIdx.x) * SCAN_CHUNK + threadIdx.x * PER;
    uint32_t v[PER], s = 0;
    for (int j = 0; j < PER; ++j) {
        v[j] = b + j < len ? hist[b + j] : 0u;
        s += v[j];
    }
    uint32_t total;
    uint32_t run = block_excl_scan_1024(s, lds, total) + partials[blockIdx.x];
    for (int j = 0; j < PER; ++j) {
        if (b + j < len) offs[b + j] = run;
        run += v[j];
    }
}

// Stable scatter of one tile.  Wave w ranks the items [w*256, (w+1)*256) of
// the tile in input order; lane l of round j holds item w*256 + j*64 + l.
__global__ __launch_bounds__(RS_THREADS) void k_scatter(const uint64_t *__restrict__ in,
                                                        uint64_t *__restrict__ out, Tiling t,
                                                        int shift, int bits,
                                                        const uint32_t *__restrict__ offs) {
    __shared__ uint64_t stage[RS_TILE];
    __shared__ uint32_t wcnt[RS_WAVES][RS_BINS];
    __shared__ uint32_t dstart[RS_BINS];
    __shared__ uint32_t goff[RS_BINS];
    __shared__ uint32_t red[RS_WAVES];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t mask = (1u << bits) - 1u;
    int64_t b, e;
    tile_range(t, blockIdx.x, b, e);
    const int count = static_cast<int>(e - b);

    for (int i = tid; i < RS_WAVES * RS_BINS; i += RS_THREADS) (&wcnt[0][0])[i] = 0;
    if (tid < RS_BINS) goff[tid] = offs[hist_index(t, blockIdx.x, tid)];
    __syncthreads();

    uint64_t item[RS_ITEMS];
    uint32_t rank[RS_ITEMS];
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int j = 0; j < RS_ITEMS; ++j) {
        const int local = wave * (RS_ITEMS * 64) + j * 64 + lane;
        const bool valid = local < count;
        uint64_t v = valid ? in[b + local] : 0ull;
        uint32_t d = static_cast<uint32_t>(v >> shift) & mask;
        uint64_t peers = __ballot(valid);
        for (int k = 0; k < bits; ++k) {
            uint64_t bal = __ballot((d >> k) & 1u);
            peers &= ((d >> k) & 1u) ? bal : ~bal;
        }
        uint32_t base = 0;
        if (valid) base = wcnt[wave][d];
        uint32_t r = static_cast<uint32_t>(__popcll(peers & lt));
        __builtin_amdgcn_wave_barrier();
        if (valid && r == 0) wcnt[wave][d] = base + static_cast<uint32_t>(__popcll(peers));
        __builtin_amdgcn_wave_barrier();
        item[j] = v;
        rank[j] = valid ? base + r : 0xFFFFFFFFu;
    }
    __syncthreads();
    // per digit (threads 0..255): prefix over waves and total; then a
    // block-wide exclusive scan of the totals (other threads contribute 0)
    {
        uint32_t run = 0;
        if (tid < RS_BINS) {
            for (int w = 0; w < RS_WAVES; ++w) {
                uint32_t c = wcnt[w][tid];
                wcnt[w][tid] = run;
                run += c;
            }
        }
        uint32_t x = run;
        for (int o = 1; o < 64; o <<= 1) {
            uint32_t y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) red[wave] = x;
        __syncthreads();
        uint32_t before = 0;
        for (int w = 0; w < wave; ++w) before += red[w];
        if (tid < RS_BINS) dstart[tid] = before + x - run;
    }
    __syncthreads();
    for (int j = 0; j < RS_ITEMS; ++j) {
        if (rank[j] == 0xFFFFFFFFu) continue;
        uint32_t d = static_cast<uint32_t>(item[j] >> shift) & mask;
        stage[dstart[d] + wcnt[wave][d] + rank[j]] = item[j];
    }
    __syncthreads();
    for (int i = tid; i < count; i += RS_THREADS) {
        uint64_t v = stage[i];
        uint32_t d = static_cast<uint32_t>(v >> shift) & mask;
        out[goff[d] + (i - dstart[d])] = v;
    }
}

}  // namespace

size_t radix_ws_bytes(int64_t n, int64_t seg) {
    Tiling t = make_tiling(n > 0 ? n : 1, seg);
    Sizer s;
    s.take<uint32_t>(RS_BINS * t.num_tiles);
    s.take<uint32_t>(RS_BINS * t.num_tiles);
    s.take<uint32_t>(ceil_div(RS_BINS * t.num_tiles, SCAN_CHUNK) + 1);
    return s.bytes();
}

int radix_carve(Carver &c, int64_t n, int64_t seg, RadixWs &w) {
    Tiling t = make_tiling(n > 0 ? n : 1, seg);
    w.hist = c.take<uint32_t>(RS_BINS * t.num_tiles);
    w.offs = c.take<uint32_t>(RS_BINS * t.num_tiles);
    w.partials = c.take<uint32_t>(ceil_div(RS_BINS * t.num_tiles, SCAN_CHUNK) + 1);
    return c.ok() ? PRT_OK : PRT_E_NOMEM;
}

int radix_sort_u64(const uint64_t *in, uint64_t *buf0, uint64_t *buf1, int64_t n, int64_t seg,
                   int bit_lo, int bit_hi, const RadixWs &w, hipStream_t s, uint64_t **result) {
    PRT_REQUIRE(n >= 0 && n < (int64_t(1) << 32), PRT_E_INVAL, "radix sort: n=%lld out of range",
                (long long)n);
    PRT_REQUIRE(bit_lo >= 0 && bit_hi <= 64 && bit_lo < bit_hi, PRT_E_INVAL,
                "radix sort: bad bit range [%d,%d)", bit_lo, bit_hi);
    if (n == 0) {
        *result = buf0;
        return PRT_OK;
    }
    Tiling t = make_tiling(n, seg);
    const int64_t len = RS_BINS * t.num_tiles;
    const int64_t np = ceil_div(len, SCAN_CHUNK);
    PRT_REQUIRE(np <= SCAN_THREADS * 4, PRT_E_UNSUPPORTED,
                "radix sort: n=%lld too large for the two-level scan", (long long)n);
    const uint64_t *src = in;
    uint64_t *dst = (in == buf0) ? buf1 : buf0;
    for (int shift = bit_lo; shift < bit_hi; shift += 8) {
        int bits = bit_hi - shift < 8 ? bit_hi - shift : 8;
        uint32_t mask = (1u << bits) - 1u;
        PRT_LAUNCH(k_hist, dim3(static_cast<unsigned>(t.num_tiles)), dim3(RS_THREADS), 0, s, src, t,
                   shift, mask, w.hist);
        PRT_LAUNCHED();
        PRT_LAUNCH(k_scan_reduce, dim3(static_cast<unsigned>(np)), dim3(SCAN_THREADS), 0, s, w.hist,
                   len, w.partials);
        PRT_LAUNCH(k_scan_partials, 1, SCAN_THREADS, 0, s, w.partials, np);
        PRT_LAUNCH(k_scan_down, dim3(static_cast<unsigned>(np)), dim3(SCAN_THREADS), 0, s, w.hist,
                   len, w.partials, w.offs);
        PRT_LAUNCHED();
        PRT_LAUNCH(k_scatter, dim3(static_cast<unsigned>(t.num_tiles)), dim3(RS_THREADS), 0, s, src,
                   dst, t, shift, bits, w.offs);
        PRT_LAUNCHED();
        src = dst;
        dst = (dst == buf0) ? buf1 : buf0;
    }
    *result = const_cast<uint64_t *>(src);
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
