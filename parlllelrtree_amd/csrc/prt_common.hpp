// prt_common.hpp -- shared host/device helpers of libprt (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/prt.h"

namespace prt {

// ---- error reporting (per host thread) ---------------------------------
void set_error(const char *fmt, ...);
int hip_fail(hipError_t e, const char *what, const char *file, int line);

#define PRT_HIP(call)                                                          \
    do {                                                                       \
        hipError_t _e = (call);                                                \
        if (_e != hipSuccess) return ::prt::hip_fail(_e, #call, __FILE__, __LINE__); \
    } while (0)

#define PRT_LAUNCHED()                                                         \
    do {                                                                       \
        hipError_t _e = hipGetLastError();                                     \
        if (_e != hipSuccess) return ::prt::hip_fail(_e, "kernel launch", __FILE__, __LINE__); \
    } while (0)

#define PRT_REQUIRE(cond, code, ...)                                           \
    do {                                                                       \
        if (!(cond)) {                                                         \
            ::prt::set_error(__VA_ARGS__);                                     \
            return code;                                                       \
        }                                                                      \
    } while (0)

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Address of the current device's status flag word (PRT_STATUS_* bits),
// passed to the kernels that raise them; null if it cannot be resolved.
uint32_t *device_status_flags();

// ---- workspace carving: 256-byte aligned sub-buffers --------------------
struct Carver {
    char *base;
    size_t used = 0;
    size_t cap;
    Carver(void *p, size_t bytes) : base(static_cast<char *>(p)), cap(bytes) {}
    template <class T> T *take(size_t count) {
        size_t off = (used + 255) & ~size_t(255);
        size_t need = off + count * sizeof(T);
        used = need;
        if (base == nullptr || need > cap) return nullptr;
        return reinterpret_cast<T *>(base + off);
    }
    bool ok() const { return base != nullptr && used <= cap; }
};
// size-only twin of Carver, used by the *_workspace_size functions
struct Sizer {
    size_t used = 0;
    template <class T> void take(size_t count) {
        used = ((used + 255) & ~size_t(255)) + count * sizeof(T);
    }
    size_t bytes() const { return used + 256; }
};

__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// A count produced on the device by an earlier launch, clamped to [0, cap];
// cap itself when n_d is null.
__device__ inline int64_t device_count(const int64_t *n_d, int64_t cap) {
    if (!n_d) return cap;
    const int64_t v = *n_d;
    return v < 0 ? 0 : (v < cap ? v : cap);
}

// ---- key helpers (host and device) --------------------------------------

// Order-preserving fp32 -> u32; -0 folds onto +0.
__host__ __device__ inline uint32_t ord32(float v) {
    uint32_t u = __builtin_bit_cast(uint32_t, v);
    if (u == 0x80000000u) u = 0u;
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// Inverse of ord32 (a zero key gives +0).
__host__ __device__ inline float unord32(uint32_t k) {
    const uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    return __builtin_bit_cast(float, u);
}

// Centre quantisation of one axis: floor(0.5f*(lo+hi) * 2^15) in [0, 32767].
__host__ __device__ inline uint32_t quant15(float lo, float hi) {
    float c = 0.5f * (lo + hi);
    float s = c * 32768.0f;
    if (!(s >= 0.0f)) return 0u;
    if (s >= 32767.0f) return 32767u;
    return static_cast<uint32_t>(s);
}

__host__ __device__ inline uint32_t spread15(uint32_t x) {
    x &= 0x7FFFu;
    x = (x | (x << 8)) & 0x00FF00FFu;
    x = (x | (x << 4)) & 0x0F0F0F0Fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}

__host__ __device__ inline uint32_t morton_center30(float4 r) {
    return spread15(quant15(r.x, r.z)) | (spread15(quant15(r.y, r.w)) << 1);
}

// Closed-interval MBR overlap (touching counts).
__host__ __device__ inline bool overlap(float4 a, float4 b) {
    return a.x <= b.z && b.x <= a.z && a.y <= b.w && b.y <= a.w;
}

__host__ __device__ inline float4 empty_box() {
    return make_float4(__builtin_huge_valf(), __builtin_huge_valf(), -__builtin_huge_valf(),
                       -__builtin_huge_valf());
}

// The rects of up to PRT_MAX_LAYERS layers of `len` rects each, read as one
// array: position i is rect i - l*len of layer l (l = i / len, found by
// compares and selects, so that no pointer is indexed by a lane value).
struct RectLayers {
    const float4 *p[PRT_MAX_LAYERS];
    int64_t len;
    int count;
    __device__ float4 at(int64_t i) const {
        const float4 *b = p[0];
        int64_t o = i;
#pragma unroll
        for (int l = 1; l < PRT_MAX_LAYERS; ++l)
            if (l < count && i >= l * len) {
                b = p[l];
                o = i - l * len;
            }
        return b[o];
    }
};
inline RectLayers one_layer(const float4 *r, int64_t n) {
    RectLayers L{};
    L.p[0] = r;
    L.len = n;
    L.count = 1;
    return L;
}

__host__ __device__ inline float4 box_union(float4 a, float4 b) {
    return make_float4(fminf(a.x, b.x), fminf(a.y, b.y), fmaxf(a.z, b.z), fmaxf(a.w, b.w));
}

// ---- wave64 digit match (stable radix ranking) --------------------------
// Lanes of the wave that do NOT share this lane's 8-bit digit d, or are not
// valid (inactive lanes count as not valid).  Per digit bit: the bit
// sign-extended to 0 / -1, one compare for the ballot, and the ballot XORed
// with it into the mismatch mask (a lane with the bit set mismatches the
// lanes whose ballot bit is clear, and the other way round).  The lanes
// equal to this one are the complement; their number below this lane is
// lane - digit_lanes_below(mismatch), their count 64 - popc(mismatch).
__device__ inline uint64_t digit_mismatch(uint32_t d, bool valid) {
    const uint64_t inval = ~__ballot(valid);
    uint32_t lo = static_cast<uint32_t>(inval), hi = static_cast<uint32_t>(inval >> 32);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const int32_t m = static_cast<int32_t>(d << (31 - b)) >> 31;
        const uint64_t bal = __ballot(m != 0);
        lo |= static_cast<uint32_t>(bal) ^ static_cast<uint32_t>(m);
        hi |= static_cast<uint32_t>(bal >> 32) ^ static_cast<uint32_t>(m);
    }
    return (static_cast<uint64_t>(hi) << 32) | lo;
}
// set bits of `mask` in the lanes below this one
__device__ inline uint32_t digit_lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

// Inclusive prefix sum over the 64 lanes of a wave (every lane active), by
// data-parallel-primitive moves instead of LDS permutes: within each row of
// 16 lanes add the lanes 1, 2 and 3 below (a 4-lane prefix), then 4 below
// for the upper three banks and 8 below for the upper two, then broadcast
// lane 15 into rows 1 and 3 and lane 31 into rows 2 and 3.
template <int CTRL, int ROW, int BANK>
__device__ inline uint32_t dpp_src(uint32_t v) {
    return static_cast<uint32_t>(
        __builtin_amdgcn_update_dpp(0, static_cast<int>(v), CTRL, ROW, BANK, true));
}
__device__ inline uint32_t wave_incl_scan(uint32_t x) {
    uint32_t v = x + dpp_src<0x111, 0xF, 0xF>(x);   // row_shr:1
    v += dpp_src<0x112, 0xF, 0xF>(x);               // row_shr:2
    v += dpp_src<0x113, 0xF, 0xF>(x);               // row_shr:3
    v += dpp_src<0x114, 0xF, 0xE>(v);               // row_shr:4, banks 1-3
    v += dpp_src<0x118, 0xF, 0xC>(v);               // row_shr:8, banks 2-3
    v += dpp_src<0x142, 0xA, 0xF>(v);               // row_bcast:15 into rows 1, 3
    v += dpp_src<0x143, 0xC, 0xF>(v);               // row_bcast:31 into rows 2, 3
    return v;
}

// ---- polygon set operations (gpc_op order: DIFF, INT, XOR, UNION) ------
// A gap between neighbouring edges belongs to the result set by the even-odd
// parities of A (subject, pp) and B (clip, pq).
__host__ __device__ inline bool op_inside(int op, int pp, int pq) {
    switch (op) {
    case PRT_OP_DIFF: return pp && !pq;
    case PRT_OP_XOR: return pp != pq;
    case PRT_OP_UNION: return pp || pq;
    default: return pp && pq;
    }
}

// y range the sweep of `op` covers, from the bounding boxes (xmin, ymin,
// xmax, ymax) of A and B; eP / eQ: the polygon has fewer than 3 vertices
// and is empty.  Returns false when the result is empty.
__host__ __device__ inline bool op_range(int op, bool eP, bool eQ, const double *bp,
                                         const double *bq, double &ylo, double &yhi) {
    if (op == PRT_OP_INT) {
        if (eP || eQ) return false;
        if (bp[2] < bq[0] || bq[2] < bp[0] || bp[3] < bq[1] || bq[3] < bp[1]) return false;
        ylo = bp[1] > bq[1] ? bp[1] : bq[1];
        yhi = bp[3] < bq[3] ? bp[3] : bq[3];
    } else if (op == PRT_OP_DIFF || eQ) {
        if (eP) return false;
        ylo = bp[1];
        yhi = bp[3];
    } else if (eP) {
        ylo = bq[1];
        yhi = bq[3];
    } else {
        ylo = bp[1] < bq[1] ? bp[1] : bq[1];
        yhi = bp[3] > bq[3] ? bp[3] : bq[3];
    }
    return ylo < yhi;
}

// ---- internal launchers shared between translation units ---------------
// Stable LSD radix sort of 64-bit items on bits [bit_lo, bit_hi), segmented:
// items never leave their segment [s*seg, (s+1)*seg).  Returns the buffer
// holding the result (one of buf0/buf1; `in` may alias buf0).
// Keys of at most 32 bits (bit_hi - bit_lo <= 32).  With layer > 0 the n
// items are n / layer layers of `layer` items, each segmented on its own
// (segments [l*layer + s*seg, l*layer + min((s+1)*seg, layer))).
struct RadixWs {
    uint32_t *ghist;     // per pass, segment and digit: item counts; + tickets
    uint64_t *status;    // per pass, tile and digit: look-back words
    int64_t ghist_cap, status_cap;  // words carved
};
size_t radix_ws_bytes(int64_t n, int64_t seg);
int radix_carve(Carver &c, int64_t n, int64_t seg, RadixWs &w);
int radix_sort_u64(const uint64_t *in, uint64_t *buf0, uint64_t *buf1, int64_t n, int64_t seg,
                   int bit_lo, int bit_hi, const RadixWs &w, hipStream_t s, uint64_t **result);
// The same sort of items made from rects by its histogram launch (written to
// buf0 there): (ord32(x1 or y1) << 32 | position), segmented, bits 32..63;
// or (morton_center30 << 32 | position), bits 32..61.  With zeroed = true the
// caller has already cleared the sort's counter words (radix_ghist_words).
// With items set (buf0 or buf1), the items are there already, written by the
// last digit pass of the level before.  With rects_out set, the last digit
// pass writes, at every item's sorted position, its rect and id (ids[i], or
// i when ids is null; i is the item's low word, the rect's index in `rects`)
// and, in place of the sorted item, the item of the next level: (ord32 of
// the other coordinate of the rect << 32 | position).  With next_items_only
// instead, it writes only the next level's items, whose low word stays the
// rect's index in `rects`.
int radix_sort_td_coord(const RectLayers &rects, int comp, const uint64_t *items, uint64_t *buf0,
                        uint64_t *buf1, int64_t n, int64_t seg, int64_t layer, const RadixWs &w,
                        hipStream_t s, uint64_t **result, bool zeroed,
                        const uint32_t *ids = nullptr, float4 *rects_out = nullptr,
                        uint32_t *ids_out = nullptr, bool next_items_only = false);
int radix_sort_morton(const float4 *rects, uint64_t *buf0, uint64_t *buf1, int64_t n,
                      const RadixWs &w, hipStream_t s, uint64_t **result);
// Counter words (cleared before a sort unless it clears them itself) and
// look-back words of a sort of n items in segments of seg.
int64_t radix_ghist_words(int64_t n, int64_t seg, int64_t layer = 0);
int64_t radix_status_words(int64_t n, int64_t seg, int64_t layer = 0);
// Stable sort by 64-bit keys (two 32-bit passes); items[k] = (key >> 32 << 32)
// | id of the k-th rect, keys_sorted[k] (optional) its key.  b0..b3 and items
// are distinct n-entry buffers.
int radix_sort_keys64(const uint64_t *keys, int64_t n, uint64_t *b0, uint64_t *b1, uint64_t *b2,
                      uint64_t *b3, const RadixWs &w, hipStream_t s, uint64_t *items,
                      uint64_t *keys_sorted);

// Exclusive scan of int64 counts; out[i] = sum_{j<i} in[j]; *total = sum.
size_t scan_i64_ws_bytes(int64_t n);
int scan_i64(const int64_t *in, int64_t *out, int64_t n, int64_t *total, void *ws,
             hipStream_t s);

}  // namespace prt
