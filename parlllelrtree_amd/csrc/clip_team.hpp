// clip_team.hpp -- the per-pair tables a team of TL lanes (a slice of one
// wave64) builds in LDS for the batched overlay kernels (the FP64 area in
// clip_batched.hip, the result contours in contours.hip), and the vertex-
// count bucketing of the pairs that both launch per bucket.
//
// The tables restate GPC's edge and scanbeam tables
// (vector_map_overlay_gpc_kernel.cu:1019-1085, 1162-1307) with the plain
// kernels' FP64 expressions: edges bottom to top with dx = (top.x - bot.x) /
// (top.y - bot.y), horizontal edges dropped, the distinct vertex heights of
// the op's y range, ascending; every edge registers itself in the lists of
// the beams it spans, with its x at the beam bottom and top.
#pragma once

#include "prt_common.hpp"

namespace prt {
namespace {

constexpr int BK_THREADS_MAX = 512;
constexpr int NBUCKET = PRT_TUNE_BUCKETS;

__device__ inline void team_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// A team of TL lanes is a slice of one wave (TL <= 64) or TL / 64 whole
// waves that make up one block (TL = 128, 256): its barrier is the wave's
// or the block's.
template <int TL> __device__ inline void tsync() {
    if constexpr (TL > 64) __syncthreads();
    else team_sync();
}

struct TeamLds {
    double *bx, *by, *tx, *ty, *dx;  // [NV] edge slot = vertex index (P: i, Q: nP + j)
    uint16_t *ekb, *ekt;             // [NV] beams [ekb, ekt) spanned by the edge
    double *ys;                      // [P2] scanbeam table (sorted, then unique)
    double *bsum;                    // [NV] beam sums (also the unique-compaction buffer)
    int32_t *cnt;                    // [NV + 1]
    int32_t *boff;                   // [NV + 1]
    uint16_t *list;                  // [SCAP] active edges of the beams of a window
    double *lxb, *lxt;               // [SCAP] their x at the beam bottom and top
    uint16_t *beam_of;               // [SCAP] beam of each list entry (then a beam queue)
    uint16_t *tlist;                 // [SCAP] the entries' edge ids in sorted order
    // team-cooperative sweep of beams with crossings (GS entries, GC heights)
    double *gxa, *gxc;               // [GS] x of the entries at the sub-beam ends
    uint16_t *gidx;                  // [GS] their edge ids
    double *giy;                     // [GC] crossing heights of the beam
    int32_t *gmisc;                  // [4] deferred-beam count, crossing count
    int GS, GC;
    // teams of several waves: the waves' partial reductions (block scratch)
    double *red;                     // [8 * waves]
    int32_t *ired;                   // [waves + 1]
    double *wgiy;                    // [64 * waves] crossing heights, 64 per wave
};

__host__ __device__ inline size_t align16(size_t v) { return (v + 15) & ~size_t(15); }

__host__ __device__ inline int pow2_ge(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

// entries (and crossing heights) of one beam the team sweeps cooperatively;
// 16 for the smallest pairs (up to 32 vertices, whose teams are limited by
// LDS: their beams rarely hold more, and the register sweep of small teams
// beat the serial one there once it existed, grid overlay 97.6 -> 94.8 ms).
// At most 64: a beam holds few active edges (jittered
// grid of lognormal stars, 82,241 beams of 3,000 pairs: 7.1 on average, 24 at
// the 99th percentile, 42 at most), and a beam with more entries or crossing
// heights takes the serial sweep (same sums), so the LDS a larger scratch
// would cost every team goes to more teams per CU instead.
__host__ __device__ inline int gen_cap(int NV) {
    if (NV <= 32) return 16;
    const int p = pow2_ge(NV);
    return p < 64 ? p : 64;
}

__host__ __device__ inline size_t team_bytes(int NV, int SCAP) {
    const int P2 = pow2_ge(NV);
    const int G = gen_cap(NV);
    return align16(5 * 8 * size_t(NV)) + 2 * align16(2 * size_t(NV)) + align16(8 * size_t(P2)) +
           align16(8 * size_t(NV)) + 2 * align16(4 * size_t(NV + 1)) + align16(2 * size_t(SCAP)) +
           2 * align16(8 * size_t(SCAP)) + 2 * align16(2 * size_t(SCAP)) +
           2 * align16(8 * size_t(G)) + align16(2 * size_t(G)) + align16(8 * size_t(G)) +
           align16(4 * 4);
}

__device__ inline TeamLds carve_team(char *p, int NV, int SCAP) {
    TeamLds t;
    const int P2 = pow2_ge(NV);
    t.bx = reinterpret_cast<double *>(p);
    t.by = t.bx + NV;
    t.tx = t.by + NV;
    t.ty = t.tx + NV;
    t.dx = t.ty + NV;
    p += align16(5 * 8 * size_t(NV));
    t.ekb = reinterpret_cast<uint16_t *>(p);
    p += align16(2 * size_t(NV));
    t.ekt = reinterpret_cast<uint16_t *>(p);
    p += align16(2 * size_t(NV));
    t.ys = reinterpret_cast<double *>(p);
    p += align16(8 * size_t(P2));
    t.bsum = reinterpret_cast<double *>(p);
    p += align16(8 * size_t(NV));
    t.cnt = reinterpret_cast<int32_t *>(p);
    p += align16(4 * size_t(NV + 1));
    t.boff = reinterpret_cast<int32_t *>(p);
    p += align16(4 * size_t(NV + 1));
    t.list = reinterpret_cast<uint16_t *>(p);
    p += align16(2 * size_t(SCAP));
    t.lxb = reinterpret_cast<double *>(p);
    p += align16(8 * size_t(SCAP));
    t.lxt = reinterpret_cast<double *>(p);
    p += align16(8 * size_t(SCAP));
    t.beam_of = reinterpret_cast<uint16_t *>(p);
    p += align16(2 * size_t(SCAP));
    t.tlist = reinterpret_cast<uint16_t *>(p);
    p += align16(2 * size_t(SCAP));
    const int G = gen_cap(NV);
    t.GS = G;
    t.GC = G;
    t.gxa = reinterpret_cast<double *>(p);
    p += align16(8 * size_t(G));
    t.gxc = reinterpret_cast<double *>(p);
    p += align16(8 * size_t(G));
    t.gidx = reinterpret_cast<uint16_t *>(p);
    p += align16(2 * size_t(G));
    t.giy = reinterpret_cast<double *>(p);
    p += align16(8 * size_t(G));
    t.gmisc = reinterpret_cast<int32_t *>(p);
    t.red = nullptr;
    t.ired = nullptr;
    t.wgiy = nullptr;
    return t;
}

__device__ inline double ex(const TeamLds &t, int e, double y) {
    if (y == t.ty[e]) return t.tx[e];
    return t.bx[e] + t.dx[e] * (y - t.by[e]);
}

// first index k in [0, n) with ys[k] >= v (n if none)
__device__ inline int lower_bound(const double *ys, int n, double v) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (ys[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// v of lane (this lane ^ J), the partner lane active: within a quad and
// across half a row a DPP move, within 32 lanes a swizzle, across the wave's
// halves a permute
template <int J> __device__ inline uint32_t xor_lane(uint32_t v) {
    const int x = static_cast<int>(v);
    if constexpr (J == 1) return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, false));
    else if constexpr (J == 2) return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, x, 0x4E, 0xF, 0xF, false));
    else if constexpr (J == 4) return static_cast<uint32_t>(__builtin_amdgcn_ds_swizzle(x, 0x101F));
    else if constexpr (J == 8) return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, x, 0x128, 0xF, 0xF, false));
    else if constexpr (J == 16) return static_cast<uint32_t>(__builtin_amdgcn_ds_swizzle(x, 0x401F));
    else return static_cast<uint32_t>(__shfl_xor(x, 32, 64));
}
template <int J> __device__ inline double xor_lane_d(double v) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = xor_lane<J>(static_cast<uint32_t>(u)), hi = xor_lane<J>(static_cast<uint32_t>(u >> 32));
    return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}

// team-wide min/max over TL lanes (butterfly inside the team's lanes, by
// the lane exchanges above: DPP moves and swizzles, a permute only across
// the wave's halves)
template <int J> __device__ inline double bfly_min(double v) { return fmin(v, xor_lane_d<J>(v)); }
template <int J> __device__ inline double bfly_max(double v) { return fmax(v, xor_lane_d<J>(v)); }
template <int TL> __device__ inline double team_min(double v) {
    if constexpr (TL > 1) v = bfly_min<1>(v);
    if constexpr (TL > 2) v = bfly_min<2>(v);
    if constexpr (TL > 4) v = bfly_min<4>(v);
    if constexpr (TL > 8) v = bfly_min<8>(v);
    if constexpr (TL > 16) v = bfly_min<16>(v);
    if constexpr (TL > 32) v = bfly_min<32>(v);
    return v;
}
template <int TL> __device__ inline double team_max(double v) {
    if constexpr (TL > 1) v = bfly_max<1>(v);
    if constexpr (TL > 2) v = bfly_max<2>(v);
    if constexpr (TL > 4) v = bfly_max<4>(v);
    if constexpr (TL > 8) v = bfly_max<8>(v);
    if constexpr (TL > 16) v = bfly_max<16>(v);
    if constexpr (TL > 32) v = bfly_max<32>(v);
    return v;
}
// inclusive scan inside the team
template <int TL> __device__ inline int team_incl_scan(int v, int tl) {
#pragma unroll
    for (int o = 1; o < TL; o <<= 1) {
        int y = __shfl_up(v, o, TL);
        if (tl >= o) v += y;
    }
    return v;
}
template <int TL> __device__ inline int team_bcast(int v, int src_lane_in_team, int team_base) {
    return __shfl(v, team_base + src_lane_in_team, 64);
}

// The team's minima of v[0, 4) and maxima of v[4, 8), on every lane: by
// butterflies inside the team's lanes of one wave, then, for teams of
// several waves, over the waves' results in LDS.
template <int TL> __device__ inline void team_minmax8(const TeamLds &t, double (&v)[8], int tl) {
    constexpr int WL = TL < 64 ? TL : 64;
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = team_min<WL>(v[i]);
#pragma unroll
    for (int i = 4; i < 8; ++i) v[i] = team_max<WL>(v[i]);
    if constexpr (TL > 64) {
        constexpr int W = TL / 64;
        const int w = tl >> 6;
        if ((tl & 63) == 0) {
#pragma unroll
            for (int i = 0; i < 8; ++i) t.red[8 * w + i] = v[i];
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            double r = t.red[i];
            for (int u = 1; u < W; ++u) r = i < 4 ? fmin(r, t.red[8 * u + i]) : fmax(r, t.red[8 * u + i]);
            v[i] = r;
        }
        __syncthreads();  // the scratch is free again
    }
}

// Inclusive scan of v over the team's lanes (in lane order) and its total,
// on every lane.
template <int TL> __device__ inline int team_scan(const TeamLds &t, int v, int tl, int base,
                                                 int &total) {
    if constexpr (TL <= 64) {
        const int incl = team_incl_scan<TL>(v, tl);
        total = team_bcast<TL>(incl, TL - 1, base);
        return incl;
    } else {
        constexpr int W = TL / 64;
        const int w = tl >> 6, lane = tl & 63;
        int incl = team_incl_scan<64>(v, lane);
        if (lane == 63) t.ired[w] = incl;
        __syncthreads();
        int before = 0, all = 0;
        for (int u = 0; u < W; ++u) {
            const int c = t.ired[u];
            if (u < w) before += c;
            all += c;
        }
        __syncthreads();  // the scratch is free again
        total = all;
        return incl + before;
    }
}

// Strict order (xa + xc, xa, id) inside a sub-beam.
__device__ inline bool key_before(double ka, double xa_a, int ea, double kb, double xa_b, int eb) {
    if (ka != kb) return ka < kb;
    if (xa_a != xa_b) return xa_a < xa_b;
    return ea < eb;
}

// The list entries of every beam of the window [k0, k1) sorted by the
// sub-beam key (xb + xt, xb, edge id), by the whole team at once: every
// entry counts the entries of its beam that come before it (the key is a
// strict order, so the counts are a permutation) and its edge id goes to
// that place with its x's at the beam ends -- moved from registers in
// windows of up to 2 TL entries, else computed again for the new order (the
// expressions of the list fill, so the same values either way).  Then
// cnt[k] = 1 for a beam whose sorted xb or xt decreases somewhere (edges
// crossing inside the beam), else 0.  Syncs the team before and after.
template <int TL> __device__ void team_sort_window(const TeamLds &t, int k0, int k1, int tl) {
    const int wbase = t.boff[k0], S = t.boff[k1] - wbase;
    tsync<TL>();
    // entries before entry i of its beam: every operand loaded unconditionally
    // and compared without branches, four entries per step, so that the
    // loads of a step are in flight together
    auto rank_of = [&](int i, double xb, double xt, int e) -> int {
        const int k = t.beam_of[i];
        const int lo = t.boff[k] - wbase, hi = t.boff[k + 1] - wbase;
        const double key = xb + xt;
        auto before = [&](double xbj, double xtj, int ej) -> int {
            const double kj = xbj + xtj;
            return static_cast<int>((kj < key) | ((kj == key) & ((xbj < xb) | ((xbj == xb) & (ej < e)))));
        };
        int r = 0, j = lo;
        for (; j + 4 <= hi; j += 4) {
            const double b0 = t.lxb[j], b1 = t.lxb[j + 1], b2 = t.lxb[j + 2], b3 = t.lxb[j + 3];
            const double c0 = t.lxt[j], c1 = t.lxt[j + 1], c2 = t.lxt[j + 2], c3 = t.lxt[j + 3];
            const int e0 = t.list[j], e1 = t.list[j + 1], e2 = t.list[j + 2], e3 = t.list[j + 3];
            r += before(b0, c0, e0) + before(b1, c1, e1) + before(b2, c2, e2) + before(b3, c3, e3);
        }
        for (; j < hi; ++j) r += before(t.lxb[j], t.lxt[j], t.list[j]);
        return lo + r;
    };
    if (S <= 2 * TL) {
        // at most two entries per lane: ranked, then (once every lane has
        // read the lists) moved to their places with their x's -- the
        // values the list fill computed, which ex() would compute again
        int pos[2] = {-1, -1}, ev[2] = {0, 0};
        double bv[2] = {0.0, 0.0}, tv[2] = {0.0, 0.0};
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int i = tl + r * TL;
            if (i < S) {
                ev[r] = t.list[i];
                bv[r] = t.lxb[i];
                tv[r] = t.lxt[i];
                pos[r] = rank_of(i, bv[r], tv[r], ev[r]);
            }
        }
        for (int k = k0 + tl; k < k1; k += TL) t.cnt[k] = 0;
        tsync<TL>();
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            if (pos[r] >= 0) {
                t.list[pos[r]] = static_cast<uint16_t>(ev[r]);
                t.lxb[pos[r]] = bv[r];
                t.lxt[pos[r]] = tv[r];
            }
        }
        tsync<TL>();
    } else {
        for (int i = tl; i < S; i += TL)
            t.tlist[rank_of(i, t.lxb[i], t.lxt[i], t.list[i])] = t.list[i];
        for (int k = k0 + tl; k < k1; k += TL) t.cnt[k] = 0;
        tsync<TL>();
        for (int i = tl; i < S; i += TL) {
            const int k = t.beam_of[i], e = t.tlist[i];
            t.list[i] = static_cast<uint16_t>(e);
            t.lxb[i] = ex(t, e, t.ys[k]);
            t.lxt[i] = ex(t, e, t.ys[k + 1]);
        }
        tsync<TL>();
    }
    for (int i = tl; i < S; i += TL) {
        const int k = t.beam_of[i];
        if (i > t.boff[k] - wbase && (t.lxb[i] < t.lxb[i - 1] || t.lxt[i] < t.lxt[i - 1]))
            t.cnt[k] = 1;
    }
    tsync<TL>();
}

// A double of lane `lane` (uniform) on every lane.
__device__ inline double readlane_d(double v, int lane) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = __builtin_amdgcn_readlane(static_cast<uint32_t>(u), lane);
    const uint32_t hi = __builtin_amdgcn_readlane(static_cast<uint32_t>(u >> 32), lane);
    return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}

// Teams of one wave: the entries [lo, lo + m) of a beam's list (m <= 64)
// with their x at the sub-beam ends ya and yc (the expressions of the serial
// sweep), sorted by the strict key (xa + xc, xa, edge id) into gxa / gxc /
// gidx[0, m): lane q holds entry q and counts the entries before it (read
// from the other lanes); the key is a strict order, so the ranks are a
// permutation.  The order the bitonic sort and the serial insertion sort
// give.  The caller syncs the team before reading the result.
__device__ inline void wave_sort_sub_beam(const TeamLds &t, int lo, int m, double yb, double yt,
                                          double ya, double yc, int q) {
    double a = 0.0, c = 0.0;
    int e = 0;
    if (q < m) {
        const int i = lo + q;
        a = ya == yb ? t.lxb[i] : (ya == yt ? t.lxt[i] : ex(t, t.list[i], ya));
        c = yc == yb ? t.lxb[i] : (yc == yt ? t.lxt[i] : ex(t, t.list[i], yc));
        e = t.list[i];
    }
    const double k = a + c;
    int rank = 0;
    for (int p = 0; p < m; ++p) {
        const double kp = readlane_d(k, p), ap = readlane_d(a, p);
        const int ep = __builtin_amdgcn_readlane(e, p);
        rank += key_before(kp, ap, ep, k, a, e) ? 1 : 0;
    }
    if (q < m) {
        t.gxa[rank] = a;
        t.gxc[rank] = c;
        t.gidx[rank] = static_cast<uint16_t>(e);
    }
}

// Teams of one wave: the crossing heights in (yb, yt) of the beam whose
// list entries are [lo, hi) -- the serial sweep's candidate set, from the
// same pairs (i, j), i < j, with the same expressions -- written to
// giy[0, min(nc, GC)) in no particular order; returns nc on every lane.  The
// pairs are spread over the lanes (pair index j (j - 1) / 2 + i), and each
// step's heights are compacted by a ballot prefix.  Every lane active.
__device__ inline int wave_crossing_heights(const TeamLds &t, int lo, int hi, double yb, double yt,
                                            int q) {
    const int m = hi - lo;
    const double dy = yt - yb;
    const int npairs = m * (m - 1) / 2;
    int nc = 0;
    for (int t0 = 0; t0 < npairs; t0 += 64) {
        const int tt = t0 + q;
        bool has = false;
        double iy = 0.0;
        if (tt < npairs) {
            int j = static_cast<int>((1.0f + __builtin_sqrtf(1.0f + 8.0f * static_cast<float>(tt))) * 0.5f);
            while (j * (j - 1) / 2 > tt) --j;
            while ((j + 1) * j / 2 <= tt) ++j;
            const int i = tt - j * (j - 1) / 2;
            const double xbi = t.lxb[lo + i], xti = t.lxt[lo + i];
            const double xbj = t.lxb[lo + j], xtj = t.lxt[lo + j];
            double xbL = 0.0, xtL = 0.0, xbR = 0.0, xtR = 0.0;
            bool cr = true;
            if (xbi < xbj && xti > xtj) { xbL = xbi; xtL = xti; xbR = xbj; xtR = xtj; }
            else if (xbj < xbi && xtj > xti) { xbL = xbj; xtL = xtj; xbR = xbi; xtR = xti; }
            else cr = false;
            if (cr) {
                const double den = (xtL - xbL) - (xtR - xbR);
                if (den > 0.0) {
                    const double r = (xbR - xbL) / den;
                    iy = r * dy + yb;
                    has = iy > yb && iy < yt;
                }
            }
        }
        const uint64_t hm = __ballot(has);
        if (has) {
            const int pos = nc + __popcll(hm & ((1ull << q) - 1ull));
            if (pos < t.GC) t.giy[pos] = iy;
        }
        nc += __popcll(hm);
    }
    return nc;
}

// Teams of one wave: giy[0, n) ascending (n <= 64), by ranks as above
// (equal heights keep their order).  The caller syncs the team afterwards.
__device__ inline void wave_sort_heights(const TeamLds &t, int n, int q) {
    const double h = q < n ? t.giy[q] : 0.0;
    int rank = 0;
    for (int p = 0; p < n; ++p) {
        const double hp = readlane_d(h, p);
        rank += (hp < h || (hp == h && p < q)) ? 1 : 0;
    }
    team_sync();  // every lane has read its height
    if (q < n) t.giy[rank] = h;
}

// team-wide bitonic sort of M2 (power of two) entries (gxa, gxc, gidx) by
// the strict key (xa + xc, xa, id), ascending
template <int TL> __device__ inline void team_sort_entries(const TeamLds &t, int M2, int tl) {
    for (int k = 2; k <= M2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int q = tl; q < M2 / 2; q += TL) {
                const int i = ((q & ~(j - 1)) << 1) | (q & (j - 1));
                const int l = i | j;
                const bool up = (i & k) == 0;
                const double ai = t.gxa[i], ci = t.gxc[i], al = t.gxa[l], cl = t.gxc[l];
                const int ei = t.gidx[i], el = t.gidx[l];
                // swap when the pair is out of order for this direction
                const bool l_first = key_before(al + cl, al, el, ai + ci, ai, ei);
                const bool i_first = key_before(ai + ci, ai, ei, al + cl, al, el);
                if (up ? l_first : i_first) {
                    t.gxa[i] = al; t.gxc[i] = cl; t.gidx[i] = static_cast<uint16_t>(el);
                    t.gxa[l] = ai; t.gxc[l] = ci; t.gidx[l] = static_cast<uint16_t>(ei);
                }
            }
            team_sync();
        }
    }
}

// team-wide bitonic sort of giy[0, M2) ascending
template <int TL> __device__ inline void team_sort_heights(const TeamLds &t, int M2, int tl) {
    for (int k = 2; k <= M2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int q = tl; q < M2 / 2; q += TL) {
                const int i = ((q & ~(j - 1)) << 1) | (q & (j - 1));
                const int l = i | j;
                const bool up = (i & k) == 0;
                const double x = t.giy[i], y = t.giy[l];
                if ((x > y) == up) {
                    t.giy[i] = y;
                    t.giy[l] = x;
                }
            }
            team_sync();
        }
    }
}

// The tables of one pair in the team's LDS: edge table (slot = vertex index,
// P: i, Q: nP + j; horizontal edges and edges outside the op's y range get
// a -inf top and are never active), the distinct vertex heights of the op's
// y range ys[0, NY) (bitonic sort across the team, then unique), the beams
// [ekb, ekt) each edge spans and the exclusive scan boff[0, NY) of the
// number of edges spanning each beam.  nP / nQ are 0 for empty operands
// (eP / eQ).  Returns NY, or 0 when the result of `op` is empty (no beam).
template <int TL>
__device__ int team_tables(const TeamLds &t, const double *__restrict__ P, int nP,
                           const double *__restrict__ Q, int nQ, bool eP, bool eQ, int op, int tl,
                           int base) {
    const int n = nP + nQ;
    // bounding boxes
    double bp0 = __builtin_huge_val(), bp1 = bp0, bp2 = -bp0, bp3 = -bp0;
    double bq0 = bp0, bq1 = bp0, bq2 = -bp0, bq3 = -bp0;
    for (int i = tl; i < nP; i += TL) {
        const double x = P[2 * i], y = P[2 * i + 1];
        bp0 = fmin(bp0, x); bp2 = fmax(bp2, x); bp1 = fmin(bp1, y); bp3 = fmax(bp3, y);
    }
    for (int i = tl; i < nQ; i += TL) {
        const double x = Q[2 * i], y = Q[2 * i + 1];
        bq0 = fmin(bq0, x); bq2 = fmax(bq2, x); bq1 = fmin(bq1, y); bq3 = fmax(bq3, y);
    }
    double mm[8] = {bp0, bp1, bq0, bq1, bp2, bp3, bq2, bq3};
    team_minmax8<TL>(t, mm, tl);
    const double bpa[4] = {mm[0], mm[1], mm[4], mm[5]}, bqa[4] = {mm[2], mm[3], mm[6], mm[7]};
    double ylo = 0.0, yhi = 0.0;
    if (!op_range(op, eP, eQ, bpa, bqa, ylo, yhi)) return 0;

    // edge table and raw scanbeam table
    const int P2 = pow2_ge(n);
    const double INF = __builtin_huge_val();
    for (int s = tl; s < P2; s += TL) {
        if (s >= n) {
            t.ys[s] = INF;
            continue;
        }
        const bool inP = s < nP;
        const double *V = inP ? P : Q;
        const int nn = inP ? nP : nQ;
        const int i = inP ? s : s - nP;
        const int j = i + 1 == nn ? 0 : i + 1;
        const double x0 = V[2 * i], y0 = V[2 * i + 1], x1 = V[2 * j], y1 = V[2 * j + 1];
        t.ys[s] = (y0 >= ylo && y0 <= yhi) ? y0 : INF;
        double bx, by, tx, ty;
        if (y0 < y1) { bx = x0; by = y0; tx = x1; ty = y1; }
        else { bx = x1; by = y1; tx = x0; ty = y0; }
        const bool valid = y0 != y1 && !(ty <= ylo || by >= yhi);
        t.bx[s] = bx;
        t.by[s] = by;
        t.tx[s] = tx;
        t.ty[s] = valid ? ty : -INF;  // -inf top: never active
        t.dx[s] = valid ? (tx - bx) / (ty - by) : 0.0;
    }
    tsync<TL>();
    // bitonic sort of ys[0, P2)
    for (int k = 2; k <= P2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int q = tl; q < P2 / 2; q += TL) {
                const int i = ((q & ~(j - 1)) << 1) | (q & (j - 1));
                const int l = i | j;
                const bool up = (i & k) == 0;
                const double x = t.ys[i], y = t.ys[l];
                if ((x > y) == up) {
                    t.ys[i] = y;
                    t.ys[l] = x;
                }
            }
            tsync<TL>();
        }
    }
    // unique (finite values) -> compacted into bsum, then back into ys
    const int per = (P2 + TL - 1) / TL;
    const int c0 = tl * per, c1 = c0 + per < P2 ? c0 + per : P2;
    int mine = 0;
    for (int s = c0; s < c1; ++s) {
        const double y = t.ys[s];
        if (y != INF && (s == 0 || y != t.ys[s - 1])) ++mine;
    }
    int NY = 0;
    const int incl = team_scan<TL>(t, mine, tl, base, NY);
    int w = incl - mine;
    for (int s = c0; s < c1; ++s) {
        const double y = t.ys[s];
        if (y != INF && (s == 0 || y != t.ys[s - 1])) t.bsum[w++] = y;
    }
    tsync<TL>();
    for (int s = tl; s < NY; s += TL) t.ys[s] = t.bsum[s];
    if (NY < 2) return 0;
    const int nb = NY - 1;
    for (int k = tl; k <= nb; k += TL) t.cnt[k] = 0;
    tsync<TL>();
    // beam span of every edge; list entries per beam
    for (int e = tl; e < n; e += TL) {
        const double ty = t.ty[e];
        int kb = 0, kt = 0;
        if (ty != -INF) {
            kb = lower_bound(t.ys, NY, t.by[e]);
            kt = lower_bound(t.ys, NY, ty);
            kt = kt < nb ? kt : nb;
        }
        t.ekb[e] = static_cast<uint16_t>(kb);
        t.ekt[e] = static_cast<uint16_t>(kt);
        for (int k = kb; k < kt; ++k) atomicAdd(&t.cnt[k], 1);
    }
    tsync<TL>();
    // exclusive scan of cnt[0, nb) -> boff, boff[nb] = S
    const int pb = (nb + TL - 1) / TL;
    const int b0 = tl * pb, b1 = b0 + pb < nb ? b0 + pb : nb;
    int sum = 0;
    for (int k = b0; k < b1; ++k) sum += t.cnt[k];
    int S = 0;
    const int sincl = team_scan<TL>(t, sum, tl, base, S);
    int run = sincl - sum;
    for (int k = b0; k < b1; ++k) {
        t.boff[k] = run;
        run += t.cnt[k];
    }
    if (tl == 0) t.boff[nb] = S;
    tsync<TL>();
    return NY;
}

// Active-edge lists of the window of beams [k0, k1): k1 is the largest beam
// end whose entries fit SCAP list slots (at least k0 + 1).  Entry pos =
// boff[k] - boff[k0] + i holds an edge spanning beam k and its x at the
// beam's bottom (lxb) and top (lxt); the order inside a beam is the atomics'
// (callers sort).  Returns k1, or -1 when one beam alone exceeds SCAP.  The
// caller syncs the team before reading the lists.
template <int TL>
__device__ int team_window(const TeamLds &t, int k0, int nb, int n, int SCAP, int tl) {
    // k1 = largest beam end with boff[k1] - boff[k0] <= SCAP (at least k0 + 1)
    int lo = k0 + 1, hi = nb;
    const int lim = t.boff[k0] + SCAP;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (t.boff[mid] <= lim) lo = mid;
        else hi = mid - 1;
    }
    const int k1 = lo;
    const int wbase = t.boff[k0];
    if (t.boff[k1] - wbase > SCAP) return -1;  // one beam alone exceeds the lists
    for (int k = k0 + tl; k < k1; k += TL) t.cnt[k] = 0;
    tsync<TL>();
    for (int e = tl; e < n; e += TL) {
        const int kb = t.ekb[e] > k0 ? t.ekb[e] : k0;
        const int kt = t.ekt[e] < k1 ? t.ekt[e] : k1;
        if (kb >= kt) continue;
        // the edge's entries in registers (the stores below would otherwise
        // make every ex() read them again); the same values ex() reads
        const double ebx = t.bx[e], eby = t.by[e], edx = t.dx[e], ety = t.ty[e], etx = t.tx[e];
        double y0 = t.ys[kb];
        double x0 = y0 == ety ? etx : ebx + edx * (y0 - eby);
        for (int k = kb; k < kt; ++k) {
            const double y1 = t.ys[k + 1];
            const double x1 = y1 == ety ? etx : ebx + edx * (y1 - eby);
            const int pos = t.boff[k] - wbase + atomicAdd(&t.cnt[k], 1);
            t.list[pos] = static_cast<uint16_t>(e);
            t.beam_of[pos] = static_cast<uint16_t>(k);
            t.lxb[pos] = x0;
            t.lxt[pos] = x1;
            x0 = x1;
        }
    }
    return k1;
}

// ---- bucketing -----------------------------------------------------------
struct Bounds {
    int32_t v[NBUCKET];
};

__device__ inline int bucket_of(int n, const Bounds &bd) {
    for (int b = 0; b < NBUCKET; ++b)
        if (n <= bd.v[b]) return b;
    return NBUCKET - 1;
}

__global__ void k_zero_u64(unsigned long long *p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0ull;
}

// per-bucket counts (block-aggregated atomics)
__global__ __launch_bounds__(256) void k_bucket_count(const uint32_t *__restrict__ pairs,
                                                      const int64_t *np_d, int64_t cap,
                                                      const int64_t *__restrict__ offA,
                                                      const int64_t *__restrict__ offB, Bounds bd,
                                                      unsigned long long *__restrict__ counts) {
    const int64_t np = device_count(np_d, cap);
    __shared__ unsigned int c[NBUCKET];
    if (threadIdx.x < NBUCKET) c[threadIdx.x] = 0;
    __syncthreads();
    for (int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; p < np;
         p += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const uint32_t a = pairs[2 * p], b = pairs[2 * p + 1];
        const int n = static_cast<int>((offA[a + 1] - offA[a]) + (offB[b + 1] - offB[b]));
        atomicAdd(&c[bucket_of(n, bd)], 1u);
    }
    __syncthreads();
    if (threadIdx.x < NBUCKET && c[threadIdx.x]) atomicAdd(&counts[threadIdx.x], c[threadIdx.x]);
}

// bucket b's pairs go to order[start[b] + ...]; start = exclusive scan
__global__ void k_bucket_starts(const unsigned long long *counts, int64_t *starts,
                                int64_t *counts_i64, unsigned long long *fill) {
    if (threadIdx.x == 0) {
        int64_t run = 0;
        for (int b = 0; b < NBUCKET; ++b) {
            starts[b] = run;
            counts_i64[b] = static_cast<int64_t>(counts[b]);
            run += static_cast<int64_t>(counts[b]);
            fill[b] = 0ull;
        }
    }
}

__global__ __launch_bounds__(256) void k_bucket_scatter(const uint32_t *__restrict__ pairs,
                                                        const int64_t *np_d, int64_t cap,
                                                        const int64_t *__restrict__ offA,
                                                        const int64_t *__restrict__ offB, Bounds bd,
                                                        const int64_t *__restrict__ starts,
                                                        unsigned long long *__restrict__ fill,
                                                        uint32_t *__restrict__ order) {
    const int64_t np = device_count(np_d, cap);
    __shared__ unsigned int c[NBUCKET];
    __shared__ unsigned long long base[NBUCKET];
    for (int64_t p0 = static_cast<int64_t>(blockIdx.x) * blockDim.x; p0 < np;
         p0 += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        if (threadIdx.x < NBUCKET) c[threadIdx.x] = 0;
        __syncthreads();
        const int64_t p = p0 + threadIdx.x;
        int bk = -1;
        unsigned int r = 0;
        if (p < np) {
            const uint32_t a = pairs[2 * p], b = pairs[2 * p + 1];
            const int n = static_cast<int>((offA[a + 1] - offA[a]) + (offB[b + 1] - offB[b]));
            bk = bucket_of(n, bd);
            r = atomicAdd(&c[bk], 1u);
        }
        __syncthreads();
        if (threadIdx.x < NBUCKET && c[threadIdx.x])
            base[threadIdx.x] = atomicAdd(&fill[threadIdx.x], static_cast<unsigned long long>(c[threadIdx.x]));
        __syncthreads();
        if (bk >= 0) order[starts[bk] + static_cast<int64_t>(base[bk]) + r] = static_cast<uint32_t>(p);
        __syncthreads();
    }
}

// list capacity per team (entries of a window of beams; a pair whose beams
// need more is processed in several windows)
int scap_for(int NV, int factor) {
    const int c = factor * NV;
    const int cap = NV > 2048 ? NV : 2048;
    return c < cap ? c : cap;
}

int device_cus() {
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

}  // namespace
}  // namespace prt
