// clip_batched.hip -- batched FP64 polygon intersection area: pairs are
// bucketed by vertex count, each bucket runs one persistent, work-stealing
// launch whose team shape (lanes per pair, waves per block) comes from the
// autotuner (prt_tune).
//
// Same scanbeam sweep and the same FP64 expressions as the plain kernel
// (clip_plain.hip) and the CPU oracle, so areas are bit-identical, but:
//   * a TEAM of TL lanes (8..64, a slice of one wave64, or 128 / 256: two
//     or four whole waves making up a block) owns a pair; the pair's edges,
//     its scanbeam table and its per-beam active-edge lists live in LDS;
//   * the scanbeam table is sorted with a bitonic network across the team;
//   * every edge registers itself in the lists of the beams it spans (LDS
//     atomics + a team scan), with its x at the beam bottom and top, so a
//     beam finds its active edges without scanning the polygon;
//   * a beam's entries are sorted by the sub-beam key; a beam whose sorted
//     entries show no crossing is summed in one linear pass;
//   * lane k sweeps the crossing-free beams k, k+TL, ...; a beam with
//     crossings is swept by one wave, its sub-beams side by side in groups
//     of lanes (one-wave and whole-wave teams) or by the team in LDS
//     (smaller teams); beams are independent, their sums are combined with
//     the fixed pairwise tree the oracle uses;
//   * teams take the next pair from a per-bucket ticket (work stealing), so
//     polygons of very uneven size do not leave lanes idle.
// The lists of a pair are built for windows of consecutive beams that fit
// the team's LDS.  Only a pair larger than its bucket's team (or a single
// beam with more active edges than the list capacity) goes to the plain
// kernel, through an overflow list (same results, slower).
#include <mutex>

#include "clip_team.hpp"
#include "prt_common.hpp"
#include "prt_profile.hpp"

namespace prt {

int plain_list_launch(const uint32_t *pairs, const uint32_t *list, const int64_t *count,
                      int64_t max_count, const double *xyA, const int64_t *offA, const double *xyB,
                      const int64_t *offB, int32_t max_verts, int op, double *area, void *ws,
                      size_t ws_bytes, hipStream_t s);
size_t plain_list_ws_bytes(int64_t max_count, int32_t max_verts);

namespace {

// Doubled area of one beam whose active edges are the list entries
// [lo, hi) (with their x at the beam bottom/top cached in lxb/lxt): the
// general sweep of the beam.  Sub-beam ends are the distinct crossing
// heights, found one minimum at a time; at every sub-beam the list region
// (still in the previous sub-beam's order, so nearly sorted) is
// insertion-sorted in place by the strict key (xa + xc, xa, edge id), with
// the x at the sub-beam ends recomputed as the entries are compared, and the
// gaps are summed left to right -- the terms and order of the plain kernel's
// sweep.  The region's order changes; its entries do not.
__device__ double beam_sum_serial(const TeamLds &t, int lo, int hi, double yb, double yt, int nP,
                                  int op) {
    const double dy = yt - yb;
    double ya = yb, bs = 0.0;
    for (;;) {
        double yc = yt;
        for (int i = lo; i < hi; ++i) {
            const double xbi = t.lxb[i], xti = t.lxt[i];
            for (int j = i + 1; j < hi; ++j) {
                const double xbj = t.lxb[j], xtj = t.lxt[j];
                double xbL, xtL, xbR, xtR;
                if (xbi < xbj && xti > xtj) { xbL = xbi; xtL = xti; xbR = xbj; xtR = xtj; }
                else if (xbj < xbi && xtj > xti) { xbL = xbj; xtL = xtj; xbR = xbi; xtR = xti; }
                else continue;
                const double den = (xtL - xbL) - (xtR - xbR);
                if (!(den > 0.0)) continue;
                const double r = (xbR - xbL) / den;
                const double iy = r * dy + yb;
                if (iy > ya && iy < yc) yc = iy;
            }
        }
        const double h = yc - ya;
        auto xat = [&](int i, double y) -> double {
            if (y == yb) return t.lxb[i];
            if (y == yt) return t.lxt[i];
            return ex(t, t.list[i], y);
        };
        for (int i = lo + 1; i < hi; ++i) {
            const int e = t.list[i];
            const double xa = xat(i, ya), k = xa + xat(i, yc);
            const double xb = t.lxb[i], xt = t.lxt[i];
            int j = i;
            while (j > lo) {
                const double xa2 = xat(j - 1, ya), k2 = xa2 + xat(j - 1, yc);
                if (!key_before(k, xa, e, k2, xa2, t.list[j - 1])) break;
                t.list[j] = t.list[j - 1];
                t.lxb[j] = t.lxb[j - 1];
                t.lxt[j] = t.lxt[j - 1];
                --j;
            }
            t.list[j] = static_cast<uint16_t>(e);
            t.lxb[j] = xb;
            t.lxt[j] = xt;
        }
        int pp = 0, pq = 0;
        double xa0 = xat(lo, ya), xc0 = xat(lo, yc);
        for (int i = lo; i + 1 < hi; ++i) {
            if (t.list[i] < nP) pp ^= 1;
            else pq ^= 1;
            const double xa1 = xat(i + 1, ya), xc1 = xat(i + 1, yc);
            if (op_inside(op, pp, pq)) bs += ((xa1 - xa0) + (xc1 - xc0)) * h;
            xa0 = xa1;
            xc0 = xc1;
        }
        if (yc == yt) break;
        ya = yc;
    }
    return bs;
}

// Doubled area of one beam whose list entries [lo, hi) are sorted by the
// sub-beam key and do not cross (cnt[k] = 0 after team_sort_window): the
// gaps summed left to right -- the same terms, in the same order, as the
// general sweep.
__device__ inline double beam_sum_sorted(const TeamLds &t, int lo, int hi, double yb, double yt,
                                         int nP, int op) {
    const double h = yt - yb;
    double bs = 0.0;
    int pp = 0, pq = 0;
    for (int i = lo; i + 1 < hi; ++i) {
        if (t.list[i] < nP) pp ^= 1;
        else pq ^= 1;
        if (op_inside(op, pp, pq)) bs += ((t.lxb[i + 1] - t.lxb[i]) + (t.lxt[i + 1] - t.lxt[i])) * h;
    }
    return bs;
}

// The general sweep of one beam (list entries [lo, hi), m <= GS) by the
// whole team, with the results of beam_sum_serial bit for bit: the
// sub-beam bounds are the distinct crossing heights in (yb, yt) in
// increasing order (the serial sweep finds them one minimum at a time from
// the same candidate set), and each sub-beam is summed by one lane in the
// strict key order (the order the serial sweep sorts the entries into).
// Returns the beam sum on every lane of the team.
template <int TL>
__device__ double team_general(const TeamLds &t, int lo, int hi, double yb, double yt, int nP,
                               int op, int tl, int base) {
    const int m = hi - lo;
    const double dy = yt - yb;
    if (tl == 0) t.gmisc[1] = 0;
    team_sync();
    for (int i = lo + tl; i < hi; i += TL) {
        const double xbi = t.lxb[i], xti = t.lxt[i];
        for (int j = i + 1; j < hi; ++j) {
            const double xbj = t.lxb[j], xtj = t.lxt[j];
            double xbL, xtL, xbR, xtR;
            if (xbi < xbj && xti > xtj) { xbL = xbi; xtL = xti; xbR = xbj; xtR = xtj; }
            else if (xbj < xbi && xtj > xti) { xbL = xbj; xtL = xtj; xbR = xbi; xtR = xti; }
            else continue;
            const double den = (xtL - xbL) - (xtR - xbR);
            if (!(den > 0.0)) continue;
            const double r = (xbR - xbL) / den;
            const double iy = r * dy + yb;
            if (iy > yb && iy < yt) {
                const int pos = atomicAdd(&t.gmisc[1], 1);
                if (pos < t.GC) t.giy[pos] = iy;
            }
        }
    }
    team_sync();
    const int nc = t.gmisc[1];
    if (nc > t.GC) {
        // more crossing heights than the scratch holds: the serial sweep
        double r = 0.0;
        if (tl == 0) r = beam_sum_serial(t, lo, hi, yb, yt, nP, op);
        double v = __shfl(r, base, 64);
        team_sync();
        return v;
    }
    const double INF = __builtin_huge_val();
    if (nc > 1) {
        if (TL == 64) {
            wave_sort_heights(t, nc, tl);
            team_sync();
        } else {
            const int C2 = pow2_ge(nc);
            for (int q = nc + tl; q < C2; q += TL) t.giy[q] = INF;
            team_sync();
            team_sort_heights<TL>(t, C2, tl);
        }
    }
    const int M2 = pow2_ge(m);
    double bs = 0.0;  // meaningful on lane 0 of the team
    double ya = yb;
    int ci = 0;
    for (;;) {
        while (ci < nc && !(t.giy[ci] > ya)) ++ci;
        const double yc = ci < nc ? t.giy[ci] : yt;
        const double h = yc - ya;
        if (TL == 64) {
            wave_sort_sub_beam(t, lo, m, yb, yt, ya, yc, tl);
            team_sync();
        } else {
            for (int q = tl; q < M2; q += TL) {
                if (q < m) {
                    const int i = lo + q;
                    t.gxa[q] = ya == yb ? t.lxb[i] : (ya == yt ? t.lxt[i] : ex(t, t.list[i], ya));
                    t.gxc[q] = yc == yb ? t.lxb[i] : (yc == yt ? t.lxt[i] : ex(t, t.list[i], yc));
                    t.gidx[q] = t.list[i];
                } else {
                    t.gxa[q] = INF;
                    t.gxc[q] = INF;
                    t.gidx[q] = 0xFFFFu;
                }
            }
            team_sync();
            team_sort_entries<TL>(t, M2, tl);
        }
        if (tl == 0) {
            int pp = 0, pq = 0;
            for (int q = 0; q + 1 < m; ++q) {
                if (t.gidx[q] < nP) pp ^= 1;
                else pq ^= 1;
                if (op_inside(op, pp, pq))
                    bs += ((t.gxa[q + 1] - t.gxa[q]) + (t.gxc[q + 1] - t.gxc[q])) * h;
            }
        }
        team_sync();
        if (yc == yt) break;
        ya = yc;
    }
    return __shfl(bs, base, 64);
}

// A double pushed to lane dst (every lane active, dst a permutation of the
// lanes): lane dst receives this lane's v.
__device__ inline double permute_d(double v, int dst) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = static_cast<uint32_t>(
        __builtin_amdgcn_ds_permute(dst << 2, static_cast<int>(static_cast<uint32_t>(u))));
    const uint32_t hi = static_cast<uint32_t>(
        __builtin_amdgcn_ds_permute(dst << 2, static_cast<int>(static_cast<uint32_t>(u >> 32))));
    return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}

// v of lane q + 1 (every lane active; lane 63 reads 0): a DPP shift of the
// whole wave by one lane
__device__ inline double next_lane_d(double v) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = static_cast<uint32_t>(
        __builtin_amdgcn_update_dpp(0, static_cast<int>(static_cast<uint32_t>(u)), 0x130, 0xF, 0xF, false));
    const uint32_t hi = static_cast<uint32_t>(
        __builtin_amdgcn_update_dpp(0, static_cast<int>(static_cast<uint32_t>(u >> 32)), 0x130, 0xF, 0xF, false));
    return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}

// v of lane q - 1 (every lane active; lane 0 reads 0): a DPP shift of the
// whole wave by one lane the other way; ints by the same two shifts
__device__ inline double prev_lane_d(double v) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = static_cast<uint32_t>(
        __builtin_amdgcn_update_dpp(0, static_cast<int>(static_cast<uint32_t>(u)), 0x138, 0xF, 0xF, false));
    const uint32_t hi = static_cast<uint32_t>(
        __builtin_amdgcn_update_dpp(0, static_cast<int>(static_cast<uint32_t>(u >> 32)), 0x138, 0xF, 0xF, false));
    return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}
__device__ inline int next_lane_i(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xF, 0xF, false); }
__device__ inline int prev_lane_i(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xF, 0xF, false); }

// The entries of every aligned group of M2 lanes (the first m of each group
// real, the rest +inf keys) sorted ascending by the strict key (xa + xc, xa,
// edge id), from an order close to sorted (the beam's, which a sub-beam's
// differs from by the swaps of its crossing pairs): rounds of odd-even
// transposition -- compare-exchange of lanes (2k, 2k + 1), then of (2k + 1,
// 2k + 2), inside the group -- until every real entry precedes the next
// (at most M2 rounds, 2 M2 >= m phases, always enough).  The key is a strict
// order, so the result is the one sorted order (a bitonic network over the
// group, which this replaces, took 6-10 dependent exchange stages for the
// usual 5-16 entries; these take 1-3 rounds).
// Every lane active; act: the lane's group holds a sub-beam.
__device__ inline void wave_oddeven(int qi, int M2, int m, bool act, double &a, double &c, int &e) {
    for (int round = 0; round < M2; ++round) {
        {  // pairs (2k, 2k + 1)
            const double oa = xor_lane_d<1>(a), oc = xor_lane_d<1>(c);
            const int oe = static_cast<int>(xor_lane<1>(static_cast<uint32_t>(e)));
            const bool o_first = key_before(oa + oc, oa, oe, a + c, a, e);
            if (o_first == ((qi & 1) == 0)) {
                a = oa;
                c = oc;
                e = oe;
            }
        }
        {  // pairs (2k + 1, 2k + 2) inside the group
            const double na = next_lane_d(a), nc = next_lane_d(c);
            const int ne = next_lane_i(e);
            const double pa = prev_lane_d(a), pc = prev_lane_d(c);
            const int pe = prev_lane_i(e);
            const bool lower = (qi & 1) == 1;
            const bool has = lower ? qi + 1 < M2 : qi > 0;
            const double oa = lower ? na : pa, oc = lower ? nc : pc;
            const int oe = lower ? ne : pe;
            const bool o_first = key_before(oa + oc, oa, oe, a + c, a, e);
            if (has && o_first == lower) {
                a = oa;
                c = oc;
                e = oe;
            }
        }
        const double na = next_lane_d(a), nc = next_lane_d(c);
        const int ne = next_lane_i(e);
        const bool out = act && qi + 1 < m && !key_before(a + c, a, e, na + nc, na, ne);
        if (__ballot(out) == 0ull) break;
    }
}

// Teams of one wave: team_general with the beam's entries in registers
// instead of LDS.  The crossing heights come from the same pairs (i < j),
// tested with the same expressions, one pair per lane per step; they are
// sorted across the lanes and their distinct values compacted.  The wave is
// then split into aligned groups of M2 lanes (M2 = pow2 >= m entries) and
// each group takes one sub-beam, 64 / M2 sub-beams per pass: lane qi of a
// group holds entry lo + qi, evaluates it at its sub-beam's ends (the
// expressions of ex() and the serial sweep), and odd-even transposition
// inside the group (wave_oddeven) sorts the entries by the strict key, so
// lane qi holds the qi-th entry of the sub-beam; its parities are prefix counts of the group's
// ballots, and its gap term is the serial sweep's ((xa1 - xa0) + (xc1 -
// xc0)) * h.  The terms of the inside gaps are added to the beam sum one at a
// time in lane order -- sub-beam by sub-beam, gap by gap -- so the sum is the
// serial sweep's bit for bit.  Returns the beam sum on every lane.
__device__ double wave_general(const TeamLds &t, int lo, int hi, double yb, double yt, int nP,
                               int op, int q) {
    const int m = hi - lo;
    const int nc = wave_crossing_heights(t, lo, hi, yb, yt, q);
    if (nc > t.GC) {
        // more crossing heights than the scratch holds: the serial sweep
        double r = 0.0;
        if (q == 0) r = beam_sum_serial(t, lo, hi, yb, yt, nP, op);
        const double v = readlane_d(r, 0);
        team_sync();
        return v;
    }
    // the distinct heights ascending across the lanes: lane i < nd holds the
    // i-th (ranks read across the wave, equal heights in lane order; the last
    // of each run of equal heights kept)
    double ds = 0.0;
    int nd = 0;
    if (nc > 0) {
        team_sync();
        const double h = q < nc ? t.giy[q] : 0.0;
        int rank = q;
        if (nc > 1) {
            int r = 0;
            for (int p = 0; p < nc; ++p) {
                const double hp = readlane_d(h, p);
                r += (hp < h || (hp == h && p < q)) ? 1 : 0;
            }
            if (q < nc) rank = r;
        }
        const double sh = permute_d(h, rank);
        const bool keep = q < nc && (q + 1 == nc || next_lane_d(sh) != sh);
        const uint64_t km = __ballot(keep);
        nd = __popcll(km);
        const int below = __popcll(km & ((1ull << q) - 1ull));
        ds = permute_d(sh, keep ? below : nd + (q - below));
    }
    const int M2 = pow2_ge(m);
    const int G = 64 / M2, g = q / M2, qi = q & (M2 - 1);
    // the entry of lane qi of every group
    const bool live = qi < m;
    int e = 0;
    double xb = 0.0, xt = 0.0, ebx = 0.0, eby = 0.0, edx = 0.0, ety = 0.0, etx = 0.0;
    if (live) {
        e = t.list[lo + qi];
        xb = t.lxb[lo + qi];
        xt = t.lxt[lo + qi];
        ebx = t.bx[e];
        eby = t.by[e];
        edx = t.dx[e];
        ety = t.ty[e];
        etx = t.tx[e];
    }
    const uint64_t gmask = M2 == 64 ? ~0ull : ((1ull << M2) - 1ull) << (g * M2);
    const uint64_t upto = ((2ull << q) - 1ull) & gmask;  // lanes of the group up to q
    double bs = 0.0;
    for (int s0 = 0; s0 <= nd; s0 += G) {
        const int s = s0 + g;  // this group's sub-beam
        const bool act = s <= nd;
        const double dlo = __shfl(ds, act && s > 0 ? s - 1 : 0, 64);
        const double dhi = __shfl(ds, act && s < nd ? s : 0, 64);
        const double ya = s == 0 ? yb : dlo, yc = s < nd ? dhi : yt;
        const double h = yc - ya;
        // lanes past the entries (and groups past the last sub-beam) hold
        // +inf keys, which sort last
        double sa = __builtin_huge_val(), sc = __builtin_huge_val();
        int se = 0xFFFF;
        if (act && live) {
            sa = ya == yb ? xb : (ya == yt ? xt : (ya == ety ? etx : ebx + edx * (ya - eby)));
            sc = yc == yb ? xb : (yc == yt ? xt : (yc == ety ? etx : ebx + edx * (yc - eby)));
            se = e;
        }
        wave_oddeven(qi, M2, m, act, sa, sc, se);
        const bool in_p = act && live && se < nP, in_q = act && live && !(se < nP);
        const int pp = __popcll(__ballot(in_p) & upto) & 1;
        const int pq = __popcll(__ballot(in_q) & upto) & 1;
        const double an = next_lane_d(sa), cn = next_lane_d(sc);
        const double term = ((an - sa) + (cn - sc)) * h;
        uint64_t inside = __ballot(act && qi + 1 < m && op_inside(op, pp, pq));
        while (inside) {
            const int gl = __builtin_ctzll(inside);
            inside &= inside - 1ull;
            bs += readlane_d(term, gl);
        }
    }
    return bs;
}

// Teams of TS = 8, 16 or 32 lanes (lanes base .. base + TS - 1 of a wave
// whose other lanes may be other teams anywhere in their own pairs), TS at
// most the team's scratch (GS, GC): wave_general's sweep of a beam of m <= TS
// entries inside the team's lanes, the same heights, orders, terms and
// summation order.  What wave_general reads across the wave with
// readlane (an index uniform over the wave) goes through the team's LDS
// scratch here: the heights' ranks are counted from giy, and the inside
// gaps' terms are added by the team's lane 0 from gxa, gap after gap.
// Ballots are shifted to the team's lanes.  Returns the beam sum on the
// team's lane 0.
template <int TS>
__device__ double small_general(const TeamLds &t, int lo, int hi, double yb, double yt, int nP,
                                int op, int q, int base) {
    constexpr uint32_t TMASK = TS == 32 ? ~0u : (1u << TS) - 1u;
    const int m = hi - lo;
    const double dy = yt - yb;
    auto half = [&](bool v) { return static_cast<uint32_t>(__ballot(v) >> base) & TMASK; };
    // the pairs (i, j), i < j, spread over the team's lanes
    const int npairs = m * (m - 1) / 2;
    int nc = 0;
    team_sync();
    for (int t0 = 0; t0 < npairs; t0 += TS) {
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
        const uint32_t hm = half(has);
        if (has) {
            const int pos = nc + __popc(hm & ((1u << q) - 1u));
            if (pos < t.GC) t.giy[pos] = iy;
        }
        nc += __popc(hm);
    }
    if (nc > t.GC || nc > TS) {
        // more crossing heights than the scratch (or the team) holds: the
        // serial sweep
        team_sync();
        double r = 0.0;
        if (q == 0) r = beam_sum_serial(t, lo, hi, yb, yt, nP, op);
        team_sync();
        return r;
    }
    // the heights ascending (ranks counted over giy, equal heights in their
    // order there), then the distinct ones compacted into giy[0, nd)
    int nd = 0;
    if (nc > 0) {
        team_sync();
        const double h = q < nc ? t.giy[q] : 0.0;
        int r = 0;
        for (int p = 0; p < nc; ++p) {
            const double hp = t.giy[p];
            r += (hp < h || (hp == h && p < q)) ? 1 : 0;
        }
        team_sync();
        if (q < nc) t.giy[r] = h;
        team_sync();
        const double sh = q < nc ? t.giy[q] : 0.0;
        const double nx = q + 1 < nc ? t.giy[q + 1] : 0.0;
        const bool keep = q < nc && (q + 1 == nc || nx != sh);
        const uint32_t km = half(keep);
        nd = __popc(km);
        team_sync();
        if (keep) t.giy[__popc(km & ((1u << q) - 1u))] = sh;
        team_sync();
    }
    const int M2 = pow2_ge(m);
    const int G = TS / M2, g = q / M2, qi = q & (M2 - 1);
    const bool live = qi < m;
    int e = 0;
    double xb = 0.0, xt = 0.0, ebx = 0.0, eby = 0.0, edx = 0.0, ety = 0.0, etx = 0.0;
    if (live) {
        e = t.list[lo + qi];
        xb = t.lxb[lo + qi];
        xt = t.lxt[lo + qi];
        ebx = t.bx[e];
        eby = t.by[e];
        edx = t.dx[e];
        ety = t.ty[e];
        etx = t.tx[e];
    }
    const uint32_t gmask = M2 == 32 ? ~0u : ((1u << M2) - 1u) << (g * M2);
    const uint32_t upto = (q == 31 ? ~0u : (2u << q) - 1u) & gmask;  // group lanes up to q
    double bs = 0.0;
    for (int s0 = 0; s0 <= nd; s0 += G) {
        const int s = s0 + g;  // this group's sub-beam
        const bool act = s <= nd;
        const double ya = act && s > 0 ? t.giy[s - 1] : yb;
        const double yc = act && s < nd ? t.giy[s] : yt;
        const double h = yc - ya;
        double sa = __builtin_huge_val(), sc = __builtin_huge_val();
        int se = 0xFFFF;
        if (act && live) {
            sa = ya == yb ? xb : (ya == yt ? xt : (ya == ety ? etx : ebx + edx * (ya - eby)));
            sc = yc == yb ? xb : (yc == yt ? xt : (yc == ety ? etx : ebx + edx * (yc - eby)));
            se = e;
        }
        wave_oddeven(qi, M2, m, act, sa, sc, se);
        const bool in_p = act && live && se < nP, in_q = act && live && !(se < nP);
        const int pp = __popc(half(in_p) & upto) & 1;
        const int pq = __popc(half(in_q) & upto) & 1;
        const double an = next_lane_d(sa), cn = next_lane_d(sc);
        const double term = ((an - sa) + (cn - sc)) * h;
        const bool ins = act && qi + 1 < m && op_inside(op, pp, pq);
        uint32_t inside = half(ins);
        team_sync();  // the scratch's earlier readers are done
        if (ins) t.gxa[q] = term;
        team_sync();
        if (q == 0) {
            while (inside) {
                const int gl = __builtin_ctz(inside);
                inside &= inside - 1u;
                bs += t.gxa[gl];
            }
        }
    }
    team_sync();
    return bs;
}

struct BatchArgs {
    const uint32_t *pairs;
    const uint32_t *order;   // pair indices grouped by bucket
    const int64_t *start;    // first index of this bucket in order (device)
    const int64_t *count;    // number of pairs in the bucket (device)
    const double *xyA, *xyB;
    const int64_t *offA, *offB;
    double *area;
    unsigned long long *ticket;
    uint32_t *ovf_list;
    unsigned long long *ovf_count;
    int NV, SCAP;
    int op;  // PRT_OP_*
};

// Returns 1 when the pair was handled, 0 when it must go to the overflow list.
template <int TL>
__device__ int team_pair(const TeamLds &t, const BatchArgs &A, uint32_t p, int tl, int base,
                         double &result) {
    const uint32_t a = A.pairs[2 * p], b = A.pairs[2 * p + 1];
    const int64_t oa = A.offA[a], ob = A.offB[b];
    int nP = static_cast<int>(A.offA[a + 1] - oa);
    int nQ = static_cast<int>(A.offB[b + 1] - ob);
    const double *P = A.xyA + 2 * oa, *Q = A.xyB + 2 * ob;
    result = 0.0;
    const bool eP = nP < 3, eQ = nQ < 3;  // empty operands
    if (eP && eQ) return 1;
    if (nP + nQ > A.NV) return 0;
    if (eP) nP = 0;
    if (eQ) nQ = 0;
    const int n = nP + nQ;

    const int NY = team_tables<TL>(t, P, nP, Q, nQ, eP, eQ, A.op, tl, base);
    if (NY == 0) return 1;
    const int nb = NY - 1;
    // windows of consecutive beams whose entries fit the list capacity
    for (int k0 = 0; k0 < nb;) {
        const int k1 = team_window<TL>(t, k0, nb, n, A.SCAP, tl);
        if (k1 < 0) return 0;
        const int wbase = t.boff[k0];
        if (tl == 0) t.gmisc[0] = 0;
        team_sort_window<TL>(t, k0, k1, tl);  // cnt[k]: the beam's edges cross
        // lane-parallel beams; beams with crossings of up to GS entries are
        // queued (in beam_of, free once the window is sorted) and swept by
        // the whole team afterwards (measured: sweeping the small ones in
        // their lanes, 16 or 32 entries at most, is 8-20 % slower), longer
        // ones by their lane
        for (int k = k0 + tl; k < k1; k += TL) {
            const int lo = t.boff[k] - wbase, hi = t.boff[k + 1] - wbase, m = hi - lo;
            if (m < 2)
                t.bsum[k] = 0.0;
            else if (t.cnt[k] == 0)
                t.bsum[k] = beam_sum_sorted(t, lo, hi, t.ys[k], t.ys[k + 1], nP, A.op);
            else if (m <= t.GS)
                t.beam_of[atomicAdd(&t.gmisc[0], 1)] = static_cast<uint16_t>(k);
            else
                t.bsum[k] = beam_sum_serial(t, lo, hi, t.ys[k], t.ys[k + 1], nP, A.op);
        }
        tsync<TL>();
        const int ndef = t.gmisc[0];
        if constexpr (TL >= 64) {
            // one wave per queued beam: wave w of a team of W waves takes
            // the queued beams w, w + W, w + 2W, ...; a wave's beams' ranges
            // and heights are loaded by its lanes at once (64 at a time),
            // then read across the wave beam after beam
            constexpr int W = TL / 64;
            const int w = tl >> 6, lane = tl & 63;
            TeamLds tw = t;
            if constexpr (W > 1) tw.giy = t.wgiy + 64 * w;
            const int nmine = ndef > w ? (ndef - w + W - 1) / W : 0;
            for (int i0 = 0; i0 < nmine; i0 += 64) {
                int kq = 0, loq = 0, hiq = 0;
                double ybq = 0.0, ytq = 0.0;
                if (i0 + lane < nmine) {
                    kq = t.beam_of[w + W * (i0 + lane)];
                    loq = t.boff[kq] - wbase;
                    hiq = t.boff[kq + 1] - wbase;
                    ybq = t.ys[kq];
                    ytq = t.ys[kq + 1];
                }
                const int cnt = nmine - i0 < 64 ? nmine - i0 : 64;
                for (int j = 0; j < cnt; ++j) {
                    const int k = __builtin_amdgcn_readlane(kq, j);
                    const double bsk = wave_general(tw, __builtin_amdgcn_readlane(loq, j),
                                                    __builtin_amdgcn_readlane(hiq, j), readlane_d(ybq, j),
                                                    readlane_d(ytq, j), nP, A.op, lane);
                    if (lane == 0) t.bsum[k] = bsk;
                }
            }
        } else {
            for (int q = 0; q < ndef; ++q) {
                const int k = t.beam_of[q];
                const int lo = t.boff[k] - wbase, hi = t.boff[k + 1] - wbase;
                double bsk;
                // the register sweep writes a term per lane of the team into
                // the team's scratch (gxa), so the team must not outnumber it
                if (TL >= 8 && hi - lo <= TL && TL <= t.GS)
                    bsk = small_general<TL>(t, lo, hi, t.ys[k], t.ys[k + 1], nP, A.op, tl, base);
                else
                    bsk = team_general<TL>(t, lo, hi, t.ys[k], t.ys[k + 1], nP, A.op, tl, base);
                if (tl == 0) t.bsum[k] = bsk;
            }
        }
        tsync<TL>();
        k0 = k1;
    }
    for (int wdt = 1; wdt < nb; wdt *= 2) {
        for (int i = tl * 2 * wdt; i + wdt < nb; i += TL * 2 * wdt) t.bsum[i] += t.bsum[i + wdt];
        tsync<TL>();
    }
    result = 0.5 * t.bsum[0];
    return 1;
}

template <int TL>
__global__ __launch_bounds__(BK_THREADS_MAX) void k_clip_batched(BatchArgs A) {
    extern __shared__ __align__(16) char lds[];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tiw = lane / TL, tl = lane % TL;
    const int base = tiw * TL;
    const int team = wave * (64 / TL) + tiw;
    const TeamLds t = carve_team(lds + team * team_bytes(A.NV, A.SCAP), A.NV, A.SCAP);
    const int64_t count = *A.count;
    const uint32_t *order = A.order + *A.start;
    for (;;) {
        unsigned long long idx = 0;
        if (tl == 0) idx = atomicAdd(A.ticket, 1ull);
        idx = __shfl(idx, base, 64);
        if (static_cast<int64_t>(idx) >= count) break;
        const uint32_t p = order[idx];
        double result;
        const int ok = team_pair<TL>(t, A, p, tl, base, result);
        if (tl == 0) {
            if (ok) A.area[p] = result;
            else A.ovf_list[atomicAdd(A.ovf_count, 1ull)] = p;
        }
        team_sync();
    }
}

// static LDS of k_clip_batched_team for teams of w waves: its reduction
// words, the waves' crossing heights and the ticket, with room for their
// alignment (the compiler reports 1,184 and 2,336 bytes for 2 and 4 waves)
constexpr size_t team_static_lds(int w) {
    return 8 * 8 * size_t(w) + 4 * size_t(w + 1) + 64 * 8 * size_t(w) + 8 + 64;
}

// Teams of TL / 64 whole waves (TL = 128 or 256), one team per block: the
// pair's tables are shared by the waves, which split the lanes' loops and
// share out the queued crossing beams (round-robin), so a pair is done by
// more lanes without more LDS.
template <int TL>
__global__ __launch_bounds__(256) void k_clip_batched_team(BatchArgs A) {
    extern __shared__ __align__(16) char lds[];
    constexpr int W = TL / 64;
    __shared__ double red[8 * W];
    __shared__ int32_t ired[W + 1];
    __shared__ double wgiy[64 * W];
    __shared__ unsigned long long s_idx;
    TeamLds t = carve_team(lds, A.NV, A.SCAP);
    t.red = red;
    t.ired = ired;
    t.wgiy = wgiy;
    const int tl = threadIdx.x;
    const int64_t count = *A.count;
    const uint32_t *order = A.order + *A.start;
    for (;;) {
        if (tl == 0) s_idx = atomicAdd(A.ticket, 1ull);
        __syncthreads();
        const unsigned long long idx = s_idx;
        __syncthreads();
        if (static_cast<int64_t>(idx) >= count) break;
        const uint32_t p = order[idx];
        double result;
        const int ok = team_pair<TL>(t, A, p, tl, 0, result);
        if (tl == 0) {
            if (ok) A.area[p] = result;
            else A.ovf_list[atomicAdd(A.ovf_count, 1ull)] = p;
        }
        __syncthreads();
    }
}

// Appends every pair of a bucket to the overflow list (bucket too large for
// the team's LDS).
__global__ void k_route_bucket(BatchArgs A) {
    const int64_t count = *A.count;
    const uint32_t *order = A.order + *A.start;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        A.ovf_list[atomicAdd(A.ovf_count, 1ull)] = order[i];
}

template <int TL>
int launch_bucket(const BatchArgs &A, int waves, size_t lds, hipStream_t s) {
    int per_cu = 1;
    if constexpr (TL > 64) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_clip_batched_team<TL>, TL, lds) !=
                hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        const unsigned grid = static_cast<unsigned>(device_cus() * per_cu);
        PRT_LAUNCH(k_clip_batched_team<TL>, grid, TL, lds, s, A);
    } else {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_clip_batched<TL>, waves * 64, lds) !=
                hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        const unsigned grid = static_cast<unsigned>(device_cus() * per_cu);
        PRT_LAUNCH(k_clip_batched<TL>, grid, waves * 64, lds, s, A);
    }
    PRT_LAUNCHED();
    return PRT_OK;
}

// Helper streams of the bucket launches, created once per device (default
// priority, non-blocking).  A call forks them from its stream with an event
// and joins them back before it returns to that stream's order, so the
// calls stay stream-ordered (and capturable: the fork and join become graph
// edges).  Calls from several host threads share them; that only orders
// their bucket launches.
constexpr int AUX_STREAMS = 2;
constexpr int AUX_DEVICES = 16;

int aux_streams(hipStream_t *out) {
    static std::mutex mu;
    static hipStream_t streams[AUX_DEVICES][AUX_STREAMS] = {};
    int dev = 0;
    PRT_HIP(hipGetDevice(&dev));
    PRT_REQUIRE(dev >= 0 && dev < AUX_DEVICES, PRT_E_UNSUPPORTED, "overlay_batched: device %d", dev);
    std::lock_guard<std::mutex> lock(mu);
    for (int i = 0; i < AUX_STREAMS; ++i) {
        if (!streams[dev][i])
            PRT_HIP(hipStreamCreateWithFlags(&streams[dev][i], hipStreamNonBlocking));
        out[i] = streams[dev][i];
    }
    return PRT_OK;
}

// The helper streams forked from the caller's stream `s` by one event, and
// joined back into `s` when this goes out of scope, on every return path: an
// error after the fork must not leave helper-stream work running on the
// caller's buffers after the call returns, nor leave a capture with work that
// never joins its origin stream.  The helper streams are shared by every
// call on the device: capture one batched overlay at a time per device.
struct StreamFork {
    hipStream_t s;
    hipStream_t aux[AUX_STREAMS] = {};
    hipEvent_t fork = nullptr;
    bool forked[AUX_STREAMS] = {};
    explicit StreamFork(hipStream_t origin) : s(origin) {}
    // helper stream i, made to wait for everything enqueued on s so far
    int take(int i, hipStream_t *out) {
        if (!fork) {
            PRT_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
            PRT_HIP(hipEventRecord(fork, s));
        }
        if (!forked[i]) {
            PRT_HIP(hipStreamWaitEvent(aux[i], fork, 0));
            forked[i] = true;
        }
        *out = aux[i];
        return PRT_OK;
    }
    // s waits for every helper stream that was forked; the first error wins
    int join() {
        int rc = PRT_OK;
        for (int i = 0; i < AUX_STREAMS; ++i) {
            if (!forked[i]) continue;
            forked[i] = false;
            hipEvent_t e = nullptr;
            hipError_t er = hipEventCreateWithFlags(&e, hipEventDisableTiming);
            if (er == hipSuccess) er = hipEventRecord(e, aux[i]);
            if (er == hipSuccess) er = hipStreamWaitEvent(s, e, 0);
            if (er != hipSuccess && rc == PRT_OK) rc = hip_fail(er, "join of a helper stream",
                                                                 __FILE__, __LINE__);
            if (e) (void)hipEventDestroy(e);
        }
        return rc;
    }
    ~StreamFork() {
        (void)join();
        if (fork) (void)hipEventDestroy(fork);
    }
};

}  // namespace
}  // namespace prt

using namespace prt;

extern "C" int prt_tune_default(prt_tune *t) {
    PRT_REQUIRE(t != nullptr, PRT_E_INVAL, "tune_default: null");
    // the fastest shapes on the grid workload (8.4M pairs of lognormal
    // stars) and on the 2 x 4M-star overlay, searched bucket by bucket over
    // the whole call on MI355X; the buckets above 64 vertices take teams of
    // two or four whole waves
    const int bounds[NBUCKET] = {32, 64, 128, 256, 512, 2 * PRT_MAX_POLY_VERTS};
    const int lanes[NBUCKET] = {16, 32, 128, 128, 256, 256};
    const int waves[NBUCKET] = {1, 2, 1, 1, 1, 1};
    const int factor[NBUCKET] = {2, 1, 2, 1, 2, 2};
    for (int b = 0; b < NBUCKET; ++b) {
        t->bucket_max_verts[b] = bounds[b];
        t->lanes_per_pair[b] = lanes[b];
        t->waves_per_block[b] = waves[b];
        t->list_factor[b] = factor[b];
    }
    return PRT_OK;
}

namespace {
int check_tune(const prt_tune *t) {
    for (int b = 0; b < NBUCKET; ++b) {
        const int L = t->lanes_per_pair[b], W = t->waves_per_block[b];
        PRT_REQUIRE(L == 8 || L == 16 || L == 32 || L == 64 || L == 128 || L == 256, PRT_E_INVAL,
                    "tune: lanes_per_pair[%d] = %d not in {8,16,32,64,128,256}", b, L);
        PRT_REQUIRE(W >= 1 && W <= 8, PRT_E_INVAL, "tune: waves_per_block[%d] = %d", b, W);
        PRT_REQUIRE(L <= 64 || W == 1, PRT_E_INVAL,
                    "tune: lanes_per_pair[%d] = %d is a block of whole waves; waves_per_block must be 1",
                    b, L);
        PRT_REQUIRE(t->list_factor[b] >= 1 && t->list_factor[b] <= 8, PRT_E_INVAL,
                    "tune: list_factor[%d] = %d not in [1, 8]", b, t->list_factor[b]);
        PRT_REQUIRE(t->bucket_max_verts[b] >= 3 && (b == 0 || t->bucket_max_verts[b] >
                                                                  t->bucket_max_verts[b - 1]),
                    PRT_E_INVAL, "tune: bucket bounds must increase");
    }
    return PRT_OK;
}

struct BatchWs {
    unsigned long long *counts, *fill, *tickets, *ovf_count;
    int64_t *starts, *counts_i64;
    uint32_t *order, *ovf_list;
    void *plain_ws;
    size_t plain_bytes;
};

template <class C> void batch_layout(C &c, int64_t np, int32_t max_verts, BatchWs *w) {
    const size_t pb = plain_list_ws_bytes(np, max_verts);
    auto u64 = c.template take<unsigned long long>(4 * NBUCKET + 1);
    auto i64 = c.template take<int64_t>(2 * NBUCKET);
    auto ord = c.template take<uint32_t>(np > 0 ? np : 1);
    auto ovf = c.template take<uint32_t>(np > 0 ? np : 1);
    auto pw = c.template take<char>(pb);
    if (w) {
        w->counts = u64;
        w->fill = u64 + NBUCKET;
        w->tickets = u64 + 2 * NBUCKET;
        w->ovf_count = u64 + 4 * NBUCKET;
        w->starts = i64;
        w->counts_i64 = i64 + NBUCKET;
        w->order = ord;
        w->ovf_list = ovf;
        w->plain_ws = pw;
        w->plain_bytes = pb;
    }
}
struct SizerT : Sizer {
    template <class T> T *take(size_t n) {
        Sizer::take<T>(n);
        return nullptr;
    }
};
}  // namespace

extern "C" size_t prt_overlay_batched_workspace_size(int64_t npairs, int32_t max_verts) {
    if (npairs < 0 || max_verts < 0) return 0;
    SizerT s;
    batch_layout(s, npairs, max_verts, nullptr);
    return s.bytes();
}

namespace {
// npairs_d == nullptr: npairs pairs; else min(*npairs_d, npairs) pairs, read
// on the device (nothing here reads device memory on the host)
int batched_impl(const uint32_t *pairs, const int64_t *npairs_d, int64_t npairs,
                 const double *xyA, const int64_t *offA, const double *xyB, const int64_t *offB,
                 int32_t max_verts, int op, const prt_tune *tune, double *area, void *ws,
                 size_t ws_bytes, void *stream) {
    PRT_REQUIRE(op >= PRT_OP_DIFF && op <= PRT_OP_UNION, PRT_E_INVAL, "overlay_batched: op %d", op);
    PRT_REQUIRE(npairs >= 0 && npairs < (int64_t(1) << 32), PRT_E_INVAL,
                "overlay_batched: npairs out of range");
    if (npairs == 0) return PRT_OK;
    PRT_REQUIRE(pairs && xyA && offA && xyB && offB && area, PRT_E_INVAL,
                "overlay_batched: null device pointer");
    PRT_REQUIRE(max_verts >= 0 && max_verts <= 2 * PRT_MAX_POLY_VERTS, PRT_E_INVAL,
                "overlay_batched: max_verts %d out of range", max_verts);
    prt_tune def;
    if (!tune) {
        prt_tune_default(&def);
        tune = &def;
    }
    int rc = check_tune(tune);
    if (rc) return rc;
    const size_t need = prt_overlay_batched_workspace_size(npairs, max_verts);
    PRT_REQUIRE(ws_bytes >= need, PRT_E_NOMEM, "overlay_batched: workspace %zu < %zu bytes",
                ws_bytes, need);
    hipStream_t s = as_stream(stream);
    Carver c(ws, ws_bytes);
    BatchWs w;
    batch_layout(c, npairs, max_verts, &w);
    PRT_REQUIRE(c.ok(), PRT_E_NOMEM, "overlay_batched: workspace too small");

    Bounds bd;
    for (int b = 0; b < NBUCKET; ++b) bd.v[b] = tune->bucket_max_verts[b];
    PRT_LAUNCH(k_zero_u64, 1, 64, 0, s, w.counts, 4 * NBUCKET + 1);
    PRT_LAUNCHED();
    const unsigned g = static_cast<unsigned>(ceil_div(npairs, 256) < 2048 ? ceil_div(npairs, 256) : 2048);
    PRT_LAUNCH(k_bucket_count, g, 256, 0, s, pairs, npairs_d, npairs, offA, offB, bd, w.counts);
    PRT_LAUNCHED();
    PRT_LAUNCH(k_bucket_starts, 1, 64, 0, s, w.counts, w.starts, w.counts_i64, w.fill);
    PRT_LAUNCHED();
    PRT_LAUNCH(k_bucket_scatter, g, 256, 0, s, pairs, npairs_d, npairs, offA, offB, bd, w.starts,
               w.fill, w.order);
    PRT_LAUNCHED();

    // The two buckets of the largest pairs run on the helper streams,
    // concurrently with the others on the caller's stream: their launches are
    // small (LDS-limited teams) and end in long tails, which the concurrent
    // launches fill.  (A third one, the 256-vertex bucket, slows workloads
    // where that bucket is heavy.)  Largest first, so the long pairs start
    // early.
    int max_lds = 0, dev = 0;
    PRT_HIP(hipGetDevice(&dev));
    PRT_HIP(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    StreamFork fk(s);
    rc = aux_streams(fk.aux);
    if (rc) return rc;
    // Launch order, eagerly and under graph capture alike: the helper
    // streams' buckets first, so that their long pairs start early and run
    // beside the others (measured on MI355X, grid workload, default shapes:
    // a replay of the captured call 128.1 ms; with the helper streams forked
    // after the caller's stream's buckets 130.2-130.8 ms; eagerly 124.2-124.8
    // ms either way).
    for (int oi = 0; oi < NBUCKET; ++oi) {
        const int b = NBUCKET - 1 - oi;
        const int lo = b == 0 ? 0 : tune->bucket_max_verts[b - 1];
        if (max_verts <= lo) continue;  // no pair can fall in this bucket
        BatchArgs A;
        A.pairs = pairs;
        A.order = w.order;
        A.start = w.starts + b;
        A.count = w.counts_i64 + b;
        A.xyA = xyA;
        A.xyB = xyB;
        A.offA = offA;
        A.offB = offB;
        A.area = area;
        A.ticket = w.tickets + b;
        A.ovf_list = w.ovf_list;
        A.ovf_count = w.ovf_count;
        A.NV = tune->bucket_max_verts[b] < max_verts ? tune->bucket_max_verts[b] : max_verts;
        A.SCAP = scap_for(A.NV, tune->list_factor[b]);
        A.op = op;
        const int slot = NBUCKET - 1 - b;  // 0, 1 for the two largest buckets
        hipStream_t bs = s;
        if (slot < AUX_STREAMS) {
            rc = fk.take(slot, &bs);
            if (rc) return rc;
        }
        const int L = tune->lanes_per_pair[b], W = tune->waves_per_block[b];
        // teams of whole waves (L > 64) are one per block, with a few static
        // LDS words per wave beside the team's tables
        const size_t lds = L > 64 ? team_bytes(A.NV, A.SCAP)
                                  : static_cast<size_t>(W) * (64 / L) * team_bytes(A.NV, A.SCAP);
        const size_t lds_static = L > 64 ? team_static_lds(L / 64) : 0;
        if (lds + lds_static > static_cast<size_t>(max_lds)) {
            // too large for LDS: the whole bucket takes the plain path
            PRT_LAUNCH(k_route_bucket, 256, 256, 0, bs, A);
            PRT_LAUNCHED();
            continue;
        }
        switch (L) {
        case 8: rc = launch_bucket<8>(A, W, lds, bs); break;
        case 16: rc = launch_bucket<16>(A, W, lds, bs); break;
        case 32: rc = launch_bucket<32>(A, W, lds, bs); break;
        case 64: rc = launch_bucket<64>(A, W, lds, bs); break;
        case 128: rc = launch_bucket<128>(A, W, lds, bs); break;
        default: rc = launch_bucket<256>(A, W, lds, bs); break;
        }
        if (rc) return rc;
    }
    rc = fk.join();
    if (rc) return rc;
    return plain_list_launch(pairs, w.ovf_list, reinterpret_cast<const int64_t *>(w.ovf_count),
                             npairs, xyA, offA, xyB, offB, max_verts, op, area, w.plain_ws,
                             w.plain_bytes, s);
}
}  // namespace

extern "C" int prt_overlay_batched(const uint32_t *pairs, int64_t npairs, const double *xyA,
                                   const int64_t *offA, const double *xyB, const int64_t *offB,
                                   int32_t max_verts, const prt_tune *tune, double *area, void *ws,
                                   size_t ws_bytes, void *stream) {
    return batched_impl(pairs, nullptr, npairs, xyA, offA, xyB, offB, max_verts, PRT_OP_INT, tune,
                        area, ws, ws_bytes, stream);
}

extern "C" int prt_overlay_batched_op(const uint32_t *pairs, int64_t npairs, const double *xyA,
                                      const int64_t *offA, const double *xyB, const int64_t *offB,
                                      int32_t max_verts, int32_t op, const prt_tune *tune,
                                      double *area, void *ws, size_t ws_bytes, void *stream) {
    return batched_impl(pairs, nullptr, npairs, xyA, offA, xyB, offB, max_verts, op, tune, area,
                        ws, ws_bytes, stream);
}

extern "C" int prt_overlay_batched_dev(const uint32_t *pairs, const int64_t *npairs_d, int64_t cap,
                                       const double *xyA, const int64_t *offA, const double *xyB,
                                       const int64_t *offB, int32_t max_verts, const prt_tune *tune,
                                       double *area, void *ws, size_t ws_bytes, void *stream) {
    PRT_REQUIRE(npairs_d != nullptr, PRT_E_INVAL, "overlay_batched_dev: null pair count");
    return batched_impl(pairs, npairs_d, cap, xyA, offA, xyB, offB, max_verts, PRT_OP_INT, tune,
                        area, ws, ws_bytes, stream);
}

// Number of pairs of the last prt_overlay_batched call on this workspace that
// took the plain path (LDS capacity exceeded); synchronises the stream.
extern "C" int prt_overlay_batched_overflow(const void *ws, size_t ws_bytes, int64_t npairs,
                                            int32_t max_verts, int64_t *count_h, void *stream) {
    PRT_REQUIRE(ws && count_h, PRT_E_INVAL, "overlay_batched_overflow: null pointer");
    Carver c(const_cast<void *>(ws), ws_bytes);
    BatchWs w;
    batch_layout(c, npairs, max_verts, &w);
    PRT_REQUIRE(c.ok(), PRT_E_NOMEM, "overlay_batched_overflow: workspace too small");
    unsigned long long v = 0;
    PRT_HIP(hipMemcpyAsync(&v, w.ovf_count, sizeof v, hipMemcpyDeviceToHost, as_stream(stream)));
    PRT_HIP(hipStreamSynchronize(as_stream(stream)));
    *count_h = static_cast<int64_t>(v);
    return PRT_OK;
}

// Largest LDS allocation one block may request on the current device.
extern "C" int prt_device_lds_limit(void) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return -1;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
        return -1;
    return v;
}

// LDS bytes of one team of the batched overlay for pairs of up to nv
// vertices (n1 + n2).
extern "C" int64_t prt_overlay_team_lds(int32_t nv, int32_t list_factor) {
    if (nv < 3) nv = 3;
    if (list_factor < 1 || list_factor > 8) return -1;
    return static_cast<int64_t>(team_bytes(nv, scap_for(nv, list_factor)));
}
