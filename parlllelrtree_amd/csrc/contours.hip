// contours.hip -- result contours of A[pairs[2t]] op B[pairs[2t+1]] (GPC's
// output polygon, vector_map_overlay_gpc_kernel.cu:527-750 and its vertex
// copy :720-728), in two passes: count, then write.  The count pass runs a
// team of lanes per pair (below, "team per pair"); the one-lane-per-pair
// kernels described first take the pairs a team cannot hold and the write
// pass of the pairs whose vertices were not kept.
//
// The sweep is the one of the plain area kernel (same edge table, beams,
// crossing heights and sub-beam order).  The contours are assembled from it
// by this rule (the CPU oracle restates it step by step):
//   * inside intervals of a sub-beam: maximal runs of neighbouring gaps that
//     are inside the operation and not of zero width (a gap whose two edges
//     have equal x at both ends is skipped and does not end a run);
//   * at each sub-beam boundary y, the 2 points of every interval below (x
//     at its top) and above (x at its bottom) are merged by x -- each list
//     keeps its own order; on equal x a right end goes first, then the point
//     below -- and paired consecutively.  Every point toggles one of the two
//     inside states, so the pairs are exactly the horizontal parts of the
//     boundary.  A pair continues a left or right chain upward, starts a
//     contour (local minimum: left-right above; hole bottom: right-left
//     above), or ends two chains (local maximum: left-right below; hole top:
//     right-left below) -- closing the contour when both chain ends belong to
//     it (outer at a maximum, hole at a hole top), else joining two contours;
//   * a contour is a path head..tail: down its left chains, along its
//     bottoms, up its right chains, so outer contours run counter-clockwise
//     and holes clockwise;
//   * a chain that goes on along the same edge adds no vertex; otherwise a
//     pair adds its two points, or one when their x's differ by no more than
//     the sum of their uncertainties: 0 at an edge end (y is the edge's
//     bottom or top), else 2^-44 * (|bx| + |tx| + |dx| * (|y| + |by|)),
//     which covers a crossing evaluated on each of its two edges at a height
//     that is itself rounded.  The one vertex is the point of smaller
//     uncertainty; on a tie the point below of a chain, else the left one.
// Contours come out in closing order; vertices, contour starts and hole
// flags land at the offsets the caller scanned from the counts.
//
// Count pass: per pair the number of contours and of vertices (no vertex is
// stored; open contours are tracked for their identity only).  Write pass:
// every emitted vertex goes to an arena slot at the pair's vertex offset,
// chained per contour; a closing contour is copied out in path order.
//
// With a staging buffer the count pass also keeps the vertices, chained per
// contour in a vertex pool that the pairs take CT_CHUNK slots at a time, and
// a record per closed contour (head, length, hole flag, next record of the
// pair).  The write pass then copies the kept pairs out in path order and
// sweeps again only the pairs that did not fit (the pools ran out), which
// the count pass listed.  A pass's time is its slowest wave's, so a stage
// pays off only when (nearly) every pair fits.
#include <mutex>

#include "clip_lane.hpp"
#include "clip_team.hpp"
#include "prt_common.hpp"
#include "prt_profile.hpp"

namespace prt {
namespace {

constexpr int CT_THREADS = 128;
constexpr int64_t CT_MAX_LANES = 262144;  // as the plain kernel's lanes

struct IvArrays {  // one interval list, icap slots per lane
    int32_t *eL, *eR, *sL, *sR, *lp, *rp;
    double *xL, *xR, *uL, *uR;  // x at the scanline and its uncertainty
};

struct ContourWs {
    PlainWs base;
    IvArrays iv[2];
    int32_t *phead, *ptail, *pcnt, *pfree;  // open contours, pcap slots per lane
    int icap, pcap;
    // team kernels only: x and uncertainty of the intervals above at the
    // sub-beam top (icap slots), and the x of a sub-beam's entries at its
    // ends for the beams one lane sweeps (cap slots)
    double *tL = nullptr, *tR = nullptr, *tuL = nullptr, *tuR = nullptr;
    double *sxa = nullptr, *sxc = nullptr;
};

struct ContourOut {  // write pass
    double *vx, *vy;
    int32_t *vnext;  // vertex arena, at the pairs' vertex offsets
    double *xy;
    int64_t *cstart;
    int32_t *chole;
};

// What the count pass keeps: sc[p] = contours of pair p kept, their records
// chained from cfirst[p]; -1 when the pair is on the spill list instead (it
// did not fit, or it failed).  ctr = {spilled pairs, vertex chunks taken,
// contour records taken, tag of the count pass's arguments}.
constexpr int CT_CHUNK = 8;
struct StageWs {
    unsigned long long *ctr;
    int64_t *spill;
    int32_t *sc, *cfirst;
    int32_t *chead, *clen, *chole, *cnext;  // ccap contour records
    int32_t *vnext;                         // vcap vertex slots
    double *vx, *vy;
    int64_t vcap = 0, ccap = 0;
};

enum BuildMode { CT_COUNT = 0, CT_STAGE = 1, CT_WRITE = 2 };

struct Pt {
    bool below, right;
    int iv;
    double x, unc;
};

// uncertainty of edge e's x at y (see the header)
__device__ inline double x_unc(const Lane &L, int e, double y) {
    const double ty = L.at(L.w.ty, e), by = L.at(L.w.by, e);
    if (y == ty || y == by) return 0.0;
    return 0x1p-44 * (fabs(L.at(L.w.bx, e)) + fabs(L.at(L.w.tx, e)) +
                      fabs(L.at(L.w.dx, e)) * (fabs(y) + fabs(by)));
}

// FIXED: the lists below and above the scanline stay in iv[0] and iv[1] and
// advance() copies the intervals above into iv[0] (the team kernels: a
// local ContourWs indexed by the runtime side would live in scratch memory);
// otherwise advance() flips which list is below.
template <int M, bool FIXED = false> struct Builder {
    static constexpr bool W = M != CT_COUNT;  // vertices go to an arena
    const Lane &L;
    const ContourWs &cw;
    const ContourOut &o;
    const StageWs *st = nullptr;  // CT_STAGE: the pools
    bool spill = false;           // CT_STAGE: the pair does not fit the pools
    int64_t chunk = 0, chunk_end = 0;
    int32_t first_rec = -1, last_rec = -1;
    int sb = 0;  // iv[sb] holds the intervals below the scanline
    int nb = 0, na = 0, nfree = 0, nnew = 0;
    int64_t nv = 0, nc = 0, written = 0;
    int64_t vbase = 0, vlim = 0, cbase = 0, clim = 0;
    bool bad = false;

    __device__ Builder(const Lane &l, const ContourWs &w, const ContourOut &out)
        : L(l), cw(w), o(out) {}

    __device__ int32_t &B(int32_t *const IvArrays::*f, int k) const {
        return L.at(cw.iv[FIXED ? 0 : sb].*f, k);
    }
    __device__ int32_t &A(int32_t *const IvArrays::*f, int k) const {
        return L.at(cw.iv[FIXED ? 1 : 1 - sb].*f, k);
    }
    __device__ double &Bx(double *const IvArrays::*f, int k) const {
        return L.at(cw.iv[FIXED ? 0 : sb].*f, k);
    }
    __device__ double &Ax(double *const IvArrays::*f, int k) const {
        return L.at(cw.iv[FIXED ? 1 : 1 - sb].*f, k);
    }

    __device__ int vertex(double x, double y) {
        const int64_t v = nv++;
        if (M == CT_STAGE) {  // a pool slot; -1 once the pair has spilled
            if (spill) return -1;
            if (chunk == chunk_end) {
                const int64_t c = static_cast<int64_t>(atomicAdd(st->ctr + 1, 1ull)) * CT_CHUNK;
                if (c + CT_CHUNK > st->vcap) {
                    spill = true;
                    return -1;
                }
                chunk = c;
                chunk_end = c + CT_CHUNK;
            }
            const int64_t g = chunk++;
            o.vx[g] = x;
            o.vy[g] = y;
            o.vnext[g] = -1;
            return static_cast<int>(g);
        }
        if (W) {
            if (v < vlim) {
                o.vx[vbase + v] = x;
                o.vy[vbase + v] = y;
                o.vnext[vbase + v] = -1;
            } else {
                bad = true;
            }
        }
        return static_cast<int>(v);
    }
    __device__ int32_t &ph(int p) const { return L.at(cw.phead, p); }
    __device__ int32_t &pt(int p) const { return L.at(cw.ptail, p); }
    __device__ int32_t &pc(int p) const { return L.at(cw.pcnt, p); }
    __device__ bool arena_ok(int v) const { return v >= 0 && v < vlim; }

    __device__ int poly_new(int h, int t, int cnt) {
        int p;
        if (nfree > 0) {
            p = L.at(cw.pfree, --nfree);
        } else if (nnew < cw.pcap) {
            p = nnew++;
        } else {
            bad = true;
            return 0;
        }
        if (W) {
            ph(p) = h;
            pt(p) = t;
            pc(p) = cnt;
        }
        return p;
    }
    __device__ void poly_free(int p) {
        if (nfree < cw.pcap) L.at(cw.pfree, nfree++) = p;
    }
    __device__ void prepend(int p, double x, double y) {
        const int v = vertex(x, y);
        if (W && arena_ok(v)) {
            o.vnext[vbase + v] = ph(p);
            ph(p) = v;
            pc(p) += 1;
        }
    }
    __device__ void append(int p, double x, double y) {
        const int v = vertex(x, y);
        if (W && arena_ok(v) && arena_ok(pt(p))) {
            o.vnext[vbase + pt(p)] = v;
            pt(p) = v;
            pc(p) += 1;
        }
    }
    // path of p, then path of q; every chain end of q now belongs to p
    __device__ __attribute__((always_inline)) void join(int p, int q) {
        if (W && arena_ok(pt(p))) {
            o.vnext[vbase + pt(p)] = ph(q);
            pt(p) = pt(q);
            pc(p) += pc(q);
        }
        for (int k = 0; k < nb; ++k) {
            if (B(&IvArrays::lp, k) == q) B(&IvArrays::lp, k) = p;
            if (B(&IvArrays::rp, k) == q) B(&IvArrays::rp, k) = p;
        }
        for (int k = 0; k < na; ++k) {
            if (A(&IvArrays::lp, k) == q) A(&IvArrays::lp, k) = p;
            if (A(&IvArrays::rp, k) == q) A(&IvArrays::rp, k) = p;
        }
        poly_free(q);
    }
    __device__ void close(int p, bool hole) {
        if (M == CT_STAGE) {
            if (!spill) {
                const unsigned long long r = atomicAdd(st->ctr + 2, 1ull);
                if (r >= static_cast<unsigned long long>(st->ccap)) {
                    spill = true;
                } else {
                    const int32_t q = static_cast<int32_t>(r);
                    st->chead[q] = ph(p);
                    st->clen[q] = pc(p);
                    st->chole[q] = hole ? 1 : 0;
                    st->cnext[q] = -1;
                    if (last_rec >= 0) st->cnext[last_rec] = q;
                    else first_rec = q;
                    last_rec = q;
                }
            }
        } else if (W) {
            if (nc < clim) {
                o.cstart[cbase + nc] = vbase + written;
                o.chole[cbase + nc] = hole ? 1 : 0;
            } else {
                bad = true;
            }
            int v = ph(p);
            for (int k = 0; k < pc(p); ++k) {
                if (!arena_ok(v) || written >= vlim) {
                    bad = true;
                    break;
                }
                o.xy[2 * (vbase + written)] = o.vx[vbase + v];
                o.xy[2 * (vbase + written) + 1] = o.vy[vbase + v];
                ++written;
                v = o.vnext[vbase + v];
            }
        }
        ++nc;
        poly_free(p);
    }
    __device__ static bool near(const Pt &u, const Pt &v) { return fabs(u.x - v.x) <= u.unc + v.unc; }
    // a chain end's contour; every interval side is assigned by the pairing
    // of the scanline it was made at, so this only guards the arena
    __device__ bool live(int p) {
        if (p >= 0 && p < cw.pcap) return true;
        bad = true;
        return false;
    }

    __device__ Pt point(bool below, int k) const {
        Pt p;
        p.below = below;
        p.right = (k & 1) != 0;
        p.iv = k >> 1;
        if (below) {
            p.x = p.right ? Bx(&IvArrays::xR, p.iv) : Bx(&IvArrays::xL, p.iv);
            p.unc = p.right ? Bx(&IvArrays::uR, p.iv) : Bx(&IvArrays::uL, p.iv);
        } else {
            p.x = p.right ? Ax(&IvArrays::xR, p.iv) : Ax(&IvArrays::xL, p.iv);
            p.unc = p.right ? Ax(&IvArrays::uR, p.iv) : Ax(&IvArrays::uL, p.iv);
        }
        return p;
    }

    __device__ void pair(double y, const Pt &u, const Pt &v) {
        if (u.below != v.below) {  // a chain continues upward
            const Pt &lo = u.below ? u : v, &hi = u.below ? v : u;
            const bool same = lo.right ? B(&IvArrays::eR, lo.iv) == A(&IvArrays::eR, hi.iv)
                                       : B(&IvArrays::eL, lo.iv) == A(&IvArrays::eL, hi.iv);
            const int p = lo.right ? B(&IvArrays::rp, lo.iv) : B(&IvArrays::lp, lo.iv);
            if (!live(p)) return;
            if (!same) {
                const bool two = !near(lo, hi);
                const double x1 = (two || !(hi.unc < lo.unc)) ? lo.x : hi.x;
                if (lo.right) {
                    append(p, x1, y);
                    if (two) append(p, hi.x, y);
                } else {
                    prepend(p, x1, y);
                    if (two) prepend(p, hi.x, y);
                }
            }
            if (lo.right) A(&IvArrays::rp, hi.iv) = p;
            else A(&IvArrays::lp, hi.iv) = p;
            return;
        }
        const bool two = !near(u, v);
        const double x1 = v.unc < u.unc ? v.x : u.x;  // the one vertex
        if (!u.below) {  // a contour starts
            int h, t;
            if (!u.right) {  // local minimum: u -> v
                h = t = vertex(two ? u.x : x1, y);
                if (two) t = vertex(v.x, y);
            } else {         // hole bottom: v -> u
                h = t = vertex(two ? v.x : x1, y);
                if (two) t = vertex(u.x, y);
            }
            if (W && two && arena_ok(h)) o.vnext[vbase + h] = t;
            const int p = poly_new(h, t, two ? 2 : 1);
            if (!u.right) {
                A(&IvArrays::lp, u.iv) = p;
                A(&IvArrays::rp, v.iv) = p;
            } else {
                A(&IvArrays::lp, v.iv) = p;
                A(&IvArrays::rp, u.iv) = p;
            }
            return;
        }
        if (!u.right) {  // local maximum: tail -> v -> u -> head
            const int p = B(&IvArrays::lp, u.iv), q = B(&IvArrays::rp, v.iv);
            if (!live(p) || !live(q)) return;
            if (two) append(q, v.x, y);
            append(q, two ? u.x : x1, y);
            if (p == q) close(q, false);
            else join(q, p);
        } else {         // hole top: tail -> u -> v -> head
            const int p = B(&IvArrays::rp, u.iv), q = B(&IvArrays::lp, v.iv);
            if (!live(p) || !live(q)) return;
            append(p, two ? u.x : x1, y);
            if (two) append(p, v.x, y);
            if (p == q) close(p, true);
            else join(p, q);
        }
    }

    __device__ __attribute__((always_inline)) void scanline(double y) {
        // merge of the points below and above by x (each list in its own
        // order; on equal x a right end first, then the point below), the
        // head of each list kept so that every point is read once
        int i = 0, j = 0;
        bool have = false;
        const int nbp = 2 * nb, nap = 2 * na;
        Pt pend{false, false, 0, 0.0, 0.0};
        Pt pb{true, false, 0, 0.0, 0.0}, pa{false, false, 0, 0.0, 0.0};
        if (nbp > 0) pb = point(true, 0);
        if (nap > 0) pa = point(false, 0);
        while (i < nbp || j < nap) {
            const bool take_b =
                j >= nap || (i < nbp && (pb.x != pa.x ? pb.x < pa.x : (pb.right || !pa.right)));
            const Pt p = take_b ? pb : pa;
            if (take_b) {
                if (++i < nbp) pb = point(true, i);
            } else {
                if (++j < nap) pa = point(false, j);
            }
            if (!have) {
                pend = p;
                have = true;
            } else {
                pair(y, pend, p);
                have = false;
            }
        }
    }

    // Team kernels (FIXED): the sub-beam's entries in order given directly --
    // edge ids (edge slot = id), x at the sub-beam bottom ya and top yc --
    // and each interval's values at both ends computed when it is made, so
    // that advance_direct() only moves them below.  Same intervals, x's and
    // uncertainties as intervals() and advance().
    __device__ void intervals_direct(const uint16_t *ids, const double *xa, const double *xc, int m,
                                     int nP, int op, double ya, double yc) {
        int pp = 0, pq = 0, start = 0, last = 0;
        bool open = false;
        na = 0;
        for (int t = 0; t + 1 < m; ++t) {
            if (ids[t] < nP) pp ^= 1;
            else pq ^= 1;
            if (xa[t] == xa[t + 1] && xc[t] == xc[t + 1]) continue;
            if (op_inside(op, pp, pq)) {
                if (!open) {
                    open = true;
                    start = t;
                }
                last = t + 1;
            } else if (open) {
                add_direct(ids, xa, xc, start, last, ya, yc);
                open = false;
            }
        }
        if (open) add_direct(ids, xa, xc, start, last, ya, yc);
    }
    __device__ void add_direct(const uint16_t *ids, const double *xa, const double *xc, int s,
                               int e, double ya, double yc) {
        if (na >= cw.icap) {
            bad = true;
            return;
        }
        const int es = ids[s], ee = ids[e];
        A(&IvArrays::eL, na) = es;
        A(&IvArrays::eR, na) = ee;
        Ax(&IvArrays::xL, na) = xa[s];
        Ax(&IvArrays::xR, na) = xa[e];
        Ax(&IvArrays::uL, na) = x_unc(L, es, ya);
        Ax(&IvArrays::uR, na) = x_unc(L, ee, ya);
        cw.tL[na] = xc[s];
        cw.tR[na] = xc[e];
        cw.tuL[na] = x_unc(L, es, yc);
        cw.tuR[na] = x_unc(L, ee, yc);
        A(&IvArrays::lp, na) = -1;
        A(&IvArrays::rp, na) = -1;
        ++na;
    }
    // Team kernels with a team of one wave: a scanline whose merge pairs
    // every interval below with the interval of the same index above --
    // left end with left end, right end with right end, as the serial merge
    // would (checked per interval by replaying its comparisons) -- is one
    // chain continuation per pair, each on its own chain end (an open
    // contour has one head and one tail), so the wave does them at once:
    // lane q the two pairs of interval q, vertex slots by a prefix sum over
    // the wave (pool chunks taken in one atomic), the contour's vertex count
    // by an LDS atomic.  Called by all 64 lanes with n = nb = na; lane 0
    // holds the Builder's state.  Returns false (nothing done) when the
    // pairing is any other.
    __device__ bool wave_continuations(double y, int n, int q) {
        const bool mine = q < n;
        Pt bl{true, false, q, 0.0, 0.0}, br{true, true, q, 0.0, 0.0};
        Pt al{false, false, q, 0.0, 0.0}, ar{false, true, q, 0.0, 0.0};
        bool ok = true;
        auto take_b = [](double bx, bool bright, double ax, bool aright) {
            return bx != ax ? bx < ax : (bright || !aright);
        };
        if (mine) {
            bl.x = Bx(&IvArrays::xL, q);
            bl.unc = Bx(&IvArrays::uL, q);
            br.x = Bx(&IvArrays::xR, q);
            br.unc = Bx(&IvArrays::uR, q);
            al.x = Ax(&IvArrays::xL, q);
            al.unc = Ax(&IvArrays::uL, q);
            ar.x = Ax(&IvArrays::xR, q);
            ar.unc = Ax(&IvArrays::uR, q);
            // first two points taken: bl and al (either first)
            ok = take_b(bl.x, false, al.x, false) ? !take_b(br.x, true, al.x, false)
                                                   : take_b(bl.x, false, ar.x, true);
            // next two: br and ar, then the next interval's left ends
            if (q + 1 < n) {
                if (take_b(br.x, true, ar.x, true))
                    ok = ok && !take_b(Bx(&IvArrays::xL, q + 1), false, ar.x, true);
                else
                    ok = ok && take_b(br.x, true, Ax(&IvArrays::xL, q + 1), false);
            }
        }
        if (__ballot(mine && !ok)) return false;
        int pL = 0, pR = 0, nL = 0, nR = 0;
        bool twoL = false, twoR = false;
        double x1L = 0.0, x1R = 0.0;
        if (mine) {
            pL = B(&IvArrays::lp, q);
            pR = B(&IvArrays::rp, q);
            if (B(&IvArrays::eL, q) != A(&IvArrays::eL, q)) {
                twoL = !near(bl, al);
                x1L = (twoL || !(al.unc < bl.unc)) ? bl.x : al.x;
                nL = twoL ? 2 : 1;
            }
            if (B(&IvArrays::eR, q) != A(&IvArrays::eR, q)) {
                twoR = !near(br, ar);
                x1R = (twoR || !(ar.unc < br.unc)) ? br.x : ar.x;
                nR = twoR ? 2 : 1;
            }
        }
        const bool dead = mine && !(pL >= 0 && pL < cw.pcap && pR >= 0 && pR < cw.pcap);
        if (__ballot(dead)) {  // as live() in pair(): the pair is discarded
            if (q == 0) bad = true;
            return true;
        }
        int incl = nL + nR;  // vertices of this lane, inclusive prefix over the wave
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int v = __shfl_up(incl, o, 64);
            if (q >= o) incl += v;
        }
        const int total = __shfl(incl, 63, 64);
        int g = incl - nL - nR;
        int64_t base = 0, rem = 0, fresh = 0;
        bool sp = false;
        if (M == CT_STAGE) {
            int64_t ch = __shfl(chunk, 0, 64), che = __shfl(chunk_end, 0, 64);
            sp = __shfl(static_cast<int>(spill), 0, 64) != 0;
            rem = che - ch;
            base = ch;
            if (!sp && total > rem) {
                const int64_t nch = (total - rem + CT_CHUNK - 1) / CT_CHUNK;
                int64_t c = 0;
                if (q == 0) c = static_cast<int64_t>(atomicAdd(st->ctr + 1, static_cast<unsigned long long>(nch))) * CT_CHUNK;
                c = __shfl(c, 0, 64);
                if (c + nch * CT_CHUNK > st->vcap) {
                    sp = true;
                } else {
                    fresh = c;
                    if (q == 0) {
                        chunk = c + (total - rem);
                        chunk_end = c + nch * CT_CHUNK;
                    }
                }
            } else if (!sp && q == 0) {
                chunk = ch + total;
            }
            if (q == 0) spill = sp;
        }
        auto slot = [&](int k) -> int {
            if (M != CT_STAGE) return k;
            if (sp) return -1;
            return static_cast<int>(k < rem ? base + k : fresh + (k - rem));
        };
        auto put = [&](int v, double x) {
            if (M == CT_STAGE && v >= 0) {
                o.vx[v] = x;
                o.vy[v] = y;
                o.vnext[v] = -1;
            }
        };
        if (W && mine) {
            if (nL) {
                const int v1 = slot(g++);
                put(v1, x1L);
                if (arena_ok(v1)) {
                    o.vnext[vbase + v1] = ph(pL);
                    ph(pL) = v1;
                    atomicAdd(&pc(pL), 1);
                }
                if (twoL) {
                    const int v2 = slot(g++);
                    put(v2, al.x);
                    if (arena_ok(v2)) {
                        o.vnext[vbase + v2] = ph(pL);
                        ph(pL) = v2;
                        atomicAdd(&pc(pL), 1);
                    }
                }
            }
            if (nR) {
                const int v1 = slot(g++);
                put(v1, x1R);
                if (arena_ok(v1) && arena_ok(pt(pR))) {
                    o.vnext[vbase + pt(pR)] = v1;
                    pt(pR) = v1;
                    atomicAdd(&pc(pR), 1);
                }
                if (twoR) {
                    const int v2 = slot(g++);
                    put(v2, ar.x);
                    if (arena_ok(v2) && arena_ok(pt(pR))) {
                        o.vnext[vbase + pt(pR)] = v2;
                        pt(pR) = v2;
                        atomicAdd(&pc(pR), 1);
                    }
                }
            }
        }
        if (mine) {
            A(&IvArrays::lp, q) = pL;
            A(&IvArrays::rp, q) = pR;
        }
        if (q == 0) nv += total;
        return true;
    }

    __device__ void advance_direct() {
        for (int k = 0; k < na; ++k) {
            B(&IvArrays::eL, k) = A(&IvArrays::eL, k);
            B(&IvArrays::eR, k) = A(&IvArrays::eR, k);
            B(&IvArrays::lp, k) = A(&IvArrays::lp, k);
            B(&IvArrays::rp, k) = A(&IvArrays::rp, k);
            Bx(&IvArrays::xL, k) = cw.tL[k];
            Bx(&IvArrays::xR, k) = cw.tR[k];
            Bx(&IvArrays::uL, k) = cw.tuL[k];
            Bx(&IvArrays::uR, k) = cw.tuR[k];
        }
        nb = na;
        na = 0;
    }

    // A <- inside intervals of the sub-beam (ord sorted), x at its bottom
    __device__ void intervals(int m, int nP, int op, double ya) {
        int pp = 0, pq = 0, start = 0, last = 0;
        bool open = false;
        na = 0;
        for (int t = 0; t + 1 < m; ++t) {
            const int l = L.at(L.w.ord, t), r = L.at(L.w.ord, t + 1);
            if (L.at(L.w.id, L.at(L.w.act, l)) < nP) pp ^= 1;
            else pq ^= 1;
            if (L.at(L.w.xa, l) == L.at(L.w.xa, r) && L.at(L.w.xc, l) == L.at(L.w.xc, r)) continue;
            if (op_inside(op, pp, pq)) {
                if (!open) {
                    open = true;
                    start = l;
                }
                last = r;
            } else if (open) {
                add_interval(start, last, ya);
                open = false;
            }
        }
        if (open) add_interval(start, last, ya);
    }
    __device__ void add_interval(int s, int e, double ya) {
        if (na >= cw.icap) {
            bad = true;
            return;
        }
        A(&IvArrays::eL, na) = L.at(L.w.id, L.at(L.w.act, s));
        A(&IvArrays::eR, na) = L.at(L.w.id, L.at(L.w.act, e));
        A(&IvArrays::sL, na) = s;
        A(&IvArrays::sR, na) = e;
        Ax(&IvArrays::xL, na) = L.at(L.w.xa, s);
        Ax(&IvArrays::xR, na) = L.at(L.w.xa, e);
        Ax(&IvArrays::uL, na) = x_unc(L, L.at(L.w.act, s), ya);
        Ax(&IvArrays::uR, na) = x_unc(L, L.at(L.w.act, e), ya);
        A(&IvArrays::lp, na) = -1;
        A(&IvArrays::rp, na) = -1;
        ++na;
    }
    // the intervals above become the ones below, x at the sub-beam top yc
    __device__ void advance(double yc) {
        if (FIXED) {
            for (int k = 0; k < na; ++k) {
                B(&IvArrays::eL, k) = A(&IvArrays::eL, k);
                B(&IvArrays::eR, k) = A(&IvArrays::eR, k);
                B(&IvArrays::sL, k) = A(&IvArrays::sL, k);
                B(&IvArrays::sR, k) = A(&IvArrays::sR, k);
                B(&IvArrays::lp, k) = A(&IvArrays::lp, k);
                B(&IvArrays::rp, k) = A(&IvArrays::rp, k);
            }
        } else {
            sb = 1 - sb;
        }
        nb = na;
        na = 0;
        for (int k = 0; k < nb; ++k) {
            const int sl = B(&IvArrays::sL, k), sr = B(&IvArrays::sR, k);
            Bx(&IvArrays::xL, k) = L.at(L.w.xc, sl);
            Bx(&IvArrays::xR, k) = L.at(L.w.xc, sr);
            Bx(&IvArrays::uL, k) = x_unc(L, L.at(L.w.act, sl), yc);
            Bx(&IvArrays::uR, k) = x_unc(L, L.at(L.w.act, sr), yc);
        }
    }
};

template <int M>
__device__ void contours_one(Builder<M> &cb, const double *__restrict__ P, int nP,
                             const double *__restrict__ Q, int nQ, int op) {
    const Lane &L = cb.L;
    double bp[4], bq[4];
    int ne = 0, ny = 0;
    if (!lane_setup(L, P, nP, Q, nQ, op, bp, bq, ne, ny)) return;
    int m = 0, next = 0;
    for (int k = 0; k + 1 < ny; ++k) {
        const double yb = L.at(L.w.ys, k), yt = L.at(L.w.ys, k + 1);
        while (next < ne && L.at(L.w.by, L.at(L.w.eord, next)) <= yb)
            L.at(L.w.act, m++) = L.at(L.w.eord, next++);
        int keep = 0;
        for (int i = 0; i < m; ++i) {
            const int e = L.at(L.w.act, i);
            if (L.at(L.w.ty, e) > yb) L.at(L.w.act, keep++) = e;
        }
        m = keep;
        if (m < 2) {  // nothing inside this beam
            cb.na = 0;
            cb.scanline(yb);
            cb.nb = 0;
            continue;
        }
        for (int i = 0; i < m; ++i) {
            const int e = L.at(L.w.act, i);
            const double b = edge_x(L, e, yb), t = edge_x(L, e, yt);
            L.at(L.w.xb, i) = b;
            L.at(L.w.xt, i) = t;
            L.at(L.w.xa, i) = b;
        }
        const double dy = yt - yb;
        double ya = yb;
        for (;;) {
            double yc = yt;  // next distinct crossing height above ya
            for (int i = 0; i < m; ++i) {
                const double xbi = L.at(L.w.xb, i), xti = L.at(L.w.xt, i);
                for (int j = i + 1; j < m; ++j) {
                    const double xbj = L.at(L.w.xb, j), xtj = L.at(L.w.xt, j);
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
            if (yc == yt) {
                for (int i = 0; i < m; ++i) L.at(L.w.xc, i) = L.at(L.w.xt, i);
            } else {
                for (int i = 0; i < m; ++i) L.at(L.w.xc, i) = edge_x(L, L.at(L.w.act, i), yc);
            }
            sort_sub_beam(L, m);
            cb.intervals(m, nP, op, ya);
            cb.scanline(ya);
            cb.advance(yc);
            if (yc == yt) break;
            for (int i = 0; i < m; ++i) L.at(L.w.xa, i) = L.at(L.w.xc, i);
            ya = yc;
        }
    }
    cb.na = 0;
    cb.scanline(L.at(L.w.ys, ny - 1));
    cb.nb = 0;
}

// Count pass: ncont[p], nvert[p]; -1 in both when the pair does not fit the
// lane (more than max_verts vertices) or its arena bounds.  CT_STAGE also
// keeps the pair's contours in st, or lists the pair in st.spill.  With a
// list (the pairs the team kernels passed on) only the listed pairs.
template <int M>
__global__ __launch_bounds__(CT_THREADS) void k_contours_count(
    const uint32_t *__restrict__ pairs, int64_t np, const double *__restrict__ xyA,
    const int64_t *__restrict__ offA, const double *__restrict__ xyB,
    const int64_t *__restrict__ offB, int op, ContourWs cw, StageWs st,
    int64_t *__restrict__ ncont, int64_t *__restrict__ nvert, const uint32_t *__restrict__ list,
    const unsigned long long *__restrict__ list_count) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= cw.base.T) return;
    Lane L(cw.base, t);
    ContourOut so{};
    if (M == CT_STAGE) {
        so.vx = st.vx;
        so.vy = st.vy;
        so.vnext = st.vnext;
    }
    int64_t cnt = np;
    if (list) {
        const unsigned long long k = *list_count;
        cnt = k < static_cast<unsigned long long>(np) ? static_cast<int64_t>(k) : np;
    }
    for (int64_t i = t; i < cnt; i += cw.base.T) {
        const int64_t p = list ? static_cast<int64_t>(list[i]) : i;
        const uint32_t a = pairs[2 * p], b = pairs[2 * p + 1];
        const int64_t oa = offA[a], ob = offB[b];
        const int nA = static_cast<int>(offA[a + 1] - oa), nB = static_cast<int>(offB[b + 1] - ob);
        Builder<M> cb(L, cw, so);
        if (M == CT_STAGE) {  // pool slots are global: no base, the pool's bound
            cb.st = &st;
            cb.vlim = st.vcap;
        }
        if (nA + nB > cw.base.cap && !(nA < 3 && nB < 3)) cb.bad = true;
        else contours_one(cb, xyA + 2 * oa, nA, xyB + 2 * ob, nB, op);
        ncont[p] = cb.bad ? -1 : cb.nc;
        nvert[p] = cb.bad ? -1 : cb.nv;
        if (M == CT_STAGE) {
            const bool kept = !cb.bad && !cb.spill;
            st.sc[p] = kept ? static_cast<int32_t>(cb.nc) : -1;
            st.cfirst[p] = cb.first_rec;
            if (!kept) st.spill[atomicAdd(st.ctr, 1ull)] = p;
        }
    }
}

// ---- team per pair ---------------------------------------------------------
//
// The count pass of pairs that fit a team: a team of TL lanes (a slice of one
// wave64) takes a pair from its vertex-count bucket's ticket and builds the
// batched overlay's tables in LDS (edge table, scanbeam table, per-beam
// active-edge lists; clip_team.hpp).  The lanes sort every beam's entries by
// the sub-beam key at once; a beam whose sorted entries cross is then swept by
// the whole team (crossing heights and each sub-beam's order by bitonic sorts
// across the team; beams with more entries or heights than the team's
// scratch by lane 0 alone).  Lane 0 assembles the contours beam after beam
// with the one-lane kernel's Builder, whose interval lists and open-contour
// table live in the team's LDS: the same sub-beams, orders, x's and
// uncertainties in the same sequence, so the same contours, vertex for
// vertex.  A pair that does not fit the team's tables (more vertices than the
// bucket's bound, a beam longer than the list capacity, more intervals or
// open contours than the LDS arrays hold) goes to the one-lane kernel
// through a list.

struct TeamContourArgs {
    const uint32_t *pairs;
    const uint32_t *order;  // pair indices grouped by bucket
    const int64_t *start;   // first index of this bucket in order (device)
    const int64_t *count;   // number of pairs in the bucket (device)
    const double *xyA, *xyB;
    const int64_t *offA, *offB;
    unsigned long long *ticket;
    uint32_t *ovf_list;  // pairs left to the one-lane kernel
    unsigned long long *ovf_count;
    int64_t *ncont, *nvert;
    StageWs st;
    int NV, SCAP, ICAP, PCAP;
    int op;
};

// interval slots and open contours a team holds (capped below the one-lane
// kernel's bounds: pairs that need more take the one-lane kernel)
__host__ __device__ inline int team_icap(int NV) { return NV / 2 + 1 < 33 ? NV / 2 + 1 : 33; }
__host__ __device__ inline int team_pcap(int NV) { return NV + 2 < 66 ? NV + 2 : 66; }

// LDS of one team: the overlay tables, then the x at the sub-beam ends of a
// beam one lane sweeps, the two interval lists (edge ids, chains, x and
// uncertainty), the values at the top of the intervals above, the open
// contours
__host__ __device__ inline size_t contour_team_bytes(int NV, int SCAP) {
    const size_t ic = team_icap(NV), pc = team_pcap(NV);
    return team_bytes(NV, SCAP) + 2 * align16(8 * size_t(NV)) +
           2 * (4 * align16(4 * ic) + 4 * align16(8 * ic)) + 4 * align16(8 * ic) +
           4 * align16(4 * pc);
}

// The Builder's view of a team's LDS: a lane arena of one lane (stride 1)
// whose edge table is the team's (edge id = edge slot); the sub-beam arrays
// of the one-lane sweep are not used (the team hands the entries over
// directly).
__device__ inline ContourWs carve_contour_team(char *p, const TeamLds &t, int NV, int SCAP) {
    ContourWs cw;
    PlainWs &w = cw.base;
    p += team_bytes(NV, SCAP);
    auto take_i = [&](size_t n) {
        int32_t *q = reinterpret_cast<int32_t *>(p);
        p += align16(4 * n);
        return q;
    };
    auto take_d = [&](size_t n) {
        double *q = reinterpret_cast<double *>(p);
        p += align16(8 * n);
        return q;
    };
    w.bx = t.bx;
    w.by = t.by;
    w.tx = t.tx;
    w.ty = t.ty;
    w.dx = t.dx;
    w.ys = t.ys;
    w.xb = w.xt = w.bsum = w.xa = w.xc = nullptr;
    w.eord = w.id = w.act = w.ord = nullptr;
    w.T = 1;
    w.cap = NV;
    cw.sxa = take_d(NV);
    cw.sxc = take_d(NV);
    cw.icap = team_icap(NV);
    cw.pcap = team_pcap(NV);
    for (IvArrays &v : cw.iv) {
        v.eL = take_i(cw.icap);
        v.eR = take_i(cw.icap);
        v.sL = v.sR = nullptr;
        v.lp = take_i(cw.icap);
        v.rp = take_i(cw.icap);
        v.xL = take_d(cw.icap);
        v.xR = take_d(cw.icap);
        v.uL = take_d(cw.icap);
        v.uR = take_d(cw.icap);
    }
    cw.tL = take_d(cw.icap);
    cw.tR = take_d(cw.icap);
    cw.tuL = take_d(cw.icap);
    cw.tuR = take_d(cw.icap);
    cw.phead = take_i(cw.pcap);
    cw.ptail = take_i(cw.pcap);
    cw.pcnt = take_i(cw.pcap);
    cw.pfree = take_i(cw.pcap);
    return cw;
}

// The intervals above a scanline built by a whole wave (a team of 64 lanes,
// m <= 64 entries): lane q holds entry q and the gap between entries q and
// q + 1.  Parities of the gaps are prefix counts of ballots; a gap with
// equal x at both ends is skipped; a run of inside gaps starts at an inside
// gap whose previous non-skipped gap is outside (or none) and ends at one
// whose next non-skipped gap is outside (or none) -- the runs the Builder's
// serial loop finds.  The k-th start and the k-th end make interval k,
// written with its values at both ends (x, uncertainty at ya and yc).
// Returns the number of intervals (on every lane); at most icap are
// written.
__device__ inline int wave_intervals(const ContourWs &cw, const Lane &L, const uint16_t *ids,
                                     const double *xa, const double *xc, int m, int nP, int op,
                                     double ya, double yc, int q) {
    const bool have = q < m, gap = q + 1 < m;
    const int e = have ? ids[q] : 0;
    const uint64_t mP = __ballot(have && e < nP), mQ = __ballot(have && !(e < nP));
    const uint64_t upto = q == 63 ? ~0ull : (2ull << q) - 1;  // lanes <= q
    const uint64_t below = (1ull << q) - 1;                   // lanes < q
    const int pp = __popcll(mP & upto) & 1, pq = __popcll(mQ & upto) & 1;
    double xa0 = 0.0, xc0 = 0.0, xa1 = 0.0, xc1 = 0.0;
    int e1 = 0;
    if (have) {
        xa0 = xa[q];
        xc0 = xc[q];
    }
    if (gap) {
        xa1 = xa[q + 1];
        xc1 = xc[q + 1];
        e1 = ids[q + 1];
    }
    const bool nz = gap && !(xa0 == xa1 && xc0 == xc1);
    const bool in = nz && op_inside(op, pp, pq);
    const uint64_t mNZ = __ballot(nz), mIN = __ballot(in);
    const uint64_t nzb = mNZ & below, nza = mNZ & ~upto;
    const bool prev_in = nzb != 0 && ((mIN >> (63 - __clzll(nzb))) & 1ull);
    const bool next_in = nza != 0 && ((mIN >> (__ffsll(static_cast<long long>(nza)) - 1)) & 1ull);
    const bool start = in && !prev_in, end = in && !next_in;
    const uint64_t mS = __ballot(start), mE = __ballot(end);
    const IvArrays &A = cw.iv[1];
    if (start) {
        const int k = __popcll(mS & below);
        if (k < cw.icap) {
            A.eL[k] = e;
            A.xL[k] = xa0;
            A.uL[k] = x_unc(L, e, ya);
            cw.tL[k] = xc0;
            cw.tuL[k] = x_unc(L, e, yc);
            A.lp[k] = -1;
            A.rp[k] = -1;
        }
    }
    if (end) {
        const int k = __popcll(mE & below);
        if (k < cw.icap) {
            A.eR[k] = e1;
            A.xR[k] = xa1;
            A.uR[k] = x_unc(L, e1, ya);
            cw.tR[k] = xc1;
            cw.tuR[k] = x_unc(L, e1, yc);
        }
    }
    return __popcll(mS);
}

// The intervals above become the ones below (x and uncertainty at the
// sub-beam top), by every lane of the team.
template <int TL> __device__ inline void team_advance(const ContourWs &cw, int na, int tl) {
    const IvArrays &B = cw.iv[0], &A = cw.iv[1];
    for (int k = tl; k < na; k += TL) {
        B.eL[k] = A.eL[k];
        B.eR[k] = A.eR[k];
        B.lp[k] = A.lp[k];
        B.rp[k] = A.rp[k];
        B.xL[k] = cw.tL[k];
        B.xR[k] = cw.tR[k];
        B.uL[k] = cw.tuL[k];
        B.uR[k] = cw.tuR[k];
    }
}

// Next sub-beam of the general sweep of one beam by lane 0 alone (more
// entries or crossing heights than the team's scratch): the lowest crossing
// height above ya (yt when none), and the region [lo, hi) insertion-sorted
// by the strict key (xa + xc, xa, edge id) at (ya, yc) -- the one-lane
// kernel's sweep step.
__device__ double serial_next(const TeamLds &t, int lo, int hi, double yb, double yt, double ya) {
    const double dy = yt - yb;
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
    return yc;
}

// Contours of one pair by the team, the Builder's state on lane 0.  Every
// sub-beam -- of a beam with no entries inside (its intervals are none), of
// a crossing-free beam (its sorted entries), of a beam with crossings (swept
// by the team, or by lane 0 when it exceeds the team's scratch), and the
// final scanline at the top height -- reaches the Builder through one call
// site.  Returns 1 when the pair was handled (the Builder holds its counts
// unless it went bad), 0 when it must go to the one-lane kernel.
template <int TL, int M>
__device__ int team_contour_pair(const TeamLds &t, const ContourWs &cw, const TeamContourArgs &A,
                                 uint32_t p, int tl, int base, Builder<M, true> &cb) {
    const uint32_t a = A.pairs[2 * p], b = A.pairs[2 * p + 1];
    const int64_t oa = A.offA[a], ob = A.offB[b];
    int nP = static_cast<int>(A.offA[a + 1] - oa);
    int nQ = static_cast<int>(A.offB[b + 1] - ob);
    const double *P = A.xyA + 2 * oa, *Q = A.xyB + 2 * ob;
    const bool eP = nP < 3, eQ = nQ < 3;  // empty operands
    if (eP && eQ) return 1;
    if (nP + nQ > A.NV) return 0;
    if (eP) nP = 0;
    if (eQ) nQ = 0;
    const int n = nP + nQ;
    const int NY = team_tables<TL>(t, P, nP, Q, nQ, eP, eQ, A.op, tl, base);
    if (NY == 0) return 1;
    const int nb = NY - 1;
    const double INF = __builtin_huge_val();
    for (int k0 = 0; k0 < nb;) {
        const int k1 = team_window<TL>(t, k0, nb, n, A.SCAP, tl);
        if (k1 < 0) return 0;
        const int wbase = t.boff[k0];
        team_sort_window<TL>(t, k0, k1, tl);  // cnt[k] = 1 when the beam's edges cross
        const int kend = k1 == nb ? nb + 1 : k1;  // k = nb: the final scanline
        for (int k = k0; k < kend; ++k) {
            int lo = 0, hi = 0;
            double yb = t.ys[k], yt = yb;
            if (k < nb) {
                lo = t.boff[k] - wbase;
                hi = t.boff[k + 1] - wbase;
                yt = t.ys[k + 1];
            }
            const int m = hi - lo;
            // 0: one sub-beam; 1: swept by the team; 2: swept by lane 0
            int mode = (m < 2 || t.cnt[k] == 0) ? 0 : (m <= t.GS ? 1 : 2);
            int nc = 0;
            if (mode == 1) {  // the beam's crossing heights
                if constexpr (TL == 64) {
                    // the pairs spread over the wave, heights compacted by
                    // ballots (lane i looping over the pairs (i, j > i) left
                    // the first lanes with the longest loops)
                    team_sync();
                    nc = wave_crossing_heights(t, lo, hi, yb, yt, tl);
                    team_sync();
                } else {
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
                    nc = t.gmisc[1];
                }
                if (nc > t.GC) {
                    mode = 2;  // more heights than the scratch holds
                } else if (nc > 1) {
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
            }
            double ya = yb;
            int ci = 0;
            for (;;) {
                double yc = yt;
                // the sub-beam's entries in order: edge ids, x at ya and yc
                const uint16_t *ids = t.list + lo;
                const double *xa = t.lxb + lo, *xc = t.lxt + lo;
                if (mode == 1) {
                    while (ci < nc && !(t.giy[ci] > ya)) ++ci;
                    yc = ci < nc ? t.giy[ci] : yt;
                    if (TL == 64) {
                        wave_sort_sub_beam(t, lo, m, yb, yt, ya, yc, tl);
                        team_sync();
                    } else {
                        const int M2 = pow2_ge(m);
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
                    ids = t.gidx;
                    xa = t.gxa;
                    xc = t.gxc;
                } else if (mode == 2) {
                    double v = 0.0;
                    if (tl == 0) {
                        v = serial_next(t, lo, hi, yb, yt, ya);
                        for (int q = 0; q < m; ++q) {
                            const int i = lo + q;
                            cw.sxa[q] = ya == yb ? t.lxb[i] : (ya == yt ? t.lxt[i] : ex(t, t.list[i], ya));
                            cw.sxc[q] = v == yb ? t.lxb[i] : (v == yt ? t.lxt[i] : ex(t, t.list[i], v));
                        }
                    }
                    yc = __shfl(v, base, 64);
                    xa = cw.sxa;
                    xc = cw.sxc;
                }
                if (TL == 64 && m <= 64) {  // intervals and their move by the wave
                    const int na = wave_intervals(cw, cb.L, ids, xa, xc, m, nP, A.op, ya, yc, tl);
                    team_sync();
                    const int nm = na < cw.icap ? na : cw.icap;
                    if (tl == 0) {
                        cb.na = nm;
                        if (na > cw.icap) cb.bad = true;
                    }
                    const int nbu = __shfl(cb.nb, base, 64);
                    bool done = nbu == 0 && nm == 0;  // no point on this scanline
                    if (!done && nbu == nm) done = cb.wave_continuations(ya, nm, tl);
                    team_sync();
                    if (!done && tl == 0) cb.scanline(ya);
                    team_sync();
                    team_advance<TL>(cw, nm, tl);
                    if (tl == 0) {
                        cb.nb = nm;
                        cb.na = 0;
                    }
                    team_sync();
                } else {
                    if (tl == 0) {
                        cb.intervals_direct(ids, xa, xc, m, nP, A.op, ya, yc);
                        cb.scanline(ya);
                        cb.advance_direct();
                    }
                    if (mode != 0) team_sync();  // the team's scratch is rewritten next
                }
                if (yc == yt) break;
                ya = yc;
            }
        }
        team_sync();  // lane 0's reads of the lists end before the next window's
        if (__shfl(static_cast<int>(cb.bad), base, 64)) return 0;
        k0 = k1;
    }
    return 1;
}

template <int TL, int M>
__global__ __launch_bounds__(BK_THREADS_MAX) void k_contours_team(TeamContourArgs A) {
    extern __shared__ __align__(16) char lds[];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tiw = lane / TL, tl = lane % TL;
    const int base = tiw * TL;
    const int team = wave * (64 / TL) + tiw;
    char *mine = lds + team * contour_team_bytes(A.NV, A.SCAP);
    const TeamLds t = carve_team(mine, A.NV, A.SCAP);
    const ContourWs cw = carve_contour_team(mine, t, A.NV, A.SCAP);
    const Lane L(cw.base, 0);
    ContourOut so{};
    if (M == CT_STAGE) {
        so.vx = A.st.vx;
        so.vy = A.st.vy;
        so.vnext = A.st.vnext;
    }
    const int64_t count = *A.count;
    const uint32_t *order = A.order + *A.start;
    for (;;) {
        unsigned long long idx = 0;
        if (tl == 0) idx = atomicAdd(A.ticket, 1ull);
        idx = __shfl(idx, base, 64);
        if (static_cast<int64_t>(idx) >= count) break;
        const uint32_t p = order[idx];
        Builder<M, true> cb(L, cw, so);
        if (M == CT_STAGE) {  // pool slots are global: no base, the pool's bound
            cb.st = &A.st;
            cb.vlim = A.st.vcap;
        }
        const int ok = team_contour_pair<TL, M>(t, cw, A, p, tl, base, cb);
        if (tl == 0) {
            if (!ok || cb.bad) {
                A.ovf_list[atomicAdd(A.ovf_count, 1ull)] = p;
            } else {
                A.ncont[p] = cb.nc;
                A.nvert[p] = cb.nv;
                if (M == CT_STAGE) {
                    A.st.sc[p] = cb.spill ? -1 : static_cast<int32_t>(cb.nc);
                    A.st.cfirst[p] = cb.first_rec;
                    if (cb.spill) A.st.spill[atomicAdd(A.st.ctr, 1ull)] = p;
                }
            }
        }
        team_sync();
    }
}

// Appends every pair of a bucket to the one-lane list (bucket too large for
// the team's LDS).
__global__ void k_contours_route(TeamContourArgs A) {
    const int64_t count = *A.count;
    const uint32_t *order = A.order + *A.start;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        A.ovf_list[atomicAdd(A.ovf_count, 1ull)] = order[i];
}

__global__ void k_stage_clear(unsigned long long *ctr, unsigned long long tag) {
    if (threadIdx.x < 3) ctr[threadIdx.x] = 0;
    if (threadIdx.x == 3) ctr[3] = tag;
}

// Write pass of the kept pairs: the pair's contour records in closing order,
// each contour's vertices along its chain in the vertex pool.  A stage whose
// tag is not this call's (the count pass ran with other arguments, or not
// into this stage) sets *status and is not read; every index read from the
// stage is checked as well, so a stage that does not belong to these counts
// sets *status instead of reading or writing out of bounds.
__global__ __launch_bounds__(256) void k_contours_copy(int64_t np, StageWs st,
                                                       unsigned long long tag,
                                                       const int64_t *__restrict__ cont_off,
                                                       const int64_t *__restrict__ vert_off,
                                                       ContourOut out, int32_t *status) {
    if (st.ctr[3] != tag) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(status, 1);
        return;
    }
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; p < np;
         p += stride) {
        const int32_t n = st.sc[p];
        if (n < 0) continue;  // swept by the write pass of the spill list
        const int64_t c0 = cont_off[p], v0 = vert_off[p], vend = vert_off[p + 1];
        bool ok = cont_off[p + 1] - c0 == n;
        int64_t w = v0;
        int32_t r = st.cfirst[p];
        for (int c = 0; ok && c < n; ++c) {
            if (r < 0 || r >= st.ccap) {
                ok = false;
                break;
            }
            out.cstart[c0 + c] = w;
            out.chole[c0 + c] = st.chole[r];
            int32_t v = st.chead[r];
            for (int32_t k = st.clen[r]; k > 0; --k) {
                if (v < 0 || v >= st.vcap || w >= vend) {
                    ok = false;
                    break;
                }
                out.xy[2 * w] = st.vx[v];
                out.xy[2 * w + 1] = st.vy[v];
                ++w;
                v = st.vnext[v];
            }
            r = st.cnext[r];
        }
        if (!ok || w != vend) atomicOr(status, 1);
    }
}

// Write pass at the scanned offsets; *status (cleared by the launch before)
// gets 1 when a pair did not produce what its counts promised.  With a list
// (the count pass's spill list) only the listed pairs are swept.
__global__ __launch_bounds__(CT_THREADS) void k_contours_write(
    const uint32_t *__restrict__ pairs, int64_t np, const double *__restrict__ xyA,
    const int64_t *__restrict__ offA, const double *__restrict__ xyB,
    const int64_t *__restrict__ offB, int op, ContourWs cw, const int64_t *__restrict__ cont_off,
    const int64_t *__restrict__ vert_off, ContourOut out, const int64_t *__restrict__ list,
    const unsigned long long *__restrict__ ctr, unsigned long long tag, int32_t *status) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= cw.base.T) return;
    Lane L(cw.base, t);
    int64_t cnt = np;
    if (list) {  // ctr[0] = listed pairs; the tag as in k_contours_copy
        const unsigned long long k = ctr[0];
        const bool ok = ctr[3] == tag && k <= static_cast<unsigned long long>(np);
        cnt = ok ? static_cast<int64_t>(k) : 0;
        if (!ok && t == 0) atomicOr(status, 1);
    }
    for (int64_t i = t; i < cnt; i += cw.base.T) {
        const int64_t p = list ? list[i] : i;
        if (p < 0 || p >= np) {
            atomicOr(status, 1);
            continue;
        }
        const uint32_t a = pairs[2 * p], b = pairs[2 * p + 1];
        const int64_t oa = offA[a], ob = offB[b];
        const int nA = static_cast<int>(offA[a + 1] - oa), nB = static_cast<int>(offB[b + 1] - ob);
        Builder<CT_WRITE> cb(L, cw, out);
        cb.vbase = vert_off[p];
        cb.vlim = vert_off[p + 1] - cb.vbase;
        cb.cbase = cont_off[p];
        cb.clim = cont_off[p + 1] - cb.cbase;
        if (nA + nB > cw.base.cap && !(nA < 3 && nB < 3)) cb.bad = true;
        else contours_one(cb, xyA + 2 * oa, nA, xyB + 2 * ob, nB, op);
        if (cb.bad || cb.nv != cb.vlim || cb.nc != cb.clim || cb.written != cb.vlim)
            atomicOr(status, 1);
    }
}

__global__ void k_status_clear(int32_t *status) {
    if (threadIdx.x == 0) *status = 0;
}

int64_t contour_lanes(int64_t np) {
    int64_t T = np < CT_MAX_LANES ? np : CT_MAX_LANES;
    return T < 1 ? 1 : T;
}

int contour_cap(int32_t max_verts) { return max_verts < 3 ? 3 : max_verts; }

// lane arrays: base, 2 interval lists of icap slots, 4 contour arrays of pcap
size_t contour_lane_bytes(int64_t T, int cap) {
    const int64_t icap = cap / 2 + 1, pcap = cap + 2;
    return lane_base_bytes(T, cap) +
           static_cast<size_t>(T) * (2 * icap * (6 * 4 + 4 * 8) + pcap * 4 * 4) + 64 * 256;
}

int carve_contours(ContourWs &cw, int64_t T, int cap, Carver &c) {
    carve_lane_base(c, cw.base, T, cap);
    cw.icap = cap / 2 + 1;
    cw.pcap = cap + 2;
    const size_t ni = static_cast<size_t>(T) * cw.icap, npc = static_cast<size_t>(T) * cw.pcap;
    for (IvArrays &v : cw.iv) {
        int32_t **ip[] = {&v.eL, &v.eR, &v.sL, &v.sR, &v.lp, &v.rp};
        for (int32_t **p : ip) *p = c.take<int32_t>(ni);
        double **dp[] = {&v.xL, &v.xR, &v.uL, &v.uR};
        for (double **p : dp) *p = c.take<double>(ni);
    }
    int32_t **pp[] = {&cw.phead, &cw.ptail, &cw.pcnt, &cw.pfree};
    for (int32_t **p : pp) *p = c.take<int32_t>(npc);
    return c.ok() ? PRT_OK : PRT_E_NOMEM;
}

// pool sizes for a vertex pool of vcap slots: a record per 3 slots (a
// contour has 3 vertices or more unless it is degenerate; a pair that runs
// out of records spills like one that runs out of slots)
constexpr int64_t CT_MAX_POOL = int64_t(1) << 31;
int64_t stage_records(int64_t vcap) { return vcap / 3 + 1; }

template <class C> void carve_stage(C &c, StageWs &st, int64_t np, int64_t vcap) {
    st.vcap = vcap;
    st.ccap = stage_records(vcap);
    st.ctr = c.template take<unsigned long long>(4);
    st.spill = c.template take<int64_t>(np);
    st.sc = c.template take<int32_t>(np);
    st.cfirst = c.template take<int32_t>(np);
    int32_t **rec[] = {&st.chead, &st.clen, &st.chole, &st.cnext};
    for (int32_t **q : rec) *q = c.template take<int32_t>(st.ccap);
    st.vnext = c.template take<int32_t>(vcap);
    st.vx = c.template take<double>(vcap);
    st.vy = c.template take<double>(vcap);
}

// Sizer twin of carve_stage
struct SizeTaker {
    Sizer s;
    template <class T> T *take(size_t n) {
        s.take<T>(n);
        return nullptr;
    }
};

size_t stage_bytes_for(int64_t np, int64_t vcap) {
    SizeTaker z;
    StageWs st;
    carve_stage(z, st, np, vcap);
    return z.s.bytes();
}

// largest vertex pool (< 2^31 slots) whose stage fits `bytes`, 0 if none
int64_t stage_pool(int64_t np, size_t bytes) {
    int64_t lo = 0, hi = CT_MAX_POOL - 1;
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo + 1) / 2;
        if (stage_bytes_for(np, mid) <= bytes) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// tag of a contour call's arguments, stamped into the stage by the count
// pass and checked by the write pass (FNV-1a over the words)
unsigned long long stage_tag(const void *pairs, int64_t np, const void *xyA, const void *offA,
                             const void *xyB, const void *offB, int op) {
    const unsigned long long w[] = {reinterpret_cast<uintptr_t>(pairs),
                                    static_cast<unsigned long long>(np),
                                    reinterpret_cast<uintptr_t>(xyA),
                                    reinterpret_cast<uintptr_t>(offA),
                                    reinterpret_cast<uintptr_t>(xyB),
                                    reinterpret_cast<uintptr_t>(offB),
                                    static_cast<unsigned long long>(op)};
    unsigned long long h = 0xcbf29ce484222325ull;
    for (unsigned long long v : w)
        for (int b = 0; b < 8; ++b) h = (h ^ ((v >> (8 * b)) & 0xFFu)) * 0x100000001b3ull;
    return h;
}

int carve_stage_checked(StageWs &st, int64_t np, const void *stage, size_t bytes) {
    const int64_t vcap = stage_pool(np, bytes);
    if (vcap < CT_CHUNK) return PRT_E_NOMEM;
    Carver c(const_cast<void *>(stage), bytes);
    carve_stage(c, st, np, vcap);
    return c.ok() ? PRT_OK : PRT_E_NOMEM;
}

// ---- the team path's host side ----------------------------------------

// bucketing arrays of the team path (after the lane arena in the workspace)
struct TeamCtWs {
    unsigned long long *counts, *fill, *tickets, *ovf_count;
    int64_t *starts, *counts_i64;
    uint32_t *order, *ovf_list;
};

template <class C> void team_ct_layout(C &c, int64_t np, TeamCtWs *w) {
    auto u64 = c.template take<unsigned long long>(4 * NBUCKET + 1);
    auto i64 = c.template take<int64_t>(2 * NBUCKET);
    auto ord = c.template take<uint32_t>(np > 0 ? np : 1);
    auto ovf = c.template take<uint32_t>(np > 0 ? np : 1);
    if (w) {
        w->counts = u64;
        w->fill = u64 + NBUCKET;
        w->tickets = u64 + 2 * NBUCKET;
        w->ovf_count = u64 + 4 * NBUCKET;
        w->starts = i64;
        w->counts_i64 = i64 + NBUCKET;
        w->order = ord;
        w->ovf_list = ovf;
    }
}

// Team shapes of the count pass per vertex-count bucket (the batched
// overlay's bounds), measured on the layer pairs and the grid workload.
prt_tune contour_tune_default() {
    prt_tune t;
    const int bounds[NBUCKET] = {32, 64, 128, 256, 512, 2 * PRT_MAX_POLY_VERTS};
    for (int b = 0; b < NBUCKET; ++b) {
        t.bucket_max_verts[b] = bounds[b];
        t.lanes_per_pair[b] = 64;  // a wave per pair: teams sharing a wave
        t.waves_per_block[b] = 1;  // diverge in lane 0's assembly
        t.list_factor[b] = 1;
    }
    return t;
}

// test hook state: 1 = every pair takes the one-lane kernel; shapes
int g_ct_one_lane = 0;
bool g_ct_tune_set = false;
prt_tune g_ct_tune;
std::mutex g_ct_mu;

// profiler labels of the team launches, per vertex-count bucket
const char *const kTeamLabels[NBUCKET] = {"k_contours_team[b0]", "k_contours_team[b1]",
                                          "k_contours_team[b2]", "k_contours_team[b3]",
                                          "k_contours_team[b4]", "k_contours_team[b5]"};

template <int TL, int M>
int launch_contour_team(const TeamContourArgs &A, int waves, size_t lds, int bucket,
                        hipStream_t s) {
    auto kern = k_contours_team<TL, M>;
    int per_cu = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, waves * 64, lds) !=
            hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    const unsigned grid = static_cast<unsigned>(device_cus() * per_cu);
    PRT_LAUNCH_AS(kTeamLabels[bucket], kern, grid, waves * 64, lds, s, A);
    PRT_LAUNCHED();
    return PRT_OK;
}

template <int M>
int launch_contour_team_any(const TeamContourArgs &A, int lanes, int waves, size_t lds,
                            int bucket, hipStream_t s) {
    switch (lanes) {
    case 8: return launch_contour_team<8, M>(A, waves, lds, bucket, s);
    case 16: return launch_contour_team<16, M>(A, waves, lds, bucket, s);
    case 32: return launch_contour_team<32, M>(A, waves, lds, bucket, s);
    default: return launch_contour_team<64, M>(A, waves, lds, bucket, s);
    }
}

// The count pass by teams: bucket the pairs by vertex count, one persistent
// launch per bucket (largest first), then the one-lane kernel over the
// pairs the teams passed on.
int contours_count_team(const uint32_t *pairs, int64_t np, const double *xyA,
                        const int64_t *offA, const double *xyB, const int64_t *offB,
                        int32_t max_verts, int op, const ContourWs &cw, const TeamCtWs &tw,
                        bool staged, const StageWs &st, int64_t *ncont, int64_t *nvert,
                        hipStream_t s) {
    prt_tune tune;
    {
        std::lock_guard<std::mutex> lk(g_ct_mu);
        tune = g_ct_tune_set ? g_ct_tune : contour_tune_default();
    }
    Bounds bd;
    for (int b = 0; b < NBUCKET; ++b) bd.v[b] = tune.bucket_max_verts[b];
    PRT_LAUNCH(k_zero_u64, 1, 64, 0, s, tw.counts, 4 * NBUCKET + 1);
    PRT_LAUNCHED();
    const unsigned g = static_cast<unsigned>(ceil_div(np, 256) < 2048 ? ceil_div(np, 256) : 2048);
    PRT_LAUNCH(k_bucket_count, g, 256, 0, s, pairs, nullptr, np, offA, offB, bd, tw.counts);
    PRT_LAUNCHED();
    PRT_LAUNCH(k_bucket_starts, 1, 64, 0, s, tw.counts, tw.starts, tw.counts_i64, tw.fill);
    PRT_LAUNCHED();
    PRT_LAUNCH(k_bucket_scatter, g, 256, 0, s, pairs, nullptr, np, offA, offB, bd, tw.starts,
               tw.fill, tw.order);
    PRT_LAUNCHED();
    int max_lds = 0, dev = 0;
    PRT_HIP(hipGetDevice(&dev));
    PRT_HIP(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    for (int b = NBUCKET - 1; b >= 0; --b) {
        const int lo = b == 0 ? 0 : tune.bucket_max_verts[b - 1];
        if (max_verts <= lo) continue;  // no pair can fall in this bucket
        TeamContourArgs A;
        A.pairs = pairs;
        A.order = tw.order;
        A.start = tw.starts + b;
        A.count = tw.counts_i64 + b;
        A.xyA = xyA;
        A.xyB = xyB;
        A.offA = offA;
        A.offB = offB;
        A.ticket = tw.tickets + b;
        A.ovf_list = tw.ovf_list;
        A.ovf_count = tw.ovf_count;
        A.ncont = ncont;
        A.nvert = nvert;
        A.st = st;
        const int nvb = tune.bucket_max_verts[b] < max_verts ? tune.bucket_max_verts[b] : max_verts;
        A.NV = nvb < 3 ? 3 : nvb;
        A.SCAP = scap_for(A.NV, tune.list_factor[b]);
        A.ICAP = team_icap(A.NV);
        A.PCAP = team_pcap(A.NV);
        A.op = op;
        const int L = tune.lanes_per_pair[b], W = tune.waves_per_block[b];
        const size_t lds = static_cast<size_t>(W) * (64 / L) * contour_team_bytes(A.NV, A.SCAP);
        int rc;
        if (lds > static_cast<size_t>(max_lds)) {
            PRT_LAUNCH(k_contours_route, 256, 256, 0, s, A);
            PRT_LAUNCHED();
            continue;
        }
        rc = staged ? launch_contour_team_any<CT_STAGE>(A, L, W, lds, b, s)
                    : launch_contour_team_any<CT_COUNT>(A, L, W, lds, b, s);
        if (rc) return rc;
    }
    const unsigned grid = static_cast<unsigned>(ceil_div(cw.base.T, CT_THREADS));
    auto one_lane = staged ? k_contours_count<CT_STAGE> : k_contours_count<CT_COUNT>;
    PRT_LAUNCH_AS("k_contours_count[list]", one_lane, grid, CT_THREADS, 0, s, pairs, np, xyA, offA,
                  xyB, offB, op, cw, st, ncont, nvert, tw.ovf_list, tw.ovf_count);
    PRT_LAUNCHED();
    return PRT_OK;
}

}  // namespace
}  // namespace prt

using namespace prt;

extern "C" size_t prt_overlay_contours_stage_size(int64_t npairs, int64_t pool_verts) {
    if (npairs < 0 || pool_verts < CT_CHUNK || pool_verts >= CT_MAX_POOL) return 0;
    return stage_bytes_for(npairs, pool_verts);
}

extern "C" size_t prt_overlay_contours_workspace_size(int64_t npairs, int32_t max_verts,
                                                      int64_t total_vertices) {
    if (npairs < 0 || max_verts < 0 || total_vertices < 0) return 0;
    Sizer s;
    s.take<char>(contour_lane_bytes(contour_lanes(npairs), contour_cap(max_verts)));
    s.take<double>(total_vertices);
    s.take<double>(total_vertices);
    s.take<int32_t>(total_vertices);
    s.take<int32_t>(64);  // status
    SizeTaker z;          // the team path's bucketing (count pass)
    team_ct_layout(z, npairs, nullptr);
    s.take<char>(z.s.bytes());
    return s.bytes();
}

namespace {
int contours_args(const uint32_t *pairs, int64_t npairs, const double *xyA, const int64_t *offA,
                  const double *xyB, const int64_t *offB, int32_t max_verts, int32_t op) {
    PRT_REQUIRE(op >= PRT_OP_DIFF && op <= PRT_OP_UNION, PRT_E_INVAL, "overlay_contours: op %d", op);
    PRT_REQUIRE(npairs >= 0, PRT_E_INVAL, "overlay_contours: npairs < 0");
    PRT_REQUIRE(npairs == 0 || (pairs && xyA && offA && xyB && offB), PRT_E_INVAL,
                "overlay_contours: null device pointer");
    PRT_REQUIRE(max_verts >= 0 && max_verts <= 2 * PRT_MAX_POLY_VERTS, PRT_E_INVAL,
                "overlay_contours: max_verts %d out of range", max_verts);
    return PRT_OK;
}
}  // namespace

extern "C" int prt_overlay_contours_count(const uint32_t *pairs, int64_t npairs, const double *xyA,
                                          const int64_t *offA, const double *xyB,
                                          const int64_t *offB, int32_t max_verts, int32_t op,
                                          int64_t *ncont, int64_t *nvert, void *stage,
                                          size_t stage_bytes, void *ws, size_t ws_bytes,
                                          void *stream) {
    int rc = contours_args(pairs, npairs, xyA, offA, xyB, offB, max_verts, op);
    if (rc) return rc;
    if (npairs == 0) return PRT_OK;
    PRT_REQUIRE(ncont && nvert && ws, PRT_E_INVAL, "overlay_contours_count: null pointer");
    const size_t need = prt_overlay_contours_workspace_size(npairs, max_verts, 0);
    PRT_REQUIRE(ws_bytes >= need, PRT_E_NOMEM, "overlay_contours_count: workspace %zu < %zu bytes",
                ws_bytes, need);
    Carver c(ws, ws_bytes);
    ContourWs cw;
    PRT_REQUIRE(carve_contours(cw, contour_lanes(npairs), contour_cap(max_verts), c) == PRT_OK,
                PRT_E_NOMEM, "overlay_contours_count: workspace too small");
    TeamCtWs tw;
    team_ct_layout(c, npairs, &tw);
    PRT_REQUIRE(c.ok(), PRT_E_NOMEM, "overlay_contours_count: workspace too small");
    const unsigned grid = static_cast<unsigned>(ceil_div(cw.base.T, CT_THREADS));
    StageWs st;
    if (stage) {
        PRT_REQUIRE(carve_stage_checked(st, npairs, stage, stage_bytes) == PRT_OK, PRT_E_NOMEM,
                    "overlay_contours_count: stage of %zu bytes too small for %lld pairs",
                    stage_bytes, static_cast<long long>(npairs));
        PRT_LAUNCH(k_stage_clear, 1, 64, 0, as_stream(stream), st.ctr,
                   stage_tag(pairs, npairs, xyA, offA, xyB, offB, op));
        PRT_LAUNCHED();
    }
    int one_lane;
    {
        std::lock_guard<std::mutex> lk(g_ct_mu);
        one_lane = g_ct_one_lane;
    }
    if (!one_lane && npairs < (int64_t(1) << 32))
        return contours_count_team(pairs, npairs, xyA, offA, xyB, offB, max_verts,
                                   static_cast<int>(op), cw, tw, stage != nullptr, st, ncont,
                                   nvert, as_stream(stream));
    if (stage)
        PRT_LAUNCH(k_contours_count<CT_STAGE>, grid, CT_THREADS, 0, as_stream(stream), pairs,
                   npairs, xyA, offA, xyB, offB, static_cast<int>(op), cw, st, ncont, nvert,
                   nullptr, nullptr);
    else
        PRT_LAUNCH(k_contours_count<CT_COUNT>, grid, CT_THREADS, 0, as_stream(stream), pairs,
                   npairs, xyA, offA, xyB, offB, static_cast<int>(op), cw, st, ncont, nvert,
                   nullptr, nullptr);
    PRT_LAUNCHED();
    return PRT_OK;
}

extern "C" int prt_overlay_contours_write(const uint32_t *pairs, int64_t npairs, const double *xyA,
                                          const int64_t *offA, const double *xyB,
                                          const int64_t *offB, int32_t max_verts, int32_t op,
                                          const int64_t *cont_off, const int64_t *vert_off,
                                          int64_t total_vertices, int64_t *cont_start,
                                          int32_t *cont_hole, double *xy, int32_t *status,
                                          const void *stage, size_t stage_bytes, void *ws,
                                          size_t ws_bytes, void *stream) {
    int rc = contours_args(pairs, npairs, xyA, offA, xyB, offB, max_verts, op);
    if (rc) return rc;
    PRT_REQUIRE(total_vertices >= 0, PRT_E_INVAL, "overlay_contours_write: total_vertices < 0");
    if (npairs == 0) return PRT_OK;
    PRT_REQUIRE(cont_off && vert_off && status && ws && (total_vertices == 0 || xy),
                PRT_E_INVAL, "overlay_contours_write: null pointer");
    const size_t need = prt_overlay_contours_workspace_size(npairs, max_verts, total_vertices);
    PRT_REQUIRE(ws_bytes >= need, PRT_E_NOMEM, "overlay_contours_write: workspace %zu < %zu bytes",
                ws_bytes, need);
    Carver c(ws, ws_bytes);
    ContourWs cw;
    PRT_REQUIRE(carve_contours(cw, contour_lanes(npairs), contour_cap(max_verts), c) == PRT_OK,
                PRT_E_NOMEM, "overlay_contours_write: workspace too small");
    ContourOut out;
    out.vx = c.take<double>(total_vertices);
    out.vy = c.take<double>(total_vertices);
    out.vnext = c.take<int32_t>(total_vertices);
    PRT_REQUIRE(c.ok(), PRT_E_NOMEM, "overlay_contours_write: workspace too small");
    out.xy = xy;
    out.cstart = cont_start;
    out.chole = cont_hole;
    const unsigned grid = static_cast<unsigned>(ceil_div(cw.base.T, CT_THREADS));
    StageWs st;
    if (stage)
        PRT_REQUIRE(carve_stage_checked(st, npairs, stage, stage_bytes) == PRT_OK, PRT_E_NOMEM,
                    "overlay_contours_write: stage of %zu bytes too small for %lld pairs",
                    stage_bytes, static_cast<long long>(npairs));
    const unsigned long long tag = stage_tag(pairs, npairs, xyA, offA, xyB, offB, op);
    PRT_LAUNCH(k_status_clear, 1, 64, 0, as_stream(stream), status);
    PRT_LAUNCHED();
    if (stage) {
        const int64_t cgrid = ceil_div(npairs, 256) < 65536 ? ceil_div(npairs, 256) : 65536;
        PRT_LAUNCH(k_contours_copy, static_cast<unsigned>(cgrid), 256, 0, as_stream(stream),
                   npairs, st, tag, cont_off, vert_off, out, status);
        PRT_LAUNCHED();
    }
    PRT_LAUNCH(k_contours_write, grid, CT_THREADS, 0, as_stream(stream), pairs, npairs, xyA, offA,
               xyB, offB, static_cast<int>(op), cw, cont_off, vert_off, out,
               stage ? st.spill : nullptr, stage ? st.ctr : nullptr, tag, status);
    PRT_LAUNCHED();
    return PRT_OK;
}

// Test hook: one_lane = 1 sends every pair of the count pass to the one-lane
// kernel; tune (null: the defaults) sets the team shapes per bucket.
extern "C" int prt_debug_contours(int32_t one_lane, const prt_tune *tune) {
    std::lock_guard<std::mutex> lk(g_ct_mu);
    if (tune) {
        for (int b = 0; b < NBUCKET; ++b) {
            const int L = tune->lanes_per_pair[b], W = tune->waves_per_block[b];
            PRT_REQUIRE(L == 8 || L == 16 || L == 32 || L == 64, PRT_E_INVAL,
                        "debug_contours: lanes_per_pair[%d] = %d", b, L);
            PRT_REQUIRE(W >= 1 && W <= 8, PRT_E_INVAL, "debug_contours: waves_per_block[%d] = %d",
                        b, W);
            PRT_REQUIRE(tune->list_factor[b] >= 1 && tune->list_factor[b] <= 8, PRT_E_INVAL,
                        "debug_contours: list_factor[%d] = %d", b, tune->list_factor[b]);
            PRT_REQUIRE(tune->bucket_max_verts[b] >= 3 &&
                            (b == 0 || tune->bucket_max_verts[b] > tune->bucket_max_verts[b - 1]),
                        PRT_E_INVAL, "debug_contours: bucket bounds must increase");
        }
        g_ct_tune = *tune;
    }
    g_ct_tune_set = tune != nullptr;
    g_ct_one_lane = one_lane ? 1 : 0;
    return PRT_OK;
}
