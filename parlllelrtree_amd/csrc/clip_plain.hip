// clip_plain.hip -- plain per-pair FP64 polygon intersection area, one lane
// per pair (the launch shape of the reference's Execute kernel,
// vector_map_overlay_gpc_kernel.cu:81-113, <<<ceil(P/128),128>>>).
//
// The reference runs GPC's scanbeam clipper (gpc_polygon_clip, :115-750)
// per thread with the device heap allocator.  This kernel runs the same
// scanbeam sweep but accumulates the area of the intersection instead of
// building contours, with no allocation: each lane owns a slice of a
// caller-provided arena (arrays interleaved by lane so that lanes touching
// the same index hit neighbouring addresses).
//
// Arithmetic (compiled with -ffp-contract=off; each expression is evaluated
// exactly as written, so results are bit-identical to the FP64 CPU oracle):
//   dx    = (top.x - bot.x) / (top.y - bot.y)                 (:1241-1242)
//   x(y)  = y == top.y ? top.x : bot.x + dx * (y - bot.y)     (:454-457)
//   den   = (xt_L - xb_L) - (xt_R - xb_R), r = (xb_R - xb_L) / den,
//   iy    = r * (yt - yb) + yb                                (:1413-1433, 480)
//   term  = ((xa_R - xa_L) + (xc_R - xc_L)) * (yc - ya), summed per beam
//   from +0 in sub-beam then left-to-right order; the beam sums s[k] are
//   added by the fixed tree s[i] += s[i + w] (w = 1, 2, 4, ...; i = 0, 2w,
//   ...), area = 0.5 * s[0].
// Sub-beam boundaries are the distinct crossing heights in (yb, yt), taken
// in increasing order; inside a sub-beam, edges are ordered by
// (xa + xc, xa, edge id); a gap is inside when both even-odd parities
// (subject = A, clip = B) are odd.
#include "clip_lane.hpp"
#include "prt_common.hpp"
#include "prt_profile.hpp"

namespace prt {
namespace {

constexpr int CP_THREADS = 128;
// one lane per pair up to 262,144 lanes (four waves per SIMD): with 65,536 (one
// wave per SIMD) the latency of the lanes' arena accesses went uncovered and
// 8.4M pairs took 1.6x as long.  The arena is lanes x max_verts x 104 bytes.
constexpr int64_t CP_MAX_LANES = 262144;

__device__ double sub_beam(const Lane &L, int m, int nP, double ya, double yc, double acc,
                           int op) {
    const double h = yc - ya;
    sort_sub_beam(L, m);
    int pp = 0, pq = 0;
    for (int t = 0; t + 1 < m; ++t) {
        int l = L.at(L.w.ord, t), r = L.at(L.w.ord, t + 1);
        if (L.at(L.w.id, L.at(L.w.act, l)) < nP) pp ^= 1;
        else pq ^= 1;
        if (op_inside(op, pp, pq))
            acc += ((L.at(L.w.xa, r) - L.at(L.w.xa, l)) + (L.at(L.w.xc, r) - L.at(L.w.xc, l))) * h;
    }
    return acc;
}

// Area of P op Q (op: PRT_OP_DIFF / INT / XOR / UNION); a polygon of fewer
// than 3 vertices is empty.
__device__ double clip_one(const Lane &L, const double *__restrict__ P, int nP,
                           const double *__restrict__ Q, int nQ, int op) {
    if (nP < 3 && nQ < 3) return 0.0;
    if (nP + nQ > L.w.cap) return __builtin_nan("");
    double bp[4], bq[4];
    int ne = 0, ny = 0;
    if (!lane_setup(L, P, nP, Q, nQ, op, bp, bq, ne, ny)) return 0.0;
    int m = 0, next = 0;
    for (int k = 0; k + 1 < ny; ++k) {
        const double yb = L.at(L.w.ys, k), yt = L.at(L.w.ys, k + 1);
        L.at(L.w.bsum, k) = 0.0;
        while (next < ne && L.at(L.w.by, L.at(L.w.eord, next)) <= yb)
            L.at(L.w.act, m++) = L.at(L.w.eord, next++);
        int keep = 0;
        for (int i = 0; i < m; ++i) {
            int e = L.at(L.w.act, i);
            if (L.at(L.w.ty, e) > yb) L.at(L.w.act, keep++) = e;
        }
        m = keep;
        if (m < 2) continue;
        for (int i = 0; i < m; ++i) {
            int e = L.at(L.w.act, i);
            double b = edge_x(L, e, yb), t = edge_x(L, e, yt);
            L.at(L.w.xb, i) = b;
            L.at(L.w.xt, i) = t;
            L.at(L.w.xa, i) = b;
        }
        const double dy = yt - yb;
        double ya = yb, bs = 0.0;
        for (;;) {
            // next distinct crossing height above ya (or yt)
            double yc = yt;
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
            bs = sub_beam(L, m, nP, ya, yc, bs, op);
            if (yc == yt) break;
            for (int i = 0; i < m; ++i) L.at(L.w.xa, i) = L.at(L.w.xc, i);
            ya = yc;
        }
        L.at(L.w.bsum, k) = bs;
    }
    // fixed pairwise tree over the beam sums (same shape as the oracle)
    const int nb = ny - 1;
    for (int wdt = 1; wdt < nb; wdt *= 2)
        for (int i = 0; i + wdt < nb; i += 2 * wdt) L.at(L.w.bsum, i) += L.at(L.w.bsum, i + wdt);
    return 0.5 * L.at(L.w.bsum, 0);
}

__global__ __launch_bounds__(CP_THREADS) void k_clip_plain(
    const uint32_t *__restrict__ pairs, int64_t np, const double *__restrict__ xyA,
    const int64_t *__restrict__ offA, const double *__restrict__ xyB,
    const int64_t *__restrict__ offB, double *__restrict__ area, PlainWs w, int op) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= w.T) return;
    Lane L(w, t);
    for (int64_t p = t; p < np; p += w.T) {
        const uint32_t a = pairs[2 * p], b = pairs[2 * p + 1];
        const int64_t oa = offA[a], ob = offB[b];
        const int nA = static_cast<int>(offA[a + 1] - oa);
        const int nB = static_cast<int>(offB[b + 1] - ob);
        area[p] = clip_one(L, xyA + 2 * oa, nA, xyB + 2 * ob, nB, op);
    }
}

// Same as k_clip_plain over the pairs list[0 .. *count) (overflow pairs of
// the batched kernel); *count is read on the device.
__global__ __launch_bounds__(CP_THREADS) void k_clip_plain_list(
    const uint32_t *__restrict__ pairs, const uint32_t *__restrict__ list,
    const int64_t *__restrict__ count, const double *__restrict__ xyA,
    const int64_t *__restrict__ offA, const double *__restrict__ xyB,
    const int64_t *__restrict__ offB, double *__restrict__ area, PlainWs w, int op) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= w.T) return;
    Lane L(w, t);
    const int64_t np = *count;
    for (int64_t i = t; i < np; i += w.T) {
        const uint32_t p = list[i];
        const uint32_t a = pairs[2 * p], b = pairs[2 * p + 1];
        const int64_t oa = offA[a], ob = offB[b];
        const int nA = static_cast<int>(offA[a + 1] - oa);
        const int nB = static_cast<int>(offB[b + 1] - ob);
        area[p] = clip_one(L, xyA + 2 * oa, nA, xyB + 2 * ob, nB, op);
    }
}

int64_t plain_lanes(int64_t np) {
    int64_t T = np < CP_MAX_LANES ? np : CP_MAX_LANES;
    return T < 1 ? 1 : T;
}

size_t plain_ws_bytes(int64_t T, int32_t max_verts) {
    return lane_base_bytes(T, max_verts < 3 ? 3 : max_verts) + 16 * 256;
}

int carve_plain(PlainWs &w, int64_t T, int32_t max_verts, void *ws, size_t ws_bytes) {
    Carver c(ws, ws_bytes);
    carve_lane_base(c, w, T, max_verts < 3 ? 3 : max_verts);
    return c.ok() ? PRT_OK : PRT_E_NOMEM;
}

// lanes of the overflow (list) launch: the list is short, a few thousand
// lanes keep the arena small
constexpr int64_t CP_LIST_LANES = 4096;

}  // namespace

size_t plain_list_ws_bytes(int64_t max_count, int32_t max_verts) {
    int64_t T = max_count < CP_LIST_LANES ? max_count : CP_LIST_LANES;
    return plain_ws_bytes(T < 1 ? 1 : T, max_verts);
}

int plain_list_launch(const uint32_t *pairs, const uint32_t *list, const int64_t *count,
                      int64_t max_count, const double *xyA, const int64_t *offA, const double *xyB,
                      const int64_t *offB, int32_t max_verts, int op, double *area, void *ws,
                      size_t ws_bytes, hipStream_t s) {
    int64_t T = max_count < CP_LIST_LANES ? max_count : CP_LIST_LANES;
    T = T < 1 ? 1 : T;
    PlainWs w;
    PRT_REQUIRE(carve_plain(w, T, max_verts, ws, ws_bytes) == PRT_OK, PRT_E_NOMEM,
                "overlay: overflow workspace too small");
    const unsigned grid = static_cast<unsigned>(ceil_div(w.T, CP_THREADS));
    PRT_LAUNCH(k_clip_plain_list, grid, CP_THREADS, 0, s, pairs, list, count, xyA, offA, xyB, offB,
               area, w, op);
    PRT_LAUNCHED();
    return PRT_OK;
}

}  // namespace prt

using namespace prt;

extern "C" size_t prt_overlay_plain_workspace_size(int64_t npairs, int32_t max_verts) {
    if (npairs < 0 || max_verts < 0) return 0;
    return plain_ws_bytes(plain_lanes(npairs), max_verts);
}

extern "C" int prt_overlay_plain(const uint32_t *pairs, int64_t npairs, const double *xyA,
                                 const int64_t *offA, const double *xyB, const int64_t *offB,
                                 int32_t max_verts, double *area, void *ws, size_t ws_bytes,
                                 void *stream) {
    return prt_overlay_plain_op(pairs, npairs, xyA, offA, xyB, offB, max_verts, PRT_OP_INT, area,
                                ws, ws_bytes, stream);
}

extern "C" int prt_overlay_plain_op(const uint32_t *pairs, int64_t npairs, const double *xyA,
                                    const int64_t *offA, const double *xyB, const int64_t *offB,
                                    int32_t max_verts, int32_t op, double *area, void *ws,
                                    size_t ws_bytes, void *stream) {
    PRT_REQUIRE(op >= PRT_OP_DIFF && op <= PRT_OP_UNION, PRT_E_INVAL, "overlay_plain: op %d", op);
    PRT_REQUIRE(npairs >= 0, PRT_E_INVAL, "overlay_plain: npairs < 0");
    if (npairs == 0) return PRT_OK;
    PRT_REQUIRE(pairs && xyA && offA && xyB && offB && area, PRT_E_INVAL,
                "overlay_plain: null device pointer");
    PRT_REQUIRE(max_verts >= 0 && max_verts <= 2 * PRT_MAX_POLY_VERTS, PRT_E_INVAL,
                "overlay_plain: max_verts %d out of range", max_verts);
    const size_t need = prt_overlay_plain_workspace_size(npairs, max_verts);
    PRT_REQUIRE(ws_bytes >= need, PRT_E_NOMEM, "overlay_plain: workspace %zu < %zu bytes", ws_bytes,
                need);
    PlainWs w;
    PRT_REQUIRE(carve_plain(w, plain_lanes(npairs), max_verts, ws, ws_bytes) == PRT_OK, PRT_E_NOMEM,
                "overlay_plain: workspace too small");
    const unsigned grid = static_cast<unsigned>(ceil_div(w.T, CP_THREADS));
    PRT_LAUNCH(k_clip_plain, grid, CP_THREADS, 0, as_stream(stream), pairs, npairs, xyA, offA, xyB,
               offB, area, w, op);
    PRT_LAUNCHED();
    return PRT_OK;
}
