// clip_lane.hpp -- the per-lane arena and sweep set-up shared by the
// one-lane-per-pair kernels (plain FP64 area, result contours).
//
// Each lane owns `cap` slots of every array; slot k of lane t lives at
// k * T + t, so lanes working on the same slot index touch neighbouring
// addresses.  The set-up restates GPC's edge and scanbeam tables
// (vector_map_overlay_gpc_kernel.cu:1019-1085, 1162-1307): edges bottom to
// top with dx = (top.x - bot.x) / (top.y - bot.y), horizontal edges dropped,
// sorted by (bottom y, edge id); the distinct vertex heights of the op's y
// range, ascending.
#pragma once

#include "prt_common.hpp"

namespace prt {
namespace {

constexpr int LANE_DOUBLES = 11;  // bx by tx ty dx ys xb xt xa xc bsum
constexpr int LANE_INTS = 4;      // id eord act ord

struct PlainWs {
    double *bx, *by, *tx, *ty, *dx, *ys, *xb, *xt, *xa, *xc, *bsum;
    int32_t *id, *eord, *act, *ord;
    int64_t T;
    int cap;
};

struct Lane {
    const PlainWs &w;
    int64_t t;
    __device__ Lane(const PlainWs &ws, int64_t lane) : w(ws), t(lane) {}
    template <class T> __device__ T &at(T *base, int64_t k) const { return base[k * w.T + t]; }
};

__device__ inline double edge_x(const Lane &L, int e, double y) {
    if (y == L.at(L.w.ty, e)) return L.at(L.w.tx, e);
    return L.at(L.w.bx, e) + L.at(L.w.dx, e) * (y - L.at(L.w.by, e));
}

// heap sort of eord[0..n) by (by, id)
__device__ inline bool edge_less(const Lane &L, int a, int b) {
    double ya = L.at(L.w.by, a), yb = L.at(L.w.by, b);
    if (ya != yb) return ya < yb;
    return L.at(L.w.id, a) < L.at(L.w.id, b);
}

__device__ void sift_edges(const Lane &L, int root, int n) {
    for (;;) {
        int c = 2 * root + 1;
        if (c >= n) return;
        if (c + 1 < n && edge_less(L, L.at(L.w.eord, c), L.at(L.w.eord, c + 1))) ++c;
        int rv = L.at(L.w.eord, root), cv = L.at(L.w.eord, c);
        if (!edge_less(L, rv, cv)) return;
        L.at(L.w.eord, root) = cv;
        L.at(L.w.eord, c) = rv;
        root = c;
    }
}

__device__ void sort_edges(const Lane &L, int n) {
    for (int i = n / 2 - 1; i >= 0; --i) sift_edges(L, i, n);
    for (int end = n - 1; end > 0; --end) {
        int v = L.at(L.w.eord, 0);
        L.at(L.w.eord, 0) = L.at(L.w.eord, end);
        L.at(L.w.eord, end) = v;
        sift_edges(L, 0, end);
    }
}

__device__ void sift_ys(const Lane &L, int root, int n) {
    for (;;) {
        int c = 2 * root + 1;
        if (c >= n) return;
        if (c + 1 < n && L.at(L.w.ys, c) < L.at(L.w.ys, c + 1)) ++c;
        double rv = L.at(L.w.ys, root), cv = L.at(L.w.ys, c);
        if (!(rv < cv)) return;
        L.at(L.w.ys, root) = cv;
        L.at(L.w.ys, c) = rv;
        root = c;
    }
}

__device__ void sort_ys(const Lane &L, int n) {
    for (int i = n / 2 - 1; i >= 0; --i) sift_ys(L, i, n);
    for (int end = n - 1; end > 0; --end) {
        double v = L.at(L.w.ys, 0);
        L.at(L.w.ys, 0) = L.at(L.w.ys, end);
        L.at(L.w.ys, end) = v;
        sift_ys(L, 0, end);
    }
}

// strict order (xa + xc, xa, id) of active slots i, j
__device__ inline bool sb_before(const Lane &L, int i, int j) {
    double ki = L.at(L.w.xa, i) + L.at(L.w.xc, i);
    double kj = L.at(L.w.xa, j) + L.at(L.w.xc, j);
    if (ki != kj) return ki < kj;
    double ai = L.at(L.w.xa, i), aj = L.at(L.w.xa, j);
    if (ai != aj) return ai < aj;
    return L.at(L.w.id, L.at(L.w.act, i)) < L.at(L.w.id, L.at(L.w.act, j));
}

// ord[0..m) <- active slots in the strict sub-beam order (insertion sort:
// the order changes little from one sub-beam to the next)
__device__ inline void sort_sub_beam(const Lane &L, int m) {
    for (int i = 0; i < m; ++i) L.at(L.w.ord, i) = i;
    for (int i = 1; i < m; ++i) {
        int v = L.at(L.w.ord, i), j = i - 1;
        while (j >= 0 && sb_before(L, v, L.at(L.w.ord, j))) {
            L.at(L.w.ord, j + 1) = L.at(L.w.ord, j);
            --j;
        }
        L.at(L.w.ord, j + 1) = v;
    }
}

// Bounding boxes bp / bq, y range of `op`, edge table (eord sorted by
// (by, id); ne edges) and distinct heights ys[0..ny).  nP / nQ become 0 for
// a polygon of fewer than 3 vertices (empty).  Returns false when the result
// is empty.  Requires nP + nQ <= cap.
__device__ inline bool lane_setup(const Lane &L, const double *__restrict__ P, int &nP,
                                  const double *__restrict__ Q, int &nQ, int op, double *bp,
                                  double *bq, int &ne, int &ny) {
    const bool eP = nP < 3, eQ = nQ < 3;
    ne = ny = 0;
    if (eP && eQ) return false;
    if (eP) nP = 0;
    if (eQ) nQ = 0;
    for (int k = 0; k < 4; ++k) bp[k] = bq[k] = 0.0;
    if (!eP) {
        bp[0] = bp[2] = P[0];
        bp[1] = bp[3] = P[1];
    }
    if (!eQ) {
        bq[0] = bq[2] = Q[0];
        bq[1] = bq[3] = Q[1];
    }
    for (int i = 1; i < nP; ++i) {
        double x = P[2 * i], y = P[2 * i + 1];
        bp[0] = x < bp[0] ? x : bp[0];
        bp[2] = x > bp[2] ? x : bp[2];
        bp[1] = y < bp[1] ? y : bp[1];
        bp[3] = y > bp[3] ? y : bp[3];
    }
    for (int i = 1; i < nQ; ++i) {
        double x = Q[2 * i], y = Q[2 * i + 1];
        bq[0] = x < bq[0] ? x : bq[0];
        bq[2] = x > bq[2] ? x : bq[2];
        bq[1] = y < bq[1] ? y : bq[1];
        bq[3] = y > bq[3] ? y : bq[3];
    }
    double ylo = 0.0, yhi = 0.0;
    if (!op_range(op, eP, eQ, bp, bq, ylo, yhi)) return false;
    for (int poly = 0; poly < 2; ++poly) {
        const double *V = poly ? Q : P;
        const int n = poly ? nQ : nP;
        for (int i = 0; i < n; ++i) {
            const int j = i + 1 == n ? 0 : i + 1;
            const double x0 = V[2 * i], y0 = V[2 * i + 1], x1 = V[2 * j], y1 = V[2 * j + 1];
            if (y0 >= ylo && y0 <= yhi) L.at(L.w.ys, ny++) = y0;
            if (y0 == y1) continue;
            double bx, by, tx, ty;
            if (y0 < y1) { bx = x0; by = y0; tx = x1; ty = y1; }
            else { bx = x1; by = y1; tx = x0; ty = y0; }
            if (ty <= ylo || by >= yhi) continue;
            L.at(L.w.bx, ne) = bx;
            L.at(L.w.by, ne) = by;
            L.at(L.w.tx, ne) = tx;
            L.at(L.w.ty, ne) = ty;
            L.at(L.w.dx, ne) = (tx - bx) / (ty - by);
            L.at(L.w.id, ne) = poly ? nP + i : i;
            L.at(L.w.eord, ne) = ne;
            ++ne;
        }
    }
    sort_edges(L, ne);
    sort_ys(L, ny);
    int u = 0;
    for (int i = 0; i < ny; ++i) {
        double y = L.at(L.w.ys, i);
        if (u == 0 || y != L.at(L.w.ys, u - 1)) L.at(L.w.ys, u++) = y;
    }
    ny = u;
    return ny >= 2;
}

// the base arrays of T lanes of cap slots
inline size_t lane_base_bytes(int64_t T, int cap) {
    return static_cast<size_t>(T) * static_cast<size_t>(cap) * (LANE_DOUBLES * 8 + LANE_INTS * 4);
}

inline void carve_lane_base(Carver &c, PlainWs &w, int64_t T, int cap) {
    w.T = T;
    w.cap = cap;
    const size_t cnt = static_cast<size_t>(T) * cap;
    double **dptr[] = {&w.bx, &w.by, &w.tx, &w.ty, &w.dx, &w.ys, &w.xb, &w.xt, &w.xa, &w.xc, &w.bsum};
    for (double **p : dptr) *p = c.take<double>(cnt);
    int32_t **iptr[] = {&w.id, &w.eord, &w.act, &w.ord};
    for (int32_t **p : iptr) *p = c.take<int32_t>(cnt);
}

}  // namespace
}  // namespace prt
