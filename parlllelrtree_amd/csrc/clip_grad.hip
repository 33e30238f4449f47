// clip_grad.hip -- adjoint of the intersection area: dArea/d(vertex) for
// every vertex of both polygons of every pair, with the clipping topology
// held fixed, FP64 or FP32.  No reference counterpart (the reference
// computes no area, vector_map_overlay_gpc_kernel.cu:45-58).
//
// The derivative of area(A cap B) moves only the parts of each boundary that
// lie inside the other polygon:  dArea = integral over (dA inside B) of V.n.
// For edge i of A (v_i -> v_{i+1}, direction d) with inside set I of t in
// [0, 1] and outward normal N = s * (d.y, -d.x) (s = +1 for a
// counter-clockwise polygon, -1 otherwise):
//   dArea/dv_i     += N * integral_I (1 - t) dt
//   dArea/dv_{i+1} += N * integral_I t dt
// Nothing is stored between forward and backward: the crossings of every
// edge with the other polygon are recomputed here.  One wave per pair; both
// polygons are staged in LDS in coordinates relative to the centre of A's
// bounding box; lane e finds edge e's crossings with the other polygon in one
// pass (up to 4 distinct ones kept sorted in registers; more fall back to a
// pass per crossing), walks them in increasing order and decides each part
// by a midpoint point-in-polygon test, exactly as the CPU oracle does, so
// the FP64 gradient matches it.
// Vertex gradients are accumulated with atomics (two contributions per
// vertex and pair, which commute exactly when each polygon is in one pair).
#include "prt_common.hpp"
#include "prt_profile.hpp"

namespace prt {
namespace {

// One wave per block, one pair per wave: a block keeps its LDS until its
// last wave ends, so blocks of one wave free it as each pair ends, and the
// dispatcher hands out pairs of very uneven cost one at a time (262,144 star
// pairs, FP64: blocks of 4 waves capped at 65,536 blocks 14.7 ms, 2 waves
// 12.7 ms, one wave per block and pair 12.4 ms).
constexpr int GR_WAVES = 1;
constexpr int GR_THREADS = 64 * GR_WAVES;

template <class T> __device__ inline T wave_min(T v) {
    for (int o = 1; o < 64; o <<= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}
template <class T> __device__ inline T wave_max(T v) {
    for (int o = 1; o < 64; o <<= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
template <class T> __device__ inline T wave_sum(T v) {
    for (int o = 1; o < 64; o <<= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// even-odd point in polygon, B in local coordinates (crossing number); B's
// vertices are (x, y) pairs (loading the next vertex a step ahead took 8 more
// VGPRs, a wave less per SIMD, and was slower)
template <class T, class T2>
__device__ inline int point_in(const T2 *b, int nB, T px, T py) {
    int in = 0;
    T2 vi = b[0];
    for (int i = 0; i < nB; ++i) {
        const T2 vj = b[i + 1 == nB ? 0 : i + 1];
        if ((vi.y > py) != (vj.y > py)) {
            const T xc = (vj.x - vi.x) * (py - vi.y) / (vj.y - vi.y) + vi.x;
            if (px < xc) in ^= 1;
        }
        vi = vj;
    }
    return in;
}

// The crossing of edge a + t r with edge c -> c + s of B, in two steps.
// Candidate: the numerators and the denominator of the oracle's quotients
// and a division-free test, without branches, that is exact in the one
// direction needed -- it rejects only pairs whose rounded quotients the
// oracle's test (t in (0, 1), u in [-1e-9, 1 + 1e-9]) would reject too: the
// sign and |t| < 1 tests on the numerators, and |u| bounds with a margin of
// 2e-9 (FP64) or 4e-7 (FP32, > 1.5 ulp of 1) of |den|, left out when |den|
// is so small that the margin would underflow.
template <class T> struct Cand {
    T den, tn, un;
    bool ok;
};
template <class T>
__device__ inline Cand<T> crossing_cand(T ax, T ay, T rx, T ry, T cx, T cy, T sx, T sy) {
    Cand<T> c;
    c.den = rx * sy - ry * sx;
    const T qx = cx - ax, qy = cy - ay;
    c.tn = qx * sy - qy * sx;
    c.un = qx * ry - qy * rx;
    const bool pos = c.den > T(0);
    const T a = pos ? c.den : -c.den, tp = pos ? c.tn : -c.tn, up = pos ? c.un : -c.un;
    constexpr bool f64 = sizeof(T) == 8;
    const T m = (f64 ? T(2e-9) : T(4e-7)) * a;
    const bool small = !(a > (f64 ? T(1e-290) : T(1e-30)));
    c.ok = (c.den != T(0)) & (tp > T(0)) & (tp < a) & (small | ((up >= -m) & (up <= a + m)));
    return c;
}
// Exact step (a candidate only): the oracle's quotients and test; t or -1.
template <class T> __device__ inline T crossing_exact(const Cand<T> &c) {
    const T t = c.tn / c.den, u = c.un / c.den;
    const T lo_u = T(-1e-9), hi_u = T(1) + T(1e-9);
    return t > T(0) && t < T(1) && u >= lo_u && u <= hi_u ? t : T(-1);
}

// Parts of edge (ax, ay) -> (ex, ey) inside polygon B: J0 = int (1-t),
// J1 = int t over the inside parameter set, green = int x dy over it.  The
// parts lie between consecutive distinct crossing parameters; each one is
// decided by a point test at its midpoint, in increasing t.
template <class T, class T2>
__device__ inline void edge_parts(T ax, T ay, T ex, T ey, const T2 *b, int nB, T &j0, T &j1,
                                  T &green) {
    const T rx = ex - ax, ry = ey - ay;
    j0 = T(0);
    j1 = T(0);
    green = T(0);
    auto part = [&](T t0, T t1) {
        const T tm = T(0.5) * (t0 + t1);
        if (point_in<T, T2>(b, nB, ax + tm * rx, ay + tm * ry)) {
            const T sq = T(0.5) * (t1 * t1 - t0 * t0);
            j0 += (t1 - t0) - sq;
            j1 += sq;
            const T x0 = ax + t0 * rx, y0 = ay + t0 * ry;
            const T x1 = ax + t1 * rx, y1 = ay + t1 * ry;
            green += T(0.5) * (x0 + x1) * (y1 - y0);
        }
    };
    // one pass over B: the distinct crossings, sorted, in registers (2 =
    // empty slot, above every parameter)
    constexpr int K = 4;
    T cut[K];
#pragma unroll
    for (int k = 0; k < K; ++k) cut[k] = T(2);
    bool more = false;  // a fifth distinct crossing
    T2 vi = b[0];
    for (int i = 0; i < nB; ++i) {
        const T2 vj = b[i + 1 == nB ? 0 : i + 1];
        const Cand<T> c = crossing_cand<T>(ax, ay, rx, ry, vi.x, vi.y, vj.x - vi.x, vj.y - vi.y);
        vi = vj;
        if (!c.ok) continue;
        const T t = crossing_exact<T>(c);
        if (t < T(0)) continue;
        bool dup = false;
#pragma unroll
        for (int k = 0; k < K; ++k) dup |= t == cut[k];
        if (dup) continue;
        more |= cut[K - 1] < T(2);
#pragma unroll
        for (int k = K - 1; k >= 0; --k) {
            const T below = k ? cut[k - 1] : T(-1);
            cut[k] = cut[k] < t ? cut[k] : (below < t ? t : below);
        }
    }
    if (!more) {
        T t0 = T(0);
#pragma unroll
        for (int k = 0; k <= K; ++k) {
            const T t1 = k < K && cut[k] < T(1) ? cut[k] : T(1);
            part(t0, t1);
            if (t1 == T(1)) break;
            t0 = t1;
        }
        return;
    }
    // more than K crossings on this edge: the next distinct crossing by a
    // pass over B per part
    T t0 = T(0);
    for (;;) {
        T t1 = T(1);
        T2 wi = b[0];
        for (int i = 0; i < nB; ++i) {
            const T2 wj = b[i + 1 == nB ? 0 : i + 1];
            const Cand<T> c = crossing_cand<T>(ax, ay, rx, ry, wi.x, wi.y, wj.x - wi.x, wj.y - wi.y);
            wi = wj;
            if (!c.ok) continue;
            const T t = crossing_exact<T>(c);
            if (t > t0 && t < t1) t1 = t;
        }
        part(t0, t1);
        if (t1 == T(1)) break;
        t0 = t1;
    }
}

template <class T> struct Vec2;
template <> struct Vec2<double> { using type = double2; };
template <> struct Vec2<float> { using type = float2; };

template <class T>
__global__ __launch_bounds__(GR_THREADS) void k_area_grad(
    const uint32_t *__restrict__ pairs, int64_t np, const T *__restrict__ xyA,
    const int64_t *__restrict__ offA, const T *__restrict__ xyB, const int64_t *__restrict__ offB,
    const T *__restrict__ w, T *__restrict__ gA, T *__restrict__ gB, T *__restrict__ green_area,
    int max_verts) {
    extern __shared__ __align__(16) char lds_raw[];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int W = blockDim.x >> 6;  // waves (= pairs in flight) per block
    using T2 = typename Vec2<T>::type;
    T2 *pv = reinterpret_cast<T2 *>(lds_raw) + static_cast<size_t>(wave) * 2 * max_verts;
    T2 *qv = pv + max_verts;
    for (int64_t p = static_cast<int64_t>(blockIdx.x) * W + wave; p < np;
         p += static_cast<int64_t>(gridDim.x) * W) {
        const uint32_t a = pairs[2 * p], b = pairs[2 * p + 1];
        const int64_t oa = offA[a], ob = offB[b];
        const int nP = static_cast<int>(offA[a + 1] - oa);
        const int nQ = static_cast<int>(offB[b + 1] - ob);
        const T *P = xyA + 2 * oa, *Q = xyB + 2 * ob;
        if (nP < 3 || nQ < 3 || nP > max_verts || nQ > max_verts) {
            if (lane == 0 && green_area) green_area[p] = nP < 3 || nQ < 3 ? T(0) : T(__builtin_nan(""));
            continue;
        }
        // centre of A's bounding box (the oracle's local origin)
        T x0 = __builtin_huge_valf(), x1 = -x0, y0 = x0, y1 = -x0;
        for (int i = lane; i < nP; i += 64) {
            x0 = fmin(x0, P[2 * i]); x1 = fmax(x1, P[2 * i]);
            y0 = fmin(y0, P[2 * i + 1]); y1 = fmax(y1, P[2 * i + 1]);
        }
        x0 = wave_min(x0); x1 = wave_max(x1); y0 = wave_min(y0); y1 = wave_max(y1);
        const T ox = T(0.5) * (x0 + x1), oy = T(0.5) * (y0 + y1);
        __builtin_amdgcn_wave_barrier();
        for (int i = lane; i < nP; i += 64) pv[i] = T2{P[2 * i] - ox, P[2 * i + 1] - oy};
        for (int i = lane; i < nQ; i += 64) qv[i] = T2{Q[2 * i] - ox, Q[2 * i + 1] - oy};
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // orientation of each polygon: sign of its shoelace area (relative to
        // its first vertex, as the oracle's orc_poly_area)
        T sp = T(0), sq = T(0);
        for (int i = lane; i < nP; i += 64) {
            const int j = i + 1 == nP ? 0 : i + 1;
            const T xi = P[2 * i] - P[0], yi = P[2 * i + 1] - P[1];
            const T xj = P[2 * j] - P[0], yj = P[2 * j + 1] - P[1];
            sp += xi * yj - xj * yi;
        }
        for (int i = lane; i < nQ; i += 64) {
            const int j = i + 1 == nQ ? 0 : i + 1;
            const T xi = Q[2 * i] - Q[0], yi = Q[2 * i + 1] - Q[1];
            const T xj = Q[2 * j] - Q[0], yj = Q[2 * j + 1] - Q[1];
            sq += xi * yj - xj * yi;
        }
        const T sgnP = wave_sum(sp) < T(0) ? T(-1) : T(1);
        const T sgnQ = wave_sum(sq) < T(0) ? T(-1) : T(1);
        const T wp = w ? w[p] : T(1);
        T green = T(0);
        for (int e = lane; e < nP + nQ; e += 64) {
            const bool inP = e < nP;
            const int n = inP ? nP : nQ;
            const int i = inP ? e : e - nP;
            const int j = i + 1 == n ? 0 : i + 1;
            const T2 *v = inP ? pv : qv, *o = inP ? qv : pv;
            const int nO = inP ? nQ : nP;
            const T2 vi = v[i], vj = v[j];
            const T ax = vi.x, ay = vi.y, bx = vj.x, by = vj.y;
            T j0, j1, gr;
            edge_parts<T, T2>(ax, ay, bx, by, o, nO, j0, j1, gr);
            const T s = inP ? sgnP : sgnQ;
            const T nx = s * (by - ay), ny = -s * (bx - ax);
            T *g = inP ? gA + 2 * oa : gB + 2 * ob;
            if (j0 != T(0)) {
                atomicAdd(&g[2 * i], wp * (j0 * nx));
                atomicAdd(&g[2 * i + 1], wp * (j0 * ny));
            }
            if (j1 != T(0)) {
                atomicAdd(&g[2 * j], wp * (j1 * nx));
                atomicAdd(&g[2 * j + 1], wp * (j1 * ny));
            }
            green += s * gr;
        }
        green = wave_sum(green);
        if (lane == 0 && green_area) green_area[p] = green;
        __builtin_amdgcn_wave_barrier();
    }
}

template <class T>
int launch_grad(const uint32_t *pairs, int64_t np, const T *xyA, const int64_t *offA, const T *xyB,
                const int64_t *offB, int32_t max_verts, const T *w, T *gA, T *gB, T *area,
                void *stream) {
    PRT_REQUIRE(np >= 0, PRT_E_INVAL, "area_grad: npairs < 0");
    if (np == 0) return PRT_OK;
    PRT_REQUIRE(pairs && xyA && offA && xyB && offB && gA && gB, PRT_E_INVAL,
                "area_grad: null device pointer");
    PRT_REQUIRE(max_verts >= 3 && max_verts <= PRT_MAX_POLY_VERTS, PRT_E_INVAL,
                "area_grad: max_verts (largest polygon) %d not in [3, %d]", max_verts,
                PRT_MAX_POLY_VERTS);
    const size_t per_wave = 4 * static_cast<size_t>(max_verts) * sizeof(T);  // P and Q as (x, y)
    int W = static_cast<int>((96 * 1024) / per_wave);
    W = W < 1 ? 1 : (W > GR_WAVES ? GR_WAVES : W);
    const size_t lds = W * per_wave;
    int64_t blocks = ceil_div(np, W);
    if (blocks > (int64_t(1) << 30)) blocks = int64_t(1) << 30;  // waves loop over the rest
    PRT_LAUNCH(k_area_grad<T>, dim3(static_cast<unsigned>(blocks)), W * 64, lds,
               as_stream(stream), pairs, np, xyA, offA, xyB, offB, w, gA, gB, area, max_verts);
    PRT_LAUNCHED();
    return PRT_OK;
}

}  // namespace
}  // namespace prt

using namespace prt;

extern "C" int prt_overlay_area_grad(const uint32_t *pairs, int64_t npairs, const double *xyA,
                                     const int64_t *offA, const double *xyB, const int64_t *offB,
                                     int32_t max_poly_verts, const double *weights, double *gA,
                                     double *gB, double *green_area, void *stream) {
    return launch_grad<double>(pairs, npairs, xyA, offA, xyB, offB, max_poly_verts, weights, gA, gB,
                               green_area, stream);
}

extern "C" int prt_overlay_area_grad_f32(const uint32_t *pairs, int64_t npairs, const float *xyA,
                                         const int64_t *offA, const float *xyB, const int64_t *offB,
                                         int32_t max_poly_verts, const float *weights, float *gA,
                                         float *gB, float *green_area, void *stream) {
    return launch_grad<float>(pairs, npairs, xyA, offA, xyB, offB, max_poly_verts, weights, gA, gB,
                              green_area, stream);
}
