// gen.cpp -- seeded synthetic inputs (host only, multi-threaded, the output
// does not depend on the thread count).
//
// The reference reads its inputs from text files that are not part of it
// (data/Rect1024.txt, BotomUp-rtree.cu:669; readData's polygon files,
// vector_map_overlay_gpc_kernel.cu:752-888).  The configurations of this
// project are synthetic; every value is a counter-based SplitMix64 draw:
//   u01(seed, i) = (mix64(seed + (i + 1) * 0x9E3779B97F4A7C15) >> 11) * 2^-53
// so element i of any stream is computable on its own.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr uint64_t GOLDEN = 0x9E3779B97F4A7C15ull;
constexpr double TWO_PI = 6.283185307179586476925286766559;

inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline double u01(uint64_t seed, uint64_t i) {
    return static_cast<double>(mix64(seed + (i + 1) * GOLDEN) >> 11) * 0x1.0p-53;
}

// per-object stream seed
inline uint64_t sub_seed(uint64_t seed, uint64_t id) { return mix64(seed ^ mix64(id + 0x632BE59BD9B4E019ull)); }

// standard normal from two uniforms (Box-Muller, cosine branch)
inline double normal(uint64_t seed, uint64_t i) {
    double u1 = 1.0 - u01(seed, 2 * i);  // (0, 1]
    double u2 = u01(seed, 2 * i + 1);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(TWO_PI * u2);
}

template <class F> void parallel_for(int64_t n, F fn) {
    unsigned hw = std::thread::hardware_concurrency();
    int64_t nt = hw == 0 ? 1 : (hw > 16 ? 16 : hw);
    if (n < 65536) nt = 1;
    if (nt == 1) {
        fn(int64_t(0), n);
        return;
    }
    std::vector<std::thread> th;
    for (int64_t k = 0; k < nt; ++k) th.emplace_back(fn, n * k / nt, n * (k + 1) / nt);
    for (auto &t : th) t.join();
}

// k vertices of a star around (cx, cy): sorted random angles, one uniform
// draw per sector of width 2pi/k (so consecutive angles are less than pi
// apart for k >= 4 and the polygon is simple and star-shaped about the
// centre), radius U[0.5, 1] times the half extent per axis;
// counter-clockwise.
void star(uint64_t s, int k, double cx, double cy, double hx, double hy, double *xy) {
    std::vector<double> ang(static_cast<size_t>(k));
    for (int j = 0; j < k; ++j) ang[j] = TWO_PI * (j + u01(s, 16 + j)) / k;
    for (int j = 0; j < k; ++j) {
        double r = 0.5 + 0.5 * u01(s, 16 + 1024 + j);
        xy[2 * j] = cx + r * hx * std::cos(ang[j]);
        xy[2 * j + 1] = cy + r * hy * std::sin(ang[j]);
    }
}

int lognormal_count(uint64_t s, double mu, double sigma, int kmin, int kmax) {
    double z = normal(s, 0);
    double v = std::exp(mu + sigma * z);
    long k = std::lround(v);
    if (k < kmin) k = kmin;
    if (k > kmax) k = kmax;
    return static_cast<int>(k);
}

}  // namespace

extern "C" {

double prt_gen_u01(uint64_t seed, uint64_t i) { return u01(seed, i); }

// Uniform rects: centre U[0,1]^2, width and height U[wmin, wmax]; fp32 from
// the FP64 corners.  Draws 4i..4i+3 of the stream.
void prt_gen_rects_uniform(uint64_t seed, int64_t n, double wmin, double wmax, float *r4) {
    parallel_for(n, [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; ++i) {
            double cx = u01(seed, 4 * i), cy = u01(seed, 4 * i + 1);
            double w = wmin + (wmax - wmin) * u01(seed, 4 * i + 2);
            double h = wmin + (wmax - wmin) * u01(seed, 4 * i + 3);
            r4[4 * i + 0] = static_cast<float>(cx - 0.5 * w);
            r4[4 * i + 1] = static_cast<float>(cy - 0.5 * h);
            r4[4 * i + 2] = static_cast<float>(cx + 0.5 * w);
            r4[4 * i + 3] = static_cast<float>(cy + 0.5 * h);
        }
    });
}

// Query windows of side `side`, centred U[0,1]^2.
void prt_gen_windows(uint64_t seed, int64_t n, double side, float *r4) {
    parallel_for(n, [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; ++i) {
            double cx = u01(seed, 2 * i), cy = u01(seed, 2 * i + 1);
            r4[4 * i + 0] = static_cast<float>(cx - 0.5 * side);
            r4[4 * i + 1] = static_cast<float>(cy - 0.5 * side);
            r4[4 * i + 2] = static_cast<float>(cx + 0.5 * side);
            r4[4 * i + 3] = static_cast<float>(cy + 0.5 * side);
        }
    });
}

// Gaussian-mixture rects: component means U[0,1]^2 (draws of sub_seed(seed,
// 0x6D6978)), rect i picks component floor(u*ncomp), centre = mean +
// sigma*N(0,1) per axis, width and height U[smin, smax].
void prt_gen_rects_gmm(uint64_t seed, int64_t n, int ncomp, double sigma, double smin, double smax,
                       float *r4) {
    const uint64_t ms = sub_seed(seed, 0x6D6978);
    parallel_for(n, [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; ++i) {
            uint64_t s = sub_seed(seed, static_cast<uint64_t>(i));
            int c = static_cast<int>(u01(s, 0) * ncomp);
            if (c >= ncomp) c = ncomp - 1;
            double cx = u01(ms, 2 * c) + sigma * normal(s, 1);
            double cy = u01(ms, 2 * c + 1) + sigma * normal(s, 2);
            double w = smin + (smax - smin) * u01(s, 10);
            double h = smin + (smax - smin) * u01(s, 11);
            r4[4 * i + 0] = static_cast<float>(cx - 0.5 * w);
            r4[4 * i + 1] = static_cast<float>(cy - 0.5 * h);
            r4[4 * i + 2] = static_cast<float>(cx + 0.5 * w);
            r4[4 * i + 3] = static_cast<float>(cy + 0.5 * h);
        }
    });
}

// ---- star polygons ------------------------------------------------------
// Two-phase CSR generation: call with xy == NULL to get the offsets
// (off[n] = total vertices), allocate, call again with xy.

// Stars inscribed in given rects: k ~ U{kmin..kmax}, half extents of the
// rect, centre of the rect.  ids[i] names the stream of polygon i.
void prt_gen_stars_in_rects(uint64_t seed, const float *r4, const uint32_t *ids, int64_t n,
                            int kmin, int kmax, int64_t *off, double *xy) {
    if (xy == nullptr) {
        off[0] = 0;
        for (int64_t i = 0; i < n; ++i) {
            uint64_t s = sub_seed(seed, ids[i]);
            int k = kmin + static_cast<int>(u01(s, 0) * (kmax - kmin + 1));
            if (k > kmax) k = kmax;
            off[i + 1] = off[i] + k;
        }
        return;
    }
    parallel_for(n, [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; ++i) {
            uint64_t s = sub_seed(seed, ids[i]);
            int k = static_cast<int>(off[i + 1] - off[i]);
            const float *r = r4 + 4 * i;
            double cx = 0.5 * (static_cast<double>(r[0]) + r[2]);
            double cy = 0.5 * (static_cast<double>(r[1]) + r[3]);
            double hx = 0.5 * (static_cast<double>(r[2]) - r[0]);
            double hy = 0.5 * (static_cast<double>(r[3]) - r[1]);
            star(s, k, cx, cy, hx, hy, xy + 2 * off[i]);
        }
    });
}

// Jittered-grid stars: polygon i sits in cell (i % grid, i / grid), centre
// ((gx + 0.5 + jx) / grid, (gy + 0.5 + jy) / grid), j ~ U[-jitter, jitter];
// radius rfac / grid; k = clamp(round(exp(mu + sigma z)), kmin, kmax).
void prt_gen_stars_grid(uint64_t seed, int64_t n, int grid, double jitter, double rfac, double mu,
                        double sigma, int kmin, int kmax, int64_t *off, double *xy) {
    if (xy == nullptr) {
        off[0] = 0;
        for (int64_t i = 0; i < n; ++i)
            off[i + 1] = off[i] + lognormal_count(sub_seed(seed, i), mu, sigma, kmin, kmax);
        return;
    }
    const double cell = 1.0 / grid;
    parallel_for(n, [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; ++i) {
            uint64_t s = sub_seed(seed, i);
            int k = static_cast<int>(off[i + 1] - off[i]);
            double gx = static_cast<double>(i % grid), gy = static_cast<double>(i / grid);
            double jx = jitter * (2.0 * u01(s, 4) - 1.0), jy = jitter * (2.0 * u01(s, 5) - 1.0);
            double R = rfac * cell;
            star(s, k, (gx + 0.5 + jx) * cell, (gy + 0.5 + jy) * cell, R, R, xy + 2 * off[i]);
        }
    });
}

// Gaussian-mixture stars (same mixture law as prt_gen_rects_gmm), radius
// U[rmin, rmax], lognormal vertex counts.
void prt_gen_stars_gmm(uint64_t seed, int64_t n, int ncomp, double sigma, double rmin, double rmax,
                       double mu, double lsigma, int kmin, int kmax, int64_t *off, double *xy) {
    if (xy == nullptr) {
        off[0] = 0;
        for (int64_t i = 0; i < n; ++i)
            off[i + 1] = off[i] + lognormal_count(sub_seed(seed, i), mu, lsigma, kmin, kmax);
        return;
    }
    const uint64_t ms = sub_seed(seed, 0x6D6978);
    parallel_for(n, [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; ++i) {
            uint64_t s = sub_seed(seed, i);
            int k = static_cast<int>(off[i + 1] - off[i]);
            int c = static_cast<int>(u01(s, 6) * ncomp);
            if (c >= ncomp) c = ncomp - 1;
            double cx = u01(ms, 2 * c) + sigma * normal(s, 4);
            double cy = u01(ms, 2 * c + 1) + sigma * normal(s, 5);
            double R = rmin + (rmax - rmin) * u01(s, 7);
            star(s, k, cx, cy, R, R, xy + 2 * off[i]);
        }
    });
}

// Independent pairs for the differentiable overlay: P centred at the
// origin, Q centred at (dx, dy) ~ U[-off_frac, off_frac]^2 (unit radius),
// k ~ U{kmin..kmax} each.  Polygon 2i is P of pair i, 2i+1 is Q.
void prt_gen_star_pairs(uint64_t seed, int64_t npairs, int kmin, int kmax, double off_frac,
                        int64_t *off, double *xy) {
    const int64_t n = 2 * npairs;
    if (xy == nullptr) {
        off[0] = 0;
        for (int64_t i = 0; i < n; ++i) {
            uint64_t s = sub_seed(seed, i);
            int k = kmin + static_cast<int>(u01(s, 0) * (kmax - kmin + 1));
            if (k > kmax) k = kmax;
            off[i + 1] = off[i] + k;
        }
        return;
    }
    parallel_for(n, [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; ++i) {
            uint64_t s = sub_seed(seed, i);
            int k = static_cast<int>(off[i + 1] - off[i]);
            double cx = 0.0, cy = 0.0;
            if (i & 1) {
                uint64_t ps = sub_seed(seed, i - 1);
                cx = off_frac * (2.0 * u01(ps, 8) - 1.0);
                cy = off_frac * (2.0 * u01(ps, 9) - 1.0);
            }
            star(s, k, cx, cy, 1.0, 1.0, xy + 2 * off[i]);
        }
    });
}

// fp32 MBR of each CSR polygon, rounded outwards so that it contains every
// FP64 vertex.
void prt_poly_mbr(const double *xy, const int64_t *off, int64_t n, float *r4) {
    parallel_for(n, [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; ++i) {
            double x0 = 1e300, y0 = 1e300, x1 = -1e300, y1 = -1e300;
            for (int64_t v = off[i]; v < off[i + 1]; ++v) {
                double x = xy[2 * v], y = xy[2 * v + 1];
                x0 = x < x0 ? x : x0;
                x1 = x > x1 ? x : x1;
                y0 = y < y0 ? y : y0;
                y1 = y > y1 ? y : y1;
            }
            float fx0 = static_cast<float>(x0), fy0 = static_cast<float>(y0);
            float fx1 = static_cast<float>(x1), fy1 = static_cast<float>(y1);
            if (static_cast<double>(fx0) > x0) fx0 = std::nextafter(fx0, -INFINITY);
            if (static_cast<double>(fy0) > y0) fy0 = std::nextafter(fy0, -INFINITY);
            if (static_cast<double>(fx1) < x1) fx1 = std::nextafter(fx1, INFINITY);
            if (static_cast<double>(fy1) < y1) fy1 = std::nextafter(fy1, INFINITY);
            r4[4 * i + 0] = fx0;
            r4[4 * i + 1] = fy0;
            r4[4 * i + 2] = fx1;
            r4[4 * i + 3] = fy1;
        }
    });
}

}  // extern "C"
