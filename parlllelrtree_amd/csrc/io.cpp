// io.cpp -- the reference's text input formats (host only).
//
// Rect files (BotomUp-rtree.cu:668-691, TopDown-rtree.cu:486-497): an
// integer n, then n records "x1 x2 y1 y2 key" -- four integer coordinates
// (read with %lld) and an integer key (%d), whitespace separated.
//
// Polygon files (vector_map_overlay_gpc_kernel.cu:766-818, 836-858): an
// integer n, then per polygon "num hole" (%d %d), a bounding box of four
// doubles (%lf; the reference keeps (int)(v * 100000) of each, OV:805-808,
// and the minimum of the first two as minX / minY, OV:809-814), then num
// vertices "x y" (%lf %lf).  Vertices are stored contiguously, polygon i at
// prefix[i] = num[0] + ... + num[i-1] (OV:796-801).
//
// The readers take two calls: with the output arrays NULL they report the
// sizes; then they fill arrays of at least those sizes.  Malformed input is
// an error (PRT_E_INVAL with the token position), where the reference's
// fscanf would leave values unset.
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "prt_common.hpp"

namespace prt {
namespace {

struct Text {
    std::vector<char> buf;
    size_t pos = 0;
    long long token = 0;  // index of the next token (for messages)

    bool load(const char *path) {
        FILE *f = std::fopen(path, "rb");
        if (!f) return false;
        char tmp[1 << 16];
        size_t got;
        while ((got = std::fread(tmp, 1, sizeof tmp, f)) > 0) buf.insert(buf.end(), tmp, tmp + got);
        std::fclose(f);
        buf.push_back('\0');
        return true;
    }
    void skip_ws() {
        while (pos < buf.size() && (buf[pos] == ' ' || buf[pos] == '\n' || buf[pos] == '\t' ||
                                    buf[pos] == '\r' || buf[pos] == '\v' || buf[pos] == '\f'))
            ++pos;
    }
    bool next_ll(long long &v) {
        skip_ws();
        char *end = nullptr;
        errno = 0;
        v = std::strtoll(buf.data() + pos, &end, 10);
        if (end == buf.data() + pos || errno) return false;
        pos = static_cast<size_t>(end - buf.data());
        ++token;
        return true;
    }
    bool next_d(double &v) {
        skip_ws();
        char *end = nullptr;
        v = std::strtod(buf.data() + pos, &end);
        if (end == buf.data() + pos) return false;
        pos = static_cast<size_t>(end - buf.data());
        ++token;
        return true;
    }
};

#define PRT_PARSE(ok, what)                                                        \
    PRT_REQUIRE(ok, PRT_E_INVAL, "%s: bad or missing %s at token %lld", path, what, \
                static_cast<long long>(tx.token))

}  // namespace
}  // namespace prt

using namespace prt;

extern "C" int prt_read_rects_text(const char *path, int64_t *n, int64_t *xyxy_h,
                                   int32_t *key_h) {
    PRT_REQUIRE(path && n, PRT_E_INVAL, "read_rects_text: null argument");
    Text tx;
    PRT_REQUIRE(tx.load(path), PRT_E_INVAL, "read_rects_text: cannot open %s", path);
    long long cnt = 0;
    PRT_PARSE(tx.next_ll(cnt) && cnt >= 0 && cnt < (1ll << 31), "record count");
    *n = cnt;
    if (!xyxy_h) return PRT_OK;
    for (long long i = 0; i < cnt; ++i) {
        long long x1, x2, y1, y2, key;
        PRT_PARSE(tx.next_ll(x1) && tx.next_ll(x2), "x1 x2");
        PRT_PARSE(tx.next_ll(y1) && tx.next_ll(y2), "y1 y2");
        PRT_PARSE(tx.next_ll(key) && key >= INT_MIN && key <= INT_MAX, "key");
        xyxy_h[4 * i + 0] = x1;
        xyxy_h[4 * i + 1] = y1;
        xyxy_h[4 * i + 2] = x2;
        xyxy_h[4 * i + 3] = y2;
        if (key_h) key_h[i] = static_cast<int32_t>(key);
    }
    return PRT_OK;
}

extern "C" int prt_write_rects_text(const char *path, int64_t n, const int64_t *xyxy_h,
                                    const int32_t *key_h) {
    PRT_REQUIRE(path && n >= 0 && (n == 0 || xyxy_h), PRT_E_INVAL, "write_rects_text: bad args");
    FILE *f = std::fopen(path, "w");
    PRT_REQUIRE(f != nullptr, PRT_E_INVAL, "write_rects_text: cannot open %s", path);
    std::fprintf(f, "%lld\n", static_cast<long long>(n));
    for (int64_t i = 0; i < n; ++i)
        std::fprintf(f, "%lld %lld %lld %lld %d\n", static_cast<long long>(xyxy_h[4 * i]),
                     static_cast<long long>(xyxy_h[4 * i + 2]),
                     static_cast<long long>(xyxy_h[4 * i + 1]),
                     static_cast<long long>(xyxy_h[4 * i + 3]),
                     key_h ? key_h[i] : static_cast<int32_t>(i));
    const bool ok = std::fclose(f) == 0;
    PRT_REQUIRE(ok, PRT_E_INVAL, "write_rects_text: write to %s failed", path);
    return PRT_OK;
}

// BotomUp-rtree.cu:682-721: shift every rect by the smallest x1 and y1 (in
// place) and find scale = 10^k, the smallest power of ten with
// max(x range, y range) / 10^k <= 1e4 by repeated integer division.
extern "C" int prt_rects_normalize(int64_t *xyxy_h, int64_t n, int64_t *min_x, int64_t *min_y,
                                   int64_t *scale) {
    PRT_REQUIRE(n >= 0 && (n == 0 || xyxy_h) && min_x && min_y && scale, PRT_E_INVAL,
                "rects_normalize: bad args");
    long long mnx = 100000000000000000ll, mxx = 0, mny = 100000000000000000ll, mxy = 0;
    for (int64_t i = 0; i < n; ++i) {
        const long long x1 = xyxy_h[4 * i], y1 = xyxy_h[4 * i + 1];
        const long long x2 = xyxy_h[4 * i + 2], y2 = xyxy_h[4 * i + 3];
        if (x1 < mnx) mnx = x1;
        if (x2 > mxx) mxx = x2;
        if (y1 < mny) mny = y1;
        if (y2 > mxy) mxy = y2;
    }
    for (int64_t i = 0; i < n; ++i) {
        xyxy_h[4 * i] -= mnx;
        xyxy_h[4 * i + 1] -= mny;
        xyxy_h[4 * i + 2] -= mnx;
        xyxy_h[4 * i + 3] -= mny;
    }
    long long range = (mxx - mnx) > (mxy - mny) ? (mxx - mnx) : (mxy - mny);
    long long sc = 1;
    while (range > 10000) {
        sc *= 10;
        range /= 10;
    }
    *min_x = mnx;
    *min_y = mny;
    *scale = sc;
    return PRT_OK;
}

extern "C" int prt_read_polys_text(const char *path, int64_t *n, int64_t *nverts,
                                   int64_t *off_h, double *xy_h, int32_t *hole_h,
                                   int32_t *box_h, int32_t *min_xy_h) {
    PRT_REQUIRE(path && n && nverts, PRT_E_INVAL, "read_polys_text: null argument");
    Text tx;
    PRT_REQUIRE(tx.load(path), PRT_E_INVAL, "read_polys_text: cannot open %s", path);
    long long cnt = 0;
    PRT_PARSE(tx.next_ll(cnt) && cnt >= 0 && cnt < (1ll << 31), "polygon count");
    const bool fill = off_h != nullptr;
    PRT_REQUIRE(!fill || xy_h, PRT_E_INVAL, "read_polys_text: off without xy");
    long long total = 0;
    int32_t mnx = INT_MAX, mny = INT_MAX;
    if (fill) off_h[0] = 0;
    for (long long i = 0; i < cnt; ++i) {
        long long k, hole;
        PRT_PARSE(tx.next_ll(k) && k >= 0 && k <= INT_MAX, "vertex count");
        PRT_PARSE(tx.next_ll(hole), "hole flag");
        double b[4];
        for (int j = 0; j < 4; ++j) PRT_PARSE(tx.next_d(b[j]), "bounding box");
        int32_t ib[4];
        for (int j = 0; j < 4; ++j) {
            const double v = b[j] * 100000;
            PRT_REQUIRE(std::isfinite(v) && v > -2147483649.0 && v < 2147483648.0, PRT_E_INVAL,
                        "%s: box value out of int range at polygon %lld", path, i);
            ib[j] = static_cast<int32_t>(v);  // truncation, as (int)(v*100000)
        }
        if (ib[0] < mnx) mnx = ib[0];
        if (ib[1] < mny) mny = ib[1];
        if (fill) {
            if (hole_h) hole_h[i] = static_cast<int32_t>(hole);
            if (box_h)
                for (int j = 0; j < 4; ++j) box_h[4 * i + j] = ib[j];
        }
        for (long long v = 0; v < k; ++v) {
            double x, y;
            PRT_PARSE(tx.next_d(x) && tx.next_d(y), "vertex");
            if (fill) {
                xy_h[2 * (total + v)] = x;
                xy_h[2 * (total + v) + 1] = y;
            }
        }
        total += k;
        if (fill) off_h[i + 1] = total;
    }
    *n = cnt;
    *nverts = total;
    if (fill && min_xy_h) {
        min_xy_h[0] = mnx;
        min_xy_h[1] = mny;
    }
    return PRT_OK;
}

extern "C" int prt_write_polys_text(const char *path, int64_t n, const int64_t *off_h,
                                    const double *xy_h, const int32_t *hole_h) {
    PRT_REQUIRE(path && n >= 0 && (n == 0 || (off_h && xy_h)), PRT_E_INVAL,
                "write_polys_text: bad args");
    FILE *f = std::fopen(path, "w");
    PRT_REQUIRE(f != nullptr, PRT_E_INVAL, "write_polys_text: cannot open %s", path);
    std::fprintf(f, "%lld\n", static_cast<long long>(n));
    for (int64_t i = 0; i < n; ++i) {
        const int64_t a = off_h[i], e = off_h[i + 1];
        double x0 = HUGE_VAL, y0 = HUGE_VAL, x1 = -HUGE_VAL, y1 = -HUGE_VAL;
        for (int64_t v = a; v < e; ++v) {
            const double x = xy_h[2 * v], y = xy_h[2 * v + 1];
            x0 = x < x0 ? x : x0;
            x1 = x > x1 ? x : x1;
            y0 = y < y0 ? y : y0;
            y1 = y > y1 ? y : y1;
        }
        if (e == a) x0 = y0 = x1 = y1 = 0.0;
        std::fprintf(f, "%lld %d\n%.17g %.17g %.17g %.17g\n", static_cast<long long>(e - a),
                     hole_h ? hole_h[i] : 0, x0, y0, x1, y1);
        for (int64_t v = a; v < e; ++v) std::fprintf(f, "%.17g %.17g\n", xy_h[2 * v], xy_h[2 * v + 1]);
    }
    const bool ok = std::fclose(f) == 0;
    PRT_REQUIRE(ok, PRT_E_INVAL, "write_polys_text: write to %s failed", path);
    return PRT_OK;
}

// ---- readData (vector_map_overlay_gpc_kernel.cu:752-888), same arguments --
namespace {
struct LayerText {
    int64_t n = 0, nv = 0;
    std::vector<int64_t> off;
    std::vector<double> xy;
    std::vector<int32_t> hole, box;
    int32_t mn[2] = {INT_MAX, INT_MAX};
};

int read_layer(const char *path, LayerText &L) {
    int rc = prt_read_polys_text(path, &L.n, &L.nv, nullptr, nullptr, nullptr, nullptr, nullptr);
    if (rc) return rc;
    L.off.resize(L.n + 1);
    L.xy.resize(2 * L.nv);
    L.hole.resize(L.n);
    L.box.resize(4 * L.n);
    return prt_read_polys_text(path, &L.n, &L.nv, L.off.data(), L.xy.data(), L.hole.data(),
                               L.box.data(), L.mn);
}

template <class T> T *host_array(int64_t n) {
    return static_cast<T *>(std::malloc(static_cast<size_t>(n > 0 ? n : 1) * sizeof(T)));
}

// the host arrays of one layer in the reference's layout; the vertices to a
// new device array
int layer_out(const LayerText &L, int32_t **r1, int32_t **r2, int32_t **r3, int32_t **r4,
              int32_t **num, int64_t **prefix, double **vertex, int32_t **hole) {
    int32_t *a[4] = {host_array<int32_t>(L.n), host_array<int32_t>(L.n), host_array<int32_t>(L.n),
                     host_array<int32_t>(L.n)};
    int32_t *nm = host_array<int32_t>(L.n), *ho = host_array<int32_t>(L.n);
    int64_t *pf = host_array<int64_t>(L.n);
    double *dv = nullptr;
    bool ok = a[0] && a[1] && a[2] && a[3] && nm && ho && pf;
    if (ok && L.nv > 0) {
        ok = hipMalloc(&dv, static_cast<size_t>(L.nv) * 2 * sizeof(double)) == hipSuccess &&
             hipMemcpy(dv, L.xy.data(), static_cast<size_t>(L.nv) * 2 * sizeof(double),
                       hipMemcpyHostToDevice) == hipSuccess;
    }
    if (!ok) {
        for (int32_t *p : a) std::free(p);
        std::free(nm);
        std::free(ho);
        std::free(pf);
        if (dv) (void)hipFree(dv);
        return PRT_E_NOMEM;
    }
    for (int64_t i = 0; i < L.n; ++i) {
        for (int j = 0; j < 4; ++j) a[j][i] = L.box[4 * i + j];
        nm[i] = static_cast<int32_t>(L.off[i + 1] - L.off[i]);
        pf[i] = L.off[i];
        ho[i] = L.hole[i];
    }
    *r1 = a[0];
    *r2 = a[1];
    *r3 = a[2];
    *r4 = a[3];
    *num = nm;
    *prefix = pf;
    *vertex = dv;
    *hole = ho;
    return PRT_OK;
}

void free_layer(int32_t **r1, int32_t **r2, int32_t **r3, int32_t **r4, int32_t **num,
                int64_t **prefix, double **vertex, int32_t **hole) {
    for (int32_t **p : {r1, r2, r3, r4, num, hole}) {
        std::free(*p);
        *p = nullptr;
    }
    std::free(*prefix);
    *prefix = nullptr;
    if (*vertex) (void)hipFree(*vertex);
    *vertex = nullptr;
}
}  // namespace

extern "C" int prt_read_data(const char *base_filename, const char *overlay_filename,
                             int32_t **rect1, int32_t **rect2, int32_t **rect3, int32_t **rect4,
                             int32_t *num_contours, int32_t **rect1_query,
                             int32_t **rect2_query, int32_t **rect3_query,
                             int32_t **rect4_query, int32_t *num_querys, int32_t **num_base,
                             int64_t **prefix_base, double **vertex_base, int32_t **num_overlay,
                             int64_t **prefix_overlay, double **vertex_overlay, int32_t *min_x,
                             int32_t *min_y, int32_t **hole_base, int32_t **hole_overlay) {
    PRT_REQUIRE(base_filename && overlay_filename && rect1 && rect2 && rect3 && rect4 &&
                    num_contours && rect1_query && rect2_query && rect3_query && rect4_query &&
                    num_querys && num_base && prefix_base && vertex_base && num_overlay &&
                    prefix_overlay && vertex_overlay && min_x && min_y && hole_base && hole_overlay,
                PRT_E_INVAL, "read_data: null argument");
    LayerText B, Q;
    int rc = read_layer(base_filename, B);
    if (rc) return rc;
    rc = read_layer(overlay_filename, Q);
    if (rc) return rc;
    rc = layer_out(B, rect1, rect2, rect3, rect4, num_base, prefix_base, vertex_base, hole_base);
    PRT_REQUIRE(rc == PRT_OK, rc, "read_data: out of memory");
    rc = layer_out(Q, rect1_query, rect2_query, rect3_query, rect4_query, num_overlay,
                   prefix_overlay, vertex_overlay, hole_overlay);
    if (rc) {
        free_layer(rect1, rect2, rect3, rect4, num_base, prefix_base, vertex_base, hole_base);
        PRT_REQUIRE(false, rc, "read_data: out of memory");
    }
    *num_contours = static_cast<int32_t>(B.n);
    *num_querys = static_cast<int32_t>(Q.n);
    // only the base layer's boxes lower the caller's minimum (OV:805-810)
    if (B.n > 0 && B.mn[0] < *min_x) *min_x = B.mn[0];
    if (B.n > 0 && B.mn[1] < *min_y) *min_y = B.mn[1];
    return PRT_OK;
}
