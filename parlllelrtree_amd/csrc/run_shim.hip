// run_shim.hip -- prt_run(): the clip stage behind the argument list of the
// reference's run() (vector_map_overlay_gpc_kernel.cu:61-79), for callers
// that hold the reference's data layout:
//   * pairs are rtree_search records; record t names the base polygon
//     zOrder[nodeid - offset].index (the leaf position mapped through the
//     Z-ordered records, OV:89) and the overlay polygon queryid (OV:90);
//   * polygons are num[] vertex counts, prefix[] (int64) first-vertex
//     offsets and {x, y} double vertices (readData, OV:752-888);
//   * hole flags are accepted and do not change the area (the clip treats
//     every contour alike; the flags only label output contours).
// It returns the reference's count1 (non-empty results, printfOutput
// OV:45-58) and, when `area` is given, the FP64 intersection area of every
// pair.  Like the reference's run() it allocates device memory and
// synchronises; the batched overlay does the clipping.
//
// The reference's run() has no polygon counts; the shim takes the largest
// referenced index of each layer from the pairs and requires prefix[] to
// be the running sum of num[] (as readData builds it) over that range.
// Every nodeid - offset must be a valid index of zorder_index and every
// index a valid polygon, as in the reference (the shim checks signs only).
#include "prt_common.hpp"
#include "prt_profile.hpp"

namespace prt {
namespace {

constexpr int RN_THREADS = 256;

// scal[0] = max base id, [1] = max overlay id, [2] = error flags,
// [3] = max base vertex count, [4] = max overlay vertex count, [5] = count1
__global__ void k_run_init(int64_t *scal) {
    if (threadIdx.x < 8) scal[threadIdx.x] = 0;
}

__global__ __launch_bounds__(RN_THREADS) void k_run_map(const prt_rtree_search *__restrict__ rs,
                                                        int64_t np, int32_t offset,
                                                        const int32_t *__restrict__ zorder,
                                                        uint32_t *__restrict__ pairs,
                                                        int64_t *scal) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * RN_THREADS + threadIdx.x;
    if (t >= np) return;
    const int64_t pos = static_cast<int64_t>(rs[t].nodeid) - offset;
    const int32_t q = rs[t].queryid;
    if (pos < 0 || q < 0) {
        atomicOr(reinterpret_cast<unsigned long long *>(&scal[2]), 1ull);
        pairs[2 * t] = 0;
        pairs[2 * t + 1] = 0;
        return;
    }
    const int32_t a = zorder[pos];
    if (a < 0) {
        atomicOr(reinterpret_cast<unsigned long long *>(&scal[2]), 1ull);
        pairs[2 * t] = 0;
        pairs[2 * t + 1] = 0;
        return;
    }
    pairs[2 * t] = static_cast<uint32_t>(a);
    pairs[2 * t + 1] = static_cast<uint32_t>(q);
    atomicMax(reinterpret_cast<unsigned long long *>(&scal[0]), static_cast<unsigned long long>(a));
    atomicMax(reinterpret_cast<unsigned long long *>(&scal[1]), static_cast<unsigned long long>(q));
}

// off[i] = prefix[i] (i < n), off[n] = prefix[n-1] + num[n-1]; flags a gap
// or overlap between consecutive polygons, and records the largest count
__global__ __launch_bounds__(RN_THREADS) void k_run_offsets(const int32_t *__restrict__ num,
                                                            const int64_t *__restrict__ prefix,
                                                            int64_t n, int64_t *__restrict__ off,
                                                            int64_t *scal, int which) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * RN_THREADS + threadIdx.x;
    if (i >= n) return;
    const int64_t p = prefix[i], k = num[i];
    off[i] = p;
    if (k < 0) atomicOr(reinterpret_cast<unsigned long long *>(&scal[2]), 4ull);
    if (i + 1 < n) {
        if (p + k != prefix[i + 1]) atomicOr(reinterpret_cast<unsigned long long *>(&scal[2]), 2ull);
    } else {
        off[n] = p + k;
    }
    atomicMax(reinterpret_cast<unsigned long long *>(&scal[3 + which]),
              static_cast<unsigned long long>(k < 0 ? 0 : k));
}

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace
}  // namespace prt

using namespace prt;

extern "C" int prt_run(const prt_rtree_search *pairs, int32_t offset, const int32_t *num_base,
                       const int64_t *prefix_base, const double *vertex_base,
                       const int32_t *num_overlay, const int64_t *prefix_overlay,
                       const double *vertex_overlay, int32_t num_output, const int32_t *hole_base,
                       const int32_t *hole_overlay, const int32_t *zorder_index, double *area,
                       int64_t *count1_h) {
    (void)hole_base;
    (void)hole_overlay;
    PRT_REQUIRE(count1_h != nullptr, PRT_E_INVAL, "run: null count1");
    PRT_REQUIRE(num_output >= 0, PRT_E_INVAL, "run: numOfOutput < 0");
    *count1_h = 0;
    if (num_output == 0) return PRT_OK;
    PRT_REQUIRE(pairs && num_base && prefix_base && vertex_base && num_overlay && prefix_overlay &&
                    vertex_overlay && zorder_index,
                PRT_E_INVAL, "run: null device pointer");
    const int64_t np = num_output;
    hipStream_t s = nullptr;  // the reference's legacy default stream
    DevBuf b_pairs, b_scal, b_area;
    PRT_HIP(hipMalloc(&b_pairs.p, static_cast<size_t>(2 * np) * sizeof(uint32_t)));
    PRT_HIP(hipMalloc(&b_scal.p, 8 * sizeof(int64_t)));
    uint32_t *pp = static_cast<uint32_t *>(b_pairs.p);
    int64_t *scal = static_cast<int64_t *>(b_scal.p);
    double *out = area;
    if (!out) {
        PRT_HIP(hipMalloc(&b_area.p, static_cast<size_t>(np) * sizeof(double)));
        out = static_cast<double *>(b_area.p);
    }
    const unsigned g = static_cast<unsigned>(ceil_div(np, RN_THREADS));
    PRT_LAUNCH(k_run_init, 1, 64, 0, s, scal);
    PRT_LAUNCHED();
    PRT_LAUNCH(k_run_map, g, RN_THREADS, 0, s, pairs, np, offset, zorder_index, pp, scal);
    PRT_LAUNCHED();
    int64_t h[8];
    PRT_HIP(hipMemcpy(h, scal, sizeof h, hipMemcpyDeviceToHost));
    PRT_REQUIRE(!(h[2] & 1), PRT_E_INVAL, "run: a record has nodeid < offset or a negative id");
    const int64_t nA = h[0] + 1, nB = h[1] + 1;
    DevBuf b_offA, b_offB;
    PRT_HIP(hipMalloc(&b_offA.p, static_cast<size_t>(nA + 1) * sizeof(int64_t)));
    PRT_HIP(hipMalloc(&b_offB.p, static_cast<size_t>(nB + 1) * sizeof(int64_t)));
    int64_t *offA = static_cast<int64_t *>(b_offA.p), *offB = static_cast<int64_t *>(b_offB.p);
    PRT_LAUNCH(k_run_offsets, static_cast<unsigned>(ceil_div(nA, RN_THREADS)), RN_THREADS, 0, s,
               num_base, prefix_base, nA, offA, scal, 0);
    PRT_LAUNCHED();
    PRT_LAUNCH(k_run_offsets, static_cast<unsigned>(ceil_div(nB, RN_THREADS)), RN_THREADS, 0, s,
               num_overlay, prefix_overlay, nB, offB, scal, 1);
    PRT_LAUNCHED();
    PRT_HIP(hipMemcpy(h, scal, sizeof h, hipMemcpyDeviceToHost));
    PRT_REQUIRE(!(h[2] & 4), PRT_E_INVAL, "run: negative vertex count");
    PRT_REQUIRE(!(h[2] & 2), PRT_E_UNSUPPORTED,
                "run: prefix[] is not the running sum of num[] over the referenced polygons");
    const int64_t mv = h[3] + h[4];
    PRT_REQUIRE(mv <= 2 * PRT_MAX_POLY_VERTS, PRT_E_UNSUPPORTED,
                "run: polygons of more than %d vertices", PRT_MAX_POLY_VERTS);
    const size_t wsb = prt_overlay_batched_workspace_size(np, static_cast<int32_t>(mv));
    const size_t tb = prt_overlay_total_workspace_size();
    DevBuf b_ws;
    PRT_HIP(hipMalloc(&b_ws.p, wsb > tb ? wsb : tb));
    int rc = prt_overlay_batched(pp, np, vertex_base, offA, vertex_overlay, offB,
                                 static_cast<int32_t>(mv), nullptr, out, b_ws.p, wsb, s);
    if (rc) return rc;
    rc = prt_overlay_total(out, nullptr, np, nullptr, scal + 5, b_ws.p, tb, s);
    if (rc) return rc;
    PRT_HIP(hipMemcpy(h, scal, sizeof h, hipMemcpyDeviceToHost));
    *count1_h = h[5];
    return PRT_OK;
}
