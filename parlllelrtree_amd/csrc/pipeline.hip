// pipeline.hip -- the map-overlay pipeline as one stream-ordered sequence:
//   polygon MBRs (A, B) -> R-tree over B's MBRs -> join with A's MBRs ->
//   batched FP64 clipping of every candidate pair -> total area and the
//   number of non-empty intersections,
// run eagerly (prt_pipeline_run) or captured once into a hipGraph and
// replayed (prt_pipeline_capture / prt_pipeline_launch).  Nothing in the
// sequence reads device memory on the host: the join writes the pair count
// to the device and the clip and the totals read it there, so the whole
// pipeline is one graph with no host round trip.
//
// The reference has no such driver: its stages are separate programs
// (BotomUp-rtree.cu main BU:642, TopDown-rtree.cu main TD:483) and the clip
// stage is entered through run() (vector_map_overlay_gpc_kernel.cu:61-79),
// which maps rtree_search pairs to polygons (OV:89-90) and counts the
// non-empty results in printfOutput (OV:45-58).  Here the join emits
// (A id, B id) directly, and the count is the number of pairs with area > 0.
#include <new>

#include "prt_common.hpp"
#include "prt_profile.hpp"

struct prt_pipeline {
    hipGraph_t graph;
    hipGraphExec_t exec;
};

namespace prt {
namespace {

// ---- polygon MBRs ---------------------------------------------------------
// fp32 boxes that contain the FP64 polygon: each bound is rounded to nearest
// and then moved one ulp outward when the rounding went inward.  A group of
// MB_GROUP lanes reads one polygon's vertices (coalesced), then reduces.
constexpr int MB_THREADS = 256;
constexpr int MB_GROUP = 8;

__device__ inline float round_down(double v) {
    float f = static_cast<float>(v);
    return static_cast<double>(f) > v ? nextafterf(f, -__builtin_huge_valf()) : f;
}
__device__ inline float round_up(double v) {
    float f = static_cast<float>(v);
    return static_cast<double>(f) < v ? nextafterf(f, __builtin_huge_valf()) : f;
}

__global__ __launch_bounds__(MB_THREADS) void k_poly_mbr(const double *__restrict__ xy,
                                                         const int64_t *__restrict__ off, int64_t n,
                                                         float4 *__restrict__ out) {
    const int sub = threadIdx.x % MB_GROUP;
    constexpr int per_block = MB_THREADS / MB_GROUP;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * per_block + threadIdx.x / MB_GROUP; i < n;
         i += static_cast<int64_t>(gridDim.x) * per_block) {
        double x0 = 1e300, y0 = 1e300, x1 = -1e300, y1 = -1e300;
        const int64_t hi = off[i + 1];
        for (int64_t v = off[i] + sub; v < hi; v += MB_GROUP) {
            const double2 p = reinterpret_cast<const double2 *>(xy)[v];
            x0 = p.x < x0 ? p.x : x0;
            x1 = p.x > x1 ? p.x : x1;
            y0 = p.y < y0 ? p.y : y0;
            y1 = p.y > y1 ? p.y : y1;
        }
        for (int o = 1; o < MB_GROUP; o <<= 1) {
            const double a = __shfl_xor(x0, o, MB_GROUP), b = __shfl_xor(x1, o, MB_GROUP);
            const double c = __shfl_xor(y0, o, MB_GROUP), d = __shfl_xor(y1, o, MB_GROUP);
            x0 = a < x0 ? a : x0;
            x1 = b > x1 ? b : x1;
            y0 = c < y0 ? c : y0;
            y1 = d > y1 ? d : y1;
        }
        if (sub == 0) out[i] = make_float4(round_down(x0), round_down(y0), round_up(x1), round_up(y1));
    }
}

int poly_mbr(const double *xy, const int64_t *off, int64_t n, float *out, hipStream_t s) {
    if (n == 0) return PRT_OK;
    constexpr int per_block = MB_THREADS / MB_GROUP;
    int64_t grid = ceil_div(n, per_block);
    if (grid > 8192) grid = 8192;
    PRT_LAUNCH(k_poly_mbr, dim3(static_cast<unsigned>(grid)), MB_THREADS, 0, s, xy, off, n,
               reinterpret_cast<float4 *>(out));
    PRT_LAUNCHED();
    return PRT_OK;
}

// ---- total area and non-empty count ---------------------------------------
// Deterministic: ST_BLOCKS fixed contiguous chunks, each summed per thread
// in index order and then by a fixed LDS tree; the chunk sums by one more
// tree.  The result depends only on the areas and their count.
constexpr int ST_BLOCKS = 256;
constexpr int ST_THREADS = 256;

__device__ inline void tree_sum(double *ls, int64_t *lc) {
    for (int w = ST_THREADS / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) {
            ls[threadIdx.x] += ls[threadIdx.x + w];
            lc[threadIdx.x] += lc[threadIdx.x + w];
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(ST_THREADS) void k_area_partial(const double *__restrict__ area,
                                                             const int64_t *n_d, int64_t cap,
                                                             double *__restrict__ psum,
                                                             int64_t *__restrict__ pcnt) {
    __shared__ double ls[ST_THREADS];
    __shared__ int64_t lc[ST_THREADS];
    const int64_t n = device_count(n_d, cap);
    const int64_t chunk = ceil_div(n, ST_BLOCKS);
    const int64_t lo = static_cast<int64_t>(blockIdx.x) * chunk;
    const int64_t hi = lo + chunk < n ? lo + chunk : n;
    double s = 0.0;
    int64_t c = 0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += ST_THREADS) {
        const double a = area[i];
        s += a;
        c += a > 0.0 ? 1 : 0;
    }
    ls[threadIdx.x] = s;
    lc[threadIdx.x] = c;
    __syncthreads();
    tree_sum(ls, lc);
    if (threadIdx.x == 0) {
        psum[blockIdx.x] = ls[0];
        pcnt[blockIdx.x] = lc[0];
    }
}

__global__ __launch_bounds__(ST_THREADS) void k_area_final(const double *__restrict__ psum,
                                                           const int64_t *__restrict__ pcnt,
                                                           double *total, int64_t *nonempty) {
    __shared__ double ls[ST_THREADS];
    __shared__ int64_t lc[ST_THREADS];
    ls[threadIdx.x] = psum[threadIdx.x];
    lc[threadIdx.x] = pcnt[threadIdx.x];
    __syncthreads();
    tree_sum(ls, lc);
    if (threadIdx.x == 0) {
        if (total) *total = ls[0];
        if (nonempty) *nonempty = lc[0];
    }
}

__global__ void k_pipe_empty(int64_t *count, double *total, int64_t *nonempty) {
    if (threadIdx.x == 0) {
        count[0] = 0;
        count[1] = 0;
        if (total) *total = 0.0;
        if (nonempty) *nonempty = 0;
    }
}

size_t total_ws_bytes() {
    Sizer z;
    z.take<double>(ST_BLOCKS);
    z.take<int64_t>(ST_BLOCKS);
    return z.bytes();
}

int area_total(const double *area, const int64_t *n_d, int64_t cap, double *total,
               int64_t *nonempty, void *ws, size_t ws_bytes, hipStream_t s) {
    Carver c(ws, ws_bytes);
    double *psum = c.take<double>(ST_BLOCKS);
    int64_t *pcnt = c.take<int64_t>(ST_BLOCKS);
    PRT_REQUIRE(c.ok(), PRT_E_NOMEM, "overlay_total: workspace too small");
    PRT_LAUNCH(k_area_partial, ST_BLOCKS, ST_THREADS, 0, s, area, n_d, cap, psum, pcnt);
    PRT_LAUNCHED();
    PRT_LAUNCH(k_area_final, 1, ST_THREADS, 0, s, psum, pcnt, total, nonempty);
    PRT_LAUNCHED();
    return PRT_OK;
}

// ---- the pipeline ----------------------------------------------------------
struct Pipe {
    float *mbrA, *mbrB;
    prt_bu_tree bu;
    prt_td_tree td;
    void *stats_ws;
    size_t stats_bytes;
    void *work;  // build, then join, then clip (stream-ordered reuse)
    size_t work_bytes;
};

int check_io(const prt_pipeline_io *io) {
    PRT_REQUIRE(io != nullptr, PRT_E_INVAL, "pipeline: null io");
    PRT_REQUIRE(io->nA >= 0 && io->nB >= 0 && io->nA < (int64_t(1) << 31) &&
                    io->nB < (int64_t(1) << 31),
                PRT_E_INVAL, "pipeline: polygon counts out of range");
    PRT_REQUIRE(io->tree == 0 || io->tree == 1, PRT_E_INVAL, "pipeline: tree must be 0 (BU) or 1 (TD)");
    PRT_REQUIRE(io->fanout >= 2 && io->fanout <= PRT_MAX_FANOUT, PRT_E_INVAL, "pipeline: fanout %d",
                io->fanout);
    PRT_REQUIRE(io->cap >= 0 && io->cap < (int64_t(1) << 32), PRT_E_INVAL, "pipeline: cap out of range");
    PRT_REQUIRE(io->count != nullptr, PRT_E_INVAL, "pipeline: null count");
    PRT_REQUIRE(io->cap == 0 || (io->pairs && io->area), PRT_E_INVAL, "pipeline: null pairs/area");
    PRT_REQUIRE(io->nA == 0 || (io->xyA && io->offA), PRT_E_INVAL, "pipeline: null A");
    PRT_REQUIRE(io->nB == 0 || (io->xyB && io->offB), PRT_E_INVAL, "pipeline: null B");
    PRT_REQUIRE(io->max_verts >= 0 && io->max_verts <= 2 * PRT_MAX_POLY_VERTS, PRT_E_INVAL,
                "pipeline: max_verts %d out of range", io->max_verts);
    return PRT_OK;
}

bool empty_io(const prt_pipeline_io *io) { return io->nA == 0 || io->nB == 0; }

// Carves (or, with a Sizer-backed call, sizes) the pipeline's buffers.
int layout(const prt_pipeline_io *io, void *ws, size_t ws_bytes, Pipe *p, size_t *need) {
    *p = Pipe{};
    size_t work = prt_join_workspace_size(io->nA);
    const size_t ov = prt_overlay_batched_workspace_size(io->cap, io->max_verts);
    work = ov > work ? ov : work;
    Sizer z;
    z.take<float>(4 * static_cast<size_t>(io->nA));
    z.take<float>(4 * static_cast<size_t>(io->nB));
    const size_t f = static_cast<size_t>(io->fanout);
    const bool empty = empty_io(io);
    if (empty) {
        // no tree, no join: only the fill of the outputs
    } else if (io->tree == 0) {
        int rc = prt_bu_plan(io->nB, io->fanout, &p->bu);
        if (rc) return rc;
        const size_t nn = static_cast<size_t>(p->bu.num_nodes), nb = static_cast<size_t>(io->nB);
        z.take<uint32_t>(nb);
        z.take<uint32_t>(nb);
        z.take<int32_t>(nn);
        z.take<int32_t>(nn);
        z.take<int32_t>(nn);
        z.take<float>(nn * f * 4);
        z.take<int32_t>(nn * f);
        const size_t b = prt_bu_workspace_size(io->nB, io->fanout);
        work = b > work ? b : work;
    } else {
        int rc = prt_td_plan(io->nB, io->fanout, &p->td);
        if (rc) return rc;
        const size_t ns = static_cast<size_t>(p->td.num_slots);
        z.take<int32_t>(ns);
        z.take<float>(ns * 4);
        z.take<uint32_t>(static_cast<size_t>(io->nB));
        const size_t b = prt_td_workspace_size(io->nB, io->fanout);
        work = b > work ? b : work;
    }
    z.take<char>(total_ws_bytes());
    z.take<char>(work);
    *need = z.bytes();
    if (!ws) return PRT_OK;
    PRT_REQUIRE(ws_bytes >= *need, PRT_E_NOMEM, "pipeline: workspace %zu < %zu bytes", ws_bytes,
                *need);
    Carver c(ws, ws_bytes);
    p->mbrA = c.take<float>(4 * static_cast<size_t>(io->nA));
    p->mbrB = c.take<float>(4 * static_cast<size_t>(io->nB));
    if (empty) {
    } else if (io->tree == 0) {
        const size_t nn = static_cast<size_t>(p->bu.num_nodes), nb = static_cast<size_t>(io->nB);
        p->bu.keys = c.take<uint32_t>(nb);
        p->bu.perm = c.take<uint32_t>(nb);
        p->bu.node_count = c.take<int32_t>(nn);
        p->bu.node_start = c.take<int32_t>(nn);
        p->bu.node_end = c.take<int32_t>(nn);
        p->bu.entry_mbr = c.take<float>(nn * f * 4);
        p->bu.entry_ref = c.take<int32_t>(nn * f);
    } else {
        const size_t ns = static_cast<size_t>(p->td.num_slots);
        p->td.slot_child = c.take<int32_t>(ns);
        p->td.slot_mbr = c.take<float>(ns * 4);
        p->td.perm = c.take<uint32_t>(static_cast<size_t>(io->nB));
    }
    p->stats_bytes = total_ws_bytes();
    p->stats_ws = c.take<char>(p->stats_bytes);
    p->work_bytes = work;
    p->work = c.take<char>(work);
    PRT_REQUIRE(c.ok(), PRT_E_NOMEM, "pipeline: workspace too small");
    return PRT_OK;
}

int enqueue(const prt_pipeline_io *io, const prt_tune *tune, Pipe &p, hipStream_t s) {
    if (empty_io(io)) {
        PRT_LAUNCH(k_pipe_empty, 1, 64, 0, s, io->count, io->total, io->nonempty);
        PRT_LAUNCHED();
        return PRT_OK;
    }
    int rc = poly_mbr(io->xyA, io->offA, io->nA, p.mbrA, s);
    if (!rc) rc = poly_mbr(io->xyB, io->offB, io->nB, p.mbrB, s);
    if (rc) return rc;
    if (io->tree == 0) {
        rc = prt_bu_build(p.mbrB, io->nB, io->fanout, &p.bu, p.work, p.work_bytes, s);
        if (!rc)
            rc = prt_join_bu(&p.bu, p.mbrA, io->nA, io->pairs, io->cap, io->count, p.work,
                             p.work_bytes, s);
    } else {
        rc = prt_td_build(p.mbrB, io->nB, io->fanout, &p.td, p.work, p.work_bytes, s);
        if (!rc)
            rc = prt_join_td(&p.td, p.mbrA, io->nA, io->pairs, io->cap, io->count, p.work,
                             p.work_bytes, s);
    }
    if (rc) return rc;
    if (io->cap > 0) {
        rc = prt_overlay_batched_dev(io->pairs, io->count, io->cap, io->xyA, io->offA, io->xyB,
                                     io->offB, io->max_verts, tune, io->area, p.work, p.work_bytes,
                                     s);
        if (rc) return rc;
    }
    if (io->total || io->nonempty)
        return area_total(io->area, io->count, io->cap, io->total, io->nonempty, p.stats_ws,
                          p.stats_bytes, s);
    return PRT_OK;
}

}  // namespace
}  // namespace prt

using namespace prt;

extern "C" int prt_poly_mbr_dev(const double *xy, const int64_t *off, int64_t n, float *mbr,
                                void *stream) {
    PRT_REQUIRE(n >= 0, PRT_E_INVAL, "poly_mbr: n < 0");
    PRT_REQUIRE(n == 0 || (xy && off && mbr), PRT_E_INVAL, "poly_mbr: null pointer");
    return poly_mbr(xy, off, n, mbr, as_stream(stream));
}

extern "C" size_t prt_overlay_total_workspace_size(void) { return total_ws_bytes(); }

extern "C" int prt_overlay_total(const double *area, const int64_t *npairs_d, int64_t cap,
                                 double *total, int64_t *nonempty, void *ws, size_t ws_bytes,
                                 void *stream) {
    PRT_REQUIRE(cap >= 0, PRT_E_INVAL, "overlay_total: cap < 0");
    PRT_REQUIRE((cap == 0 || area) && ws, PRT_E_INVAL, "overlay_total: null pointer");
    return area_total(area, npairs_d, cap, total, nonempty, ws, ws_bytes, as_stream(stream));
}

extern "C" size_t prt_pipeline_workspace_size(const prt_pipeline_io *io) {
    if (check_io(io) != PRT_OK) return 0;
    Pipe p;
    size_t need = 0;
    if (layout(io, nullptr, 0, &p, &need) != PRT_OK) return 0;
    return need;
}

extern "C" int prt_pipeline_run(const prt_pipeline_io *io, const prt_tune *tune, void *ws,
                                size_t ws_bytes, void *stream) {
    int rc = check_io(io);
    if (rc) return rc;
    Pipe p;
    size_t need = 0;
    PRT_REQUIRE(ws != nullptr, PRT_E_INVAL, "pipeline: null workspace");
    rc = layout(io, ws, ws_bytes, &p, &need);
    if (rc) return rc;
    return enqueue(io, tune, p, as_stream(stream));
}

extern "C" int prt_pipeline_capture(const prt_pipeline_io *io, const prt_tune *tune, void *ws,
                                    size_t ws_bytes, void *stream, prt_pipeline **out) {
    PRT_REQUIRE(out != nullptr, PRT_E_INVAL, "pipeline_capture: null out");
    *out = nullptr;
    PRT_REQUIRE(stream != nullptr, PRT_E_INVAL,
                "pipeline_capture: needs a created stream (the null stream cannot be captured)");
    int rc = check_io(io);
    if (rc) return rc;
    Pipe p;
    size_t need = 0;
    PRT_REQUIRE(ws != nullptr, PRT_E_INVAL, "pipeline: null workspace");
    rc = layout(io, ws, ws_bytes, &p, &need);
    if (rc) return rc;
    hipStream_t s = as_stream(stream);
    ProfPause pause;  // no event nodes in the graph
    PRT_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    rc = enqueue(io, tune, p, s);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (rc != PRT_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
    }
    hipGraphExec_t x = nullptr;
    const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        (void)hipGraphDestroy(g);
        return hip_fail(ei, "hipGraphInstantiate", __FILE__, __LINE__);
    }
    prt_pipeline *h = new (std::nothrow) prt_pipeline{g, x};
    if (!h) {
        (void)hipGraphExecDestroy(x);
        (void)hipGraphDestroy(g);
        PRT_REQUIRE(false, PRT_E_NOMEM, "pipeline_capture: out of host memory");
    }
    *out = h;
    return PRT_OK;
}

extern "C" int prt_pipeline_launch(prt_pipeline *p, void *stream) {
    PRT_REQUIRE(p != nullptr, PRT_E_INVAL, "pipeline_launch: null pipeline");
    PRT_HIP(hipGraphLaunch(p->exec, as_stream(stream)));
    return PRT_OK;
}

extern "C" int prt_pipeline_destroy(prt_pipeline *p) {
    if (!p) return PRT_OK;
    const hipError_t a = hipGraphExecDestroy(p->exec);
    const hipError_t b = hipGraphDestroy(p->graph);
    delete p;
    PRT_HIP(a);
    PRT_HIP(b);
    return PRT_OK;
}
