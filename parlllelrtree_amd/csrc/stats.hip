// stats.hip -- the statistics the reference prints after its searches: the
// per-query counters of its CPU search (entries tested `num`, the search's
// return value, leaf hits `num1`; BU:455-483, TD:56-76, summed and maxed
// as in BU:990-1027 / TD:760-798) and the count of overlapping sibling
// entries (RtreeSearch1, BU:574-590, TD:78-93).
//
// One thread per query walks the tree depth first with an explicit stack
// of (node, next entry, running value) frames, which is the reference's
// recursion unrolled: a frame's value starts at -1, an overlapping entry
// adds one plus (internal) the child's result, and a finished frame hands
// max(0, value) to its parent.  Per-query results are reduced in the block
// and added to six 64-bit totals with atomics (sums are order-independent
// integers, so the totals are exact).
#include "prt_common.hpp"
#include "prt_profile.hpp"

namespace prt {
namespace {

constexpr int ST_THREADS = 256;

__global__ void k_zero_stats(int64_t *p, int n) {
    if (static_cast<int>(threadIdx.x) < n) p[threadIdx.x] = 0;
}

__device__ inline void block_add_max(int64_t v_num, int64_t v_cnt, int64_t v_hit, int64_t *out) {
    __shared__ int64_t s[3][ST_THREADS / 64];
    __shared__ int64_t m[3][ST_THREADS / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int64_t a = v_num, b = v_cnt, c = v_hit, ma = v_num, mb = v_cnt, mc = v_hit;
    for (int o = 1; o < 64; o <<= 1) {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
        c += __shfl_xor(c, o, 64);
        const int64_t xa = __shfl_xor(ma, o, 64), xb = __shfl_xor(mb, o, 64), xc = __shfl_xor(mc, o, 64);
        ma = xa > ma ? xa : ma;
        mb = xb > mb ? xb : mb;
        mc = xc > mc ? xc : mc;
    }
    if (lane == 0) {
        s[0][wave] = a; s[1][wave] = b; s[2][wave] = c;
        m[0][wave] = ma; m[1][wave] = mb; m[2][wave] = mc;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        int64_t t = 0, mx = 0;
        for (int w = 0; w < ST_THREADS / 64; ++w) {
            t += s[threadIdx.x][w];
            mx = m[threadIdx.x][w] > mx ? m[threadIdx.x][w] : mx;
        }
        atomicAdd(reinterpret_cast<unsigned long long *>(&out[2 * threadIdx.x]),
                  static_cast<unsigned long long>(t));
        atomicMax(reinterpret_cast<unsigned long long *>(&out[2 * threadIdx.x + 1]),
                  static_cast<unsigned long long>(mx));
    }
}

// out[0..5] = sum num, max num, sum count, max count, sum num1, max num1
__global__ __launch_bounds__(ST_THREADS) void k_search_stats_bu(
    const int32_t *__restrict__ cnt, const float4 *__restrict__ mbr, const int32_t *__restrict__ ref,
    int64_t root, int64_t num_leaves, int f, int levels, const float4 *__restrict__ q4, int64_t nq,
    int64_t *out) {
    const int64_t qi = static_cast<int64_t>(blockIdx.x) * ST_THREADS + threadIdx.x;
    int64_t num = 0, count = 0, num1 = 0;
    if (qi < nq) {
        const float4 q = q4[qi];
        int32_t node[PRT_MAX_LEVELS], idx[PRT_MAX_LEVELS];
        int64_t val[PRT_MAX_LEVELS];
        int d = 0;
        node[0] = static_cast<int32_t>(root);
        idx[0] = 0;
        val[0] = -1;
        for (;;) {
            const int32_t nd = node[d];
            if (idx[d] < cnt[nd]) {
                const int i = idx[d]++;
                ++num;
                const int64_t e = static_cast<int64_t>(nd) * f + i;
                if (!overlap(q, mbr[e])) continue;
                ++val[d];
                if (nd < num_leaves) {
                    ++num1;
                } else if (d + 1 < levels) {
                    ++d;
                    node[d] = ref[e];
                    idx[d] = 0;
                    val[d] = -1;
                }
            } else {
                const int64_t r = val[d] > 0 ? val[d] : 0;
                if (d == 0) {
                    count = r;
                    break;
                }
                --d;
                val[d] += r;
            }
        }
    }
    block_add_max(num, count, num1, out);
}

__global__ __launch_bounds__(ST_THREADS) void k_search_stats_td(
    const int32_t *__restrict__ child, const float4 *__restrict__ mbr, int f, int levels,
    const float4 *__restrict__ q4, int64_t nq, int64_t *out) {
    const int64_t qi = static_cast<int64_t>(blockIdx.x) * ST_THREADS + threadIdx.x;
    int64_t num = 0, count = 0, num1 = 0;
    if (qi < nq) {
        const float4 q = q4[qi];
        int32_t node[PRT_MAX_LEVELS], idx[PRT_MAX_LEVELS];
        int64_t val[PRT_MAX_LEVELS];
        int d = 0;
        node[0] = 0;
        idx[0] = 0;
        val[0] = -1;
        for (;;) {
            if (idx[d] < f) {
                const int64_t s = static_cast<int64_t>(node[d]) + idx[d]++;
                ++num;
                const int32_t c = child[s];
                if (c == 0 || !overlap(mbr[s], q)) continue;
                ++val[d];
                if (c < 0) {
                    ++num1;
                } else if (d + 1 < levels) {
                    ++d;
                    node[d] = c;
                    idx[d] = 0;
                    val[d] = -1;
                }
            } else {
                const int64_t r = val[d] > 0 ? val[d] : 0;
                if (d == 0) {
                    count = r;
                    break;
                }
                --d;
                val[d] += r;
            }
        }
    }
    block_add_max(num, count, num1, out);
}

__device__ inline void block_add(int64_t v, int64_t *out) {
    __shared__ int64_t s[ST_THREADS / 64];
    for (int o = 1; o < 64; o <<= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t t = 0;
        for (int w = 0; w < ST_THREADS / 64; ++w) t += s[w];
        atomicAdd(reinterpret_cast<unsigned long long *>(out), static_cast<unsigned long long>(t));
    }
}

// every node of a bottom-up tree is reachable from the root
__global__ __launch_bounds__(ST_THREADS) void k_sibling_bu(const int32_t *__restrict__ cnt,
                                                           const float4 *__restrict__ mbr,
                                                           int64_t num_nodes, int f,
                                                           int64_t *out) {
    const int64_t nd = static_cast<int64_t>(blockIdx.x) * ST_THREADS + threadIdx.x;
    int64_t c = 0;
    if (nd < num_nodes) {
        const int k = cnt[nd];
        for (int i = 1; i < k; ++i) {
            const float4 bi = mbr[nd * f + i];
            for (int j = i - 1; j >= 0; --j)
                if (overlap(mbr[nd * f + j], bi)) ++c;
        }
    }
    block_add(c, out);
}

// blocks whose first slot is empty are unreachable and hold no entries
// (entries fill every level from the left)
__global__ __launch_bounds__(ST_THREADS) void k_sibling_td(const int32_t *__restrict__ child,
                                                           const float4 *__restrict__ mbr,
                                                           int64_t num_blocks, int f,
                                                           int64_t *out) {
    const int64_t b = static_cast<int64_t>(blockIdx.x) * ST_THREADS + threadIdx.x;
    int64_t c = 0;
    if (b < num_blocks) {
        const int64_t s0 = b * f;
        for (int i = 1; i < f && child[s0 + i] != 0; ++i) {
            const float4 bi = mbr[s0 + i];
            for (int j = i - 1; j >= 0; --j)
                if (overlap(bi, mbr[s0 + j])) ++c;
        }
    }
    block_add(c, out);
}

unsigned grid_for(int64_t n) {
    const int64_t g = ceil_div(n > 0 ? n : 1, ST_THREADS);
    return static_cast<unsigned>(g);
}

}  // namespace
}  // namespace prt

using namespace prt;

extern "C" int prt_search_stats_bu(const prt_bu_tree *tree, const float *queries, int64_t nq,
                                   int64_t *stats, void *stream) {
    PRT_REQUIRE(tree && tree->num_nodes > 0 && tree->node_count && tree->entry_mbr &&
                    tree->entry_ref && stats,
                PRT_E_INVAL, "search_stats_bu: tree not built or null output");
    PRT_REQUIRE(nq >= 0 && nq < (int64_t(1) << 40) && (nq == 0 || queries), PRT_E_INVAL,
                "search_stats_bu: bad queries");
    hipStream_t s = as_stream(stream);
    PRT_LAUNCH(k_zero_stats, 1, 64, 0, s, stats, 6);
    PRT_LAUNCHED();
    if (nq == 0) return PRT_OK;
    PRT_LAUNCH(k_search_stats_bu, grid_for(nq), ST_THREADS, 0, s, tree->node_count,
               reinterpret_cast<const float4 *>(tree->entry_mbr), tree->entry_ref,
               tree->num_nodes - 1, tree->level_count[0], tree->fanout, tree->num_levels,
               reinterpret_cast<const float4 *>(queries), nq, stats);
    PRT_LAUNCHED();
    return PRT_OK;
}

extern "C" int prt_search_stats_td(const prt_td_tree *tree, const float *queries, int64_t nq,
                                   int64_t *stats, void *stream) {
    PRT_REQUIRE(tree && tree->num_slots > 0 && tree->slot_child && tree->slot_mbr && stats,
                PRT_E_INVAL, "search_stats_td: tree not built or null output");
    PRT_REQUIRE(nq >= 0 && nq < (int64_t(1) << 40) && (nq == 0 || queries), PRT_E_INVAL,
                "search_stats_td: bad queries");
    hipStream_t s = as_stream(stream);
    PRT_LAUNCH(k_zero_stats, 1, 64, 0, s, stats, 6);
    PRT_LAUNCHED();
    if (nq == 0) return PRT_OK;
    PRT_LAUNCH(k_search_stats_td, grid_for(nq), ST_THREADS, 0, s, tree->slot_child,
               reinterpret_cast<const float4 *>(tree->slot_mbr), tree->fanout, tree->num_levels,
               reinterpret_cast<const float4 *>(queries), nq, stats);
    PRT_LAUNCHED();
    return PRT_OK;
}

extern "C" int prt_sibling_overlap_bu(const prt_bu_tree *tree, int64_t *count, void *stream) {
    PRT_REQUIRE(tree && tree->num_nodes > 0 && tree->node_count && tree->entry_mbr && count,
                PRT_E_INVAL, "sibling_overlap_bu: tree not built or null output");
    hipStream_t s = as_stream(stream);
    PRT_LAUNCH(k_zero_stats, 1, 64, 0, s, count, 1);
    PRT_LAUNCHED();
    PRT_LAUNCH(k_sibling_bu, grid_for(tree->num_nodes), ST_THREADS, 0, s, tree->node_count,
               reinterpret_cast<const float4 *>(tree->entry_mbr), tree->num_nodes, tree->fanout,
               count);
    PRT_LAUNCHED();
    return PRT_OK;
}

extern "C" int prt_sibling_overlap_td(const prt_td_tree *tree, int64_t *count, void *stream) {
    PRT_REQUIRE(tree && tree->num_slots > 0 && tree->slot_child && tree->slot_mbr && count,
                PRT_E_INVAL, "sibling_overlap_td: tree not built or null output");
    hipStream_t s = as_stream(stream);
    const int64_t blocks = tree->num_slots / tree->fanout;
    PRT_LAUNCH(k_zero_stats, 1, 64, 0, s, count, 1);
    PRT_LAUNCHED();
    PRT_LAUNCH(k_sibling_td, grid_for(blocks), ST_THREADS, 0, s, tree->slot_child,
               reinterpret_cast<const float4 *>(tree->slot_mbr), blocks, tree->fanout, count);
    PRT_LAUNCHED();
    return PRT_OK;
}
