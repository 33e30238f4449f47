// bu_rtree.cpp -- the bottom-up program of the reference (BotomUp-rtree.cu
// main, :642-1027) over libprt: read the rect file, normalise, build the
// tree in the reference's key order on the GPU, join every rect against it,
// and print the reference's search counters.
//
//   prt_bu_rtree RECTS.txt [FANOUT=10]
#include <algorithm>
#include <cstring>

#include "common.hpp"

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s RECTS.txt [FANOUT=10]\n", argv[0]);
        return 2;
    }
    const int f = argc > 2 ? std::atoi(argv[2]) : 10;
    int64_t n = 0;
    DRV_CHECK(prt_read_rects_text(argv[1], &n, nullptr, nullptr));
    std::vector<int64_t> r(4 * n);
    std::vector<int32_t> key(n);
    DRV_CHECK(prt_read_rects_text(argv[1], &n, r.data(), key.data()));
    int64_t maxx = 0, maxy = 0;  // the reference starts its maxima at 0 (BU:664, 666)
    for (int64_t i = 0; i < n; ++i) {
        maxx = std::max(maxx, r[4 * i + 2]);
        maxy = std::max(maxy, r[4 * i + 3]);
    }
    int64_t minx = 0, miny = 0, scale = 1;
    DRV_CHECK(prt_rects_normalize(r.data(), n, &minx, &miny, &scale));
    std::printf("%lld %lld %lld %lld\n", (long long)minx, (long long)maxx, (long long)miny,
                (long long)maxy);
    std::printf("%lld\n", (long long)scale);

    prt_bu_tree t;
    std::memset(&t, 0, sizeof t);
    DRV_CHECK(prt_bu_plan(n, f, &t));
    t.keys = dev_alloc<uint32_t>(n);
    t.perm = dev_alloc<uint32_t>(n);
    t.node_count = dev_alloc<int32_t>(t.num_nodes);
    t.node_start = dev_alloc<int32_t>(t.num_nodes);
    t.node_end = dev_alloc<int32_t>(t.num_nodes);
    t.entry_mbr = dev_alloc<float>(t.num_nodes * f * 4);
    t.entry_ref = dev_alloc<int32_t>(t.num_nodes * f);
    int64_t *rd = dev_copy(r);
    const size_t wsb = prt_bu_ref4d_workspace_size(n, f);
    void *ws = dev_alloc<char>(wsb);
    Timer tm;
    tm.start();
    DRV_CHECK(prt_bu_build_ref4d(rd, n, f, scale, &t, nullptr, ws, wsb, nullptr));
    std::printf("GPU Build time %0.2fms\n", tm.stop_ms());

    // the reference searches every data rect against the tree
    float *q = dev_copy(rects_f32(r));
    const size_t jws = prt_join_workspace_size(n);
    void *jw = dev_alloc<char>(jws);
    int64_t *cnt = dev_alloc<int64_t>(2);
    tm.start();
    DRV_CHECK(prt_join_bu(&t, q, n, nullptr, 0, cnt, jw, jws, nullptr));
    std::printf("GPU Query time %0.2fms\n", tm.stop_ms());
    int64_t hc[2];
    DRV_HIP(hipMemcpy(hc, cnt, sizeof hc, hipMemcpyDeviceToHost));
    std::printf("pairs=%lld\n", (long long)hc[0]);

    int64_t *st = dev_alloc<int64_t>(7);
    DRV_CHECK(prt_search_stats_bu(&t, q, n, st, nullptr));
    DRV_CHECK(prt_sibling_overlap_bu(&t, st + 6, nullptr));
    int64_t hs[7];
    DRV_HIP(hipMemcpy(hs, st, sizeof hs, hipMemcpyDeviceToHost));
    // ave3 as BU:1026 prints it: sum / num_nodes / m, with num_nodes the
    // reference's node budget 3 * num_contours / m (integer division, BU:674)
    const int64_t ref_nodes = 3 * n / f;
    print_stats(hs, n, hs[6], static_cast<double>(hs[6]) / static_cast<double>(ref_nodes) / f,
                false);
    return 0;
}
