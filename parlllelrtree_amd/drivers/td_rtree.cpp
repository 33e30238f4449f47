// td_rtree.cpp -- the top-down program of the reference (TopDown-rtree.cu
// main, :483-798) over libprt: read the rect file, build the top-down tree
// on the GPU, join every rect against it and print the reference's search
// counters.  Like the reference, no normalisation.
//
//   prt_td_rtree RECTS.txt [FANOUT=10]
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
    float *rd = dev_copy(rects_f32(r));

    prt_td_tree t;
    std::memset(&t, 0, sizeof t);
    DRV_CHECK(prt_td_plan(n, f, &t));
    t.slot_child = dev_alloc<int32_t>(t.num_slots);
    t.slot_mbr = dev_alloc<float>(4 * t.num_slots);
    t.perm = dev_alloc<uint32_t>(n);
    const size_t wsb = prt_td_workspace_size(n, f);
    void *ws = dev_alloc<char>(wsb);
    Timer tm;
    tm.start();
    DRV_CHECK(prt_td_build(rd, n, f, &t, ws, wsb, nullptr));
    std::printf("GPU Build time %0.2fms\n", tm.stop_ms());

    const size_t jws = prt_join_workspace_size(n);
    void *jw = dev_alloc<char>(jws);
    int64_t *cnt = dev_alloc<int64_t>(2);
    tm.start();
    DRV_CHECK(prt_join_td(&t, rd, n, nullptr, 0, cnt, jw, jws, nullptr));
    std::printf("GPU Query time %0.2fms\n", tm.stop_ms());
    int64_t hc[2];
    DRV_HIP(hipMemcpy(hc, cnt, sizeof hc, hipMemcpyDeviceToHost));
    std::printf("pairs=%lld\n", (long long)hc[0]);

    int64_t *st = dev_alloc<int64_t>(7);
    DRV_CHECK(prt_search_stats_td(&t, rd, n, st, nullptr));
    DRV_CHECK(prt_sibling_overlap_td(&t, st + 6, nullptr));
    int64_t hs[7];
    DRV_HIP(hipMemcpy(hs, st, sizeof hs, hipMemcpyDeviceToHost));
    print_stats(hs, n, hs[6], static_cast<double>(hs[6]) / t.num_slots, true);
    return 0;
}
