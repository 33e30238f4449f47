// common.hpp -- helpers of the command-line drivers (host C++ over the
// C-ABI of prt.h; the only HIP calls are allocation, copies and events).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "prt.h"

#define DRV_CHECK(call)                                                              \
    do {                                                                             \
        int _rc = (call);                                                            \
        if (_rc != PRT_OK) {                                                         \
            std::fprintf(stderr, "%s failed (%d): %s\n", #call, _rc, prt_last_error()); \
            std::exit(1);                                                            \
        }                                                                            \
    } while (0)

#define DRV_HIP(call)                                                                \
    do {                                                                             \
        hipError_t _e = (call);                                                      \
        if (_e != hipSuccess) {                                                      \
            std::fprintf(stderr, "%s failed: %s\n", #call, hipGetErrorString(_e));   \
            std::exit(1);                                                            \
        }                                                                            \
    } while (0)

template <class T> T *dev_alloc(size_t n) {
    void *p = nullptr;
    DRV_HIP(hipMalloc(&p, (n ? n : 1) * sizeof(T)));
    return static_cast<T *>(p);
}

template <class T> T *dev_copy(const std::vector<T> &h) {
    T *d = dev_alloc<T>(h.size());
    if (!h.empty()) DRV_HIP(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

struct Timer {
    hipEvent_t a, b;
    Timer() {
        DRV_HIP(hipEventCreate(&a));
        DRV_HIP(hipEventCreate(&b));
    }
    void start() { DRV_HIP(hipEventRecord(a, nullptr)); }
    float stop_ms() {
        DRV_HIP(hipEventRecord(b, nullptr));
        DRV_HIP(hipEventSynchronize(b));
        float ms = 0.f;
        DRV_HIP(hipEventElapsedTime(&ms, a, b));
        return ms;
    }
};

// integer rects (x1, y1, x2, y2) -> fp32 boxes; refuses values that are not
// exact in fp32
inline std::vector<float> rects_f32(const std::vector<int64_t> &r) {
    std::vector<float> f(r.size());
    for (size_t i = 0; i < r.size(); ++i) {
        if (r[i] > (int64_t(1) << 24) || r[i] < -(int64_t(1) << 24)) {
            std::fprintf(stderr, "coordinate %lld is not exact in fp32\n", (long long)r[i]);
            std::exit(1);
        }
        f[i] = static_cast<float>(r[i]);
    }
    return f;
}

// the counters the reference prints after its CPU search; `td` selects the
// top-down file's format (TopDown-rtree.cu:793-797) over the bottom-up one
// (BotomUp-rtree.cu:1021-1026)
inline void print_stats(const int64_t *st, int64_t nq, int64_t sum3, double ave3, bool td) {
    const double d = nq > 0 ? static_cast<double>(nq) : 1.0;
    const long long sum2 = st[4] / 2;
    if (td) {
        std::printf("sum = %lld  ave = %0.2lf max = %lld\n", (long long)st[0], st[0] / d, (long long)st[1]);
        std::printf("sum1 = %lld ave1 = %0.2lf max1 = %lld \n", (long long)st[2], st[2] / d, (long long)st[3]);
        std::printf("sum2 = %lld ave2 = %0.2lf max2 = %lld \n", sum2, sum2 / d, (long long)st[5]);
    } else {
        std::printf("sum=%lld  ave = %0.2lf max = %lld\n", (long long)st[0], st[0] / d, (long long)st[1]);
        std::printf("sum1=%lld  ave = %0.2lf max1 = %lld\n", (long long)st[2], st[2] / d, (long long)st[3]);
        std::printf("sum2=%lld  ave = %0.2lf max2 = %lld\n", sum2, sum2 / d, (long long)st[5]);
    }
    std::printf("sum3 = %lld  ave3 = %0.2lf\n", (long long)sum3, ave3);
}
