// status.hip -- per-device status flags that kernels raise when a result
// they wrote cannot be trusted, and the query that reports them.
//
// The entry points are stream-ordered and return before their kernels run,
// so a condition found on the device cannot become their return value.  The
// kernels OR a bit into this device's flag word instead (PRT_STATUS_* of
// prt.h); prt_device_status() waits for the stream, reads the word, and
// turns a raised bit into PRT_E_DEVICE with a message that names it.
#include <mutex>

#include "prt_common.hpp"

namespace prt {
namespace {

__device__ __align__(256) uint32_t g_status_flags[64];

constexpr int MAX_DEVICES = 64;
uint32_t *g_flags_ptr[MAX_DEVICES] = {};
std::mutex g_flags_mu;

__global__ void k_status_read_clear(uint32_t *flags, uint32_t *out, int clear) {
    if (threadIdx.x == 0) {
        *out = *flags;
        if (clear) *flags = 0u;
    }
}

}  // namespace

// Address of the current device's flag word (resolved once per device).
uint32_t *device_status_flags() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_flags_mu);
    if (!g_flags_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(g_status_flags)) != hipSuccess) return nullptr;
        g_flags_ptr[dev] = static_cast<uint32_t *>(p);
    }
    return g_flags_ptr[dev];
}

}  // namespace prt

using namespace prt;

extern "C" int prt_device_status(int32_t *flags_h, int clear, void *stream) {
    uint32_t *flags = device_status_flags();
    PRT_REQUIRE(flags != nullptr, PRT_E_HIP, "device_status: no flag word on this device");
    hipStream_t s = as_stream(stream);
    // the value travels through the word after the flags (the array has 64)
    uint32_t *out = flags + 1;
    k_status_read_clear<<<1, 64, 0, s>>>(flags, out, clear ? 1 : 0);
    PRT_LAUNCHED();
    uint32_t v = 0;
    PRT_HIP(hipMemcpyAsync(&v, out, sizeof v, hipMemcpyDeviceToHost, s));
    PRT_HIP(hipStreamSynchronize(s));
    if (flags_h) *flags_h = static_cast<int32_t>(v);
    if (v == 0) return PRT_OK;
    set_error("device status 0x%x:%s%s", v,
              (v & PRT_STATUS_LOOKBACK)
                  ? " a radix sort's look-back gave up waiting for an earlier tile (the sorted "
                    "order, and every tree built from it, is invalid);"
                  : "",
              (v & PRT_STATUS_INEXACT)
                  ? " int64 coordinates not exact in fp32 were rounded outward (joins over that "
                    "tree return a superset of the exact hit set);"
                  : "");
    return PRT_E_DEVICE;
}
