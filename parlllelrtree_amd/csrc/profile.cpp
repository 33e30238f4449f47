// profile.cpp -- optional per-kernel timing with HIP events.
//
// When enabled (prt_profile_enable(1)), every kernel launch site of libprt
// launches through hipExtLaunchKernel with a start and a stop event, which
// take the kernel's own start and end times; prt_profile_read() synchronises
// those events and reports, per kernel name, the number of launches and the
// summed duration.  Disabled (the default) it costs one branch per launch.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "prt_common.hpp"
#include "prt_profile.hpp"

namespace prt {
namespace {

struct Rec {
    const char *name;
    hipEvent_t a, b;
};

std::mutex g_mu;
bool g_on = false;
std::vector<Rec> g_recs;
std::vector<hipEvent_t> g_free;
thread_local int t_paused = 0;

hipEvent_t get_event() {
    if (!g_free.empty()) {
        hipEvent_t e = g_free.back();
        g_free.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

}  // namespace

bool profiling() { return g_on && t_paused == 0; }

ProfPause::ProfPause() { ++t_paused; }
ProfPause::~ProfPause() { --t_paused; }

ProfScope::ProfScope(const char *name) : name_(name), a_(nullptr), b_(nullptr) {
    if (!g_on || t_paused) return;
    std::lock_guard<std::mutex> lk(g_mu);
    hipEvent_t a = get_event();
    hipEvent_t b = a ? get_event() : nullptr;
    if (!b) {
        if (a) g_free.push_back(a);
        return;
    }
    a_ = a;
    b_ = b;
}

ProfScope::~ProfScope() {
    if (!a_) return;
    std::lock_guard<std::mutex> lk(g_mu);
    g_recs.push_back({name_, a_, b_});
}

}  // namespace prt

using namespace prt;

extern "C" void prt_profile_enable(int on) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_on = on != 0;
}

// Aggregates the recorded launches by name into names (nul-separated,
// buffer of names_bytes), counts[] and total_ms[] (max_kernels entries) and
// clears the record.  Returns the number of distinct names (or < 0).
extern "C" int prt_profile_read(char *names, int names_bytes, int64_t *counts, double *total_ms,
                                int max_kernels) {
    std::lock_guard<std::mutex> lk(g_mu);
    std::vector<std::string> keys;
    std::vector<int64_t> cnt;
    std::vector<double> ms;
    for (const Rec &r : g_recs) {
        if (hipEventSynchronize(r.b) != hipSuccess) return PRT_E_HIP;
        float t = 0.f;
        if (hipEventElapsedTime(&t, r.a, r.b) != hipSuccess) return PRT_E_HIP;
        size_t k = 0;
        while (k < keys.size() && keys[k] != r.name) ++k;
        if (k == keys.size()) {
            keys.emplace_back(r.name);
            cnt.push_back(0);
            ms.push_back(0.0);
        }
        cnt[k] += 1;
        ms[k] += t;
        g_free.push_back(r.a);
        g_free.push_back(r.b);
    }
    g_recs.clear();
    int n = static_cast<int>(keys.size()) < max_kernels ? static_cast<int>(keys.size()) : max_kernels;
    int pos = 0;
    for (int k = 0; k < n; ++k) {
        int len = static_cast<int>(keys[k].size()) + 1;
        if (pos + len > names_bytes) return PRT_E_NOMEM;
        std::memcpy(names + pos, keys[k].c_str(), static_cast<size_t>(len));
        pos += len;
        counts[k] = cnt[k];
        total_ms[k] = ms[k];
    }
    return n;
}
