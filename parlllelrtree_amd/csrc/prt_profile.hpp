// prt_profile.hpp -- scope object that brackets one kernel launch with HIP
// events when profiling is enabled (see profile.cpp).
#pragma once

#include <hip/hip_runtime.h>

namespace prt {

bool profiling();

// While one of these lives, launches on the calling thread record no events
// (used around stream capture, where the events would become graph nodes).
class ProfPause {
  public:
    ProfPause();
    ~ProfPause();
    ProfPause(const ProfPause &) = delete;
    ProfPause &operator=(const ProfPause &) = delete;
};

class ProfScope {
  public:
    ProfScope(const char *name, hipStream_t s);
    ~ProfScope();
    ProfScope(const ProfScope &) = delete;
    ProfScope &operator=(const ProfScope &) = delete;

  private:
    const char *name_;
    hipStream_t s_;
    void *a_;
};

}  // namespace prt

// Launch `kernel<<<grid, block, lds, stream>>>(args...)` inside a profiling
// scope named after the kernel.
#define PRT_LAUNCH(kernel, grid, block, lds, stream, ...)                       \
    do {                                                                        \
        ::prt::ProfScope _ps(#kernel, (stream));                                \
        kernel<<<(grid), (block), (lds), (stream)>>>(__VA_ARGS__);              \
    } while (0)
