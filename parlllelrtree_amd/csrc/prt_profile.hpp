// prt_profile.hpp -- scope object that times one kernel launch with HIP
// events when profiling is enabled (see profile.cpp).
#pragma once

#include <hip/hip_ext.h>
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

// While profiling, holds the start and stop events of one launch; they are
// handed to hipExtLaunchKernel, which stamps them with the kernel's own
// start and end on the GPU (not with the times of markers queued around it).
class ProfScope {
  public:
    explicit ProfScope(const char *name);
    ~ProfScope();
    ProfScope(const ProfScope &) = delete;
    ProfScope &operator=(const ProfScope &) = delete;
    bool active() const { return a_ != nullptr; }
    hipEvent_t start() const { return a_; }
    hipEvent_t stop() const { return b_; }

  private:
    const char *name_;
    hipEvent_t a_, b_;
};

}  // namespace prt

// Launch `kernel<<<grid, block, lds, stream>>>(args...)`; while profiling,
// through hipExtLaunchKernelGGL with the scope's events, recorded under the
// kernel's name (PRT_LAUNCH) or under `name` (PRT_LAUNCH_AS: template
// instances of one kernel recorded together).
#define PRT_LAUNCH(kernel, grid, block, lds, stream, ...)                              \
    PRT_LAUNCH_AS(#kernel, kernel, grid, block, lds, stream, __VA_ARGS__)
#define PRT_LAUNCH_AS(name, kernel, grid, block, lds, stream, ...)                     \
    do {                                                                               \
        ::prt::ProfScope _ps(name);                                                    \
        if (_ps.active())                                                              \
            hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(block),                     \
                                  static_cast<std::uint32_t>(lds), (stream), _ps.start(), \
                                  _ps.stop(), 0u, __VA_ARGS__);                        \
        else                                                                           \
            kernel<<<(grid), (block), (lds), (stream)>>>(__VA_ARGS__);                 \
    } while (0)
