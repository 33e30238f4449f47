// api.cpp -- error reporting and version of libprt.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>

#include "prt_common.hpp"

namespace prt {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

int hip_fail(hipError_t e, const char *what, const char *file, int line) {
    set_error("HIP error %d (%s) in %s at %s:%d", static_cast<int>(e), hipGetErrorString(e), what,
              file, line);
    return PRT_E_HIP;
}

}  // namespace prt

extern "C" const char *prt_version(void) { return "prt 0.1.0 (gfx950)"; }

extern "C" const char *prt_last_error(void) { return prt::g_err; }
