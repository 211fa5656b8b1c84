// api.cpp — error reporting and small host-only entry points of the C ABI.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <functional>
#include <string>
#include <thread>

#include "prt_internal.h"

namespace prt {

static thread_local std::string g_last_error;

// PRT_LOG_ERRORS=1 also prints every error to stderr with the thread that raised it: errors raised on the
// serving loop's scheduler thread reach callers only as error codes in events.
static const bool g_log_errors = [] {
    const char* v = std::getenv("PRT_LOG_ERRORS");
    return v && v[0] == '1';
}();

void set_error(const std::string& msg) {
    g_last_error = msg;
    if (g_log_errors) {
        std::fprintf(stderr, "[prt error, thread %zx] %s\n", std::hash<std::thread::id>{}(std::this_thread::get_id()),
                     msg.c_str());
        std::fflush(stderr);
    }
}

int fail(int code, const std::string& msg) {
    set_error(msg);
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    set_error(m);
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return PRT_ERR_OUT_OF_MEMORY;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return PRT_ERR_NO_DEVICE;
    return PRT_ERR_DEVICE;
}

}  // namespace prt

extern "C" {

const char* prt_last_error(void) { return prt::g_last_error.c_str(); }

int prt_version(void) { return 10000; }

int prt_device_count(int32_t* count) {
    if (!count) return prt::fail(PRT_ERR_INVALID_ARGUMENT, "NULL argument");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return PRT_OK;
}

// getSamplesForPass, pkg/renderer/progressive.go:98-123
int prt_samples_for_pass(const prt_render_params* p, int32_t pass, int32_t* target) {
    if (!p || !target) return prt::fail(PRT_ERR_INVALID_ARGUMENT, "NULL argument");
    if (pass < 1) return prt::fail(PRT_ERR_INVALID_ARGUMENT, "pass numbers start at 1");
    if (p->max_passes < 1) return prt::fail(PRT_ERR_INVALID_ARGUMENT, "max_passes must be >= 1");
    if (p->max_passes == 1) {
        *target = p->max_samples_per_pixel;
        return PRT_OK;
    }
    if (pass == 1) {
        *target = p->initial_samples;
        return PRT_OK;
    }
    int32_t rem = p->max_samples_per_pixel - p->initial_samples;
    int32_t rp = p->max_passes - 1;
    int32_t per = rem / rp;
    int32_t t = p->initial_samples + (pass - 1) * per;
    if (pass == p->max_passes) t = p->max_samples_per_pixel;
    *target = t;
    return PRT_OK;
}

}  // extern "C"
