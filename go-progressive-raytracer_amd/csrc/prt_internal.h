// prt_internal.h — host-side objects behind the opaque C handles.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/prt.h"
#include "prt_common.h"

namespace prt {

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define PRT_HIP(call)                                      \
    do {                                                   \
        hipError_t _e = (call);                            \
        if (_e != hipSuccess) return hip_fail(_e, #call);  \
    } while (0)

// Device allocation owned by a scene or film.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

}  // namespace prt

struct prt_scene {
    int device = 0;
    prt::DevScene dev{};
    prt_scene_info info{};
    std::vector<prt::DevBuf> bufs;
    uint2* gstack = nullptr;     // traversal stack overflow for the intersection kernels
    uint32_t gstack_slots = 0;   // thread slots covered by gstack
    int32_t max_stack = 0;
    int32_t num_cus = 0;
    std::vector<uint32_t> orig_id;  // host copy: leaf order -> original primitive id
    // float64 camera/scene facts needed by the host-side renderer
    int32_t image_width = 0, image_height = 0;
    double aspect = 1.0;
};

namespace prt {
// Allocates device memory owned by `s` and copies `bytes` from host (if src != nullptr).
int scene_alloc(prt_scene* s, size_t bytes, const void* src, void** out);
// Grid size used by persistent kernels (work-group count).
int persistent_blocks(const prt_scene* s, int block, int per_cu);
}  // namespace prt
