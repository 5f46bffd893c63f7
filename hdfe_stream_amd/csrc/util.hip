// Error reporting and device facts for the C ABI.
#include "common.h"

#include <cstring>

namespace hdfe {
static thread_local std::string g_error;
void set_error(const std::string &msg) { g_error = msg; }
const char *get_error() { return g_error.c_str(); }
}  // namespace hdfe

extern "C" {

const char *hdfe_last_error(void) { return hdfe::get_error(); }

int hdfe_version(void) { return 1; }

// Name, free and total memory of the current device; HDFE_STATUS_NODEV
// without one.
int hdfe_device_info(char *name, int len, int64_t *free_bytes, int64_t *total_bytes) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
        hdfe::set_error("no HIP device visible");
        return HDFE_STATUS_NODEV;
    }
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    hipDeviceProp_t p;
    HIP_TRY(hipGetDeviceProperties(&p, dev));
    if (name && len > 0) {
        const char *src = p.name[0] ? p.name : p.gcnArchName;
        if (!src[0]) src = "AMD GPU";
        std::strncpy(name, src, len - 1);
        name[len - 1] = 0;
    }
    size_t f = 0, t = 0;
    HIP_TRY(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = (int64_t)f;
    if (total_bytes) *total_bytes = (int64_t)t;
    return HDFE_STATUS_OK;
}

}  // extern "C"
