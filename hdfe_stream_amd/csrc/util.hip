// Error reporting, device facts and the device-memory cache's release for
// the C ABI.
#include "common.h"
#include "devbuf.h"

#include <cstring>

namespace hdfe {
static thread_local std::string g_error;

// A plain streaming copy, 16 bytes per lane per step: the calibration
// kernel for the HBM byte counters (its traffic is known exactly).
__global__ void k_copy_calibrate(int64_t n16, const uint4 *__restrict__ src,
                                 uint4 *__restrict__ dst) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n16;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}
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

// Returns the cache's free blocks to the driver (slabs of small blocks
// stay); *freed: bytes returned.  Synchronises the device first.
int hdfe_cache_release(int64_t *freed) {
    hdfe::BigCache &c = hdfe::big_cache();
    std::lock_guard<std::mutex> lock(c.mu);
    int64_t b = 0;
    for (const auto &kv : c.free_list)
        if (!kv.second.slab) b += (int64_t)kv.second.bytes;
    HIP_TRY(hdfe::big_release_free(c));
    if (freed) *freed = b;
    return HDFE_STATUS_OK;
}

// Copies `bytes` (a multiple of 16) between two fresh device buffers `reps`
// times on the null stream; *avg_us: the mean launch duration from HIP
// events.  Reads `bytes` and writes `bytes` per launch.
int hdfe_copy_calibrate(int64_t bytes, int reps, double *avg_us) {
    ARG_CHECK(bytes > 0 && bytes % 16 == 0 && reps >= 1 && avg_us,
              "hdfe_copy_calibrate: bytes must be a positive multiple of 16");
    void *a = nullptr, *b = nullptr;
    HIP_TRY(hipMalloc(&a, bytes));
    if (hipMalloc(&b, bytes) != hipSuccess) {
        (void)hipFree(a);
        hdfe::set_error("hdfe_copy_calibrate: out of device memory");
        return HDFE_STATUS_NOMEM;
    }
    const int64_t n16 = bytes / 16;
    (void)hipMemset(a, 1, bytes);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const int grid = 256 * 8 * 4;  // 4 workgroups of 256 per CU on 256 CUs, grid-stride
    hipLaunchKernelGGL(hdfe::k_copy_calibrate, dim3(grid), dim3(256), 0, 0, n16,
                       (const uint4 *)a, (uint4 *)b);
    (void)hipEventRecord(e0, 0);
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(hdfe::k_copy_calibrate, dim3(grid), dim3(256), 0, 0, n16,
                           (const uint4 *)a, (uint4 *)b);
    (void)hipEventRecord(e1, 0);
    hipError_t err = hipEventSynchronize(e1);
    float ms = 0.f;
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    HIP_TRY(err);
    *avg_us = 1e3 * ms / reps;
    return HDFE_STATUS_OK;
}

}  // extern "C"
