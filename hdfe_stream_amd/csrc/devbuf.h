// Device buffers and hipCUB wrappers shared by the pipeline sources.
#pragma once
#include <hipcub/hipcub.hpp>

#include <memory>
#include <mutex>
#include <vector>

#include "common.h"

namespace hdfe {

// Device memory comes from the device's stream-ordered pool
// (hipMallocAsync / hipFreeAsync).  The pool keeps up to kPoolKeep bytes of
// freed memory for reuse instead of returning it at every synchronisation, so
// the many large temporaries of a fit cost no driver allocation after the
// first fit, and a free neither blocks the host nor synchronises the device.
// A buffer is freed on the stream it was allocated on: reuse is ordered after
// the work queued there.  Owners that hand their memory to other streams
// synchronise the device before freeing it.
constexpr uint64_t kPoolKeep = 16ull << 30;

inline hipError_t pool_prepare() {
    static std::mutex mu;
    static std::vector<int> done;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lock(mu);
    if ((int)done.size() > dev && done[dev]) return hipSuccess;
    hipMemPool_t pool;
    if ((e = hipDeviceGetDefaultMemPool(&pool, dev)) != hipSuccess) return e;
    uint64_t keep = kPoolKeep;
    if ((e = hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep)) != hipSuccess)
        return e;
    if ((int)done.size() <= dev) done.resize(dev + 1, 0);
    done[dev] = 1;
    return hipSuccess;
}

// An owning device allocation from the stream-ordered pool.
template <typename T>
struct DevBuf {
    T *p = nullptr;
    int64_t n = 0;
    hipStream_t st = nullptr;  // the stream it was allocated and is freed on
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n), st(o.st) {
        o.p = nullptr;
        o.n = 0;
    }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) {
            reset();
            p = o.p;
            n = o.n;
            st = o.st;
            o.p = nullptr;
            o.n = 0;
        }
        return *this;
    }
    ~DevBuf() { reset(); }
    void reset() {
        if (p) (void)hipFreeAsync(p, st);
        p = nullptr;
        n = 0;
    }
    hipError_t alloc(int64_t count, hipStream_t s) {
        reset();
        hipError_t e = pool_prepare();
        if (e != hipSuccess) return e;
        st = s;
        e = hipMallocAsync((void **)&p, sizeof(T) * (size_t)std::max<int64_t>(count, 1), s);
        if (e != hipSuccess) {
            p = nullptr;
            return e;
        }
        n = count;
        return hipSuccess;
    }
    T *get() const { return p; }
};

#define DEV_ALLOC(buf, count, stream) HIP_TRY((buf).alloc((count), (stream)))

inline int grid1(int64_t n, int threads = 256, int cap = 65536) {
    int64_t g = (n + threads - 1) / threads;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

// Reusable temporary storage for hipCUB calls.
struct Temp {
    DevBuf<char> buf;
    hipError_t ensure(size_t bytes, hipStream_t s) {
        if ((int64_t)bytes <= buf.n) return hipSuccess;
        return buf.alloc((int64_t)bytes + (1 << 20), s);
    }
};

template <typename K, typename V>
inline hipError_t sort_pairs(Temp &t, const K *kin, K *kout, const V *vin, V *vout, int64_t n,
                             int begin_bit, int end_bit, hipStream_t s) {
    size_t bytes = 0;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kin, kout, vin, vout,
                                                      (int)n, begin_bit, end_bit, s);
    if (e != hipSuccess) return e;
    if ((e = t.ensure(bytes, s)) != hipSuccess) return e;
    bytes = (size_t)t.buf.n;
    return hipcub::DeviceRadixSort::SortPairs(t.buf.p, bytes, kin, kout, vin, vout, (int)n,
                                              begin_bit, end_bit, s);
}

template <typename T, typename U>
inline hipError_t inclusive_sum(Temp &t, const T *in, U *out, int64_t n, hipStream_t s) {
    size_t bytes = 0;
    hipError_t e = hipcub::DeviceScan::InclusiveSum(nullptr, bytes, in, out, (int)n, s);
    if (e != hipSuccess) return e;
    if ((e = t.ensure(bytes, s)) != hipSuccess) return e;
    bytes = (size_t)t.buf.n;
    return hipcub::DeviceScan::InclusiveSum(t.buf.p, bytes, in, out, (int)n, s);
}

template <typename T, typename U>
inline hipError_t exclusive_sum(Temp &t, const T *in, U *out, int64_t n, hipStream_t s) {
    size_t bytes = 0;
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, s);
    if (e != hipSuccess) return e;
    if ((e = t.ensure(bytes, s)) != hipSuccess) return e;
    bytes = (size_t)t.buf.n;
    return hipcub::DeviceScan::ExclusiveSum(t.buf.p, bytes, in, out, (int)n, s);
}

template <typename T>
inline hipError_t to_host(std::vector<T> &h, const T *d, int64_t n, hipStream_t s) {
    h.resize((size_t)n);
    if (n == 0) return hipSuccess;
    hipError_t e = hipMemcpyAsync(h.data(), d, sizeof(T) * n, hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

template <typename T>
inline hipError_t scalar_to_host(T *h, const T *d, hipStream_t s) {
    hipError_t e = hipMemcpyAsync(h, d, sizeof(T), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

inline int bits_for(uint64_t maxval) {
    int b = 0;
    while (b < 64 && (maxval >> b) != 0) ++b;
    return b;
}

}  // namespace hdfe

namespace hdfe {
// build.hip: codes of an int64 column in sorted level order.
int factorize_col(hipStream_t s, Temp &tmp, const int64_t *col, int64_t n, DevBuf<int> &code,
                  DevBuf<int64_t> &levels, int64_t &nl);
}  // namespace hdfe
