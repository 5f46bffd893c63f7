// Device buffers and hipCUB wrappers shared by the pipeline sources.
#pragma once
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <map>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "common.h"

namespace hdfe {

// Device buffers come from the library's own cache (big_alloc below), which
// keeps freed blocks for reuse instead of returning them to the driver, so
// the many temporaries of a fit cost no driver allocation after the first
// fit, and a free neither blocks the host nor synchronises the device.  A
// buffer is freed on the stream it was allocated on: reuse is ordered after
// the work queued there.  Owners that hand their memory to other streams
// synchronise the device before freeing it.
//
// Buffers are kept by the library once allocated: a freed block waits in a
// free list with an event recorded on the stream it was freed on, and a later
// allocation that it fits (at most twice its size class, same device) takes
// it, making its own stream wait for that event unless it is the same stream.
// Blocks are sized in powers of two (from 256 bytes) up to 2 MB and in 2 MB
// steps above; classes under 256 KB are cut 64 at a time from one hipMalloc
// (a slab), so a cold start maps few allocations.  The
// stream-ordered pool was measured spending milliseconds inside hipFreeAsync
// for buffers of hundreds of megabytes, and 10-25 us per call for small ones
// (about a hundred per fit); with the cache a repeated fit maps no memory and
// frees none.  Blocks go back to the driver only when an allocation fails for
// want of memory.
struct BigCache {
    struct Block {
        void *p = nullptr;
        size_t bytes = 0;
        hipEvent_t ev = nullptr;
        hipStream_t freed_on = nullptr;
        int dev = 0;
        bool slab = false;  // cut from a slab: never returned alone
    };
    std::mutex mu;
    std::multimap<size_t, Block> free_list;
    std::unordered_map<void *, Block> live;
};

inline BigCache &big_cache() {
    static BigCache *c = new BigCache();  // lives to process exit (used by static owners)
    return *c;
}

inline hipError_t big_release_free(BigCache &c) {
    // caller holds c.mu
    if (c.free_list.empty()) return hipSuccess;
    hipError_t e = hipDeviceSynchronize();
    for (auto it = c.free_list.begin(); it != c.free_list.end();) {
        if (it->second.slab) {
            ++it;
            continue;
        }
        if (it->second.ev) (void)hipEventDestroy(it->second.ev);
        (void)hipFree(it->second.p);
        it = c.free_list.erase(it);
    }
    return e;
}

inline size_t big_class(size_t bytes) {
    const size_t grain = size_t(2) << 20;
    if (bytes >= grain) return (bytes + grain - 1) / grain * grain;
    size_t b = 256;
    while (b < bytes) b <<= 1;
    return b;
}

inline hipError_t big_alloc(void **p, size_t bytes, hipStream_t s) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    BigCache &c = big_cache();
    std::lock_guard<std::mutex> lock(c.mu);
    const size_t cls = big_class(bytes);
    for (auto it = c.free_list.lower_bound(cls);
         it != c.free_list.end() && it->first <= 2 * cls; ++it) {
        if (it->second.dev != dev) continue;
        BigCache::Block b = it->second;
        c.free_list.erase(it);
        // a block never freed yet has no event: nothing to wait for
        if (b.ev && b.freed_on != s && (e = hipStreamWaitEvent(s, b.ev, 0)) != hipSuccess)
            return e;
        c.live[b.p] = b;
        *p = b.p;
        return hipSuccess;
    }
    BigCache::Block b;
    b.bytes = cls;
    b.dev = dev;
    const bool slab = cls < (size_t(256) << 10);
    const size_t total = slab ? 64 * cls : cls;
    e = hipMalloc(&b.p, total);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        if ((e = big_release_free(c)) != hipSuccess) return e;
        e = hipMalloc(&b.p, total);
    }
    if (e != hipSuccess) return e;
    if (slab) {
        // the slab's first block is returned; the other 63 wait in the list
        b.slab = true;
        for (int q = 1; q < 64; ++q) {
            BigCache::Block f = b;
            f.p = (char *)b.p + (size_t)q * cls;
            c.free_list.emplace(cls, f);
        }
    }
    c.live[b.p] = b;
    *p = b.p;
    return hipSuccess;
}

// Returns a cached block to the free list, usable after the work queued on
// s; false when p is not the cache's.
inline bool big_free(void *p, hipStream_t s) {
    BigCache &c = big_cache();
    std::lock_guard<std::mutex> lock(c.mu);
    auto it = c.live.find(p);
    if (it == c.live.end()) return false;
    BigCache::Block b = it->second;
    c.live.erase(it);
    if (!b.ev && hipEventCreateWithFlags(&b.ev, hipEventDisableTiming) != hipSuccess) {
        // no event to order a later reuse by: wait for the block's work now
        (void)hipGetLastError();
        b.ev = nullptr;
        (void)hipStreamSynchronize(s);
    } else {
        (void)hipEventRecord(b.ev, s);
    }
    b.freed_on = s;
    c.free_list.emplace(b.bytes, b);
    return true;
}

// Device memory of any size on stream s, from the cache.
inline hipError_t dev_malloc(void **p, size_t bytes, hipStream_t s) {
    return big_alloc(p, std::max<size_t>(bytes, 1), s);
}

inline void dev_free(void *p, hipStream_t s) {
    if (p && !big_free(p, s)) (void)hipFree(p);  // not the cache's: a plain allocation
}

// Streams of the library's handles (panels, solver matrices, row passes) come
// from a per-device free list: creating a HIP stream was measured at 0.1 to
// 5 ms of host time and destroying one at 0.1 to 0.5 ms, twice per fit.  A
// returned stream is synchronised first, so the next owner finds it idle.
// All are non-blocking streams of the default priority.
struct StreamPool {
    std::mutex mu;
    std::vector<std::pair<int, hipStream_t>> free_list;
};

inline StreamPool &stream_pool() {
    static StreamPool *sp = new StreamPool();  // lives to process exit
    return *sp;
}

inline hipError_t stream_get(hipStream_t *s) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    StreamPool &sp = stream_pool();
    {
        std::lock_guard<std::mutex> lock(sp.mu);
        for (size_t i = sp.free_list.size(); i-- > 0;) {
            if (sp.free_list[i].first != dev) continue;
            *s = sp.free_list[i].second;
            sp.free_list.erase(sp.free_list.begin() + (std::ptrdiff_t)i);
            return hipSuccess;
        }
    }
    return hipStreamCreateWithFlags(s, hipStreamNonBlocking);
}

// Returns s (made on the current device's list when dev < 0) to the pool.
inline void stream_put(hipStream_t s, int dev = -1) {
    if (!s) return;
    if (hipStreamSynchronize(s) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipStreamDestroy(s);
        return;
    }
    if (dev < 0 && hipGetDevice(&dev) != hipSuccess) return;
    StreamPool &sp = stream_pool();
    std::lock_guard<std::mutex> lock(sp.mu);
    sp.free_list.emplace_back(dev, s);
}

// An owning device allocation (dev_malloc: the cache).
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
        dev_free(p, st);
        p = nullptr;
        n = 0;
    }
    hipError_t alloc(int64_t count, hipStream_t s) {
        reset();
        st = s;
        const size_t b = sizeof(T) * (size_t)std::max<int64_t>(count, 1);
        hipError_t e = dev_malloc((void **)&p, b, s);
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
