// Shared helpers for the hdfe-stream HIP library (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string>

namespace hdfe {

// Last error text, per thread, for hdfe_last_error().
void set_error(const std::string &msg);
const char *get_error();

#define HDFE_STATUS_OK 0
#define HDFE_STATUS_ARG 1
#define HDFE_STATUS_HIP 2
#define HDFE_STATUS_NOMEM 3
#define HDFE_STATUS_NODEV 4

// Evaluate a HIP call; on failure record the error and return HDFE_STATUS_HIP
// (HDFE_STATUS_NOMEM for allocation failures) from the enclosing function.
#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t _e = (expr);                                                         \
        if (_e != hipSuccess) {                                                         \
            hdfe::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));         \
            return _e == hipErrorOutOfMemory ? HDFE_STATUS_NOMEM : HDFE_STATUS_HIP;     \
        }                                                                               \
    } while (0)

#define ARG_CHECK(cond, msg)                                                            \
    do {                                                                                \
        if (!(cond)) {                                                                  \
            hdfe::set_error(msg);                                                       \
            return HDFE_STATUS_ARG;                                                     \
        }                                                                               \
    } while (0)

// Propagate a non-OK status from a call that returns one.
#define TRY(expr)                              \
    do {                                       \
        int _st = (expr);                      \
        if (_st != HDFE_STATUS_OK) return _st; \
    } while (0)

constexpr int kWave = 64;

// Sum over the 64 lanes of a wavefront (butterfly: every lane gets the sum).
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Sum over aligned groups of G lanes (G a power of two <= 64).
template <int G, typename T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Block-wide sum of one value per thread, fixed tree (deterministic).
// `lds` must hold blockDim.x / 64 values.  Result valid in every thread.
template <typename T>
__device__ __forceinline__ T block_sum(T v, T *lds) {
    v = wave_sum(v);
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    __syncthreads();
    if (lane == 0) lds[w] = v;
    __syncthreads();
    T s = T(0);
    for (int i = 0; i < nw; ++i) s += lds[i];
    return s;
}

// Runs of equal keys among consecutive lanes of a wave (rows sorted by group
// repeat a level over neighbouring rows): the sum of v over the caller's run,
// valid in the run's last lane (`last` is true there), so that one atomic per
// run replaces one per lane.  Every lane of the wave calls it; a lane with
// nothing to add passes v = 0, and one past the end a key no row has.
template <typename T>
__device__ __forceinline__ T run_total(int key, T v, bool &last) {
    const int lane = threadIdx.x & (kWave - 1);
    const int up = __shfl_up(key, 1, kWave);
    const bool head = lane == 0 || up != key;
    const unsigned long long hm = __ballot(head);
    T p = v;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const T t = __shfl_up(p, d, kWave);
        if (lane >= d) p += t;
    }
    last = lane == kWave - 1 || ((hm >> (lane + 1)) & 1ull);
    const unsigned long long mine = lane == kWave - 1 ? hm : (hm & ((2ull << lane) - 1ull));
    const int hl = 63 - __clzll((long long)mine);
    const T before = __shfl(p, hl > 0 ? hl - 1 : 0, kWave);
    return hl > 0 ? p - before : p;
}

// Lock-free union-find over int parent links with parent[x] <= x: the root
// of a set is its smallest member whatever the order of the unions.
// uf_root halves the path it walks (a CAS, so a link only moves up).
__device__ __forceinline__ int uf_root(int *parent, int x) {
    while (true) {
        int p = parent[x];
        if (p == x) return x;
        int gp = parent[p];
        if (gp != p) atomicCAS(&parent[x], p, gp);
        x = p;
    }
}

// Joins the sets of x and y: hooks the larger root under the smaller with a
// CAS, retrying from the values a failed CAS returns until both walks meet
// (each retry lowers the larger root, so the loop ends).
__device__ __forceinline__ void uf_join(int *parent, int x, int y) {
    int a = uf_root(parent, x), b = uf_root(parent, y);
    while (a != b) {
        if (a > b) {
            int t = a;
            a = b;
            b = t;
        }
        int old = atomicCAS(&parent[b], b, a);
        if (old == b) break;
        b = uf_root(parent, old);
        a = uf_root(parent, a);
    }
}

// Row range of part `w` of `nparts` equal contiguous parts of [0, n).
__host__ __device__ __forceinline__ int64_t part_begin(int64_t n, int64_t w, int64_t nparts) {
    return (n * w) / nparts;
}

}  // namespace hdfe
