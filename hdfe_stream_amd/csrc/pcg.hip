// Jacobi-preconditioned block CG on an explicit CSR matrix, for gfx950.
//
// Replaces the reference's device solve (core/gpu.py:186-253, explicit_solver
// and pcg) and its CPU twin (core/solve.py:66-140, _SolveMixin._pcg).  The
// algorithm is the reference's, column by column in lockstep: zero columns are
// solved by 0, a warm start x0 may already solve a column, each column stops
// when ||r|| <= tol ||b||, alpha is guarded by pAp > 0 and beta by rz != 0.
//
// Layout: every (L, k) block is row major, the reference's C order, so the
// gather of one neighbour row in the SpMM reads k contiguous values.
//
// One iteration is three or four launches:
//   [long rows]  rows longer than kLongThresh (the levels of small dimensions,
//                e.g. years, which touch every firm) are split in chunks of
//                kLongChunk nonzeros over many workgroups; the last workgroup
//                adds the chunks of each row in a fixed order.
//   spmm_dot     AP = S P for the other rows, with the pAp column sums fused;
//                the last workgroup to finish forms alpha.
//   step         X += alpha P, R -= alpha AP, with the sums of r^2 and r^2 dinv;
//                the last workgroup forms rel, the active mask, beta, and the
//                stop flag.
//   direction    P = dinv R + beta P.
// Column sums are per-workgroup partials over fixed row ranges added in a
// fixed order by the last workgroup, so results do not depend on scheduling.
// The host runs iterations in batches and reads the stop flag once per batch;
// kernels launched after the stop return at once.
//
// precision 1 ("mixed"): the CG above runs in FP32 (values, vectors, SpMM)
// with FP64 column sums, as the inner solver of an iterative refinement whose
// residual b - S x is recomputed in FP64 with the FP64 matrix; columns stop
// when that true FP64 residual meets tol.
#include "common.h"

#include <climits>
#include <cmath>
#include <algorithm>
#include <array>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "devbuf.h"
#include "solver.h"

namespace hdfe {

constexpr int kThreads = 256;
constexpr int kLongThresh = 4096;  // nonzeros: rows above this are split
constexpr int kLongChunk = 2048;   // nonzeros per workgroup of a split row
constexpr int kMaxK = 1024;        // widest block of right-hand sides per call
constexpr int kSmallK = 8;         // k <= 8: lanes over nonzeros; wider: over columns
constexpr int64_t kL2Budget = 3 << 20;  // bytes of P an XCD's 4 MB L2 should hold (3 MB)
constexpr int kBlkParts = 128;     // row parts per XCD of the blocked SpMM
constexpr int kDenseLongMax = 16;  // split rows the blocked SpMM takes as dense columns
constexpr int kNB = 16;            // column blocks of the blocked layout
constexpr int kCombineMax = 512;   // workgroups of the blocked SpMM's combine

// Device-side solver state (one allocation; see State layout in Work).
struct CgState {
    double *rz, *alpha, *beta, *bnorm, *rel;  // (k) each
    int *active;                              // (k)
    int *done, *it, *it_final;                // scalars
    unsigned *counter;                        // unused (kept for the layout)
    double *dtol;                             // tol, on the device (k_set_params)
    int *dmaxiter;                            // maxiter, on the device
    int k, maxiter;
    double tol;
};

// ----------------------------------------------------------------- kernels

constexpr int kFinThreads = 1024;  // threads of the one-workgroup finishing kernels
// CG batches are captured as graphs only below this many bytes moved per
// iteration (about 16 us of device time; six launches cost ~20 us of host time)
constexpr int64_t kGraphMaxBytes = int64_t(48) << 20;

// out[j] = the sum over w < nw of part[w * stride + j], for every column
// j < k, by the whole workgroup: TJ threads across columns and TR = blockDim
// / TJ across the partials (two accumulators each), then a fixed tree in LDS.
// out may be shared or global memory.
__device__ void reduce_parts(const double *part, int nw, int k, int stride, double *out) {
    __shared__ double red[kFinThreads];
    int tj = 1;
    while (tj < k && tj < (int)blockDim.x) tj <<= 1;
    const int TR = blockDim.x / tj;
    const int c = threadIdx.x % tj, r = threadIdx.x / tj;
    for (int j0 = 0; j0 < k; j0 += tj) {
        const int j = j0 + c;
        double a0 = 0.0, a1 = 0.0;
        if (j < k) {
            int w = r;
            for (; w + TR < nw; w += 2 * TR) {
                a0 += part[(int64_t)w * stride + j];
                a1 += part[(int64_t)(w + TR) * stride + j];
            }
            if (w < nw) a0 += part[(int64_t)w * stride + j];
        }
        __syncthreads();
        red[r * tj + c] = a0 + a1;
        __syncthreads();
        for (int h = TR / 2; h > 0; h >>= 1) {
            if (r < h) red[r * tj + c] += red[(r + h) * tj + c];
            __syncthreads();
        }
        if (r == 0 && j < k) out[j] = red[c];
        __syncthreads();
    }
}

// Split rows, wide k: waves over nonzeros, lanes over columns.  I is the
// row pointers' type (int, or int64_t for 2^31 nonzeros and more), here and
// in the other plain CSR kernels.
template <typename T, typename I>
__global__ __launch_bounds__(kThreads) void k_long_wide(
    const I *__restrict__ indptr, const int *__restrict__ indices, const T *__restrict__ val,
    const int *__restrict__ chunk_row, const I *__restrict__ chunk_beg, int nchunk,
    const int *__restrict__ long_rows, const int *__restrict__ long_cptr, int nlong,
    const T *__restrict__ P, T *__restrict__ AP, double *__restrict__ lpart, CgState st) {
    if (st.done && *st.done) return;
    extern __shared__ double wlds[];  // [4][k]
    const int k = st.k;
    const int c = blockIdx.x;
    const int row = chunk_row[c];
    const I s = chunk_beg[c], e = min(s + (I)kLongChunk, indptr[row + 1]);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int j0 = 0; j0 < k; j0 += 64) {
        const int j = j0 + lane;
        T acc = T(0);
        if (j < k)
            for (I p = s + wv; p < e; p += 4) acc += val[p] * P[(int64_t)indices[p] * k + j];
        if (j < k) wlds[wv * k + j] = (double)acc;
    }
    __syncthreads();
    for (int j = threadIdx.x; j < k; j += blockDim.x)
        lpart[(int64_t)c * k + j] = wlds[j] + wlds[k + j] + wlds[2 * k + j] + wlds[3 * k + j];
}

// One workgroup, after k_long_wide: each split row's AP from its chunk
// partials in chunk order.
template <typename T>
__global__ __launch_bounds__(kFinThreads) void k_fin_long_wide(
    const int *__restrict__ long_rows, const int *__restrict__ long_cptr, int nlong,
    const double *__restrict__ lpart, T *__restrict__ AP, CgState st) {
    if (st.done && *st.done) return;
    const int k = st.k;
    for (int t = threadIdx.x; t < nlong * k; t += blockDim.x) {
        const int l = t / k, j = t % k;
        double s2 = 0.0;
        for (int q = long_cptr[l]; q < long_cptr[l + 1]; ++q) s2 += lpart[(int64_t)q * k + j];
        AP[(int64_t)long_rows[l] * k + j] = (T)s2;
    }
}

// Split rows, wide k, when they are dense (every level of a small dimension
// -- years -- couples to most others): AP[long t] = sum_r dw[r, t] P[r], with
// dw (L, nlong) the split rows' values by column, row major.  Workgroup (b, c)
// takes row block b of kDwBlocks and the 64 columns of tile c: each wave a
// stride of the block's rows, each lane one column, every split row's sum in
// registers, so P is read once for all split rows (k_long_wide gathers the
// whole of P once per split row).  The four waves' sums are added in LDS in
// wave order into dwpart[(b nlong + t) k + j].
constexpr int kDwBlocks = 256;
template <typename T>
__global__ __launch_bounds__(kThreads) void k_dense_long_wide(
    int64_t L, int nlong, const T *__restrict__ dw, const T *__restrict__ P,
    double *__restrict__ dwpart, CgState st) {
    if (st.done && *st.done) return;
    __shared__ double red[kThreads / kWave][kDenseLongMax][kWave];
    const int k = st.k;
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const int b = blockIdx.x, j = blockIdx.y * kWave + lane;
    const int64_t r0 = part_begin(L, b, kDwBlocks), r1 = part_begin(L, b + 1, kDwBlocks);
    double acc[kDenseLongMax];
#pragma unroll
    for (int t = 0; t < kDenseLongMax; ++t) acc[t] = 0.0;
    if (j < k) {
        for (int64_t r = r0 + wv; r < r1; r += kThreads / kWave) {
            const double pv = (double)P[r * k + j];
#pragma unroll
            for (int t = 0; t < kDenseLongMax; ++t)
                if (t < nlong) acc[t] += (double)dw[r * nlong + t] * pv;
        }
    }
#pragma unroll
    for (int t = 0; t < kDenseLongMax; ++t)
        if (t < nlong) red[wv][t][lane] = acc[t];
    __syncthreads();
    if (wv == 0 && j < k)
        for (int t = 0; t < nlong; ++t)
            dwpart[((int64_t)b * nlong + t) * k + j] =
                ((red[0][t][lane] + red[1][t][lane]) + red[2][t][lane]) + red[3][t][lane];
}

// After k_dense_long_wide: each split row's AP from the row blocks' sums in
// block order, a thread per (split row, column).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_fin_dense_long(
    int nlong, const int *__restrict__ long_rows, const double *__restrict__ dwpart,
    T *__restrict__ AP, CgState st) {
    if (st.done && *st.done) return;
    const int k = st.k;
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= (int64_t)nlong * k) return;
    const int t = (int)(q / k), j = (int)(q % k);
    double s0 = 0.0, s1 = 0.0;
    int b = 0;
    for (; b + 1 < kDwBlocks; b += 2) {
        s0 += dwpart[((int64_t)b * nlong + t) * k + j];
        s1 += dwpart[((int64_t)(b + 1) * nlong + t) * k + j];
    }
    for (; b < kDwBlocks; ++b) s0 += dwpart[((int64_t)b * nlong + t) * k + j];
    AP[(int64_t)long_rows[t] * k + j] = (T)(s0 + s1);
}

// The dense split rows' values by column: dw[col, t] = S[long t, col].
__global__ void k_dw_fill(int nlong, int64_t L, const int *__restrict__ long_rows,
                          const int *__restrict__ indptr, const int *__restrict__ indices,
                          const double *__restrict__ v64, const float *__restrict__ v32,
                          double *__restrict__ dw64, float *__restrict__ dw32) {
    const int t = blockIdx.y;
    const int row = long_rows[t];
    const int s = indptr[row], e = indptr[row + 1];
    for (int64_t p = s + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < e;
         p += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = indices[p];
        dw64[c * nlong + t] = v64[p];
        dw32[c * nlong + t] = v32[p];
    }
}

// One workgroup, after a wide SpMM: alpha = rz / pAp where active and
// pAp > 0, else 0 (core/solve.py:120), from the nw workgroups' pAp sums.
__global__ __launch_bounds__(kFinThreads) void k_fin_alpha(const double *__restrict__ part, int nw,
                                                        CgState st) {
    if (st.done && *st.done) return;
    __shared__ double tmp[kMaxK];
    reduce_parts(part, nw, st.k, st.k, tmp);
    __syncthreads();
    for (int j = threadIdx.x; j < st.k; j += blockDim.x) {
        const double pAp = tmp[j];
        st.alpha[j] = (st.active[j] && pAp > 0.0) ? st.rz[j] / pAp : 0.0;
    }
}

// One workgroup after a small-k SpMM (k_fin_small): each split row's AP from
// its chunk partials (a wave per row and column: lanes over the chunks, then
// the wave's fixed tree) and its P AP; with with_dot, the pAp sums of the nw
// workgroups and alpha = rz / pAp where active and pAp > 0, else 0
// (core/solve.py:120).
// With ll (the split rows' block among themselves, nlong x nlong) the
// partials hold only the split rows' products with the other rows, and
// sum_t' ll[l, t'] P[long t'] is added last.
// Every load is issued up front -- the chunk partials, each thread's share of
// the pAp partials, and (up to kFinStage split rows) the split rows' P and ll,
// held in registers until the partials are in -- so the kernel waits on
// memory about twice, not once per split row (loops over the split rows that
// each chased long_rows[u] to P cost ~10 us per product at eight rows).
constexpr int kFinStage = 64;
template <typename T, int K>
__device__ void spmm_finish(const T *__restrict__ P, T *__restrict__ AP,
                            const double *__restrict__ part, int nw, int with_dot,
                            const int *__restrict__ long_rows, const int *__restrict__ long_cptr,
                            int nlong, const double *__restrict__ lpart, CgState st,
                            const T *__restrict__ ll = nullptr) {
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1), wv = tid / kWave;
    const int nwv = blockDim.x / kWave;
    const bool staged = nlong <= kFinStage;
    const bool lls = ll && nlong <= kDenseLongMax;
    __shared__ double spl[kFinStage * K], sap[kFinStage * K];
    __shared__ double sll[kDenseLongMax * kDenseLongMax];
    // the split rows' P and ll, into registers first
    double plv = 0.0, llv = 0.0;
    if (staged && tid < nlong * K) plv = (double)P[(int64_t)long_rows[tid / K] * K + tid % K];
    if (lls && tid < nlong * nlong) llv = (double)ll[tid];
    // the pAp partials: thread (r, c) sums rows r, r + TR, ... of column c
    constexpr int TJ = K <= 1 ? 1 : (K <= 2 ? 2 : (K <= 4 ? 4 : 8));
    const int TR = blockDim.x / TJ, pc = tid % TJ, pr = tid / TJ;
    double pa0 = 0.0, pa1 = 0.0, pa2 = 0.0, pa3 = 0.0;
    if (with_dot && pc < K) {
        int w = pr;
        for (; w + 3 * TR < nw; w += 4 * TR) {
            pa0 += part[(int64_t)w * K + pc];
            pa1 += part[(int64_t)(w + TR) * K + pc];
            pa2 += part[(int64_t)(w + 2 * TR) * K + pc];
            pa3 += part[(int64_t)(w + 3 * TR) * K + pc];
        }
        for (; w < nw; w += TR) pa0 += part[(int64_t)w * K + pc];
    }
    // split row l's partials (its range of partial rows, each K values) in
    // H stretches, one wave each: every lane adds whole partial rows (all K
    // values in one contiguous load), four in flight; the stretches' sums
    // then in order
    __shared__ double hsum[kFinThreads / kWave][K];
    // more split rows than waves (only without ll): a wave per row, written
    // directly
    const bool direct = nlong > nwv;
    const int H = nlong > 0 && !direct ? max(1, min(nwv / nlong, 4)) : 1;
    for (int q = wv; q < nlong * H; q += nwv) {
        const int l = q / H, h = q % H;
        // without long_cptr, split row l's partials are [l nw, (l + 1) nw)
        const int cb = long_cptr ? long_cptr[l] : l * nw;
        const int ce = long_cptr ? long_cptr[l + 1] : (l + 1) * nw;
        const int c1 = cb + (int)((int64_t)(ce - cb) * (h + 1) / H);
        int c = cb + (int)((int64_t)(ce - cb) * h / H) + lane;
        double a0[K], a1[K], a2[K], a3[K];
#pragma unroll
        for (int j = 0; j < K; ++j) a0[j] = a1[j] = a2[j] = a3[j] = 0.0;
        for (; c + 3 * kWave < c1; c += 4 * kWave) {
#pragma unroll
            for (int j = 0; j < K; ++j) {
                a0[j] += lpart[(int64_t)c * K + j];
                a1[j] += lpart[(int64_t)(c + kWave) * K + j];
                a2[j] += lpart[(int64_t)(c + 2 * kWave) * K + j];
                a3[j] += lpart[(int64_t)(c + 3 * kWave) * K + j];
            }
        }
        if (c < c1)
#pragma unroll
            for (int j = 0; j < K; ++j) a0[j] += lpart[(int64_t)c * K + j];
        if (c + kWave < c1)
#pragma unroll
            for (int j = 0; j < K; ++j) a1[j] += lpart[(int64_t)(c + kWave) * K + j];
        if (c + 2 * kWave < c1)
#pragma unroll
            for (int j = 0; j < K; ++j) a2[j] += lpart[(int64_t)(c + 2 * kWave) * K + j];
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const double v = wave_sum((a0[j] + a1[j]) + (a2[j] + a3[j]));
            if (lane == 0) {
                if (!direct) {
                    hsum[q][j] = v;
                } else {
                    double s2 = v;
                    if (ll)
                        for (int u = 0; u < nlong; ++u)
                            s2 += (double)ll[l * nlong + u] *
                                  (double)P[(int64_t)long_rows[u] * K + j];
                    AP[(int64_t)long_rows[l] * K + j] = (T)s2;
                    if (staged) sap[l * K + j] = (double)(T)s2;
                }
            }
        }
    }
    if (staged && tid < nlong * K) spl[tid] = plv;
    if (lls && tid < nlong * nlong) sll[tid] = llv;
    __shared__ double red[kFinThreads];
    red[tid] = (pa0 + pa1) + (pa2 + pa3);
    __syncthreads();
    for (int t = tid; t < (direct ? 0 : nlong * K); t += blockDim.x) {
        const int l = t / K, j = t % K;
        double s2 = 0.0;
        for (int h = 0; h < H; ++h) s2 += hsum[l * H + h][j];
        if (lls) {
            for (int u = 0; u < nlong; ++u) s2 += sll[l * nlong + u] * spl[u * K + j];
        } else if (ll) {
            for (int u = 0; u < nlong; ++u)
                s2 += (double)ll[l * nlong + u] * (double)P[(int64_t)long_rows[u] * K + j];
        }
        AP[(int64_t)long_rows[l] * K + j] = (T)s2;
        if (staged) sap[t] = (double)(T)s2;
    }
    if (!with_dot) return;
    // the pAp partials' tree over the TR row lanes of each column
    for (int h = TR / 2; h > 0; h >>= 1) {
        __syncthreads();
        if (pr < h) red[pr * TJ + pc] += red[(pr + h) * TJ + pc];
    }
    __syncthreads();  // the split rows' AP, written above, are visible here
    if (tid < K) {
        const int j = tid;
        double d = red[j];
        if (staged) {
            for (int l = 0; l < nlong; ++l) d += spl[l * K + j] * sap[l * K + j];
        } else {
            for (int l = 0; l < nlong; ++l) {
                const int64_t row = long_rows[l];
                d += (double)P[row * K + j] * (double)AP[row * K + j];
            }
        }
        st.alpha[j] = (st.active[j] && d > 0.0) ? st.rz[j] / d : 0.0;
    }
}

template <typename T, int K>
__global__ __launch_bounds__(kFinThreads) void k_fin_small(
    const T *__restrict__ P, T *__restrict__ AP, const double *__restrict__ part, int nw,
    int with_dot, const int *__restrict__ long_rows, const int *__restrict__ long_cptr, int nlong,
    const double *__restrict__ lpart, CgState st, const T *__restrict__ ll = nullptr) {
    if (st.done && *st.done) return;
    spmm_finish<T, K>(P, AP, part, nw, with_dot, long_rows, long_cptr, nlong, lpart, st, ll);
}

// AP = S P with, when with_dot, part[w, j] = sum over the workgroup's rows of
// P AP and alpha formed by the last workgroup.  One launch: workgroups
// [0, nchunk) take the chunks of the split rows (rows over kLongThresh
// nonzeros: the levels of small dimensions), the others ranges of ordinary
// rows with G lanes per row.  The last workgroup to finish adds each split
// row's chunk partials in a fixed order, adds their P AP, and forms alpha.
template <typename T, int K, int G, typename I>
__global__ __launch_bounds__(kThreads) void k_spmm_small(
    const I *__restrict__ indptr, const int *__restrict__ indices, const T *__restrict__ val,
    int64_t L, const T *__restrict__ P, T *__restrict__ AP, double *__restrict__ part,
    int with_dot, const int *__restrict__ chunk_row, const I *__restrict__ chunk_beg,
    int nchunk, const int *__restrict__ long_rows, const int *__restrict__ long_cptr, int nlong,
    double *__restrict__ lpart, CgState st) {
    if (st.done && *st.done) return;
    __shared__ double lds[kThreads / kWave];
    const int nw = gridDim.x - nchunk;
    double dot[K];
#pragma unroll
    for (int j = 0; j < K; ++j) dot[j] = 0.0;
    if ((int)blockIdx.x < nchunk) {
        const int c = blockIdx.x;
        const int row = chunk_row[c];
        const I s = chunk_beg[c], e = min(s + (I)kLongChunk, indptr[row + 1]);
        T acc[K];
#pragma unroll
        for (int j = 0; j < K; ++j) acc[j] = T(0);
        for (I p = s + threadIdx.x; p < e; p += blockDim.x) {
            const int col = indices[p];
            const T v = val[p];
#pragma unroll
            for (int j = 0; j < K; ++j) acc[j] += v * P[(int64_t)col * K + j];
        }
#pragma unroll
        for (int j = 0; j < K; ++j) {
            double t = block_sum<double>((double)acc[j], lds);
            if (threadIdx.x == 0) lpart[(int64_t)c * K + j] = t;
        }
    } else {
        const int w = blockIdx.x - nchunk;
        const int64_t r0 = part_begin(L, w, nw), r1 = part_begin(L, w + 1, nw);
        const int sub = threadIdx.x % G, slot = threadIdx.x / G, nslot = blockDim.x / G;
        for (int64_t base = r0; base < r1; base += nslot) {
            const int64_t row = base + slot;
            const bool ok = row < r1;
            I s = 0, e = 0;
            if (ok) {
                s = indptr[row];
                e = indptr[row + 1];
            }
            const bool is_long = (e - s) > kLongThresh;
            T acc[K];
#pragma unroll
            for (int j = 0; j < K; ++j) acc[j] = T(0);
            if (ok && !is_long) {
                // four nonzeros per lane in flight: their column indices
                // first, then all their P gathers
                I p = s + sub;
                for (; p + 3 * G < e; p += 4 * G) {
                    const int c0 = indices[p], c1 = indices[p + G];
                    const int c2 = indices[p + 2 * G], c3 = indices[p + 3 * G];
                    const T v0 = val[p], v1 = val[p + G], v2 = val[p + 2 * G], v3 = val[p + 3 * G];
#pragma unroll
                    for (int j = 0; j < K; ++j)
                        acc[j] += v0 * P[(int64_t)c0 * K + j] + v1 * P[(int64_t)c1 * K + j] +
                                  v2 * P[(int64_t)c2 * K + j] + v3 * P[(int64_t)c3 * K + j];
                }
                for (; p < e; p += G) {
                    const int col = indices[p];
                    const T v = val[p];
#pragma unroll
                    for (int j = 0; j < K; ++j) acc[j] += v * P[(int64_t)col * K + j];
                }
            }
#pragma unroll
            for (int j = 0; j < K; ++j) acc[j] = group_sum<G>(acc[j]);
            if (ok && sub == 0 && !is_long) {
#pragma unroll
                for (int j = 0; j < K; ++j) {
                    AP[row * K + j] = acc[j];
                    dot[j] += (double)P[row * K + j] * (double)acc[j];
                }
            }
        }
        if (with_dot) {
#pragma unroll
            for (int j = 0; j < K; ++j) {
                double t = block_sum<double>(dot[j], lds);
                if (threadIdx.x == 0) part[(int64_t)w * K + j] = t;
            }
        }
    }
}

// Blocked small-k SpMM, launch 1 of 2: the non-split rows' nonzeros are
// stored again by column block (kNB contiguous column ranges of equal
// nonzero counts, block-major, rows ascending within a block, the row
// pointers of all blocks in one array bptr (kNB L + 1)).  Workgroup w < nblk
// runs on XCD w % 8 (the dispatcher deals workgroups to the eight XCDs in
// turn) and takes one group of ngrp groups of consecutive blocks -- the
// XCD's group is xcd % ngrp -- over one part of the rows, so the rows of P
// that an XCD gathers are its group's slice, which stays in that XCD's L2.
// It writes that group's partial AP (L, K) for its rows.  Workgroups nblk..
// take the split rows' chunks as k_spmm_small does.
template <typename T, int K, int G>
__global__ __launch_bounds__(kThreads) void k_spmm_blk(
    const int *__restrict__ bptr, const int *__restrict__ bidx, const T *__restrict__ bval,
    int64_t L, int ngrp, int nblk, const T *__restrict__ P, T *__restrict__ gpart,
    const int *__restrict__ indptr, const int *__restrict__ indices, const T *__restrict__ val,
    const int *__restrict__ chunk_row, const int *__restrict__ chunk_beg,
    double *__restrict__ lpart, CgState st) {
    if (st.done && *st.done) return;
    __shared__ double lds[kThreads / kWave];
    const int w = blockIdx.x;
    if (w >= nblk) {
        const int c = w - nblk;
        const int row = chunk_row[c];
        const int s = chunk_beg[c], e = min(s + kLongChunk, indptr[row + 1]);
        T acc[K];
#pragma unroll
        for (int j = 0; j < K; ++j) acc[j] = T(0);
        for (int p = s + threadIdx.x; p < e; p += blockDim.x) {
            const int col = indices[p];
            const T v = val[p];
#pragma unroll
            for (int j = 0; j < K; ++j) acc[j] += v * P[(int64_t)col * K + j];
        }
#pragma unroll
        for (int j = 0; j < K; ++j) {
            double t = block_sum<double>((double)acc[j], lds);
            if (threadIdx.x == 0) lpart[(int64_t)c * K + j] = t;
        }
        return;
    }
    const int xcd = w & 7;
    const int grp = xcd % ngrp;
    const int share = 8 / ngrp;               // XCDs working on one group
    const int nparts = (nblk / 8) * share;    // row parts of one group
    const int rp = (w >> 3) * share + xcd / ngrp;
    const int64_t r0 = part_begin(L, rp, nparts), r1 = part_begin(L, rp + 1, nparts);
    const int b0 = grp * (kNB / ngrp), b1 = b0 + kNB / ngrp;
    T *out = gpart + (int64_t)grp * L * K;
    const int sub = threadIdx.x % G, slot = threadIdx.x / G, nslot = kThreads / G;
    for (int64_t base = r0; base < r1; base += nslot) {
        const int64_t row = base + slot;
        T acc[K];
#pragma unroll
        for (int j = 0; j < K; ++j) acc[j] = T(0);
        if (row < r1) {
            for (int b = b0; b < b1; ++b) {
                const int s = bptr[(int64_t)b * L + row], e = bptr[(int64_t)b * L + row + 1];
                int p = s + sub;
                for (; p + G < e; p += 2 * G) {
                    const int c0 = bidx[p], c1 = bidx[p + G];
                    const T v0 = bval[p], v1 = bval[p + G];
#pragma unroll
                    for (int j = 0; j < K; ++j)
                        acc[j] += v0 * P[(int64_t)c0 * K + j] + v1 * P[(int64_t)c1 * K + j];
                }
                if (p < e) {
                    const int c0 = bidx[p];
                    const T v0 = bval[p];
#pragma unroll
                    for (int j = 0; j < K; ++j) acc[j] += v0 * P[(int64_t)c0 * K + j];
                }
            }
        }
#pragma unroll
        for (int j = 0; j < K; ++j) acc[j] = group_sum<G>(acc[j]);
        if (row < r1 && sub == 0) {
#pragma unroll
            for (int j = 0; j < K; ++j) out[row * K + j] = acc[j];
        }
    }
}

// Blocked small-k SpMM, launch 2 of 2: AP = the sum of the groups' partials
// in group order; with with_dot the pAp sums.  With nld dense long columns
// (the split rows taken by symmetry: dt[t * L + row] = S[row, long t] for the
// other rows, column-major), each row also adds sum_t dt[t, row] P[long t],
// and each wave then forms, for its split rows t, the workgroup's partial of
// AP[long t] = sum_row dt[t, row] P[row] into
// dlpart[(t * nw + w) * K + j] (added in workgroup order by k_fin_small).
template <typename T, int K>
__global__ __launch_bounds__(kThreads) void k_spmm_combine(
    int64_t L, int ngrp, const T *__restrict__ gpart, const T *__restrict__ P, T *__restrict__ AP,
    double *__restrict__ part, int with_dot, int nld, const T *__restrict__ dt,
    const int *__restrict__ long_rows, double *__restrict__ dlpart, CgState st) {
    if (st.done && *st.done) return;
    __shared__ double lds[kThreads / kWave];
    __shared__ T plong[kDenseLongMax * K];
    const int nw = gridDim.x, w = blockIdx.x;
    const int64_t r0 = part_begin(L, w, nw), r1 = part_begin(L, w + 1, nw);
    for (int q = threadIdx.x; q < nld * K; q += blockDim.x)
        plong[q] = P[(int64_t)long_rows[q / K] * K + q % K];
    if (nld) __syncthreads();
    double dot[K];
#pragma unroll
    for (int j = 0; j < K; ++j) dot[j] = 0.0;
    const int64_t LK = L * K;
    for (int64_t row = r0 + threadIdx.x; row < r1; row += blockDim.x) {
        T acc[K];
#pragma unroll
        for (int j = 0; j < K; ++j) acc[j] = gpart[row * K + j];
        for (int g = 1; g < ngrp; ++g) {
#pragma unroll
            for (int j = 0; j < K; ++j) acc[j] += gpart[g * LK + row * K + j];
        }
        for (int t = 0; t < nld; ++t) {
            const T d = dt[t * L + row];
#pragma unroll
            for (int j = 0; j < K; ++j) acc[j] += d * plong[t * K + j];
        }
#pragma unroll
        for (int j = 0; j < K; ++j) {
            AP[row * K + j] = acc[j];
            dot[j] += (double)P[row * K + j] * (double)acc[j];
        }
    }
    if (nld) {
        const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
        for (int t = wv; t < nld; t += kThreads / kWave) {
            double ta[K];
#pragma unroll
            for (int j = 0; j < K; ++j) ta[j] = 0.0;
            for (int64_t row = r0 + lane; row < r1; row += kWave) {
                const double d = (double)dt[t * L + row];
#pragma unroll
                for (int j = 0; j < K; ++j) ta[j] += d * (double)P[row * K + j];
            }
#pragma unroll
            for (int j = 0; j < K; ++j) {
                const double v = wave_sum(ta[j]);
                if (lane == 0) dlpart[((int64_t)t * nw + w) * K + j] = v;
            }
        }
    }
    if (with_dot) {
#pragma unroll
        for (int j = 0; j < K; ++j) {
            double t = block_sum<double>(dot[j], lds);
            if (threadIdx.x == 0) part[(int64_t)w * K + j] = t;
        }
    }
}

// Tiled small-k SpMM (one launch before k_fin_small), the default form of
// the column-blocked product.  Workgroup w owns a tile of R x 256
// consecutive rows (R fixed by L, k and the precision, never by timing); in
// column block b the tile's nonzeros are one
// contiguous range of the blocked arrays, which the workgroup streams in
// chunks of CH nonzeros: every thread loads V consecutive nonzeros' column
// indices and values (one vector load each) and gathers their P rows, all V
// in flight, writing val x P[col] to LDS; then each row's thread adds its
// row's products from LDS in ascending order.  No per-row pointer chase sits between a load and
// its gather, and no lane waits on a longer row.  The blocks are taken in
// order, so the workgroups, started together, gather from one block's slice
// of P at a time (kNB blocks: a slice of P small enough to stay in an XCD's
// L2 while S streams through it).  Per tile: the pAp partial part[w, j] and,
// for every split row t, the tile's partial of AP[long t] = sum_row dt[t, row]
// P[row] in dlpart[(t nw + w) K + j] (k_fin_small adds them in tile order).
// Measured against the groups-of-blocks form (k_spmm_blk + k_spmm_combine)
// on the 5M-worker system: 107 vs 123 us at k = 3 (FP64), 77 vs 93 (FP32),
// 172 vs 315 at k = 8; non-temporal loads of S made it slower (135 us).  The
// counters put it at the L2's limit (TCC busy 82% of the kernel, 11.5M
// requests: about one per gathered nonzero); one vector load per thread for
// its V indices and values (instead of V strided loads of each) took the FP64
// product from 105 to 99 us, two column blocks a round (half the barriers)
// or six nonzeros a thread made it slower (114 us).
// positions of the blocked arrays a thread of the tiled SpMM takes per chunk
// (its column indices and values in one vector load each); the blocked
// arrays hold kTilePad spare entries past their end for the last vectors
template <int K>
constexpr int tile_vec() {
    return K <= 4 ? 4 : 2;
}
constexpr int kTilePad = 4;
template <typename T, int N>
struct VecN {
    typedef T type __attribute__((ext_vector_type(N)));
};
template <typename T, int K, int R>
__global__ __launch_bounds__(kThreads) void k_spmm_tiles(
    const int *__restrict__ bptr, const int *__restrict__ bidx, const T *__restrict__ bval,
    int64_t L, const T *__restrict__ P, T *__restrict__ AP, double *__restrict__ part,
    int with_dot, int nld, const T *__restrict__ dt, const int *__restrict__ long_rows,
    double *__restrict__ dlpart, CgState st, const int *__restrict__ tbp,
    const uint16_t *__restrict__ off16, const uint16_t *__restrict__ bidx16,
    const int *__restrict__ col_lo) {
    if (st.done && *st.done) return;
    const bool C16 = off16 != nullptr;   // the compact indices (tbp, off16, bidx16)
    constexpr int V = tile_vec<K>();
    constexpr int CH = V * kThreads;
    __shared__ T prod[CH * K];
    __shared__ double lds[kThreads / kWave];
    __shared__ T plong[kDenseLongMax * K];
    // the rows' starts in the current block: a row's end is the next row's
    // start (read after the chunk's barrier), one load per row and block
    __shared__ int rsl[R * kThreads + 1];
    const int nw = gridDim.x, w = blockIdx.x, tid = threadIdx.x;
    const int64_t r0 = (int64_t)w * R * kThreads, r1 = min(L, r0 + R * kThreads);
    for (int q = tid; q < nld * K; q += blockDim.x)
        plong[q] = P[(int64_t)long_rows[q / K] * K + q % K];
    T acc[R][K];
#pragma unroll
    for (int q = 0; q < R; ++q)
#pragma unroll
        for (int j = 0; j < K; ++j) acc[q][j] = T(0);
    const int64_t nt = (L + kThreads - 1) / kThreads;   // 256-row stretches (tbp's rows)
    for (int b = 0; b < kNB; ++b) {
        const int64_t boff = (int64_t)b * L;
        int s0, s1, cb = 0;
        int rs[R], re[R];
        if (C16) {
            // the stretch starts in tbp, each row's offset from its stretch's
            // start in 16 bits; column indices relative to the block's first
            const int *tb = tbp + b * (nt + 1);
            s0 = tb[r0 / kThreads];
            s1 = tb[(r1 + kThreads - 1) / kThreads];
            cb = col_lo[b];
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const int64_t row = r0 + q * kThreads + tid;
                const int64_t si = r0 / kThreads + q;   // row's stretch
                rs[q] = row < r1 ? tb[si] + off16[boff + row] : s1;
                rsl[q * kThreads + tid] = rs[q];
            }
            if (tid == 0) rsl[R * kThreads] = s1;
        } else {
            s0 = bptr[boff + r0];
            s1 = bptr[boff + r1];
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const int64_t row = r0 + q * kThreads + tid;
                rs[q] = row < r1 ? bptr[boff + row] : 0;
                re[q] = row < r1 ? bptr[boff + row + 1] : 0;
            }
        }
        // chunks of CH positions from s0 rounded down to a multiple of V:
        // thread t takes positions base + V t .. + V - 1 (one aligned vector
        // load of their indices and one of their values) and gathers the P
        // rows of those in [s0, s1), all V in flight
        for (int base = s0 & ~(V - 1); base < s1; base += CH) {
            const int p0 = base + V * tid;
            int cc[V];
            T vv[V];
            if (p0 < s1) {
                if (C16) {
                    const typename VecN<uint16_t, V>::type ix =
                        *(const typename VecN<uint16_t, V>::type *)(bidx16 + p0);
#pragma unroll
                    for (int v = 0; v < V; ++v) cc[v] = cb + (int)ix[v];
                } else {
                    const typename VecN<int, V>::type ix =
                        *(const typename VecN<int, V>::type *)(bidx + p0);
#pragma unroll
                    for (int v = 0; v < V; ++v) cc[v] = ix[v];
                }
                const typename VecN<T, V>::type xv = *(const typename VecN<T, V>::type *)(bval + p0);
#pragma unroll
                for (int v = 0; v < V; ++v) vv[v] = xv[v];
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    const int p = p0 + v;
                    if (p >= s0 && p < s1) {
#pragma unroll
                        for (int j = 0; j < K; ++j)
                            prod[(V * tid + v) * K + j] = vv[v] * P[(int64_t)cc[v] * K + j];
                    }
                }
            }
            __syncthreads();
            if (C16) {
#pragma unroll
                for (int q = 0; q < R; ++q) re[q] = rsl[q * kThreads + tid + 1];
            }
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const int a = max(rs[q], base) - base, e = min(re[q], base + CH) - base;
                for (int p = a; p < e; ++p)
#pragma unroll
                    for (int j = 0; j < K; ++j) acc[q][j] += prod[p * K + j];
            }
            __syncthreads();
        }
    }
    if (nld) __syncthreads();  // plong (no chunk may have run)
    double dot[K];
#pragma unroll
    for (int j = 0; j < K; ++j) dot[j] = 0.0;
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int64_t row = r0 + q * kThreads + tid;
        if (row < r1) {
            for (int t = 0; t < nld; ++t) {
                const T d = dt[t * L + row];
#pragma unroll
                for (int j = 0; j < K; ++j) acc[q][j] += d * plong[t * K + j];
            }
#pragma unroll
            for (int j = 0; j < K; ++j) {
                AP[row * K + j] = acc[q][j];
                dot[j] += (double)P[row * K + j] * (double)acc[q][j];
            }
        }
    }
    if (nld) {
        const int lane = tid & (kWave - 1), wv = tid / kWave;
        for (int t = wv; t < nld; t += kThreads / kWave) {
            double ta[K];
#pragma unroll
            for (int j = 0; j < K; ++j) ta[j] = 0.0;
            for (int64_t row = r0 + lane; row < r1; row += kWave) {
                const double d = (double)dt[t * L + row];
#pragma unroll
                for (int j = 0; j < K; ++j) ta[j] += d * (double)P[row * K + j];
            }
#pragma unroll
            for (int j = 0; j < K; ++j) {
                const double v = wave_sum(ta[j]);
                if (lane == 0) dlpart[((int64_t)t * nw + w) * K + j] = v;
            }
        }
    }
    if (with_dot) {
#pragma unroll
        for (int j = 0; j < K; ++j) {
            double t = block_sum<double>(dot[j], lds);
            if (tid == 0) part[(int64_t)w * K + j] = t;
        }
    }
}

// a wave-uniform value in a scalar register (64-bit: its halves in turn)
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int64_t uniform(int64_t v) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Wide k: one wave per row, lanes over columns (j = lane + 64 t, t < CPL).
// Each nonzero's column index and value are loaded once (wave-uniform) and
// feed CPL coalesced gathers of the neighbour's row of P; two nonzeros are in
// flight at a time.  Rows are dealt to the waves grid-stride (wave v of
// workgroup w takes rows 4 w + v, 4 (w + nw) + v, ...), so the rows in flight
// at any moment form one narrow window that sweeps the matrix: their
// neighbours' rows of P -- for a block-diagonal S of many models, one model's
// block -- stay resident in the memory-side cache instead of the whole of P
// (a contiguous row range per workgroup keeps every part of the matrix in
// flight at once).
template <typename T, int CPL, typename I>
__global__ __launch_bounds__(kThreads) void k_spmm_wide(
    const I *__restrict__ indptr, const int *__restrict__ indices, const T *__restrict__ val,
    int64_t L, const T *__restrict__ P, T *__restrict__ AP, double *__restrict__ part,
    int with_dot, CgState st) {
    if (st.done && *st.done) return;
    extern __shared__ double wlds[];  // [4][k]
    const int k = st.k;
    const int w = blockIdx.x, nw = gridDim.x;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    double dot[CPL];
#pragma unroll
    for (int t = 0; t < CPL; ++t) dot[t] = 0.0;
    for (int64_t row = (int64_t)w * 4 + wv; row < L; row += (int64_t)nw * 4) {
        const I s = uniform(indptr[row]);
        const I e = uniform(indptr[row + 1]);
        const bool is_long = (e - s) > kLongThresh;
        T acc[CPL];
#pragma unroll
        for (int t = 0; t < CPL; ++t) acc[t] = T(0);
        if (!is_long) {
            I p = s;
            for (; p + 1 < e; p += 2) {
                const int c0 = indices[p], c1 = indices[p + 1];
                const T v0 = val[p], v1 = val[p + 1];
                const T *p0 = P + (int64_t)c0 * k + lane;
                const T *p1 = P + (int64_t)c1 * k + lane;
#pragma unroll
                for (int t = 0; t < CPL; ++t)
                    if (lane + 64 * t < k) acc[t] += v0 * p0[64 * t] + v1 * p1[64 * t];
            }
            if (p < e) {
                const int c0 = indices[p];
                const T v0 = val[p];
                const T *p0 = P + (int64_t)c0 * k + lane;
#pragma unroll
                for (int t = 0; t < CPL; ++t)
                    if (lane + 64 * t < k) acc[t] += v0 * p0[64 * t];
            }
        }
#pragma unroll
        for (int t = 0; t < CPL; ++t) {
            const int j = lane + 64 * t;
            if (j < k) {
                T a = acc[t];
                if (is_long) a = AP[row * k + j];
                else AP[row * k + j] = a;
                dot[t] += (double)P[row * k + j] * (double)a;
            }
        }
    }
    if (!with_dot) return;
#pragma unroll
    for (int t = 0; t < CPL; ++t) {
        const int j = lane + 64 * t;
        if (j < k) wlds[wv * k + j] = dot[t];
    }
    __syncthreads();
    for (int j = threadIdx.x; j < k; j += blockDim.x)
        part[(int64_t)w * k + j] = wlds[j] + wlds[k + j] + wlds[2 * k + j] + wlds[3 * k + j];
}

// Wide k in column slabs, a slab per XCD: the same product AP = S P with the
// columns cut into slabs of kSlabW, and in each launch ("round") XCD x (the
// workgroups w with w % 8 == x) working on one slab only, several XCDs
// sharing a slab when a round has fewer than eight.  Within an XCD the rows
// go grid-stride, kSlabW lanes per row (64 / kSlabW rows per wave), so an XCD's
// gathers in flight are one slab of its window's neighbour rows: for a
// block-diagonal S of many models a slab of one model's block (kSlabW x 8 B
// per level), which its 4 MB L2 can hold where the full rows (k x 8 B per
// level) could not.  Each row's sum runs over its nonzeros in order and in
// pairs, as in k_spmm_wide; the pAp partials are per (round, workgroup), zero
// outside the workgroup's slab.
// a slab's row is one 128-byte line of P: 16 columns in FP64, 32 in FP32,
// always 16 lanes per row (a lane takes two FP32 columns): at 16 FP32 columns
// a gather moved the same lines as in FP64 (0.93 ms against FP64's 1.22 on
// the 256-column JLA block), and at 32 lanes per row a wave had half FP64's
// gathers in flight (FP32 0.86 ms against FP64's 0.65)
template <typename T>
constexpr int slab_v() {
    return sizeof(T) == 4 ? 2 : 1;
}
constexpr int kSlabLanes = 16;
template <typename T>
constexpr int slab_w() {
    return kSlabLanes * slab_v<T>();
}
constexpr int kSlabGroups = 128;  // workgroups per XCD
constexpr int kPartRows = 8192;   // rows of k partials in the work buffer

template <typename T, typename I>
__global__ __launch_bounds__(kThreads) void k_spmm_slab(
    const I *__restrict__ indptr, const int *__restrict__ indices, const T *__restrict__ val,
    int64_t L, const T *__restrict__ P, T *__restrict__ AP, double *__restrict__ part,
    int with_dot, int slab0, int nslab, CgState st) {
    if (st.done && *st.done) return;
    constexpr int V = slab_v<T>();
    constexpr int kSlabW = slab_w<T>();
    __shared__ double lds[kThreads * V];
    const int k = st.k;
    const int w = blockIdx.x, x = w & 7, lw = w >> 3, gx = (int)gridDim.x >> 3;
    const int si = x % nslab, q = x / nslab, nq = (8 - si + nslab - 1) / nslab;
    const int lane = threadIdx.x & 63;
    const int jb = (slab0 + si) * kSlabW + (lane & (kSlabLanes - 1)) * V;
    bool colok[V];
    int jc[V];   // the column read (clamped into the row when past k)
#pragma unroll
    for (int v = 0; v < V; ++v) {
        colok[v] = jb + v < k;
        jc[v] = colok[v] ? jb + v : 0;
    }
    constexpr int kRowsPerIter = kThreads / kSlabLanes;
    const int64_t stride = (int64_t)gx * nq * kRowsPerIter;
    double dot[V];
#pragma unroll
    for (int v = 0; v < V; ++v) dot[v] = 0.0;
    for (int64_t base = ((int64_t)lw * nq + q) * kRowsPerIter; base < L; base += stride) {
        const int64_t row = base + threadIdx.x / kSlabLanes;
        if (row >= L) continue;
        const I s = indptr[row], e = indptr[row + 1];
        T acc[V];
#pragma unroll
        for (int v = 0; v < V; ++v) acc[v] = T(0);
        if (e - s > kLongThresh) {
            // a split row: its product is already in AP
#pragma unroll
            for (int v = 0; v < V; ++v)
                if (colok[v]) acc[v] = AP[row * k + jc[v]];
        } else {
            // the nonzeros in pairs, as k_spmm_wide adds them; four in flight
            I p = s;
            for (; p + 3 < e; p += 4) {
                const int c0 = indices[p], c1 = indices[p + 1], c2 = indices[p + 2],
                          c3 = indices[p + 3];
                const T v0 = val[p], v1 = val[p + 1], v2 = val[p + 2], v3 = val[p + 3];
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    const T *Pj = P + jc[v];
                    const T a0 = Pj[(int64_t)c0 * k], a1 = Pj[(int64_t)c1 * k];
                    const T a2 = Pj[(int64_t)c2 * k], a3 = Pj[(int64_t)c3 * k];
                    acc[v] += v0 * a0 + v1 * a1;
                    acc[v] += v2 * a2 + v3 * a3;
                }
            }
            if (p + 1 < e) {
                const int c0 = indices[p], c1 = indices[p + 1];
                const T v0 = val[p], v1 = val[p + 1];
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    const T *Pj = P + jc[v];
                    acc[v] += v0 * Pj[(int64_t)c0 * k] + v1 * Pj[(int64_t)c1 * k];
                }
                p += 2;
            }
            if (p < e)
#pragma unroll
                for (int v = 0; v < V; ++v) acc[v] += val[p] * P[(int64_t)indices[p] * k + jc[v]];
#pragma unroll
            for (int v = 0; v < V; ++v)
                if (colok[v]) AP[row * k + jc[v]] = acc[v];
        }
#pragma unroll
        for (int v = 0; v < V; ++v)
            if (colok[v]) dot[v] += (double)P[row * k + jc[v]] * (double)acc[v];
    }
    if (!with_dot) return;
    // lds[r * kSlabW + cs]: row lane r's sum for slab column cs
    const int r = threadIdx.x / kSlabLanes, cl = (threadIdx.x & (kSlabLanes - 1)) * V;
#pragma unroll
    for (int v = 0; v < V; ++v) lds[r * kSlabW + cl + v] = colok[v] ? dot[v] : 0.0;
    __syncthreads();
    // the workgroup's partial: the kSlabW columns of its slab summed over its
    // row lanes in a fixed order, every other column zero
    for (int c = threadIdx.x; c < k; c += blockDim.x) {
        const int cs = c - (slab0 + si) * kSlabW;
        double vsum = 0.0;
        if (cs >= 0 && cs < kSlabW)
            for (int rr = 0; rr < kRowsPerIter; ++rr) vsum += lds[rr * kSlabW + cs];
        part[(int64_t)w * k + c] = vsum;
    }
}

// Column-blocked mapping for the vector kernels: TJ threads across columns,
// TR = blockDim / TJ across rows; a thread owns columns tj + TJ c, c < CJ.
struct ColMap {
    int TJ, TR, CJ;
};
__host__ __device__ inline ColMap col_map(int k) {
    int tj = 1;
    while (tj < k && tj < kThreads) tj <<= 1;
    ColMap m;
    m.TJ = tj;
    m.TR = kThreads / tj;
    m.CJ = (k + tj - 1) / tj;
    return m;
}
constexpr int kMaxCJ = kMaxK / kThreads;

// Per-column sums over the workgroup's rows, written as part[w * stride + j].
__device__ void store_col_parts(double (&acc)[kMaxCJ], ColMap m, int k, double *part,
                                int64_t stride) {
    __shared__ double lds[kThreads];
    const int tj = threadIdx.x % m.TJ, tr = threadIdx.x / m.TJ;
    for (int c = 0; c < m.CJ; ++c) {
        __syncthreads();
        lds[tr * m.TJ + tj] = acc[c];
        __syncthreads();
        const int j = tj + m.TJ * c;
        if (tr == 0 && j < k) {
            double s = 0.0;
            for (int r = 0; r < m.TR; ++r) s += lds[r * m.TJ + tj];
            part[(int64_t)blockIdx.x * stride + j] = s;
        }
    }
}

// X += alpha P; R -= alpha AP; per-workgroup sums of r^2 and r^2 dinv
// (core/kernels_base.py:444-462), finished by k_fin_step.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_step(
    int64_t L, T *__restrict__ X, T *__restrict__ R, const T *__restrict__ P,
    const T *__restrict__ AP, const T *__restrict__ dinv, double *__restrict__ part, CgState st) {
    if (*st.done) return;
    const int k = st.k;
    const ColMap m = col_map(k);
    const int tj = threadIdx.x % m.TJ, tr = threadIdx.x / m.TJ;
    const int w = blockIdx.x, nw = gridDim.x;
    const int64_t r0 = part_begin(L, w, nw), r1 = part_begin(L, w + 1, nw);
    double rr[kMaxCJ], rz[kMaxCJ];
    for (int c = 0; c < kMaxCJ; ++c) rr[c] = rz[c] = 0.0;
    for (int c = 0; c < m.CJ; ++c) {
        const int j = tj + m.TJ * c;
        if (j >= k) break;
        const T a = (T)st.alpha[j];
        for (int64_t i = r0 + tr; i < r1; i += m.TR) {
            const int64_t e = i * k + j;
            X[e] += a * P[e];
            const T r = R[e] - a * AP[e];
            R[e] = r;
            const double r2 = (double)r * (double)r;
            rr[c] += r2;
            rz[c] += r2 * (double)dinv[i];
        }
    }
    store_col_parts(rr, m, k, part, 2 * (int64_t)k);
    store_col_parts(rz, m, k, part + k, 2 * (int64_t)k);
}

// The solve's tolerance and iteration cap into the state, so that the
// iterations' kernels take the same arguments for every solve (a captured
// graph of them is replayed with new ones).
__global__ void k_set_params(CgState st, double tol, int maxiter) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        *st.dtol = tol;
        *st.dmaxiter = maxiter;
    }
}

// One workgroup after k_step: rel = sqrt(rr) / bnorm, active = rel > tol, stop
// when no column is active or at maxiter; beta = rz_new / rz
// (core/solve.py:123-139), from the nw workgroups' sums.
__global__ __launch_bounds__(kFinThreads) void k_fin_step(const double *__restrict__ part, int nw,
                                                       CgState st) {
    if (*st.done) return;
    const int k = st.k;
    __shared__ double s_rr[kMaxK];
    __shared__ double s_rz[kMaxK];
    __shared__ int any;
    if (threadIdx.x == 0) any = 0;
    reduce_parts(part, nw, k, 2 * k, s_rr);
    reduce_parts(part + k, nw, k, 2 * k, s_rz);
    __syncthreads();
    const double tol = *st.dtol;
    const int maxiter = *st.dmaxiter;
    for (int j = threadIdx.x; j < k; j += blockDim.x) {
        const double rel = sqrt(s_rr[j]) / st.bnorm[j];
        st.rel[j] = rel;
        const int act = rel > tol;
        st.active[j] = act;
        if (act) atomicOr(&any, 1);
    }
    __syncthreads();
    const int it = *st.it + 1;
    if (any && it < maxiter) {
        for (int j = threadIdx.x; j < k; j += blockDim.x) {
            const double rzn = s_rz[j], rzo = st.rz[j];
            st.beta[j] = st.active[j] ? rzn / (rzo != 0.0 ? rzo : 1.0) : 0.0;
            st.rz[j] = rzn;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        *st.it = it;
        if (!any || it >= maxiter) {
            *st.done = any ? 2 : 1;  // 1 converged, 2 hit maxiter
            *st.it_final = it;
        }
    }
}

// P = dinv R + beta P (core/kernels_base.py:465-474).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_direction(int64_t L, T *__restrict__ P,
                                                        const T *__restrict__ R,
                                                        const T *__restrict__ dinv, CgState st) {
    if (*st.done) return;
    const int k = st.k;
    const int64_t n = L * k;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e / k;
        const int j = (int)(e - i * k);
        P[e] = dinv[i] * R[e] + (T)st.beta[j] * P[e];
    }
}

// out[j] = sum_i A_ij B_ij (d_i if d), fixed-order; mode 1 also stores
// sqrt(out) with 0 -> 1 (the column norms of b, core/solve.py:81-82).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_colsum(int64_t L, int k, const T *__restrict__ A,
                                                     const T *__restrict__ B,
                                                     const T *__restrict__ d,
                                                     double *__restrict__ part) {
    const ColMap m = col_map(k);
    const int tj = threadIdx.x % m.TJ, tr = threadIdx.x / m.TJ;
    const int w = blockIdx.x, nw = gridDim.x;
    const int64_t r0 = part_begin(L, w, nw), r1 = part_begin(L, w + 1, nw);
    double acc[kMaxCJ];
    for (int c = 0; c < kMaxCJ; ++c) acc[c] = 0.0;
    for (int c = 0; c < m.CJ; ++c) {
        const int j = tj + m.TJ * c;
        if (j >= k) break;
        for (int64_t i = r0 + tr; i < r1; i += m.TR) {
            const int64_t e = i * k + j;
            double v = (double)A[e] * (double)B[e];
            if (d) v *= (double)d[i];
            acc[c] += v;
        }
    }
    store_col_parts(acc, m, k, part, k);
}

// One workgroup after k_colsum: out[j] from the nw workgroups' sums.
__global__ __launch_bounds__(kFinThreads) void k_fin_colsum(int k, const double *__restrict__ part,
                                                         int nw, double *__restrict__ out,
                                                         int mode) {
    __shared__ double tmp[kMaxK];
    reduce_parts(part, nw, k, k, tmp);
    __syncthreads();
    for (int j = threadIdx.x; j < k; j += blockDim.x) {
        double v = tmp[j];
        if (mode == 1) {
            v = sqrt(v);
            if (v == 0.0) v = 1.0;
        }
        out[j] = v;
    }
}

// Elementwise helpers.
template <typename T>
__global__ void k_set_start(int64_t L, int k, const double *__restrict__ b,
                            const double *__restrict__ Sx, const int *__restrict__ zero,
                            T *__restrict__ X, T *__restrict__ R, T *__restrict__ P,
                            const T *__restrict__ dinv) {
    // R = b - Sx (Sx may be null: R = b); zero columns get X = R = 0; P = dinv R.
    const int64_t n = L * k;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e / k;
        const int j = (int)(e - i * k);
        double r = b[e] - (Sx ? Sx[e] : 0.0);
        if (zero[j]) {
            r = 0.0;
            X[e] = T(0);
        }
        R[e] = (T)r;
        P[e] = dinv[i] * (T)r;
    }
}

__global__ void k_cast_f64_to_f32(int64_t n, const double *__restrict__ a, float *__restrict__ b) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x)
        b[e] = (float)a[e];
}

__global__ void k_widen(int64_t n, const int *__restrict__ a, int64_t *__restrict__ b) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x)
        b[e] = a[e];
}

// Mixed precision: R32 = r64 / s_j for active columns (0 otherwise), X32 = 0,
// P32 = dinv R32.  s_j = ||r64_j|| keeps the FP32 inner system well scaled.
__global__ void k_inner_start(int64_t L, int k, const double *__restrict__ r64,
                              const double *__restrict__ scale, const int *__restrict__ act,
                              float *__restrict__ X, float *__restrict__ R, float *__restrict__ P,
                              const float *__restrict__ dinv) {
    const int64_t n = L * k;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e / k;
        const int j = (int)(e - i * k);
        const float r = act[j] ? (float)(r64[e] / scale[j]) : 0.0f;
        X[e] = 0.0f;
        R[e] = r;
        P[e] = dinv[i] * r;
    }
}

// x64 += s_j X32 (active columns); r64 = b - Sx64 after the caller's SpMM.
__global__ void k_refine_update(int64_t L, int k, double *__restrict__ x64,
                                const float *__restrict__ X, const double *__restrict__ scale,
                                const int *__restrict__ act) {
    const int64_t n = L * k;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(e % k);
        if (act[j]) x64[e] += scale[j] * (double)X[e];
    }
}

__global__ void k_sub(int64_t n, const double *__restrict__ b, const double *__restrict__ Sx,
                      double *__restrict__ r) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x)
        r[e] = b[e] - Sx[e];
}

__global__ void k_zero_cols(int64_t L, int k, const int *__restrict__ zero, double *__restrict__ X) {
    const int64_t n = L * k;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x)
        if (zero[e % k]) X[e] = 0.0;
}

}  // namespace hdfe

namespace hdfe {

// Blocked layout of the non-split rows (see k_spmm_blk); with slot (the
// dense-long slot of each column, -1 for the others), a row's entries at the
// split rows' columns are left out of the blocks.
// The blocked layout's counts and fill, a wave per row: lanes take the row's
// entries 64 at a time (coalesced), each entry's column block found from
// col_lo; per block a ballot ranks the wave's entries in it.  The fill puts
// each entry at its block's position for the row plus the block's entries on
// lower lanes and in earlier passes, so every block keeps the row's order.
// Split rows (over kLongThresh nonzeros) are left out; entries at dense long
// columns (slot >= 0) go to dt instead.
__device__ inline int blk_of(int col, const int *__restrict__ col_lo) {
    int b = 0;
    while (b < kNB - 1 && col >= col_lo[b + 1]) ++b;
    return b;
}

__global__ __launch_bounds__(kThreads) void k_blk_rowcount(
    int64_t L, const int *__restrict__ indptr, const int *__restrict__ indices,
    const int *__restrict__ col_lo, const int *__restrict__ slot, int *__restrict__ cnt) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t nwave = (int64_t)gridDim.x * (blockDim.x / kWave);
    for (int64_t r = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / kWave; r < L; r += nwave) {
        const int s = indptr[r], e = indptr[r + 1];
        int c = 0;  // lane b < kNB: block b's entries
        if (e - s <= kLongThresh) {
            for (int p0 = s; p0 < e; p0 += kWave) {
                const int p = p0 + lane;
                int b = -1;
                if (p < e) {
                    const int col = indices[p];
                    if (!slot || slot[col] < 0) b = blk_of(col, col_lo);
                }
#pragma unroll
                for (int bb = 0; bb < kNB; ++bb) {
                    const uint64_t m = __ballot(b == bb);
                    if (lane == bb) c += __popcll(m);
                }
            }
        }
        if (lane < kNB) cnt[(int64_t)lane * L + r] = c;
    }
}

__global__ __launch_bounds__(kThreads) void k_blk_fill(
    int64_t L, const int *__restrict__ indptr, const int *__restrict__ indices,
    const double *__restrict__ v64, const float *__restrict__ v32, const int *__restrict__ col_lo,
    const int *__restrict__ bptr, int *__restrict__ bidx, double *__restrict__ b64,
    float *__restrict__ b32, const int *__restrict__ slot, double *__restrict__ dt64,
    float *__restrict__ dt32) {
    const int lane = threadIdx.x & (kWave - 1);
    const uint64_t below = lane ? (~0ull >> (kWave - lane)) : 0ull;
    const int64_t nwave = (int64_t)gridDim.x * (blockDim.x / kWave);
    for (int64_t r = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / kWave; r < L; r += nwave) {
        const int s = indptr[r], e = indptr[r + 1];
        if (e - s > kLongThresh) continue;
        int next = lane < kNB ? bptr[(int64_t)lane * L + r] : 0;  // lane b: block b's next slot
        for (int p0 = s; p0 < e; p0 += kWave) {
            const int p = p0 + lane;
            int col = 0, b = -1, sl = -1;
            if (p < e) {
                col = indices[p];
                sl = slot ? slot[col] : -1;
                if (sl < 0) b = blk_of(col, col_lo);
            }
            int q = 0;
#pragma unroll
            for (int bb = 0; bb < kNB; ++bb) {
                const uint64_t m = __ballot(b == bb);
                const int at = __shfl(next, bb);
                if (b == bb) q = at + __popcll(m & below);
                if (lane == bb) next += __popcll(m);
            }
            if (sl >= 0) {
                dt64[(int64_t)sl * L + r] = v64[p];
                dt32[(int64_t)sl * L + r] = v32[p];
            } else if (b >= 0) {
                bidx[q] = col;
                b64[q] = v64[p];
                b32[q] = v32[p];
            }
        }
    }
}

// The tiled form's compact indices from the blocked arrays: each block's
// 256-row stretch starts (tbp, kNB x (nt + 1), the last the block's end),
// each row's start less its stretch's (off16), each column less its block's
// first (bidx16).  The caller checks that both fit 16 bits.
__global__ void k_tbp_fill(int64_t L, int64_t nt, const int *__restrict__ bptr,
                           int *__restrict__ tbp) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < kNB * (nt + 1);
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = i / (nt + 1), t = i - b * (nt + 1);
        tbp[i] = bptr[b * L + min(t * kThreads, L)];
    }
}

__global__ void k_off16_fill(int64_t L, int64_t nt, const int *__restrict__ bptr,
                             const int *__restrict__ tbp, uint16_t *__restrict__ off16) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < kNB * L;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = i / L, r = i - b * L;
        off16[i] = (uint16_t)(bptr[i] - tbp[b * (nt + 1) + r / kThreads]);
    }
}

__global__ void k_bidx16_fill(int64_t n, const int *__restrict__ bidx,
                              const int *__restrict__ col_lo, uint16_t *__restrict__ bidx16) {
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int col = bidx[q];
        bidx16[q] = (uint16_t)(col - col_lo[blk_of(col, col_lo)]);
    }
}

// The split rows' block among themselves: ll[t, slot(col)] = S[long t, col]
// for the split rows' entries at split rows' columns (kLlParts workgroups per
// row, each a stride of its entries).
constexpr int kLlParts = 64;
__global__ void k_ll_fill(int nlong, const int *__restrict__ long_rows,
                          const int *__restrict__ indptr, const int *__restrict__ indices,
                          const double *__restrict__ v64, const float *__restrict__ v32,
                          const int *__restrict__ slot, double *__restrict__ ll64,
                          float *__restrict__ ll32) {
    const int t = blockIdx.x / kLlParts, part = blockIdx.x % kLlParts;
    if (t >= nlong) return;
    const int r = long_rows[t];
    for (int p = indptr[r] + part * blockDim.x + threadIdx.x; p < indptr[r + 1];
         p += kLlParts * blockDim.x) {
        const int sl = slot[indices[p]];
        if (sl >= 0) {
            ll64[t * nlong + sl] = v64[p];
            ll32[t * nlong + sl] = v32[p];
        }
    }
}

}  // namespace hdfe

// ====================================================================== host

namespace hdfe {

static int grid_for(int64_t n, int64_t per, int lo, int hi) {
    int64_t g = (n + per - 1) / per;
    return (int)std::max<int64_t>(lo, std::min<int64_t>(hi, g));
}

// The matrix's device arrays and work buffers come from dev_malloc on the
// matrix's stream (a repeated fit reuses the same memory instead of mapping
// new pages), and go back in stream order.
static hipError_t palloc(void *p, size_t bytes, hipStream_t s) {
    return dev_malloc((void **)p, bytes, s);
}

static void pfree(void *p, hipStream_t s) { dev_free(p, s); }

// A pair of scratch allocations freed on scope exit.
struct DevBuf2 {
    double *p = nullptr;
    void *extra = nullptr;
    ~DevBuf2() {
        if (p) (void)hipFree(p);
        if (extra) (void)hipFree(extra);
    }
};

// Work buffers, cached on the matrix for the last block width used.
// ---------------------------------------------- matrix-free S P (stream_cg)
//
// The reference's fallback when S does not fit (`_nb_stream_matvec`,
// core/kernels_base.py:477-507; core/solve.py:251-271): S P from the
// identifying cells without S.  With a_c the cell's levels (one per
// non-streamed dimension), n_c its weight and N_g its group's total,
//   (S P)_l = sum over the cells c containing l of v_c,
//   v_c = n_c (t_c - m_g),  t_c = sum_d P[a_cd],  m_g = sum_c n_c t_c / N_g.
// v is formed per group (a thread per group and column, its cells in order),
// then summed per level through the level-to-cells CSR in chunks of
// kMfChunk cells (a wave per chunk, a fixed shuffle tree) and the chunks in
// order: no atomics, the same result every run.

constexpr int kMfMaxD = 8;
constexpr int kMfChunk = 2048;
struct MfDims {
    int D;
    int64_t off[kMfMaxD];
};

template <typename T>
__global__ __launch_bounds__(kThreads) void k_mf_cells(int64_t Gi, const int *__restrict__ igstart,
                                                     const int *__restrict__ icodes, int64_t ld,
                                                     MfDims dm, const double *__restrict__ cn,
                                                     const T *__restrict__ P, int k,
                                                     T *__restrict__ v, CgState st) {
    if (st.done && *st.done) return;
    const int64_t nth = (int64_t)gridDim.x * blockDim.x;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < Gi * k; q += nth) {
        const int64_t g = q / k;
        const int j = (int)(q - g * k);
        const int s = igstart[g], e = igstart[g + 1];
        double N = 0.0, S = 0.0;
        for (int c = s; c < e; ++c) {
            double t = 0.0;
            for (int d = 0; d < dm.D; ++d)
                t += (double)P[(dm.off[d] + icodes[d * ld + c]) * k + j];
            N += cn[c];
            S += cn[c] * t;
        }
        const double mg = N > 0.0 ? S / N : 0.0;
        for (int c = s; c < e; ++c) {
            double t = 0.0;
            for (int d = 0; d < dm.D; ++d)
                t += (double)P[(dm.off[d] + icodes[d * ld + c]) * k + j];
            v[(int64_t)c * k + j] = (T)(cn[c] * (t - mg));
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_mf_chunks(int64_t nchunk, const int *__restrict__ ch_beg,
                                                      const int *__restrict__ ch_end,
                                                      const int *__restrict__ lcell,
                                                      const T *__restrict__ v, int k,
                                                      double *__restrict__ cpart, CgState st) {
    if (st.done && *st.done) return;
    const int lane = threadIdx.x & 63;
    const int64_t nwave = ((int64_t)gridDim.x * blockDim.x) >> 6;
    int tj = 1;
    while (tj < k && tj < 64) tj <<= 1;
    const int tr = 64 / tj, cj = lane % tj, rj = lane / tj;
    for (int64_t q = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; q < nchunk;
         q += nwave) {
        const int b = ch_beg[q], e = ch_end[q];
        for (int j0 = 0; j0 < k; j0 += tj) {
            const int j = j0 + cj;
            double acc = 0.0;
            if (j < k)
                for (int i = b + rj; i < e; i += tr) acc += (double)v[(int64_t)lcell[i] * k + j];
            for (int o = tj; o < 64; o <<= 1) acc += __shfl_xor(acc, o, 64);
            if (rj == 0 && j < k) cpart[q * k + j] = acc;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_mf_finish(int64_t L, const int *__restrict__ lch,
                                                      const double *__restrict__ cpart, int k,
                                                      T *__restrict__ AP, CgState st) {
    if (st.done && *st.done) return;
    const int64_t nth = (int64_t)gridDim.x * blockDim.x;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < L * k; q += nth) {
        const int64_t l = q / k;
        const int j = (int)(q - l * k);
        double acc = 0.0;
        for (int c = lch[l]; c < lch[l + 1]; ++c) acc += cpart[(int64_t)c * k + j];
        AP[q] = (T)acc;
    }
}

// The level key of every (dimension, cell) pair and its cell, for the sort
// that makes the level-to-cells CSR.
__global__ void k_mf_pairs(int64_t Mc, const int *__restrict__ icodes, int64_t ld, MfDims dm,
                           int *__restrict__ key, int *__restrict__ cell,
                           int *__restrict__ count) {
    const int64_t n = Mc * dm.D;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int d = (int)(q / Mc);
        const int64_t c = q - d * Mc;
        const int l = (int)(dm.off[d] + icodes[d * ld + c]);
        key[q] = l;
        cell[q] = (int)c;
        atomicAdd(count + l, 1);
    }
}

// ----------------------------------- matrix-free S P for wide blocks (k > 8)
//
// With hundreds of right-hand-side columns (the GLM's [z, X], the JLA draws)
// the per-cell vector v of k_mf_cells would be M x k (45 GB at 18.7M cells
// and 301 columns).  The wide form keeps only the group means:
//   m_g = sum_{c in g} n_c t_c / N_g,   t_c = sum_d P[a_cd]       (G x k)
// and forms each level's row from them.  Levels of a *sparse* dimension
// (firms: tens of cells each) are rows of their own:
//   (S P)_l = sum_{c contains l} n_c (t_c - m_g(c))               k_mfw_levels
// over the level-to-cells CSR, one workgroup per level, lanes over columns,
// gathering the cells' other P rows and their groups' m rows.  Levels of a
// *dense* dimension (years: many thousands of cells each, whose rows would
// each read most of m) are summed where m_g is made: k_mfw_groups takes a
// planned range of consecutive groups per workgroup, every lane one column
// for all of the range's groups in order, and adds n_c (t_c - m_g) of each
// cell into a per-lane LDS slot of the cell's dense level (the range's
// levels of a dense dimension fall in a window of at most kMfwSlots); the
// range's slots go to rows of dpart, and k_mfw_dense adds each dense level's
// rows over the ranges that cover it, in range order.  Every sum runs in a
// fixed order in FP64: results do not depend on scheduling.

__global__ void k_fill_u64(int64_t n, unsigned long long v, unsigned long long *__restrict__ x) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        x[i] = v;
}

// An order-independent 64-bit fingerprint of int arrays: the sum over
// elements of a mix of (position, value) -- integer sums, so the same arrays
// give the same fingerprint whatever the schedule.  Row r of `rows` rows
// starts at a + r * ld.
__global__ void k_fingerprint(int64_t n, int rows, int64_t ld, const int *__restrict__ a,
                              unsigned long long *__restrict__ out) {
    unsigned long long h = 0;
    for (int r = 0; r < rows; ++r)
        for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
             i += (int64_t)gridDim.x * blockDim.x) {
            unsigned long long x = ((unsigned long long)(r * n + i) << 32) ^
                                   (unsigned int)a[r * ld + i];
            x ^= x >> 31;
            x *= 0x7fb5d329728ea185ull;
            x ^= x >> 27;
            x *= 0x81dadef4bc2dd44dull;
            x ^= x >> 33;
            h += x;
        }
    for (int o = 32; o > 0; o >>= 1) h += __shfl_xor(h, o, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(out, h);
}

struct MfList {
    int n = 0;
    int d[kMfMaxD] = {0};
};

constexpr int kMfwSlots = 16;           // LDS slots of dense levels per range
constexpr int kMfwRanges = 8192;        // workgroups of k_mfw_groups (target)
constexpr int64_t kMfwDenseAvg = 1024;  // cells per level that make a dimension dense

struct MatFree {
    int64_t Gi = 0, Mc = 0, ld = 0, nchunk = 0;
    MfDims dm{};
    const int *igstart = nullptr, *icodes = nullptr;
    const double *cn = nullptr;
    int *lcell = nullptr, *ch_beg = nullptr, *ch_end = nullptr, *lch = nullptr;
    void *v = nullptr;          // (Mc, k) of the current k
    double *cpart = nullptr;    // (nchunk, k)
    int vk = 0;
    // the wide form (k > kSmallK), when a plan exists (wide)
    bool wide = false;
    MfList dd;                          // the dense dimensions
    MfList sd;                          // the others (gathered per cell)
    int64_t nsl = 0;                    // levels of the sparse dimensions
    int *slev = nullptr;                // those levels
    int *lptr = nullptr;                // (L + 1) each level's cells in lcell
    int *cgroup = nullptr;              // (Mc) each cell's group
    int64_t nr = 0, dprows = 0, ndl = 0;
    int maxslot = 0;                    // the widest range's slots
    int *rng = nullptr;                 // (nr + 1) first group of each range
    int *rbase = nullptr, *rsoff = nullptr;  // (nr, nds) window start, slot offset
    int *dpoff = nullptr;               // (nr + 1) first dpart row of each range
    int *dlev = nullptr, *dcptr = nullptr, *dcrow = nullptr;  // dense levels, their rows
    int64_t *offs_d = nullptr;          // (D) the dimensions' level offsets, on the device
    void *mg = nullptr;                 // (Gi, k) group means, of the current k
    double *dpart = nullptr;            // (dprows, k)
    unsigned long long fp = 0;          // the cells' fingerprint (igstart, icodes)
};

// Cells of a range or a level are staged kMfwStage at a time in LDS (their
// levels, weights and dense slots or groups, loaded by all lanes at once),
// then every lane walks them in order with kMfwIlp cells' gathers in flight.
// A lane adds into its dense slots with LDS atomics that return nothing
// (ds_add_f64: the lane goes on without waiting; one lane's adds to a slot
// land in program order, so the sums are the same every run).
constexpr int kMfwStage = 256;
constexpr int kMfwIlp = 4;
// k_mfw_groups' cells in flight across group ends (16 FP64 cells cost
// occupancy: 12.4 -> 16.2 ms at C5, where FP32 went 11.8 -> 10.3)
template <typename T>
constexpr int mfw_ilp_groups() {
    return sizeof(T) == 4 ? 16 : 8;
}
constexpr int kMfwLanes = 512;      // k_mfw_groups' workgroup at most (a column each)
constexpr int kMfwMaxDense = 2;     // dense dimensions summed per range (more: as sparse)
constexpr int kMfwStageBytes = kMfwStage * (4 * (kMfMaxD + kMfwMaxDense) + 8);
constexpr int64_t kMfwLds = 96 * 1024;   // k_mfw_groups' LDS at most (gfx950: 160 KB a workgroup)

// The matrix-free kernels take the cells' dimension count DT as a template
// parameter (0: the runtime D, up to kMfMaxD): with the loops over the
// dimensions unrolled and every gather's address valid (cells and columns
// past the end clamped to the last, their terms weighted 0), a lane's loads
// carry no branches and fly together -- with a runtime loop and guarded loads
// the compiler waited on each gather in turn (k_mfw_groups 22.8 ms per FP32
// product at C5).
template <int DT>
constexpr int mfw_dmax() {
    return DT > 0 ? DT : kMfMaxD;
}

// Stages cells [p0, p0 + np) of range r: their levels in the sparse
// dimensions (sd, ns of them), their dense slots and weights, and the ends of
// groups [g, g2) relative to p0 (every thread of the workgroup calls it).
template <int DS>
__device__ __forceinline__ void mfw_stage_cells(int p0, int np, const int *__restrict__ icodes,
                                                int64_t ld, const MfList &sd,
                                                const int64_t *__restrict__ offs, int nds,
                                                int dd0, int dd1, int r,
                                                const int *__restrict__ rbase,
                                                const int *__restrict__ rsoff,
                                                const double *__restrict__ cn,
                                                int (*s_lv)[kMfwStage],
                                                int (*s_sl)[kMfwStage], double *s_w,
                                                const int *__restrict__ igstart, int g, int g2,
                                                int *s_gend) {
    constexpr int DM = mfw_dmax<DS>();
    const int ns = DS > 0 ? DS : sd.n;
    __syncthreads();
    for (int q = threadIdx.x; q < g2 - g; q += blockDim.x) s_gend[q] = igstart[g + 1 + q] - p0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) {
        const int c = p0 + i;
        int code[DM];
#pragma unroll
        for (int e = 0; e < DM; ++e) code[e] = e < ns ? icodes[sd.d[e] * ld + c] : 0;
        const int c0 = nds > 0 ? icodes[dd0 * ld + c] : 0;
        const int c1 = nds > 1 ? icodes[dd1 * ld + c] : 0;
        const double wc = cn[c];
#pragma unroll
        for (int e = 0; e < DM; ++e)
            if (e < ns) s_lv[e][i] = (int)(offs[sd.d[e]] + code[e]);
        if (nds > 0) s_sl[0][i] = rsoff[r * nds] + c0 - rbase[r * nds];
        if (nds > 1) s_sl[1][i] = rsoff[r * nds + 1] + c1 - rbase[r * nds + 1];
        s_w[i] = wc;
    }
    __syncthreads();
}

// t_c for staged cell i and the lane's column: its sparse dimensions' rows of
// P gathered (column jc, the lane's clamped into the row), its dense levels'
// from the LDS copy pw (slot-major, bd lanes)
template <typename T, int DS, int NDS>
__device__ __forceinline__ double mfw_cell_t(int i, int jc, int k, int ns,
                                             const T *__restrict__ P, int (*s_lv)[kMfwStage],
                                             int (*s_sl)[kMfwStage], const T *pw, int bd) {
    constexpr int DM = mfw_dmax<DS>();
    double x = 0.0;
#pragma unroll
    for (int e = 0; e < DM; ++e)
        if (DS > 0 || e < ns) x += (double)P[(int64_t)s_lv[e][i] * k + jc];
#pragma unroll
    for (int q = 0; q < NDS; ++q) x += (double)pw[s_sl[q][i] * bd + threadIdx.x];
    return x;
}

// Staged cells [a, e): N += n_c, S += n_c t_c and n_c t_c into each cell's
// dense slots, kMfwIlp cells' gathers in flight.
template <typename T, int DS, int NDS>
__device__ __forceinline__ void mfw_add_cells(int a, int e, int jc, int k, int ns,
                                              const T *__restrict__ P, int (*s_lv)[kMfwStage],
                                              int (*s_sl)[kMfwStage], const double *s_w,
                                              double *slot, const T *pw, int bd, double &N,
                                              double &S) {
    for (int i0 = a; i0 < e; i0 += kMfwIlp) {
        const int nu = min(kMfwIlp, e - i0);
        double t[kMfwIlp];
#pragma unroll
        for (int u = 0; u < kMfwIlp; ++u)
            t[u] = mfw_cell_t<T, DS, NDS>(min(i0 + u, e - 1), jc, k, ns, P, s_lv, s_sl, pw, bd);
#pragma unroll
        for (int u = 0; u < kMfwIlp; ++u)
            if (u < nu) {
                const double w = s_w[i0 + u];
                N += w;
                S += w * t[u];
#pragma unroll
                for (int q = 0; q < NDS; ++q)
                    unsafeAtomicAdd(&slot[s_sl[q][i0 + u] * bd + threadIdx.x], w * t[u]);
            }
    }
}

// Range r's groups [rng[r], rng[r+1]): m_g into mg, and each dense level's
// part into dpart rows dpoff[r] + rsoff[r, q] + (code - rbase[r, q]).  One
// workgroup per range; lane t takes columns j = t, t + blockDim, ...  The
// groups are staged whole, as many as fit kMfwStage cells (a larger group
// alone, in pieces).  A cell adds n_c t_c to its dense slots as it is read;
// once the group's mean is known, each of its cells takes n_c m_g back out.
// The range's dense levels' rows of P (its window: as many as its slots) are
// copied to LDS (pw) at the start, so only the sparse dimensions' rows are
// gathered per cell: with the years gathered too, every cell's 1.2 KB year
// row went through the L2 as well, which ran 98% busy (C5).  Dynamic LDS:
// slot (ms x bd doubles), then pw (ms x bd T).
template <typename T, int DS, int NDS>
__global__ __launch_bounds__(kMfwLanes) void k_mfw_groups(const int *__restrict__ rng, const int *__restrict__ rbase,
                             const int *__restrict__ rsoff, const int *__restrict__ dpoff,
                             const int *__restrict__ igstart, const int *__restrict__ icodes,
                             int64_t ld, MfList sd, const int64_t *__restrict__ offs,
                             int dd0, int dd1, int ms, const double *__restrict__ cn,
                             const T *__restrict__ P, int k, T *__restrict__ mg,
                             double *__restrict__ dpart, CgState st) {
    if (st.done && *st.done) return;
    constexpr int DM = mfw_dmax<DS>();
    __shared__ int s_lv[DM][kMfwStage];
    __shared__ int s_sl[kMfwMaxDense][kMfwStage];
    __shared__ double s_w[kMfwStage];
    __shared__ int s_gend[kMfwStage];
    extern __shared__ double slot[];
    const int r = blockIdx.x, tid = threadIdx.x, bd = blockDim.x;
    T *pw = (T *)(slot + (int64_t)ms * bd);
    const int ns = DS > 0 ? DS : sd.n;
    constexpr int nds = NDS;
    const int gb = rng[r], ge = rng[r + 1];
    const int nslot = dpoff[r + 1] - dpoff[r];
    constexpr int kIlp = mfw_ilp_groups<T>();
    for (int j0 = 0; j0 < k; j0 += bd) {
        const int j = j0 + tid;
        const bool on = j < k;
        const int jc = min(j, k - 1);
        for (int q = 0; q < nslot; ++q) slot[q * bd + tid] = 0.0;
        // the window's rows of P, each lane its column (no lane reads
        // another's, so no barrier)
        for (int q = 0; q < nds; ++q) {
            const int qs = rsoff[r * nds + q];
            const int qe = q + 1 < nds ? rsoff[r * nds + q + 1] : nslot;
            const int64_t row0 = offs[q == 0 ? dd0 : dd1] + rbase[r * nds + q] - qs;
            for (int sl = qs; sl < qe; ++sl) pw[sl * bd + tid] = P[(row0 + sl) * k + jc];
        }
        for (int g = gb; g < ge;) {
            // the whole groups [g, g2) that fit the stage (at least group g)
            const int c0 = igstart[g];
            int g2 = g + 1;
            while (g2 < ge && igstart[g2 + 1] - c0 <= kMfwStage) ++g2;
            const int cend = igstart[g2];
            if (cend - c0 > kMfwStage) {
                // one group larger than the stage: its pieces once for the
                // mean (and n_c t_c into the slots), again for n_c m_g out
                double N = 0.0, S = 0.0;
                for (int p0 = c0; p0 < cend; p0 += kMfwStage) {
                    const int np = min(kMfwStage, cend - p0);
                    mfw_stage_cells<DS>(p0, np, icodes, ld, sd, offs, nds, dd0, dd1, r, rbase,
                                        rsoff, cn, s_lv, s_sl, s_w, igstart, g, g + 1, s_gend);
                    mfw_add_cells<T, DS, NDS>(0, np, jc, k, ns, P, s_lv, s_sl, s_w, slot, pw, bd,
                                              N, S);
                }
                const double m = N > 0.0 ? S / N : 0.0;
                if (on) mg[(int64_t)g * k + j] = (T)m;
                if (nds)
                    for (int p0 = c0; p0 < cend; p0 += kMfwStage) {
                        const int np = min(kMfwStage, cend - p0);
                        mfw_stage_cells<DS>(p0, np, icodes, ld, sd, offs, nds, dd0, dd1, r, rbase,
                                            rsoff, cn, s_lv, s_sl, s_w, igstart, g, g + 1,
                                            s_gend);
                        for (int i = 0; i < np; ++i)
                            for (int q = 0; q < nds; ++q)
                                unsafeAtomicAdd(&slot[s_sl[q][i] * bd + tid], -(s_w[i] * m));
                    }
            } else {
                mfw_stage_cells<DS>(c0, cend - c0, icodes, ld, sd, offs, nds, dd0, dd1, r, rbase,
                                    rsoff, cn, s_lv, s_sl, s_w, igstart, g, g2, s_gend);
                // every staged cell in order, kIlp cells' gathers in flight
                // across group ends; a group is finished when its end passes
                const int ncell = cend - c0;
                int gi = g, gbeg = 0, gend = s_gend[0];
                double N = 0.0, S = 0.0;
                auto finish = [&]() {
                    const double m = N > 0.0 ? S / N : 0.0;
                    if (on) mg[(int64_t)gi * k + j] = (T)m;
                    for (int i = gbeg; i < gend; ++i)
                        for (int q = 0; q < nds; ++q)
                            unsafeAtomicAdd(&slot[s_sl[q][i] * bd + tid], -(s_w[i] * m));
                    N = 0.0;
                    S = 0.0;
                    ++gi;
                    gbeg = gend;
                    gend = gi < g2 ? s_gend[gi - g] : ncell;
                };
                for (int i0 = 0; i0 < ncell; i0 += kIlp) {
                    const int nu = min(kIlp, ncell - i0);
                    double t[kIlp];
#pragma unroll
                    for (int u = 0; u < kIlp; ++u)
                        t[u] = mfw_cell_t<T, DS, NDS>(min(i0 + u, ncell - 1), jc, k, ns, P, s_lv,
                                                      s_sl, pw, bd);
#pragma unroll
                    for (int u = 0; u < kIlp; ++u)
                        if (u < nu) {
                            const int i = i0 + u;
                            while (i == gend) finish();
                            const double w = s_w[i];
                            N += w;
                            S += w * t[u];
#pragma unroll
                            for (int q = 0; q < NDS; ++q)
                                unsafeAtomicAdd(&slot[s_sl[q][i] * bd + tid], w * t[u]);
                        }
                }
                while (gi < g2) finish();
            }
            g = g2;
        }
        if (on)
            for (int q = 0; q < nslot; ++q)
                dpart[(int64_t)(dpoff[r] + q) * k + j] = slot[q * bd + tid];
        __syncthreads();
    }
}

// Four consecutive columns of a row of T (the row's start need only be
// aligned to T: the vector type's alignment is T's, so the compiler emits
// the wide loads the target allows at that alignment).
template <typename T>
struct Vec4 {
    typedef T type __attribute__((ext_vector_type(4), aligned(sizeof(T))));
};

// The sparse dimensions' rows: (S P)_l = sum over l's cells of
// n_c (t_c - m_g(c)) = N_l P[l] + sum_c n_c (sum over the cell's other
// dimensions of P[a_cd] - m_g(c)), the cells in level order (the
// level-to-cells CSR), staged as in k_mfw_groups.  A lane takes four
// consecutive columns (one vector load per row it gathers; the last lane
// of a row whose width is not a multiple of four loads its columns one by
// one).
template <typename T, int DT>
__global__ void k_mfw_levels(int64_t nsl, const int *__restrict__ slev,
                             const int *__restrict__ lptr, const int *__restrict__ lcell,
                             const int *__restrict__ cgroup, const int *__restrict__ icodes,
                             int64_t ld, MfDims dm, const double *__restrict__ cn,
                             const T *__restrict__ P, const T *__restrict__ mg, int k,
                             T *__restrict__ AP, CgState st) {
    if (st.done && *st.done) return;
    typedef typename Vec4<T>::type V4;
    constexpr int DM = mfw_dmax<DT>();
    // the cells' levels in the dimensions other than the level's own
    __shared__ int s_lv[kMfMaxD - 1][kMfwStage];
    __shared__ int s_g[kMfwStage];
    __shared__ double s_w[kMfwStage];
    const int tid = threadIdx.x, bd = blockDim.x, D = DT > 0 ? DT : dm.D;
    for (int64_t li = blockIdx.x; li < nsl; li += gridDim.x) {
        const int64_t l = slev[li];
        int own = 0;   // the level's own dimension: the highest offset at or below l
        for (int d = 1; d < D; ++d)
            if (dm.off[d] <= l && dm.off[d] > dm.off[own]) own = d;
        const int q0 = lptr[l], q1 = lptr[l + 1];
        for (int j0 = 0; j0 < k; j0 += 4 * bd) {
            const int j = j0 + 4 * tid;
            const int nj = j < k ? min(4, k - j) : 0;   // this lane's columns
            double out[4] = {0.0, 0.0, 0.0, 0.0};
            double N = 0.0;
            for (int p0 = q0; p0 < q1; p0 += kMfwStage) {
                const int np = min(kMfwStage, q1 - p0);
                __syncthreads();
                for (int i = tid; i < np; i += bd) {
                    const int c = lcell[p0 + i];
                    int code[DM];
#pragma unroll
                    for (int d = 0; d < DM; ++d) code[d] = d < D ? icodes[d * ld + c] : 0;
                    const int gc = cgroup[c];
                    const double wc = cn[c];
#pragma unroll
                    for (int d = 0; d < DM; ++d)
                        if (d < D && d != own)
                            s_lv[d < own ? d : d - 1][i] = (int)(dm.off[d] + code[d]);
                    s_g[i] = gc;
                    s_w[i] = wc;
                }
                __syncthreads();
                for (int i = 0; i < np; ++i) N += s_w[i];
                if (!nj) continue;
                for (int i0 = 0; i0 < np; i0 += kMfwIlp) {
                    double t[kMfwIlp][4];
                    if (nj == 4) {
#pragma unroll
                        for (int u = 0; u < kMfwIlp; ++u) {
                            const int iu = min(i0 + u, np - 1);
                            const V4 m4 = *(const V4 *)(mg + (int64_t)s_g[iu] * k + j);
#pragma unroll
                            for (int c = 0; c < 4; ++c) t[u][c] = -(double)m4[c];
#pragma unroll
                            for (int e = 0; e < DM - 1; ++e)
                                if (DT > 0 || e < D - 1) {
                                    const V4 p4 = *(const V4 *)(P + (int64_t)s_lv[e][iu] * k + j);
#pragma unroll
                                    for (int c = 0; c < 4; ++c) t[u][c] += (double)p4[c];
                                }
                        }
                    } else {
                        // the row's last lane: its columns one by one, clamped
#pragma unroll
                        for (int u = 0; u < kMfwIlp; ++u) {
                            const int iu = min(i0 + u, np - 1);
#pragma unroll
                            for (int c = 0; c < 4; ++c) {
                                const int jj = j + min(c, nj - 1);
                                double x = -(double)mg[(int64_t)s_g[iu] * k + jj];
#pragma unroll
                                for (int e = 0; e < DM - 1; ++e)
                                    if (DT > 0 || e < D - 1)
                                        x += (double)P[(int64_t)s_lv[e][iu] * k + jj];
                                t[u][c] = x;
                            }
                        }
                    }
#pragma unroll
                    for (int u = 0; u < kMfwIlp; ++u)
                        if (i0 + u < np)
#pragma unroll
                            for (int c = 0; c < 4; ++c) out[c] += s_w[i0 + u] * t[u][c];
                }
            }
            for (int c = 0; c < nj; ++c)
                AP[l * k + j + c] = (T)(out[c] + N * (double)P[l * k + j + c]);
        }
    }
}

// The dense levels' rows: each one's dpart rows over the ranges covering
// it, in range order.
template <typename T>
__global__ void k_mfw_dense(int64_t ndl, const int *__restrict__ dlev,
                            const int *__restrict__ dcptr, const int *__restrict__ dcrow,
                            const double *__restrict__ dpart, int k, T *__restrict__ AP,
                            CgState st) {
    if (st.done && *st.done) return;
    const int tid = threadIdx.x, bd = blockDim.x;
    for (int64_t i = blockIdx.x; i < ndl; i += gridDim.x) {
        const int64_t l = dlev[i];
        for (int j = tid; j < k; j += bd) {
            double out = 0.0;
            for (int q = dcptr[i]; q < dcptr[i + 1]; ++q) out += dpart[(int64_t)dcrow[q] * k + j];
            AP[l * k + j] = (T)out;
        }
    }
}

// Each identifying cell's group.
__global__ void k_mfw_cell_groups(int64_t Gi, const int *__restrict__ igstart,
                                  int *__restrict__ cgroup) {
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < Gi;
         g += (int64_t)gridDim.x * blockDim.x)
        for (int c = igstart[g]; c < igstart[g + 1]; ++c) cgroup[c] = (int)g;
}

// Each group's lowest and highest code in dense dimension q (out (Gi, nds, 2);
// a group with no cells gets (INT_MAX, -1)).
__global__ void k_mfw_group_window(int64_t Gi, const int *__restrict__ igstart,
                                   const int *__restrict__ icodes, int64_t ld, MfList dd,
                                   int *__restrict__ out) {
    const int nds = dd.n;
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < Gi;
         g += (int64_t)gridDim.x * blockDim.x)
        for (int q = 0; q < nds; ++q) {
            const int d = dd.d[q];
            int lo = INT_MAX, hi = -1;
            for (int c = igstart[g]; c < igstart[g + 1]; ++c) {
                const int a = icodes[d * ld + c];
                lo = min(lo, a);
                hi = max(hi, a);
            }
            out[(g * nds + q) * 2] = lo;
            out[(g * nds + q) * 2 + 1] = hi;
        }
}

struct Work {
    int k = 0;
    int64_t L = 0;
    double *X64 = nullptr, *R64 = nullptr, *P64 = nullptr, *AP64 = nullptr, *B64 = nullptr;
    float *X32 = nullptr, *R32 = nullptr, *P32 = nullptr, *AP32 = nullptr;
    double *part = nullptr, *lpart = nullptr, *colbuf = nullptr;
    void *gpart = nullptr;  // blocked SpMM: the groups' partial AP (8, L, k)
    double *dwpart = nullptr;  // dense split rows, wide k: (kDwBlocks, nlong, k)
    double *sd = nullptr;  // rz alpha beta bnorm rel scale : 6k
    int *si = nullptr;     // active(k) zero(k) outer_active(k) done it it_final
    unsigned *cnt = nullptr;
    void release(hipStream_t s) {
        void *ps[] = {X64, R64, P64, AP64, B64, X32, R32, P32, AP32, part, lpart, colbuf, gpart,
                      dwpart, sd, si, cnt};
        for (void *p : ps) pfree(p, s);
        *this = Work();
    }
};

struct Matrix {
    int device = 0;
    int64_t L = 0, nnz = 0;
    int *indptr = nullptr, *indices = nullptr;
    // 64-bit row pointers (2^31 nonzeros and more, or asked for): indptr64
    // and chunk_beg64 replace indptr and chunk_beg, and only the plain CSR
    // forms run (no column-blocked copy, no dense split rows)
    bool p64 = false;
    int64_t *indptr64 = nullptr, *chunk_beg64 = nullptr;
    double *val64 = nullptr, *dinv64 = nullptr;
    float *val32 = nullptr, *dinv32 = nullptr;
    int nlong = 0, nchunk = 0;
    int *long_rows = nullptr, *long_cptr = nullptr, *chunk_row = nullptr, *chunk_beg = nullptr;
    // blocked copy of the non-split rows for the small-k SpMM (k_spmm_blk)
    bool blk = false, blk_tried = false;
    std::vector<int> ip_host;  // host row pointers (the blocked layout's balance)
    int col_lo_h[kNB + 1] = {0};
    int *col_lo = nullptr, *bptr = nullptr, *bidx = nullptr;
    // the tiled form's compact indices (when every block spans at most 65,536
    // columns and every 256-row stretch of a block holds under 65,536
    // nonzeros): tbp (kNB, stretches + 1) the stretches' starts, off16 (kNB L)
    // each row's start less its stretch's, bidx16 the column less the block's
    // first
    int *tbp = nullptr;
    uint16_t *off16 = nullptr, *bidx16 = nullptr;
    double *bval64 = nullptr;
    float *bval32 = nullptr;
    int64_t bnnz = 0;
    // dense long columns: the split rows taken by symmetry in the blocked SpMM
    bool dl = false;
    // dense split rows for wide k (k_dense_long_wide): -1 not yet looked at,
    // 0 not dense, 1 built
    int dw = -1;
    double *dw64 = nullptr;  // (L, nlong)
    float *dw32 = nullptr;
    int *dl_slot = nullptr, *dl_cptr = nullptr;  // (L) slot or -1; (kSmallK, nlong + 1)
    double *dt64 = nullptr, *ll64 = nullptr;     // (nlong, L) column-major; (nlong, nlong)
    float *dt32 = nullptr, *ll32 = nullptr;
    // the blocked SpMM's row parts per XCD, chosen by timing (0: the default),
    // per precision (0: FP64, 1: FP32) and k
    int blk_parts[2][kSmallK + 1] = {};
    // the column-blocked product's form per precision and k: 0 the default
    // (tiled where it applies), 1 groups of blocks + combine, 2 tiled
    int small_form[2][kSmallK + 1] = {};
    // the wide SpMM's form per (precision, k), chosen by timing: 0 a wave per
    // row over all columns (k_spmm_wide), 1 column slabs per XCD (k_spmm_slab)
    std::map<std::pair<int, int>, int> wide_mode;
    // captured graphs of batches of CG iterations, per (buffers, k, precision,
    // batch); a key's first use runs eagerly, the second captures
    std::map<std::array<uintptr_t, 9>, hipGraphExec_t> cg_graphs;
    std::map<std::array<uintptr_t, 9>, int> cg_seen;
    // CG solves run to the end on this matrix, per (k, precision): a large
    // system's batches become graphs once it has been solved twice with them
    std::map<std::pair<int, int>, int> cg_solves;
    void drop_graphs() {
        for (auto &kv : cg_graphs) (void)hipGraphExecDestroy(kv.second);
        cg_graphs.clear();
        cg_seen.clear();
    }
    hipStream_t stream = nullptr;  // the solver's own stream
    hipEvent_t ev = nullptr;
    MatFree *mf = nullptr;         // set: S P from the cells (no explicit S)
    Work w;
    std::mutex mu;
    // optional HIP-event timing of the SpMM launches of the CG iterations
    bool timing = false;
    std::vector<hipEvent_t> evpool;
    double spmm_ms = 0.0;
    int64_t spmm_n = 0;
    ~Matrix() {
        for (hipEvent_t e : evpool) (void)hipEventDestroy(e);
    }
};

static int build_blocked(Matrix &M, const std::vector<int> &ip32);
static int spmm_grid(int64_t L, int k);
static int tune_blocked_both(Matrix &M, int k);

static int alloc_work_buffers(Matrix &M, int k);

static int alloc_work(Matrix &M, int k) {
    // the blocked layout for small k, once, when P outgrows an XCD's L2
    if (k <= kSmallK && !M.blk_tried) {
        M.blk_tried = true;
        if (M.p64) return alloc_work_buffers(M, k);
        if (M.L * (int64_t)kSmallK * 8 > kL2Budget && M.L < (INT32_MAX - 1) / kNB) {
            M.drop_graphs();
            M.w.release(M.stream);
            const int st = build_blocked(M, M.ip_host);
            if (st) return st;
        }
    }
    // the split rows by column for wide k, once, when they are dense
    if (k > kSmallK && M.nlong && M.dw < 0) {
        M.dw = 0;
        if ((int64_t)M.ip_host.size() != M.L + 1) return alloc_work_buffers(M, k);
        int64_t long_nnz = 0;
        std::vector<int> lr(M.nlong);
        HIP_TRY(hipMemcpy(lr.data(), M.long_rows, M.nlong * 4, hipMemcpyDeviceToHost));
        for (int t = 0; t < M.nlong; ++t)
            long_nnz += (int64_t)M.ip_host[lr[t] + 1] - M.ip_host[lr[t]];
        if (M.nlong <= kDenseLongMax && 4 * long_nnz >= (int64_t)M.nlong * M.L) {
            M.drop_graphs();
            M.w.release(M.stream);
            const int64_t n = (int64_t)M.nlong * M.L;
            HIP_TRY(palloc(&M.dw64, n * 8, M.stream));
            HIP_TRY(palloc(&M.dw32, n * 4, M.stream));
            HIP_TRY(hipMemsetAsync(M.dw64, 0, n * 8, M.stream));
            HIP_TRY(hipMemsetAsync(M.dw32, 0, n * 4, M.stream));
            hipLaunchKernelGGL(k_dw_fill, dim3(64, M.nlong), dim3(kThreads), 0, M.stream, M.nlong,
                               M.L, M.long_rows, M.indptr, M.indices, M.val64, M.val32, M.dw64,
                               M.dw32);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipStreamSynchronize(M.stream));
            M.dw = 1;
        }
    }
    return alloc_work_buffers(M, k);
}

static int alloc_work_buffers(Matrix &M, int k) {
    if (M.w.k == k && M.w.L == M.L) return HDFE_STATUS_OK;
    M.drop_graphs();  // their kernels point at the work buffers
    M.w.release(M.stream);
    const int64_t Lk = std::max<int64_t>(1, M.L * (int64_t)k);
    Work &W = M.w;
    HIP_TRY(palloc(&W.X64, Lk * 8, M.stream));
    HIP_TRY(palloc(&W.R64, Lk * 8, M.stream));
    HIP_TRY(palloc(&W.P64, Lk * 8, M.stream));
    HIP_TRY(palloc(&W.AP64, Lk * 8, M.stream));
    HIP_TRY(palloc(&W.B64, Lk * 8, M.stream));
    HIP_TRY(palloc(&W.X32, Lk * 4, M.stream));
    HIP_TRY(palloc(&W.R32, Lk * 4, M.stream));
    HIP_TRY(palloc(&W.P32, Lk * 4, M.stream));
    HIP_TRY(palloc(&W.AP32, Lk * 4, M.stream));
    HIP_TRY(palloc(&W.part, (int64_t)kPartRows * k * 8, M.stream));
    const int64_t dlp = M.dl ? (int64_t)M.nlong * spmm_grid(M.L, k) : 0;
    HIP_TRY(palloc(&W.lpart, std::max<int64_t>(1, std::max<int64_t>(M.nchunk, dlp) * k) * 8, M.stream));
    HIP_TRY(palloc(&W.colbuf, (int64_t)k * 8, M.stream));
    if (M.blk && k <= kSmallK) HIP_TRY(palloc(&W.gpart, (int64_t)8 * Lk * 8, M.stream));
    if (M.dw == 1 && k > kSmallK)
        HIP_TRY(palloc(&W.dwpart, (int64_t)kDwBlocks * M.nlong * k * 8, M.stream));
    HIP_TRY(palloc(&W.sd, ((int64_t)6 * k + 1) * 8, M.stream));
    HIP_TRY(palloc(&W.si, ((int64_t)3 * k + 8) * 4, M.stream));
    HIP_TRY(palloc(&W.cnt, 8 * sizeof(unsigned), M.stream));
    HIP_TRY(hipMemsetAsync(W.cnt, 0, 8 * sizeof(unsigned), M.stream));
    HIP_TRY(hipMemsetAsync(W.sd, 0, ((int64_t)6 * k + 1) * 8, M.stream));
    W.k = k;
    W.L = M.L;
    if (M.mf) {
        MatFree &F = *M.mf;
        void *old[] = {F.v, F.cpart, F.mg, F.dpart};
        for (void *q : old) pfree(q, M.stream);
        F.v = nullptr;
        F.cpart = nullptr;
        F.mg = nullptr;
        F.dpart = nullptr;
        if (F.wide && k > kSmallK) {
            HIP_TRY(palloc(&F.mg, std::max<int64_t>(F.Gi * k, 1) * 8, M.stream));
            HIP_TRY(palloc(&F.dpart, std::max<int64_t>(F.dprows * k, 1) * 8, M.stream));
        } else {
            HIP_TRY(palloc(&F.v, std::max<int64_t>(F.Mc * k, 1) * 8, M.stream));
            HIP_TRY(palloc(&F.cpart, std::max<int64_t>(F.nchunk * k, 1) * 8, M.stream));
        }
        F.vk = k;
        return HDFE_STATUS_OK;
    }
    return tune_blocked_both(M, k);
}

static CgState make_state(Work &W, int k, double tol, int maxiter) {
    CgState st;
    st.rz = W.sd;
    st.alpha = W.sd + k;
    st.beta = W.sd + 2 * k;
    st.bnorm = W.sd + 3 * k;
    st.rel = W.sd + 4 * k;
    st.active = W.si;
    st.done = W.si + 3 * k;
    st.it = W.si + 3 * k + 1;
    st.it_final = W.si + 3 * k + 2;
    st.counter = W.cnt;
    st.dtol = W.sd + 6 * k;
    st.dmaxiter = W.si + 3 * k + 3;
    st.k = k;
    st.maxiter = maxiter;
    st.tol = tol;
    return st;
}

static int spmm_grid(int64_t L, int k) {
    return k <= kSmallK ? grid_for(L, 256, 1, 2048) : grid_for(L, 128, 1, 512);
}

// The tiled form's shape (k_spmm_tiles) from L, k and the precision alone
// (so that its sums never depend on timing): R x 256 rows a tile, R = 2 for
// FP64 with k >= 2 (measured faster on the 5M-worker system: 107 vs 118 us
// at k = 3, 172 vs 183 at k = 8), else 1 (77 vs 80 us FP32 at k = 3), raised
// to 4 if the tiles would outnumber the work buffers' partials (spmm_grid's
// count); 0 when even that does not fit.
static int tiles_shape(int64_t L, int k, bool fp64, int *r) {
    for (int R = fp64 && k >= 2 ? 2 : 1; R <= 4; R *= 2) {
        const int64_t nw = (L + R * (int64_t)kThreads - 1) / (R * (int64_t)kThreads);
        if (nw <= spmm_grid(L, k)) {
            *r = R;
            return (int)nw;
        }
    }
    return 0;
}
static int vec_grid(int64_t L, int k) {
    return grid_for(L * (int64_t)k, 2048, 1, k <= kSmallK ? 2048 : 512);
}

template <typename T>
static const T *blk_values(const Matrix &M);
template <>
const double *blk_values<double>(const Matrix &M) { return M.bval64; }
template <>
const float *blk_values<float>(const Matrix &M) { return M.bval32; }
template <typename T>
static const T *dl_values(const Matrix &M);
template <>
const double *dl_values<double>(const Matrix &M) { return M.dt64; }
template <>
const float *dl_values<float>(const Matrix &M) { return M.dt32; }
template <typename T>
static const T *ll_values(const Matrix &M);
template <>
const double *ll_values<double>(const Matrix &M) { return M.ll64; }
template <>
const float *ll_values<float>(const Matrix &M) { return M.ll32; }

// The blocked copy of the non-split rows (k_spmm_blk): kNB column blocks
// of about equal nonzero counts, on the matrix's stream.  S is symmetric, so
// a column's nonzeros are counted by its row's length (ip32, host).
static int build_blocked(Matrix &M, const std::vector<int> &ip32) {
    const int64_t L = M.L;
    hipStream_t ms = M.stream;
    Temp tmp;
    const int64_t total = (int64_t)ip32[L] - ip32[0];
    if (total == 0) return HDFE_STATUS_OK;
    // the split rows as dense long columns: few of them, each coupled to most
    // of the other rows (a small dimension's levels: years)
    std::vector<int> lr(M.nlong), slot_h(L, -1);
    int64_t long_nnz = 0;
    if (M.nlong) {
        HIP_TRY(hipMemcpy(lr.data(), M.long_rows, M.nlong * 4, hipMemcpyDeviceToHost));
        for (int t = 0; t < M.nlong; ++t) {
            slot_h[lr[t]] = t;
            long_nnz += (int64_t)ip32[lr[t] + 1] - ip32[lr[t]];
        }
    }
    M.dl = M.nlong >= 1 && M.nlong <= kDenseLongMax &&
           4 * long_nnz >= (int64_t)M.nlong * (L - M.nlong) && (int64_t)M.nlong * L < INT32_MAX;
    // col_lo[b] = the first column whose prefix count reaches b / kNB of the
    // total (S is symmetric: a column's count is its row's length); dense
    // long columns are not in the blocks
    int64_t btotal = total;
    if (M.dl) btotal -= long_nnz;
    M.col_lo_h[0] = 0;
    int64_t acc = 0, c = 0;
    for (int b = 1; b < kNB; ++b) {
        const int64_t target = btotal * b / kNB;
        while (c < L && acc < target) {
            if (!(M.dl && slot_h[c] >= 0)) acc += (int64_t)ip32[c + 1] - ip32[c];
            ++c;
        }
        M.col_lo_h[b] = (int)c;
    }
    if (M.dl) {
        // per k <= kSmallK, the combine's partials of split row t start at
        // t * (the combine's workgroups)
        std::vector<int> cptr((int64_t)kSmallK * (M.nlong + 1));
        for (int k = 1; k <= kSmallK; ++k)
            for (int t = 0; t <= M.nlong; ++t)
                cptr[(k - 1) * (M.nlong + 1) + t] = t * std::min(spmm_grid(L, k), kCombineMax);
        HIP_TRY(palloc(&M.dl_cptr, cptr.size() * 4, ms));
        HIP_TRY(hipMemcpyAsync(M.dl_cptr, cptr.data(), cptr.size() * 4, hipMemcpyHostToDevice, ms));
        HIP_TRY(palloc(&M.dl_slot, L * 4, ms));
        HIP_TRY(hipMemcpyAsync(M.dl_slot, slot_h.data(), L * 4, hipMemcpyHostToDevice, ms));
        HIP_TRY(palloc(&M.dt64, (int64_t)M.nlong * L * 8, ms));
        HIP_TRY(palloc(&M.dt32, (int64_t)M.nlong * L * 4, ms));
        HIP_TRY(palloc(&M.ll64, (int64_t)M.nlong * M.nlong * 8, ms));
        HIP_TRY(palloc(&M.ll32, (int64_t)M.nlong * M.nlong * 4, ms));
        HIP_TRY(hipMemsetAsync(M.dt64, 0, (int64_t)M.nlong * L * 8, ms));
        HIP_TRY(hipMemsetAsync(M.dt32, 0, (int64_t)M.nlong * L * 4, ms));
        HIP_TRY(hipMemsetAsync(M.ll64, 0, (int64_t)M.nlong * M.nlong * 8, ms));
        HIP_TRY(hipMemsetAsync(M.ll32, 0, (int64_t)M.nlong * M.nlong * 4, ms));
    }
    M.col_lo_h[kNB] = (int)L;
    HIP_TRY(palloc(&M.col_lo, (kNB + 1) * 4, ms));
    HIP_TRY(hipMemcpyAsync(M.col_lo, M.col_lo_h, (kNB + 1) * 4, hipMemcpyHostToDevice, ms));
    DevBuf<int> cnt;
    DEV_ALLOC(cnt, kNB * L + 1, ms);
    HIP_TRY(hipMemsetAsync(cnt.p + kNB * L, 0, 4, ms));
    const int gwave = (int)std::min<int64_t>((L + kThreads / kWave - 1) / (kThreads / kWave), 65536);
    hipLaunchKernelGGL(k_blk_rowcount, dim3(gwave), dim3(kThreads), 0, ms, L, M.indptr,
                       M.indices, M.col_lo, M.dl ? (const int *)M.dl_slot : nullptr, cnt.p);
    HIP_TRY(palloc(&M.bptr, (kNB * L + 1) * 4, ms));
    HIP_TRY(exclusive_sum(tmp, cnt.p, M.bptr, kNB * L + 1, ms));
    int last = 0;
    HIP_TRY(scalar_to_host(&last, M.bptr + kNB * L, ms));
    M.bnnz = last;
    // kTilePad spare entries: the tiled SpMM's last vector loads
    HIP_TRY(palloc(&M.bidx, (M.bnnz + kTilePad) * 4, ms));
    HIP_TRY(palloc(&M.bval64, (M.bnnz + kTilePad) * 8, ms));
    HIP_TRY(palloc(&M.bval32, (M.bnnz + kTilePad) * 4, ms));
    hipLaunchKernelGGL(k_blk_fill, dim3(gwave), dim3(kThreads), 0, ms, L, M.indptr, M.indices,
                       M.val64, M.val32, M.col_lo, M.bptr, M.bidx, M.bval64, M.bval32,
                       M.dl ? (const int *)M.dl_slot : nullptr, M.dt64, M.dt32);
    if (M.dl)
        hipLaunchKernelGGL(k_ll_fill, dim3(M.nlong * kLlParts), dim3(kThreads), 0, ms, M.nlong,
                           M.long_rows, M.indptr, M.indices, M.val64, M.val32, M.dl_slot, M.ll64, M.ll32);
    HIP_TRY(hipGetLastError());
    // the tiled form's 16-bit indices, when every block spans at most 65,536
    // columns and every stretch of 256 rows holds under 65,536 of a block's
    // nonzeros (38 of the 64 MB of 32-bit indices and row pointers at 5M
    // workers)
    bool narrow = M.bnnz > 0;
    for (int b = 0; b < kNB; ++b)
        if (M.col_lo_h[b + 1] - M.col_lo_h[b] > 65536) narrow = false;
    if (narrow) {
        const int64_t nt = (L + kThreads - 1) / kThreads;
        HIP_TRY(palloc(&M.tbp, kNB * (nt + 1) * 4, ms));
        hipLaunchKernelGGL(k_tbp_fill, dim3(grid_for(kNB * (nt + 1), 256, 1, 4096)), dim3(256), 0, ms,
                           L, nt, M.bptr, M.tbp);
        std::vector<int> tb(kNB * (nt + 1));
        HIP_TRY(hipMemcpyAsync(tb.data(), M.tbp, tb.size() * 4, hipMemcpyDeviceToHost, ms));
        HIP_TRY(hipStreamSynchronize(ms));
        for (int b = 0; b < kNB && narrow; ++b)
            for (int64_t t = 0; t < nt; ++t)
                if (tb[b * (nt + 1) + t + 1] - tb[b * (nt + 1) + t] > 65535) {
                    narrow = false;
                    break;
                }
        if (narrow) {
            HIP_TRY(palloc(&M.off16, kNB * L * 2, ms));
            HIP_TRY(palloc(&M.bidx16, (M.bnnz + kTilePad) * 2, ms));
            hipLaunchKernelGGL(k_off16_fill, dim3(grid_for(kNB * L, 256, 1, 16384)), dim3(256), 0,
                               ms, L, nt, M.bptr, M.tbp, M.off16);
            hipLaunchKernelGGL(k_bidx16_fill, dim3(grid_for(M.bnnz, 256, 1, 16384)), dim3(256), 0,
                               ms, M.bnnz, M.bidx, M.col_lo, M.bidx16);
            HIP_TRY(hipGetLastError());
        } else {
            pfree(M.tbp, ms);
            M.tbp = nullptr;
        }
    }
    HIP_TRY(hipStreamSynchronize(ms));
    M.blk = true;
    return HDFE_STATUS_OK;
}

// The plain CSR forms (no column-blocked copy): k <= 8 a few lanes per row,
// wide k a wave per row or column slabs, split rows in chunks; I the row
// pointers' type.
template <typename T, typename I>
static void launch_plain(Matrix &M, const I *ip, const I *cb, const T *val, const T *P, T *AP,
                         double *part, double *lpart, int with_dot, CgState st, hipStream_t s,
                         int force_wide, int g) {
    const int k = st.k;
    if (k <= kSmallK) {
#define HDFE_SPMM(KK)                                                                          \
    case KK:                                                                                   \
        hipLaunchKernelGGL((k_spmm_small<T, KK, 8, I>), dim3(g + M.nchunk), dim3(kThreads), 0, s, \
                           ip, M.indices, val, M.L, P, AP, part, with_dot, M.chunk_row,  \
                           cb, M.nchunk, M.long_rows, M.long_cptr, M.nlong, lpart,    \
                           st);                                                                \
        if (with_dot || M.nlong)                                                               \
            hipLaunchKernelGGL((k_fin_small<T, KK>), dim3(1), dim3(kFinThreads), 0, s, P, AP,  \
                               part, g, with_dot, M.long_rows, M.long_cptr, M.nlong, lpart,    \
                               st);                                                            \
        break;
        switch (k) {
            HDFE_SPMM(1) HDFE_SPMM(2) HDFE_SPMM(3) HDFE_SPMM(4)
            HDFE_SPMM(5) HDFE_SPMM(6) HDFE_SPMM(7) HDFE_SPMM(8)
        }
#undef HDFE_SPMM
        return;
    }
    if (M.nlong && M.dw == 1 && M.w.dwpart) {
        const T *dw = sizeof(T) == 4 ? (const T *)M.dw32 : (const T *)M.dw64;
        hipLaunchKernelGGL((k_dense_long_wide<T>), dim3(kDwBlocks, (k + kWave - 1) / kWave),
                           dim3(kThreads), 0, s, M.L, M.nlong, dw, P, M.w.dwpart, st);
        hipLaunchKernelGGL((k_fin_dense_long<T>), dim3((M.nlong * k + kThreads - 1) / kThreads),
                           dim3(kThreads), 0, s, M.nlong, M.long_rows, M.w.dwpart, AP, st);
    } else if (M.nlong) {
        hipLaunchKernelGGL((k_long_wide<T, I>), dim3(M.nchunk), dim3(kThreads), 4 * k * 8, s,
                           ip, M.indices, val, M.chunk_row, cb, M.nchunk,
                           M.long_rows, M.long_cptr, M.nlong, P, AP, lpart, st);
        hipLaunchKernelGGL((k_fin_long_wide<T>), dim3(1), dim3(kFinThreads), 0, s, M.long_rows,
                           M.long_cptr, M.nlong, lpart, AP, st);
    }
    int mode = force_wide;
    if (mode < 0) {
        auto it = M.wide_mode.find({sizeof(T) == 4 ? 1 : 0, k});
        mode = it == M.wide_mode.end() ? 0 : it->second;
    }
    const int nslab = (k + slab_w<T>() - 1) / slab_w<T>(), rounds = (nslab + 7) / 8;
    // the slab form's pAp partials, one row of k per (round, workgroup), must
    // fit the work buffer's kPartRows rows: fewer workgroups per XCD for wide
    // k, and the wave-per-row form when not even eight per XCD would fit
    const int gs = std::min(8 * kSlabGroups, (kPartRows / rounds) / 8 * 8);
    if (mode == 1 && gs >= 64) {
        for (int r = 0; r < rounds; ++r) {
            const int slab0 = 8 * r, ns = std::min(8, nslab - slab0);
            hipLaunchKernelGGL((k_spmm_slab<T, I>), dim3(gs), dim3(kThreads), 0, s, ip,
                               M.indices, val, M.L, P, AP, part + (int64_t)r * gs * k, with_dot,
                               slab0, ns, st);
        }
        if (with_dot)
            hipLaunchKernelGGL(k_fin_alpha, dim3(1), dim3(kFinThreads), 0, s, part, rounds * gs,
                               st);
        return;
    }
    const int cpl = (k + 63) / 64;
#define HDFE_WIDE(C)                                                                        \
    hipLaunchKernelGGL((k_spmm_wide<T, C, I>), dim3(g), dim3(kThreads), 4 * k * 8, s, ip, \
                       M.indices, val, M.L, P, AP, part, with_dot, st)
    if (cpl <= 1) HDFE_WIDE(1);
    else if (cpl <= 2) HDFE_WIDE(2);
    else if (cpl <= 3) HDFE_WIDE(3);
    else if (cpl <= 4) HDFE_WIDE(4);
    else if (cpl <= 5) HDFE_WIDE(5);
    else if (cpl <= 6) HDFE_WIDE(6);
    else if (cpl <= 8) HDFE_WIDE(8);
    else HDFE_WIDE(16);
#undef HDFE_WIDE
    if (with_dot)
        hipLaunchKernelGGL(k_fin_alpha, dim3(1), dim3(kFinThreads), 0, s, part, g, st);
}

// AP = S P (and, with with_dot, pAp sums and alpha).  `st.done` may be null.
template <typename T>
static void launch_spmm(Matrix &M, const T *val, const T *P, T *AP, double *part, double *lpart,
                        int with_dot, CgState st, hipStream_t s, int force_parts = 0,
                        int force_wide = -1) {
    const int k = st.k;
    if (M.mf && M.mf->wide && k > kSmallK) {
        MatFree &F = *M.mf;
        T *mg = (T *)F.mg;
        const int bd = (int)std::min<int64_t>(1024, (k + kWave - 1) / kWave * kWave);
        // k_mfw_groups' lanes: its slots (double) and the window's rows of P
        // (T) in dynamic LDS, with the stage, within kMfwLds; the columns
        // split evenly over the passes a workgroup makes
        int bg = std::min(bd, kMfwLanes);
        while (bg > kWave &&
               (int64_t)F.maxslot * bg * (8 + sizeof(T)) + kMfwStageBytes > kMfwLds)
            bg -= kWave;
        {
            const int passes = (k + bg - 1) / bg;
            bg = (int)std::min<int64_t>(bg, ((k + passes - 1) / passes + kWave - 1) / kWave * kWave);
        }
        const size_t dyn = (size_t)F.maxslot * bg * (8 + sizeof(T));
        // k_mfw_levels: four columns a lane
        const int bl = (int)std::min<int64_t>(1024, ((k + 3) / 4 + kWave - 1) / kWave * kWave);
        // the sparse dimensions' count as a template parameter up to four
#define HDFE_MFW(SS, NN)                                                                           \
    do {                                                                                           \
        if (dyn > 65536)                                                                           \
            (void)hipFuncSetAttribute((const void *)k_mfw_groups<T, SS, NN>,                       \
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);       \
        hipLaunchKernelGGL((k_mfw_groups<T, SS, NN>), dim3(F.nr), dim3(bg), dyn, s, F.rng, F.rbase, \
                           F.rsoff, F.dpoff, F.igstart, F.icodes, F.ld, F.sd, F.offs_d, F.dd.d[0], \
                           F.dd.d[1], F.maxslot, F.cn, P, k, mg, F.dpart, st);                     \
    } while (0)
#define HDFE_MFWN(SS)                          \
    switch (F.dd.n) {                          \
        case 0: HDFE_MFW(SS, 0); break;        \
        case 1: HDFE_MFW(SS, 1); break;        \
        default: HDFE_MFW(SS, 2); break;       \
    }
        // the sparse and the dense dimensions' counts as template parameters
        if (F.nr) {
            switch (F.sd.n) {
                case 1: HDFE_MFWN(1); break;
                case 2: HDFE_MFWN(2); break;
                case 3: HDFE_MFWN(3); break;
                default: HDFE_MFWN(0); break;
            }
        }
#undef HDFE_MFWN
#undef HDFE_MFW
#define HDFE_MFL(DD)                                                                              \
    hipLaunchKernelGGL((k_mfw_levels<T, DD>), dim3((int)std::min<int64_t>(F.nsl, 16384)), dim3(bl), \
                       0, s, F.nsl, F.slev, F.lptr, F.lcell, F.cgroup, F.icodes, F.ld, F.dm, F.cn,  \
                       P, (const T *)mg, k, AP, st)
        if (F.nsl) {
            switch (F.dm.D) {
                case 1: HDFE_MFL(1); break;
                case 2: HDFE_MFL(2); break;
                case 3: HDFE_MFL(3); break;
                case 4: HDFE_MFL(4); break;
                default: HDFE_MFL(0); break;
            }
        }
#undef HDFE_MFL
        if (F.ndl)
            hipLaunchKernelGGL((k_mfw_dense<T>), dim3((int)std::min<int64_t>(F.ndl, 4096)), dim3(bd),
                               0, s, F.ndl, F.dlev, F.dcptr, F.dcrow, F.dpart, k, AP, st);
        if (with_dot) {
            const int gv = vec_grid(M.L, k);
            hipLaunchKernelGGL((k_colsum<T>), dim3(gv), dim3(kThreads), 0, s, M.L, k, P, AP,
                               (const T *)nullptr, part);
            hipLaunchKernelGGL(k_fin_alpha, dim3(1), dim3(kFinThreads), 0, s, part, gv, st);
        }
        (void)force_parts;
        (void)force_wide;
        (void)val;
        (void)lpart;
        return;
    }
    if (M.mf) {
        MatFree &F = *M.mf;
        T *v = (T *)F.v;
        const int gc = grid_for(std::max<int64_t>(F.Gi * k, 1), 1024, 1, 8192);
        hipLaunchKernelGGL((k_mf_cells<T>), dim3(gc), dim3(kThreads), 0, s, F.Gi, F.igstart,
                           F.icodes, F.ld, F.dm, F.cn, P, k, v, st);
        const int gk = (int)std::min<int64_t>(std::max<int64_t>((F.nchunk + 3) / 4, 1), 16384);
        hipLaunchKernelGGL((k_mf_chunks<T>), dim3(gk), dim3(kThreads), 0, s, F.nchunk, F.ch_beg,
                           F.ch_end, F.lcell, v, k, F.cpart, st);
        const int gf = grid_for(std::max<int64_t>(M.L * k, 1), 1024, 1, 8192);
        hipLaunchKernelGGL((k_mf_finish<T>), dim3(gf), dim3(kThreads), 0, s, M.L, F.lch, F.cpart,
                           k, AP, st);
        if (with_dot) {
            const int gv = vec_grid(M.L, k);
            hipLaunchKernelGGL((k_colsum<T>), dim3(gv), dim3(kThreads), 0, s, M.L, k, P, AP,
                               (const T *)nullptr, part);
            hipLaunchKernelGGL(k_fin_alpha, dim3(1), dim3(kFinThreads), 0, s, part, gv, st);
        }
        (void)force_parts;
        (void)force_wide;
        (void)val;
        (void)lpart;
        return;
    }
    const int g = spmm_grid(M.L, k);
    if (k <= kSmallK && M.blk && M.L * (int64_t)k * (int64_t)sizeof(T) > kL2Budget &&
        M.w.gpart) {
        // the fewest groups whose widest P slice fits the L2 budget
        int ngrp = 8;
        for (int ng = 2; ng <= 8; ng *= 2) {
            int64_t widest = 0;
            for (int q = 0; q < ng; ++q)
                widest = std::max<int64_t>(
                    widest, (int64_t)M.col_lo_h[(q + 1) * kNB / ng] - M.col_lo_h[q * kNB / ng]);
            if (widest * k * (int64_t)sizeof(T) <= kL2Budget) {
                ngrp = ng;
                break;
            }
        }
        const double per_row = (double)M.bnnz / ((double)std::max<int64_t>(M.L, 1) * ngrp);
        const int G = per_row < 4.0 ? 1 : (per_row < 12.0 ? 2 : 4);
        const int tuned = M.blk_parts[sizeof(T) == 4 ? 1 : 0][k];
        const int nblk = 8 * (force_parts ? force_parts : (tuned ? tuned : kBlkParts));
        const T *bval = blk_values<T>(M);
        T *gp = (T *)M.w.gpart;
        // dense long columns: the combine's partials of the split rows, g per row
        const int nld = M.dl ? M.nlong : 0;
        const T *dt = dl_values<T>(M), *ll = M.dl ? ll_values<T>(M) : nullptr;
        const int *lcp = M.dl ? M.dl_cptr + (k - 1) * (M.nlong + 1) : M.long_cptr;
        const int nchunk = M.dl ? 0 : M.nchunk;
        // the tiled form (the default): one pass over the rows, blocks in order
        const int prec = sizeof(T) == 4 ? 1 : 0;
        int rt = 0;
        const int nwt = M.small_form[prec][k] != 1 && (M.dl || M.nlong == 0)
                            ? tiles_shape(M.L, k, sizeof(T) == 8, &rt)
                            : 0;
        // form 3: the tiled form with the 32-bit indices (for comparison)
        const bool c16 = M.small_form[prec][k] != 3;
        if (nwt > 0) {
#define HDFE_TL(KK, RR)                                                                          \
    hipLaunchKernelGGL((k_spmm_tiles<T, KK, RR>), dim3(nwt), dim3(kThreads), 0, s, M.bptr,       \
                       M.bidx, bval, M.L, P, AP, part, with_dot, nld, dt, M.long_rows, lpart, st,  \
                       M.tbp, c16 ? M.off16 : nullptr, M.bidx16, M.col_lo)
#define HDFE_TLG(KK)                                                                               \
    case KK:                                                                                       \
        if (rt == 1) { HDFE_TL(KK, 1); }                                                           \
        else if (rt == 2) { HDFE_TL(KK, 2); }                                                      \
        else { HDFE_TL(KK, 4); }                                                                   \
        if (with_dot || M.nlong)                                                                   \
            hipLaunchKernelGGL((k_fin_small<T, KK>), dim3(1), dim3(kFinThreads), 0, s, P, AP, part, \
                               nwt, with_dot, M.long_rows, (const int *)nullptr, M.nlong, lpart,   \
                               st, ll);                                                           \
        break;
            switch (k) {
                HDFE_TLG(1) HDFE_TLG(2) HDFE_TLG(3) HDFE_TLG(4)
                HDFE_TLG(5) HDFE_TLG(6) HDFE_TLG(7) HDFE_TLG(8)
            }
#undef HDFE_TLG
#undef HDFE_TL
            return;
        }
        // the combine in at most 512 workgroups (a few rows per thread): the
        // one-workgroup finish then adds 512 partials per split row, not 2,048
        const int gc = std::min(g, kCombineMax);
#define HDFE_BLK(KK, GG)                                                                          \
    hipLaunchKernelGGL((k_spmm_blk<T, KK, GG>), dim3(nblk + nchunk), dim3(kThreads), 0, s,        \
                       M.bptr, M.bidx, bval, M.L, ngrp, nblk, P, gp, M.indptr, M.indices, val,   \
                       M.chunk_row, M.chunk_beg, lpart, st)
#define HDFE_BLKG(KK)                   \
    case KK:                            \
        if (G == 1) HDFE_BLK(KK, 1);    \
        else if (G == 2) HDFE_BLK(KK, 2); \
        else HDFE_BLK(KK, 4);           \
        hipLaunchKernelGGL((k_spmm_combine<T, KK>), dim3(gc), dim3(kThreads), 0, s, M.L, ngrp, gp, \
                           P, AP, part, with_dot, nld, dt, M.long_rows, lpart, st);             \
        if (with_dot || M.nlong)                                                                  \
            hipLaunchKernelGGL((k_fin_small<T, KK>), dim3(1), dim3(kFinThreads), 0, s, P, AP, part, \
                               gc, with_dot, M.long_rows, lcp, M.nlong, lpart, st, ll);           \
        break;
        switch (k) {
            HDFE_BLKG(1) HDFE_BLKG(2) HDFE_BLKG(3) HDFE_BLKG(4)
            HDFE_BLKG(5) HDFE_BLKG(6) HDFE_BLKG(7) HDFE_BLKG(8)
        }
#undef HDFE_BLKG
#undef HDFE_BLK
        return;
    }
    if (M.p64)
        launch_plain<T, int64_t>(M, M.indptr64, M.chunk_beg64, val, P, AP, part, lpart, with_dot,
                                 st, s, force_wide, g);
    else
        launch_plain<T, int>(M, M.indptr, M.chunk_beg, val, P, AP, part, lpart, with_dot, st, s,
                             force_wide, g);
}

template <typename T>
static void launch_colsum(int64_t L, int k, const T *A, const T *B, const T *d, double *part,
                          double *out, int mode, unsigned *counter, hipStream_t s) {
    (void)counter;
    const int g = vec_grid(L, k);
    hipLaunchKernelGGL((k_colsum<T>), dim3(g), dim3(kThreads), 0, s, L, k, A, B, d, part);
    hipLaunchKernelGGL(k_fin_colsum, dim3(1), dim3(kFinThreads), 0, s, k, part, g, out, mode);
}

// Iterate the CG already set up in (X, R, P, st.rz, st.active, st.bnorm);
// returns the iterations run and whether every column converged.
template <typename T>
static int run_cg(Matrix &M, int k, const T *val, const T *dinv, T *X, T *R, T *P, T *AP,
                  CgState st, hipStream_t s, int64_t *it_out, int *conv_out) {
    Work &W = M.w;
    HIP_TRY(hipMemsetAsync(st.done, 0, 3 * sizeof(int), s));
    hipLaunchKernelGGL(k_set_params, dim3(1), dim3(64), 0, s, st, st.tol, st.maxiter);
    const int gv = vec_grid(M.L, k);
    const int gd = grid_for(M.L * (int64_t)k, 1024, 1, 4096);
    // one iteration: SpMM (+ pAp, alpha), step, its finish (rel, stop, beta),
    // direction -- every kernel returns at once after the stop
    auto iteration = [&](int b) {
        if (M.timing) (void)hipEventRecord(M.evpool[2 * b], s);
        launch_spmm<T>(M, val, P, AP, W.part, W.lpart, 1, st, s);
        if (M.timing) (void)hipEventRecord(M.evpool[2 * b + 1], s);
        hipLaunchKernelGGL((k_step<T>), dim3(gv), dim3(kThreads), 0, s, M.L, X, R, P, AP, dinv,
                           W.part, st);
        hipLaunchKernelGGL(k_fin_step, dim3(1), dim3(kFinThreads), 0, s, W.part, gv, st);
        hipLaunchKernelGGL((k_direction<T>), dim3(gd), dim3(kThreads), 0, s, M.L, P, R, dinv,
                           st);
    };
    // Graphs pay where launching an iteration's six kernels costs more host
    // time than they run on the device: small systems.  A large one keeps
    // the device busy from eager launches, and capturing and instantiating a
    // graph there idles it (0.3-0.4 ms per capture measured on 400,000
    // levels): its batches run eagerly until the matrix has been solved twice
    // with this k and precision (the leave-out paths solve one matrix
    // hundreds of times; a fit builds a new one each time), then as graphs.
    const int64_t iter_bytes = M.nnz * (int64_t)(4 + sizeof(T)) + M.L * (int64_t)k * sizeof(T) * 6;
    const std::pair<int, int> sk = {k, (int)sizeof(T)};
    const bool use_graphs = !M.timing && (iter_bytes < kGraphMaxBytes || M.cg_solves[sk] >= 2);
    int host[2] = {0, 0};
    int batch = 8;
    int64_t done_before = 0;  // iterations launched before this batch
    // Eager batches are sized from the convergence so far: the worst
    // column's relative residual at the last two polls (1 before the first
    // iteration) gives a rate per iteration, and the next batch is the
    // iterations that rate needs to reach the tolerance, plus one.  Every
    // iteration after the stop returns at once but still costs its launches,
    // so this trims the overshoot that doubling leaves.  Graph batches keep
    // the doubling sizes, whose graphs are reused.
    std::vector<double> relh(k, 0.0);
    double rel_prev = 1.0;
    for (int64_t launched = 0; launched < st.maxiter;) {
        const int nb = (int)std::min<int64_t>(batch, st.maxiter - launched);
        if (M.timing)
            while ((int)M.evpool.size() < 2 * nb) {
                hipEvent_t e;
                HIP_TRY(hipEventCreate(&e));
                M.evpool.push_back(e);
            }
        hipGraphExec_t ex = nullptr;
        if (use_graphs) {
            // a batch of nb iterations as a graph: its kernels' arguments are
            // fixed for these buffers (tol and maxiter are read on the device)
            const std::array<uintptr_t, 9> key = {
                (uintptr_t)val, (uintptr_t)dinv, (uintptr_t)X, (uintptr_t)R, (uintptr_t)P,
                (uintptr_t)AP, (uintptr_t)k, sizeof(T), (uintptr_t)nb};
            auto it = M.cg_graphs.find(key);
            if (it != M.cg_graphs.end()) {
                ex = it->second;
            } else if (M.cg_seen[key]++ >= 1) {
                hipGraph_t g = nullptr;
                HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                for (int b = 0; b < nb; ++b) iteration(b);
                const hipError_t ce = hipStreamEndCapture(s, &g);
                HIP_TRY(ce);
                HIP_TRY(hipGetLastError());
                const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                HIP_TRY(ie);
                M.cg_graphs[key] = ex;
            }
        }
        if (ex) {
            HIP_TRY(hipGraphLaunch(ex, s));
        } else {
            for (int b = 0; b < nb; ++b) iteration(b);
        }
        done_before = launched;
        launched += nb;
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(host, st.done, sizeof(int), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(host + 1, st.it_final, sizeof(int), hipMemcpyDeviceToHost, s));
        if (!use_graphs)
            HIP_TRY(hipMemcpyAsync(relh.data(), st.rel, k * sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (M.timing) {
            // iterations that ran (the ones after the stop return at once)
            const int64_t ran = host[0] ? std::min<int64_t>(nb, host[1] - done_before) : nb;
            for (int64_t b = 0; b < ran; ++b) {
                float ms = 0.f;
                HIP_TRY(hipEventElapsedTime(&ms, M.evpool[2 * b], M.evpool[2 * b + 1]));
                M.spmm_ms += ms;
                M.spmm_n += 1;
            }
        }
        if (host[0]) break;
        if (use_graphs) {
            if (batch < 64) batch *= 2;
        } else {
            double r = 0.0;
            for (int j = 0; j < k; ++j) r = std::max(r, relh[j]);
            int next = std::min(64, 2 * batch);
            if (r > 0.0 && r < rel_prev && st.tol > 0.0 && st.tol < r) {
                const double rate = std::pow(r / rel_prev, 1.0 / nb);
                const double need = std::log(st.tol / r) / std::log(rate);
                if (std::isfinite(need)) next = (int)std::min(64.0, std::max(2.0, std::ceil(need) + 1.0));
            }
            rel_prev = r;
            batch = next;
        }
    }
    ++M.cg_solves[sk];
    if (!host[0]) {  // maxiter reached without the flag (maxiter == 0)
        *it_out = 0;
        *conv_out = 0;
        return HDFE_STATUS_OK;
    }
    *it_out = host[1];
    *conv_out = host[0] == 1;
    return HDFE_STATUS_OK;
}

static int read_max(const double *dev, int k, hipStream_t s, double *out,
                    std::vector<double> *all = nullptr) {
    std::vector<double> h(k);
    HIP_TRY(hipMemcpyAsync(h.data(), dev, k * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    double m = 0.0;
    for (double v : h) m = std::max(m, v);
    *out = m;
    if (all) *all = h;
    return HDFE_STATUS_OK;
}

// The blocked SpMM's launch shape by timing: the row parts per XCD (how many
// workgroups share an XCD's column group) that give the shortest product on
// this matrix, once per precision and k.  Only the row partition changes --
// every row's sum, and every partial's order, is the same whatever is chosen
// -- so the results are bit-identical for any choice.  Kept process-wide per
// matrix shape, so repeated fits of one panel time it once.
static std::mutex g_tune_mu;
static std::map<std::array<int64_t, 6>, int> g_tune;
constexpr int kTuneParts[] = {64, 128, 256, 512, 1024};

template <typename T>
static int tune_blocked(Matrix &M, int k) {
    if (!(k <= kSmallK && M.blk && M.L * (int64_t)k * (int64_t)sizeof(T) > kL2Budget &&
          M.w.gpart))
        return HDFE_STATUS_OK;
    const int prec = sizeof(T) == 4 ? 1 : 0;
    // the tiled form, where it is taken, has no launch shape to choose
    int rt = 0;
    if (M.small_form[prec][k] != 1 && (M.dl || M.nlong == 0) &&
        tiles_shape(M.L, k, sizeof(T) == 8, &rt) > 0)
        return HDFE_STATUS_OK;
    const std::array<int64_t, 6> key = {M.L, M.nnz, M.bnnz, M.nlong, k, prec};
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        auto it = g_tune.find(key);
        if (it != g_tune.end()) {
            M.blk_parts[prec][k] = it->second;
            return HDFE_STATUS_OK;
        }
    }
    hipStream_t s = M.stream;
    T *P = sizeof(T) == 4 ? (T *)M.w.P32 : (T *)M.w.P64;
    T *AP = sizeof(T) == 4 ? (T *)M.w.AP32 : (T *)M.w.AP64;
    const T *val = sizeof(T) == 4 ? (const T *)M.val32 : (const T *)M.val64;
    HIP_TRY(hipMemsetAsync(P, 0, M.L * (int64_t)k * sizeof(T), s));
    CgState cs = make_state(M.w, k, 0.0, 1);
    cs.done = nullptr;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    float best = 1e30f;
    int choice = kBlkParts;
    for (int parts : kTuneParts) {
        launch_spmm<T>(M, val, P, AP, M.w.part, M.w.lpart, 1, cs, s, parts);
        HIP_TRY(hipEventRecord(e0, s));
        for (int r = 0; r < 3; ++r) launch_spmm<T>(M, val, P, AP, M.w.part, M.w.lpart, 1, cs, s, parts);
        HIP_TRY(hipEventRecord(e1, s));
        HIP_TRY(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) {
            best = ms;
            choice = parts;
        }
    }
    HIP_TRY(hipEventDestroy(e0));
    HIP_TRY(hipEventDestroy(e1));
    HIP_TRY(hipGetLastError());
    M.blk_parts[prec][k] = choice;
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tune[key] = choice;
    return HDFE_STATUS_OK;
}

// The wide SpMM's form (k > kSmallK), per precision: each candidate timed
// with HIP events over three launches after one untimed, the fastest kept,
// process-wide per matrix shape (size, nonzeros, split rows, k).
static std::map<std::array<int64_t, 5>, int> g_wide_tune;

template <typename T>
static int tune_wide(Matrix &M, int k) {
    if (k <= kSmallK) return HDFE_STATUS_OK;
    const int prec = sizeof(T) == 4 ? 1 : 0;
    const std::array<int64_t, 5> key = {M.L, M.nnz, M.nlong, k, prec};
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        auto it = g_wide_tune.find(key);
        if (it != g_wide_tune.end()) {
            M.wide_mode[{prec, k}] = it->second;
            return HDFE_STATUS_OK;
        }
    }
    hipStream_t s = M.stream;
    T *P = sizeof(T) == 4 ? (T *)M.w.P32 : (T *)M.w.P64;
    T *AP = sizeof(T) == 4 ? (T *)M.w.AP32 : (T *)M.w.AP64;
    const T *val = sizeof(T) == 4 ? (const T *)M.val32 : (const T *)M.val64;
    HIP_TRY(hipMemsetAsync(P, 0, M.L * (int64_t)k * sizeof(T), s));
    CgState cs = make_state(M.w, k, 0.0, 1);
    cs.done = nullptr;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    float best = 1e30f;
    int choice = 0;
    for (int mode = 0; mode < 2; ++mode) {
        launch_spmm<T>(M, val, P, AP, M.w.part, M.w.lpart, 1, cs, s, 0, mode);
        HIP_TRY(hipEventRecord(e0, s));
        for (int r = 0; r < 3; ++r)
            launch_spmm<T>(M, val, P, AP, M.w.part, M.w.lpart, 1, cs, s, 0, mode);
        HIP_TRY(hipEventRecord(e1, s));
        HIP_TRY(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) {
            best = ms;
            choice = mode;
        }
    }
    HIP_TRY(hipEventDestroy(e0));
    HIP_TRY(hipEventDestroy(e1));
    HIP_TRY(hipGetLastError());
    M.wide_mode[{prec, k}] = choice;
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_wide_tune[key] = choice;
    return HDFE_STATUS_OK;
}

static int tune_blocked_both(Matrix &M, int k) {
    TRY(tune_blocked<double>(M, k));
    TRY(tune_blocked<float>(M, k));
    TRY(tune_wide<double>(M, k));
    return tune_wide<float>(M, k);
}

// FP64 block PCG (core/solve.py:66-140).
static int solve_f64(Matrix &M, int k, const uint8_t *zero_h, bool has_x0, double tol,
                     int64_t maxiter, int64_t *it_out, int *conv_out, double *rel_out,
                     hipStream_t s) {
    Work &W = M.w;
    CgState st = make_state(W, k, tol, (int)std::min<int64_t>(maxiter, INT32_MAX));
    const int64_t Lk = M.L * (int64_t)k;
    std::vector<int> zero(k), act(k);
    bool all_zero = true;
    for (int j = 0; j < k; ++j) {
        zero[j] = zero_h ? zero_h[j] != 0 : 0;
        act[j] = !zero[j];
        all_zero &= (bool)zero[j];
    }
    int *zero_d = W.si + k;  // the zero mask lives after `active`
    HIP_TRY(hipMemcpyAsync(zero_d, zero.data(), k * 4, hipMemcpyHostToDevice, s));
    launch_colsum<double>(M.L, k, W.B64, W.B64, nullptr, W.part, st.bnorm, 1, W.cnt + 4, s);
    const int ge = grid_for(Lk, 1024, 1, 4096);
    CgState plain = st;
    plain.done = nullptr;
    if (has_x0) {
        launch_spmm<double>(M, M.val64, W.X64, W.AP64, W.part, W.lpart, 0, plain, s);
        hipLaunchKernelGGL(k_set_start<double>, dim3(ge), dim3(kThreads), 0, s, M.L, k, W.B64,
                           W.AP64, zero_d, W.X64, W.R64, W.P64, M.dinv64);
    } else {
        HIP_TRY(hipMemsetAsync(W.X64, 0, Lk * 8, s));
        hipLaunchKernelGGL(k_set_start<double>, dim3(ge), dim3(kThreads), 0, s, M.L, k, W.B64,
                           (const double *)nullptr, zero_d, W.X64, W.R64, W.P64, M.dinv64);
    }
    HIP_TRY(hipGetLastError());
    *it_out = 0;
    *conv_out = 1;
    *rel_out = 0.0;
    if (all_zero) return HDFE_STATUS_OK;
    if (has_x0) {
        // rel = ||R|| / ||b||; columns already solved take no iterations
        launch_colsum<double>(M.L, k, W.R64, W.R64, nullptr, W.part, W.colbuf, 0, W.cnt + 4, s);
        std::vector<double> rr;
        std::vector<double> bn(k);
        double dummy;
        TRY(read_max(W.colbuf, k, s, &dummy, &rr));
        HIP_TRY(hipMemcpyAsync(bn.data(), st.bnorm, k * 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        bool any = false;
        double rmax = 0.0;
        for (int j = 0; j < k; ++j) {
            double rel = std::sqrt(rr[j]) / bn[j];
            rmax = std::max(rmax, rel);
            act[j] = act[j] && rel > tol;
            any |= (bool)act[j];
        }
        if (!any) {
            *rel_out = rmax;
            return HDFE_STATUS_OK;
        }
    }
    HIP_TRY(hipMemcpyAsync(st.active, act.data(), k * 4, hipMemcpyHostToDevice, s));
    launch_colsum<double>(M.L, k, W.R64, W.P64, nullptr, W.part, st.rz, 0, W.cnt + 4, s);
    if (maxiter <= 0) {
        *conv_out = 0;
        return HDFE_STATUS_OK;
    }
    TRY(run_cg<double>(M, k, M.val64, M.dinv64, W.X64, W.R64, W.P64, W.AP64, st, s, it_out,
                       conv_out));
    return read_max(st.rel, k, s, rel_out);
}

// Mixed precision: FP32 CG inside FP64 iterative refinement.
static int solve_mixed(Matrix &M, int k, const uint8_t *zero_h, bool has_x0, double tol,
                       int64_t maxiter, double inner_tol, int64_t *it_out, int *conv_out,
                       double *rel_out, int64_t *outer_out, hipStream_t s) {
    Work &W = M.w;
    const int64_t Lk = M.L * (int64_t)k;
    std::vector<int> zero(k), act(k);
    bool all_zero = true;
    for (int j = 0; j < k; ++j) {
        zero[j] = zero_h ? zero_h[j] != 0 : 0;
        all_zero &= (bool)zero[j];
    }
    int *zero_d = W.si + k;
    HIP_TRY(hipMemcpyAsync(zero_d, zero.data(), k * 4, hipMemcpyHostToDevice, s));
    double *bnorm = W.sd + 3 * k, *scale = W.sd + 5 * k;
    launch_colsum<double>(M.L, k, W.B64, W.B64, nullptr, W.part, bnorm, 1, W.cnt + 4, s);
    std::vector<double> bn(k), rr(k), sc(k);
    HIP_TRY(hipMemcpyAsync(bn.data(), bnorm, k * 8, hipMemcpyDeviceToHost, s));
    const int ge = grid_for(Lk, 1024, 1, 4096);
    if (!has_x0) HIP_TRY(hipMemsetAsync(W.X64, 0, Lk * 8, s));
    hipLaunchKernelGGL(k_zero_cols, dim3(ge), dim3(kThreads), 0, s, M.L, k, zero_d, W.X64);
    *it_out = 0;
    *conv_out = 1;
    *rel_out = 0.0;
    *outer_out = 0;
    if (all_zero) return HDFE_STATUS_OK;
    *conv_out = 0;
    int *outer_act = W.si + 2 * k;  // the refinement's own mask of open columns
    CgState plain = make_state(W, k, tol, 0);
    plain.done = nullptr;
    int64_t total = 0;
    bool first = true;
    double prev = 1e300;
    int64_t first_inner = 0;
    for (int outer = 0; outer < 64; ++outer) {
        // the true residual, in FP64
        if (first && !has_x0) {
            HIP_TRY(hipMemcpyAsync(W.R64, W.B64, Lk * 8, hipMemcpyDeviceToDevice, s));
        } else {
            launch_spmm<double>(M, M.val64, W.X64, W.AP64, W.part, W.lpart, 0, plain, s);
            hipLaunchKernelGGL(k_sub, dim3(ge), dim3(kThreads), 0, s, Lk, W.B64, W.AP64, W.R64);
        }
        first = false;
        hipLaunchKernelGGL(k_zero_cols, dim3(ge), dim3(kThreads), 0, s, M.L, k, zero_d, W.R64);
        launch_colsum<double>(M.L, k, W.R64, W.R64, nullptr, W.part, W.colbuf, 0, W.cnt + 4, s);
        HIP_TRY(hipMemcpyAsync(rr.data(), W.colbuf, k * 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        bool any = false;
        double rmax = 0.0;
        for (int j = 0; j < k; ++j) {
            const double rn = std::sqrt(rr[j]);
            const double rel = rn / bn[j];
            rmax = std::max(rmax, rel);
            act[j] = !zero[j] && rel > tol;
            any |= (bool)act[j];
            sc[j] = rn > 0.0 ? rn : 1.0;
        }
        *rel_out = rmax;
        if (!any) {
            *conv_out = 1;
            break;
        }
        if (total >= maxiter) {
            *conv_out = 0;
            break;
        }
        // the true residual has reached its rounding floor: another step
        // would not lower it by half
        if (outer >= 2 && rmax > 0.5 * prev) {
            *conv_out = 0;
            break;
        }
        prev = rmax;
        *outer_out = outer + 1;
        // the FP32 inner solve of S d = r / ||r||; after the first, each may
        // take at most three times the first's iterations (and 20): once the
        // residual nears its rounding floor the inner target can be out of
        // reach, and the outer test then ends the refinement
        int64_t cap = maxiter - total;
        if (first_inner > 0) cap = std::min<int64_t>(cap, std::max<int64_t>(20, 3 * first_inner));
        CgState st = make_state(W, k, inner_tol, (int)std::min<int64_t>(cap, INT32_MAX));
        HIP_TRY(hipMemcpyAsync(scale, sc.data(), k * 8, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(st.active, act.data(), k * 4, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(outer_act, act.data(), k * 4, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_inner_start, dim3(ge), dim3(kThreads), 0, s, M.L, k, W.R64, scale,
                           st.active, W.X32, W.R32, W.P32, M.dinv32);
        launch_colsum<float>(M.L, k, W.R32, W.R32, nullptr, W.part, st.bnorm, 1, W.cnt + 4, s);
        launch_colsum<float>(M.L, k, W.R32, W.P32, nullptr, W.part, st.rz, 0, W.cnt + 4, s);
        int64_t it = 0;
        int conv = 0;
        TRY(run_cg<float>(M, k, M.val32, M.dinv32, W.X32, W.R32, W.P32, W.AP32, st, s, &it, &conv));
        total += std::max<int64_t>(it, 1);
        if (first_inner == 0) first_inner = std::max<int64_t>(it, 1);
        hipLaunchKernelGGL(k_refine_update, dim3(ge), dim3(kThreads), 0, s, M.L, k, W.X64, W.X32,
                           scale, outer_act);
        HIP_TRY(hipGetLastError());
    }
    *it_out = total;
    return HDFE_STATUS_OK;
}

// A device matrix from host row pointers (int32, or int64: 64-bit row
// pointers on the device) and host or device indices, values and inverse
// diagonal.
template <typename I>
static int matrix_create_t(int64_t L, int64_t nnz, const std::vector<I> &ip32,
                           const int *indices, const double *data, const double *dinv,
                           bool on_device, void **out, int *own_indptr, int *own_indices,
                           double *own_data) {
    constexpr bool p64 = sizeof(I) == 8;
    std::unique_ptr<Matrix> M(new Matrix());
    M->p64 = p64;
    HIP_TRY(hipGetDevice(&M->device));
    M->L = L;
    M->nnz = nnz;
    std::vector<int> lrows, lcptr(1, 0), crow;
    std::vector<I> cbeg;
    for (int64_t i = 0; i < L; ++i) {
        const int64_t len = (int64_t)ip32[i + 1] - ip32[i];
        if (len > kLongThresh) {
            lrows.push_back((int)i);
            for (int64_t p = ip32[i]; p < ip32[i + 1]; p += kLongChunk) {
                crow.push_back((int)i);
                cbeg.push_back((I)p);
            }
            lcptr.push_back((int)crow.size());
        }
    }
    M->nlong = (int)lrows.size();
    M->nchunk = (int)crow.size();
    HIP_TRY(stream_get(&M->stream));
    HIP_TRY(hipEventCreateWithFlags(&M->ev, hipEventDisableTiming));
    hipStream_t ms = M->stream;
    const bool adopt = !p64 && own_indptr && own_indices && own_data;
    if (p64) {
        HIP_TRY(palloc(&M->indptr64, (L + 1) * 8, ms));
        HIP_TRY(palloc(&M->indices, std::max<int64_t>(nnz, 1) * 4, ms));
        HIP_TRY(palloc(&M->val64, std::max<int64_t>(nnz, 1) * 8, ms));
    } else if (adopt) {
        M->indptr = own_indptr;
        M->indices = own_indices;
        M->val64 = own_data;
    } else {
        HIP_TRY(palloc(&M->indptr, (L + 1) * 4, ms));
        HIP_TRY(palloc(&M->indices, std::max<int64_t>(nnz, 1) * 4, ms));
        HIP_TRY(palloc(&M->val64, std::max<int64_t>(nnz, 1) * 8, ms));
    }
    HIP_TRY(palloc(&M->val32, std::max<int64_t>(nnz, 1) * 4, ms));
    HIP_TRY(palloc(&M->dinv64, std::max<int64_t>(L, 1) * 8, ms));
    HIP_TRY(palloc(&M->dinv32, std::max<int64_t>(L, 1) * 4, ms));
    // every upload and the FP32 casts in order on the matrix's own stream (a
    // device-to-device hipMemcpy may return before the copy has finished)
    if (!adopt)
        HIP_TRY(hipMemcpyAsync(p64 ? (void *)M->indptr64 : (void *)M->indptr, ip32.data(),
                               (L + 1) * sizeof(I), hipMemcpyHostToDevice, ms));
    const hipMemcpyKind up = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (nnz && !adopt) {
        HIP_TRY(hipMemcpyAsync(M->indices, indices, nnz * 4, up, ms));
        HIP_TRY(hipMemcpyAsync(M->val64, data, nnz * 8, up, ms));
    }
    if (L) HIP_TRY(hipMemcpyAsync(M->dinv64, dinv, L * 8, up, ms));
    const int g1 = grid_for(std::max<int64_t>(nnz, 1), 1024, 1, 4096);
    hipLaunchKernelGGL(k_cast_f64_to_f32, dim3(g1), dim3(kThreads), 0, ms, nnz, M->val64,
                       M->val32);
    hipLaunchKernelGGL(k_cast_f64_to_f32, dim3(grid_for(std::max<int64_t>(L, 1), 1024, 1, 4096)),
                       dim3(kThreads), 0, ms, L, M->dinv64, M->dinv32);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(ms));
    if (M->nlong) {
        HIP_TRY(palloc(&M->long_rows, M->nlong * 4, ms));
        HIP_TRY(palloc(&M->long_cptr, (M->nlong + 1) * 4, ms));
        HIP_TRY(palloc(&M->chunk_row, M->nchunk * 4, ms));
        if (p64) HIP_TRY(palloc(&M->chunk_beg64, M->nchunk * 8, ms));
        else HIP_TRY(palloc(&M->chunk_beg, M->nchunk * 4, ms));
        HIP_TRY(hipMemcpyAsync(M->long_rows, lrows.data(), M->nlong * 4, hipMemcpyHostToDevice, ms));
        HIP_TRY(hipMemcpyAsync(M->long_cptr, lcptr.data(), (M->nlong + 1) * 4,
                               hipMemcpyHostToDevice, ms));
        HIP_TRY(hipMemcpyAsync(M->chunk_row, crow.data(), M->nchunk * 4, hipMemcpyHostToDevice, ms));
        HIP_TRY(hipMemcpyAsync(p64 ? (void *)M->chunk_beg64 : (void *)M->chunk_beg, cbeg.data(),
                               M->nchunk * sizeof(I), hipMemcpyHostToDevice, ms));
    }
    HIP_TRY(hipStreamSynchronize(M->stream));
    // kept for the blocked layout, built at the first small-k use
    if constexpr (!p64) M->ip_host = ip32;
    *out = M.release();
    return HDFE_STATUS_OK;
}

int matrix_create(int64_t L, int64_t nnz, const std::vector<int> &ip32, const int *indices,
                  const double *data, const double *dinv, bool on_device, void **out,
                  int *own_indptr, int *own_indices, double *own_data) {
    return matrix_create_t<int>(L, nnz, ip32, indices, data, dinv, on_device, out, own_indptr,
                                own_indices, own_data);
}

// The wide form's plan (matfree_create): the dense dimensions, the sparse
// levels, each cell's group, the ranges of consecutive groups with their
// dense windows (at most kMfwSlots slots in all), and for every dense level
// the dpart rows of the ranges that cover it, in range order.  A group whose
// own window exceeds the slots leaves the wide form off (k_mf_cells serves).
static int mfw_plan(Matrix &M, MatFree &F, const std::vector<int> &cnt, hipStream_t ms) {
    const int64_t L = M.L, Gi = F.Gi, Mc = F.Mc;
    const int D = F.dm.D;
    std::vector<int> lptr(L + 1, 0);
    for (int64_t l = 0; l < L; ++l) lptr[l + 1] = lptr[l] + cnt[l];
    // a dimension's levels run from its offset to the next offset above it
    int64_t dlo[kMfMaxD], dhi[kMfMaxD];
    bool dense[kMfMaxD] = {false};
    F.dd.n = 0;
    for (int d = 0; d < D; ++d) {
        dlo[d] = F.dm.off[d];
        dhi[d] = L;
        for (int e = 0; e < D; ++e)
            if (F.dm.off[e] > dlo[d]) dhi[d] = std::min<int64_t>(dhi[d], F.dm.off[e]);
        const int64_t nl = dhi[d] - dlo[d];
        if (nl > 0 && Mc / nl >= kMfwDenseAvg && F.dd.n < kMfwMaxDense) {
            dense[d] = true;
            F.dd.d[F.dd.n++] = d;
        }
    }
    const int nds = F.dd.n;
    F.sd.n = 0;
    for (int d = 0; d < D; ++d)
        if (!dense[d]) F.sd.d[F.sd.n++] = d;
    std::vector<int> slev, dlev;
    for (int d = 0; d < D; ++d)
        for (int64_t l = dlo[d]; l < dhi[d]; ++l) (dense[d] ? dlev : slev).push_back((int)l);
    std::vector<int> win;
    if (nds) {
        int *wd = nullptr;
        HIP_TRY(palloc(&wd, std::max<int64_t>(Gi * nds * 2, 1) * 4, ms));
        hipLaunchKernelGGL(k_mfw_group_window, dim3(grid_for(std::max<int64_t>(Gi, 1), 256, 1, 8192)),
                           dim3(kThreads), 0, ms, Gi, F.igstart, F.icodes, F.ld, F.dd, wd);
        win.resize(Gi * nds * 2);
        if (Gi) HIP_TRY(hipMemcpyAsync(win.data(), wd, Gi * nds * 2 * 4, hipMemcpyDeviceToHost, ms));
        HIP_TRY(hipStreamSynchronize(ms));
        pfree(wd, ms);
    }
    const int64_t R = std::max<int64_t>(1, (Gi + kMfwRanges - 1) / kMfwRanges);
    std::vector<int> rng{0}, rbase, rsoff, dpoff{0};
    F.maxslot = 0;
    // a range's window is cut at the widest single group's (when the dense
    // levels are per model, the years of C5, a range stays within a model:
    // fewer slots, less LDS)
    int soft = 0;
    for (int64_t g = 0; g < Gi; ++g) {
        int tot = 0;
        for (int q = 0; q < nds; ++q) {
            const int lo = win[(g * nds + q) * 2], hi = win[(g * nds + q) * 2 + 1];
            tot += hi >= lo ? hi - lo + 1 : 0;
        }
        soft = std::max(soft, tot);
    }
    if (soft > kMfwSlots) return HDFE_STATUS_OK;   // one group alone is too wide: no wide form
    for (int64_t g = 0; g < Gi;) {
        int wl[kMfMaxD], wh[kMfMaxD];
        for (int q = 0; q < nds; ++q) {
            wl[q] = INT_MAX;
            wh[q] = -1;
        }
        int64_t e = g;
        while (e < Gi && e - g < R) {
            int nl[kMfMaxD], nh[kMfMaxD], tot = 0;
            for (int q = 0; q < nds; ++q) {
                nl[q] = std::min(wl[q], win[(e * nds + q) * 2]);
                nh[q] = std::max(wh[q], win[(e * nds + q) * 2 + 1]);
                tot += nh[q] >= nl[q] ? nh[q] - nl[q] + 1 : 0;
            }
            if (tot > soft) break;
            for (int q = 0; q < nds; ++q) {
                wl[q] = nl[q];
                wh[q] = nh[q];
            }
            ++e;
        }
        if (e == g) return HDFE_STATUS_OK;   // one group alone is too wide: no wide form
        int off = 0;
        for (int q = 0; q < nds; ++q) {
            rbase.push_back(wh[q] >= wl[q] ? wl[q] : 0);
            rsoff.push_back(off);
            off += wh[q] >= wl[q] ? wh[q] - wl[q] + 1 : 0;
        }
        dpoff.push_back(dpoff.back() + off);
        F.maxslot = std::max(F.maxslot, off);
        rng.push_back((int)e);
        g = e;
    }
    const int64_t nr = (int64_t)rng.size() - 1;
    // every dense level's rows, range by range
    std::vector<int> dindex(L, -1);
    for (size_t i = 0; i < dlev.size(); ++i) dindex[dlev[i]] = (int)i;
    std::vector<int> dcnt(dlev.size() + 1, 0), dcptr(dlev.size() + 1, 0), dcrow;
    for (int pass = 0; pass < 2; ++pass) {
        std::vector<int> fill;
        if (pass) {
            for (size_t i = 0; i < dlev.size(); ++i) dcptr[i + 1] = dcptr[i] + dcnt[i];
            dcrow.assign(dcptr.back(), 0);
            fill.assign(dcptr.begin(), dcptr.end() - 1);
        }
        for (int64_t r = 0; r < nr; ++r)
            for (int q = 0; q < nds; ++q) {
                const int qs = rsoff[r * nds + q];
                const int span = (q + 1 < nds ? rsoff[r * nds + q + 1] : dpoff[r + 1] - dpoff[r]) - qs;
                for (int c = 0; c < span; ++c) {
                    const int l = (int)dlo[F.dd.d[q]] + rbase[r * nds + q] + c;
                    const int i = dindex[l];
                    if (pass) dcrow[fill[i]++] = dpoff[r] + qs + c;
                    else ++dcnt[i];
                }
            }
    }
    F.nsl = (int64_t)slev.size();
    F.ndl = (int64_t)dlev.size();
    F.nr = nr;
    F.dprows = dpoff.back();
    auto up = [&](int **dst, const std::vector<int> &src) -> hipError_t {
        hipError_t e = palloc(dst, std::max<size_t>(src.size(), 1) * 4, ms);
        if (e != hipSuccess || src.empty()) return e;
        return hipMemcpyAsync(*dst, src.data(), src.size() * 4, hipMemcpyHostToDevice, ms);
    };
    HIP_TRY(palloc(&F.offs_d, kMfMaxD * 8, ms));
    HIP_TRY(hipMemcpyAsync(F.offs_d, F.dm.off, D * 8, hipMemcpyHostToDevice, ms));
    HIP_TRY(up(&F.lptr, lptr));
    HIP_TRY(up(&F.slev, slev));
    HIP_TRY(up(&F.rng, rng));
    HIP_TRY(up(&F.rbase, rbase));
    HIP_TRY(up(&F.rsoff, rsoff));
    HIP_TRY(up(&F.dpoff, dpoff));
    HIP_TRY(up(&F.dlev, dlev));
    HIP_TRY(up(&F.dcptr, dcptr));
    HIP_TRY(up(&F.dcrow, dcrow));
    HIP_TRY(palloc(&F.cgroup, std::max<int64_t>(Mc, 1) * 4, ms));
    hipLaunchKernelGGL(k_mfw_cell_groups, dim3(grid_for(std::max<int64_t>(Gi, 1), 256, 1, 8192)),
                       dim3(kThreads), 0, ms, Gi, F.igstart, F.cgroup);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(ms));   // the host vectors above go out of scope
    F.wide = true;
    return HDFE_STATUS_OK;
}

// The fingerprint of a cell structure: igstart (Gi + 1) and icodes (D rows).
static int cells_fingerprint(int64_t Gi, int64_t Mc, int D, const int *igstart,
                             const int *icodes, int64_t ld, hipStream_t s,
                             unsigned long long *fp) {
    unsigned long long *d = nullptr;
    HIP_TRY(palloc(&d, 8, s));
    // a fill kernel, not hipMemsetAsync: a caller may capture this stream
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, (int64_t)1, 0ull, d);
    hipLaunchKernelGGL(k_fingerprint, dim3(grid_for(Gi + 1, 1024, 1, 1024)), dim3(kThreads), 0, s,
                       Gi + 1, 1, (int64_t)0, igstart, d);
    if (Mc)
        hipLaunchKernelGGL(k_fingerprint, dim3(grid_for(Mc, 1024, 1, 2048)), dim3(kThreads), 0, s,
                           Mc, D, ld, icodes, d);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(fp, d, 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    pfree(d, s);
    return HDFE_STATUS_OK;
}

// A matrix-free operator over a panel's identifying cells (see k_mf_cells):
// the level-to-cells CSR by a stable radix sort of the (level, cell) pairs
// (each level's cells ascending), its chunks, dinv64 copied and cast.
int matfree_create(int64_t L, int64_t Gi, int64_t Mc, int D, const int *igstart,
                   const int *icodes, int64_t ld, const int64_t *offs_h, const double *cell_n,
                   const double *dinv, void **out) {
    std::unique_ptr<Matrix> M(new Matrix());
    HIP_TRY(hipGetDevice(&M->device));
    M->L = L;
    M->nnz = 0;
    M->blk_tried = true;
    HIP_TRY(stream_get(&M->stream));
    HIP_TRY(hipEventCreateWithFlags(&M->ev, hipEventDisableTiming));
    hipStream_t ms = M->stream;
    auto *F = new MatFree();
    M->mf = F;
    F->Gi = Gi;
    F->Mc = Mc;
    F->ld = ld;
    F->dm.D = D;
    for (int d = 0; d < D; ++d) F->dm.off[d] = offs_h[d];
    F->igstart = igstart;
    F->icodes = icodes;
    F->cn = cell_n;
    HIP_TRY(palloc(&M->dinv64, std::max<int64_t>(L, 1) * 8, ms));
    HIP_TRY(palloc(&M->dinv32, std::max<int64_t>(L, 1) * 4, ms));
    if (L) HIP_TRY(hipMemcpyAsync(M->dinv64, dinv, L * 8, hipMemcpyDeviceToDevice, ms));
    hipLaunchKernelGGL(k_cast_f64_to_f32, dim3(grid_for(std::max<int64_t>(L, 1), 1024, 1, 4096)),
                       dim3(kThreads), 0, ms, L, M->dinv64, M->dinv32);
    const int64_t n = Mc * D;
    int *key = nullptr, *cell = nullptr, *key2 = nullptr, *count = nullptr;
    void *tmp = nullptr;
    HIP_TRY(palloc(&key, std::max<int64_t>(n, 1) * 4, ms));
    HIP_TRY(palloc(&cell, std::max<int64_t>(n, 1) * 4, ms));
    HIP_TRY(palloc(&key2, std::max<int64_t>(n, 1) * 4, ms));
    HIP_TRY(palloc(&F->lcell, std::max<int64_t>(n, 1) * 4, ms));
    HIP_TRY(palloc(&count, (L + 1) * 4, ms));
    HIP_TRY(hipMemsetAsync(count, 0, (L + 1) * 4, ms));
    hipLaunchKernelGGL(k_mf_pairs, dim3(grid_for(std::max<int64_t>(n, 1), 1024, 1, 8192)),
                       dim3(kThreads), 0, ms, Mc, icodes, ld, F->dm, key, cell, count);
    size_t tb = 0;
    const int bits = std::max(1, 64 - __builtin_clzll((unsigned long long)std::max<int64_t>(L, 1)));
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, key, key2, cell, F->lcell, (int)n, 0,
                                               bits, ms));
    HIP_TRY(palloc(&tmp, std::max<size_t>(tb, 1), ms));
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, tb, key, key2, cell, F->lcell, (int)n, 0,
                                               bits, ms));
    std::vector<int> cnt(L);
    if (L) HIP_TRY(hipMemcpyAsync(cnt.data(), count, L * 4, hipMemcpyDeviceToHost, ms));
    HIP_TRY(hipStreamSynchronize(ms));
    pfree(key, ms);
    pfree(cell, ms);
    pfree(key2, ms);
    pfree(count, ms);
    pfree(tmp, ms);
    // chunks of at most kMfChunk cells, level by level
    std::vector<int> cb, ce, lch(L + 1, 0);
    int64_t at = 0;
    for (int64_t l = 0; l < L; ++l) {
        lch[l] = (int)cb.size();
        for (int64_t b = 0; b < cnt[l]; b += kMfChunk) {
            cb.push_back((int)(at + b));
            ce.push_back((int)(at + std::min<int64_t>(cnt[l], b + kMfChunk)));
        }
        at += cnt[l];
    }
    lch[L] = (int)cb.size();
    F->nchunk = (int64_t)cb.size();
    HIP_TRY(palloc(&F->ch_beg, std::max<int64_t>(F->nchunk, 1) * 4, ms));
    HIP_TRY(palloc(&F->ch_end, std::max<int64_t>(F->nchunk, 1) * 4, ms));
    HIP_TRY(palloc(&F->lch, (L + 1) * 4, ms));
    if (F->nchunk) {
        HIP_TRY(hipMemcpyAsync(F->ch_beg, cb.data(), F->nchunk * 4, hipMemcpyHostToDevice, ms));
        HIP_TRY(hipMemcpyAsync(F->ch_end, ce.data(), F->nchunk * 4, hipMemcpyHostToDevice, ms));
    }
    HIP_TRY(hipMemcpyAsync(F->lch, lch.data(), (L + 1) * 4, hipMemcpyHostToDevice, ms));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(ms));
    TRY(mfw_plan(*M, *F, cnt, ms));
    TRY(cells_fingerprint(Gi, Mc, D, igstart, icodes, ld, ms, &F->fp));
    *out = M.release();
    return HDFE_STATUS_OK;
}

}  // namespace hdfe

// ================================================================== C ABI

using namespace hdfe;

extern "C" {

typedef struct hdfe_solver hdfe_solver;

int hdfe_csr_create(int64_t L, int64_t nnz, const int64_t *indptr, const int32_t *indices,
                    const double *data, const double *dinv, int on_device, void **out) {
    ARG_CHECK(out != nullptr, "hdfe_csr_create: null output handle");
    ARG_CHECK(L >= 0 && nnz >= 0, "hdfe_csr_create: negative size");
    ARG_CHECK(L < INT32_MAX, "hdfe_csr_create: L must be below 2^31 (int32 column indices)");
    const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToHost : hipMemcpyHostToHost;
    std::vector<int64_t> ip(L + 1);
    HIP_TRY(hipMemcpy(ip.data(), indptr, (L + 1) * 8, kind));
    ARG_CHECK(ip[0] == 0 && ip[L] == nnz, "hdfe_csr_create: indptr must run from 0 to nnz");
    for (int64_t i = 0; i < L; ++i)
        ARG_CHECK(ip[i + 1] >= ip[i], "hdfe_csr_create: indptr must be non-decreasing");
    // 2^31 nonzeros and more: 64-bit row pointers on the device (the
    // reference's int64 index arrays, core/solve.py:169, 188-189)
    if (nnz >= INT32_MAX)
        return hdfe::matrix_create_t<int64_t>(L, nnz, ip, indices, data, dinv, on_device != 0,
                                              out, nullptr, nullptr, nullptr);
    std::vector<int> ip32(ip.begin(), ip.end());
    return hdfe::matrix_create(L, nnz, ip32, indices, data, dinv, on_device != 0, out);
}

// The row pointers' width on the device (32 or 64 bits); set = 64 widens a
// matrix's 32-bit row pointers (its column-blocked and dense split-row copies
// are dropped: the 64-bit form runs the plain CSR kernels only); set = -1
// reads.  A matrix of 2^31 nonzeros or more is created with 64-bit ones.
int hdfe_csr_row_pointer_bits(void *h, int set, int *bits) {
    ARG_CHECK(h && bits && (set == -1 || set == 32 || set == 64),
              "hdfe_csr_row_pointer_bits: bad argument");
    Matrix &M = *(Matrix *)h;
    std::lock_guard<std::mutex> lock(M.mu);
    ARG_CHECK(!M.mf, "hdfe_csr_row_pointer_bits: a matrix-free operator has no row pointers");
    ARG_CHECK(!(set == 32 && M.p64), "hdfe_csr_row_pointer_bits: row pointers only widen");
    if (set == 64 && !M.p64) {
        HIP_TRY(hipSetDevice(M.device));
        hipStream_t ms = M.stream;
        M.drop_graphs();
        M.w.release(ms);
        void *ps[] = {M.tbp, M.off16, M.bidx16, M.col_lo, M.bptr, M.bidx, M.bval64, M.bval32,
                      M.dl_slot, M.dl_cptr, M.dt64, M.dt32, M.ll64, M.ll32, M.dw64, M.dw32};
        for (void *q : ps) pfree(q, ms);
        M.tbp = M.col_lo = M.bptr = M.bidx = M.dl_slot = M.dl_cptr = nullptr;
        M.off16 = M.bidx16 = nullptr;
        M.bval64 = M.dt64 = M.ll64 = M.dw64 = nullptr;
        M.bval32 = M.dt32 = M.ll32 = M.dw32 = nullptr;
        M.blk = M.dl = false;
        M.blk_tried = true;
        M.dw = 0;
        M.bnnz = 0;
        HIP_TRY(palloc(&M.indptr64, (M.L + 1) * 8, ms));
        hipLaunchKernelGGL(k_widen, dim3(grid_for(M.L + 1, 1024, 1, 4096)), dim3(kThreads), 0, ms,
                           M.L + 1, M.indptr, M.indptr64);
        if (M.nchunk) {
            HIP_TRY(palloc(&M.chunk_beg64, (int64_t)M.nchunk * 8, ms));
            hipLaunchKernelGGL(k_widen, dim3(grid_for(M.nchunk, 1024, 1, 4096)), dim3(kThreads), 0,
                               ms, (int64_t)M.nchunk, M.chunk_beg, M.chunk_beg64);
        }
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(ms));
        pfree(M.indptr, ms);
        pfree(M.chunk_beg, ms);
        M.indptr = M.chunk_beg = nullptr;
        M.ip_host.clear();
        M.ip_host.shrink_to_fit();
        M.p64 = true;
    }
    *bits = M.p64 ? 64 : 32;
    return HDFE_STATUS_OK;
}

int hdfe_csr_info(void *h, int64_t *L, int64_t *nnz, int *nlong) {
    ARG_CHECK(h != nullptr, "hdfe_csr_info: null handle");
    Matrix *M = (Matrix *)h;
    if (L) *L = M->L;
    if (nnz) *nnz = M->nnz;
    if (nlong) *nlong = M->nlong;
    return HDFE_STATUS_OK;
}

int hdfe_pcg_solve(void *h, int64_t k, const double *b, const uint8_t *zero, const double *x0,
                   double tol, int64_t maxiter, int precision, double inner_tol, double *X,
                   int64_t *iterations, int *converged, double *max_rel, int64_t *outer,
                   int on_device, void *stream) {
    ARG_CHECK(h != nullptr, "hdfe_pcg_solve: null handle");
    ARG_CHECK(k >= 1 && k <= kMaxK, "hdfe_pcg_solve: k must be in [1, 1024]");
    ARG_CHECK(precision == 0 || precision == 1, "hdfe_pcg_solve: precision must be 0 or 1");
    ARG_CHECK(b != nullptr && X != nullptr && iterations && converged && max_rel,
              "hdfe_pcg_solve: null argument");
    Matrix &M = *(Matrix *)h;
    std::lock_guard<std::mutex> lock(M.mu);
    HIP_TRY(hipSetDevice(M.device));
    int st = alloc_work(M, (int)k);
    if (st) return st;
    hipStream_t s = M.stream;
    hipStream_t caller = (hipStream_t)stream;
    const int64_t Lk = M.L * k;
    if (on_device) {  // order after the caller's stream
        HIP_TRY(hipEventRecord(M.ev, caller));
        HIP_TRY(hipStreamWaitEvent(s, M.ev, 0));
    }
    const hipMemcpyKind up = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (Lk) HIP_TRY(hipMemcpyAsync(M.w.B64, b, Lk * 8, up, s));
    if (x0 && Lk) HIP_TRY(hipMemcpyAsync(M.w.X64, x0, Lk * 8, up, s));
    int64_t outer_n = 0;
    if (precision == 0)
        st = solve_f64(M, (int)k, zero, x0 != nullptr, tol, maxiter, iterations, converged, max_rel, s);
    else
        st = solve_mixed(M, (int)k, zero, x0 != nullptr, tol, maxiter,
                         inner_tol > 0 ? inner_tol : 1e-5, iterations, converged, max_rel,
                         &outer_n, s);
    if (st) return st;
    if (outer) *outer = outer_n;
    const hipMemcpyKind down = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    if (Lk) HIP_TRY(hipMemcpyAsync(X, M.w.X64, Lk * 8, down, s));
    if (on_device) {
        HIP_TRY(hipEventRecord(M.ev, s));
        HIP_TRY(hipStreamWaitEvent(caller, M.ev, 0));
    } else {
        HIP_TRY(hipStreamSynchronize(s));
    }
    return HDFE_STATUS_OK;
}

// y = S x (k columns, FP64), on the device; for tests of the operator.
int hdfe_csr_matmat(void *h, int64_t k, const double *x, double *y, int on_device) {
    ARG_CHECK(h != nullptr && x != nullptr && y != nullptr, "hdfe_csr_matmat: null argument");
    ARG_CHECK(k >= 1 && k <= kMaxK, "hdfe_csr_matmat: k must be in [1, 1024]");
    Matrix &M = *(Matrix *)h;
    std::lock_guard<std::mutex> lock(M.mu);
    HIP_TRY(hipSetDevice(M.device));
    int st = alloc_work(M, (int)k);
    if (st) return st;
    hipStream_t s = M.stream;
    const int64_t Lk = M.L * k;
    const hipMemcpyKind up = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    HIP_TRY(hipMemcpyAsync(M.w.P64, x, Lk * 8, up, s));
    CgState plain = make_state(M.w, (int)k, 0.0, 0);
    plain.done = nullptr;
    launch_spmm<double>(M, M.val64, M.w.P64, M.w.AP64, M.w.part, M.w.lpart, 0, plain, s);
    HIP_TRY(hipGetLastError());
    const hipMemcpyKind down = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    HIP_TRY(hipMemcpyAsync(y, M.w.AP64, Lk * 8, down, s));
    HIP_TRY(hipStreamSynchronize(s));
    return HDFE_STATUS_OK;
}

// Copy S, as the device holds it, to host arrays: indptr int64 (L+1),
// indices int32 (nnz), data float64 (nnz), dinv float64 (L); any may be NULL.
int hdfe_csr_export(void *h, int64_t *indptr, int32_t *indices, double *data, double *dinv) {
    ARG_CHECK(h != nullptr, "hdfe_csr_export: null handle");
    Matrix &M = *(Matrix *)h;
    ARG_CHECK(!M.mf || (!indptr && !indices && !data),
              "hdfe_csr_export: a matrix-free operator has no explicit S");
    HIP_TRY(hipSetDevice(M.device));
    if (indptr && M.p64) {
        HIP_TRY(hipMemcpy(indptr, M.indptr64, (M.L + 1) * 8, hipMemcpyDeviceToHost));
    } else if (indptr) {
        std::vector<int> ip(M.L + 1);
        HIP_TRY(hipMemcpy(ip.data(), M.indptr, (M.L + 1) * 4, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i <= M.L; ++i) indptr[i] = ip[i];
    }
    if (indices && M.nnz) HIP_TRY(hipMemcpy(indices, M.indices, M.nnz * 4, hipMemcpyDeviceToHost));
    if (data && M.nnz) HIP_TRY(hipMemcpy(data, M.val64, M.nnz * 8, hipMemcpyDeviceToHost));
    if (dinv && M.L) HIP_TRY(hipMemcpy(dinv, M.dinv64, M.L * 8, hipMemcpyDeviceToHost));
    return HDFE_STATUS_OK;
}

// Turn HIP-event timing of the CG's SpMM launches on or off (resets the
// totals); read the totals (milliseconds, launches).
int hdfe_csr_timing(void *h, int enable) {
    ARG_CHECK(h != nullptr, "hdfe_csr_timing: null handle");
    Matrix &M = *(Matrix *)h;
    std::lock_guard<std::mutex> lock(M.mu);
    M.timing = enable != 0;
    M.spmm_ms = 0.0;
    M.spmm_n = 0;
    return HDFE_STATUS_OK;
}

int hdfe_csr_timing_read(void *h, double *ms, int64_t *launches) {
    ARG_CHECK(h != nullptr, "hdfe_csr_timing_read: null handle");
    Matrix &M = *(Matrix *)h;
    if (ms) *ms = M.spmm_ms;
    if (launches) *launches = M.spmm_n;
    return HDFE_STATUS_OK;
}

// Average duration (microseconds) of the CG's SpMM kernel on this matrix:
// `reps` back-to-back launches of the same launch configuration as in the
// solve (k right-hand sides, with the fused pAp sums), between two HIP events
// on the solver's stream.  precision 0: FP64, 1: FP32.
// The blocked SpMM's row parts per XCD for k columns in one precision (0 FP64,
// 1 FP32): *parts gets the tuned choice (0: not tuned, the default of 128),
// after set >= 1 overrides it.  The choice never changes a result.
int hdfe_csr_blk_parts(void *h, int k, int precision, int set, int *parts) {
    ARG_CHECK(h && parts && k >= 1 && k <= kSmallK && (precision == 0 || precision == 1),
              "hdfe_csr_blk_parts: bad argument");
    Matrix &M = *(Matrix *)h;
    std::lock_guard<std::mutex> lock(M.mu);
    if (set >= 1) {
        ARG_CHECK(set <= 4096, "hdfe_csr_blk_parts: at most 4096 parts per XCD");
        M.blk_parts[precision][k] = set;
    }
    *parts = M.blk_parts[precision][k];
    return HDFE_STATUS_OK;
}

int hdfe_csr_small_form(void *h, int k, int precision, int set, int *form) {
    ARG_CHECK(h && form && k >= 1 && k <= kSmallK && (precision == 0 || precision == 1) &&
                  set >= -1 && set <= 3,
              "hdfe_csr_small_form: bad argument");
    Matrix &M = *(Matrix *)h;
    std::lock_guard<std::mutex> lock(M.mu);
    if (set >= 0 && set != M.small_form[precision][k]) {
        M.drop_graphs();  // their kernels are the old form's
        M.small_form[precision][k] = set;
    }
    *form = M.small_form[precision][k];
    return HDFE_STATUS_OK;
}

int hdfe_csr_wide_mode(void *h, int k, int precision, int set, int *mode) {
    ARG_CHECK(h && mode && k > kSmallK && k <= kMaxK && (precision == 0 || precision == 1) &&
                  set >= -1 && set <= 1,
              "hdfe_csr_wide_mode: bad argument");
    Matrix &M = *(Matrix *)h;
    std::lock_guard<std::mutex> lock(M.mu);
    if (set >= 0) M.wide_mode[{precision, k}] = set;
    auto it = M.wide_mode.find({precision, k});
    *mode = it == M.wide_mode.end() ? -1 : it->second;
    return HDFE_STATUS_OK;
}

int hdfe_csr_graphs(void *h, int64_t *count) {
    ARG_CHECK(h && count, "hdfe_csr_graphs: null argument");
    Matrix &M = *(Matrix *)h;
    std::lock_guard<std::mutex> lock(M.mu);
    *count = (int64_t)M.cg_graphs.size();
    return HDFE_STATUS_OK;
}

int hdfe_csr_spmm_time(void *h, int64_t k, int reps, int precision, double *avg_us) {
    ARG_CHECK(h != nullptr && avg_us != nullptr, "hdfe_csr_spmm_time: null argument");
    ARG_CHECK(k >= 1 && k <= kMaxK && reps >= 1, "hdfe_csr_spmm_time: bad k or reps");
    Matrix &M = *(Matrix *)h;
    std::lock_guard<std::mutex> lock(M.mu);
    HIP_TRY(hipSetDevice(M.device));
    int st = alloc_work(M, (int)k);
    if (st) return st;
    hipStream_t s = M.stream;
    const int64_t Lk = M.L * k;
    CgState cs = make_state(M.w, (int)k, 0.0, 1);
    HIP_TRY(hipMemsetAsync(cs.done, 0, 3 * sizeof(int), s));
    HIP_TRY(hipMemsetAsync(cs.active, 0, k * 4, s));
    HIP_TRY(hipMemsetAsync(M.w.P64, 0, Lk * 8, s));
    HIP_TRY(hipMemsetAsync(M.w.P32, 0, Lk * 4, s));
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    for (int r = 0; r < 2; ++r) {  // warm
        if (precision) launch_spmm<float>(M, M.val32, M.w.P32, M.w.AP32, M.w.part, M.w.lpart, 1, cs, s);
        else launch_spmm<double>(M, M.val64, M.w.P64, M.w.AP64, M.w.part, M.w.lpart, 1, cs, s);
    }
    HIP_TRY(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) {
        if (precision) launch_spmm<float>(M, M.val32, M.w.P32, M.w.AP32, M.w.part, M.w.lpart, 1, cs, s);
        else launch_spmm<double>(M, M.val64, M.w.P64, M.w.AP64, M.w.part, M.w.lpart, 1, cs, s);
    }
    HIP_TRY(hipEventRecord(e1, s));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    HIP_TRY(hipEventDestroy(e0));
    HIP_TRY(hipEventDestroy(e1));
    HIP_TRY(hipGetLastError());
    *avg_us = 1e3 * ms / reps;
    return HDFE_STATUS_OK;
}

// Column norms of a device (L, k) block, to the host: sqrt(sum_i X_ij^2),
// reduced in a fixed order.
int hdfe_col_norms(const double *X, int64_t L, int64_t k, double *out) {
    ARG_CHECK(X && out && k >= 1 && k <= kMaxK, "hdfe_col_norms: bad argument");
    DevBuf2 part, res;
    unsigned *cnt = nullptr;
    HIP_TRY(hipMalloc(&part.p, (int64_t)4096 * k * 8));
    HIP_TRY(hipMalloc(&res.p, k * 8));
    HIP_TRY(hipMalloc(&cnt, 8));
    part.extra = cnt;
    HIP_TRY(hipMemset(cnt, 0, 8));
    launch_colsum<double>(L, (int)k, X, X, nullptr, part.p, res.p, 0, cnt, nullptr);
    HIP_TRY(hipGetLastError());
    std::vector<double> h(k);
    HIP_TRY(hipMemcpy(h.data(), res.p, k * 8, hipMemcpyDeviceToHost));
    for (int64_t j = 0; j < k; ++j) out[j] = std::sqrt(h[j]);
    return HDFE_STATUS_OK;
}

// dst <- src, `bytes` bytes; kind 0 device to host, 1 host to device,
// 2 device to device.
int hdfe_memcpy(void *dst, const void *src, int64_t bytes, int kind) {
    ARG_CHECK(kind >= 0 && kind <= 2, "hdfe_memcpy: kind must be 0, 1 or 2");
    if (bytes <= 0) return HDFE_STATUS_OK;
    const hipMemcpyKind k = kind == 0 ? hipMemcpyDeviceToHost
                            : kind == 1 ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
    HIP_TRY(hipMemcpy(dst, src, (size_t)bytes, k));
    if (kind == 2) HIP_TRY(hipDeviceSynchronize());  // a D2D copy may return early
    return HDFE_STATUS_OK;
}

// y = S x in FP32 (the inner solver's operator: FP32 values, FP32 vectors),
// host or device float32 blocks.
int hdfe_csr_matmat_f32(void *h, int64_t k, const float *x, float *y, int on_device) {
    ARG_CHECK(h != nullptr && x != nullptr && y != nullptr, "hdfe_csr_matmat_f32: null argument");
    ARG_CHECK(k >= 1 && k <= kMaxK, "hdfe_csr_matmat_f32: k must be in [1, 1024]");
    Matrix &M = *(Matrix *)h;
    std::lock_guard<std::mutex> lock(M.mu);
    HIP_TRY(hipSetDevice(M.device));
    int st = alloc_work(M, (int)k);
    if (st) return st;
    hipStream_t s = M.stream;
    const int64_t Lk = M.L * k;
    const hipMemcpyKind up = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    HIP_TRY(hipMemcpyAsync(M.w.P32, x, Lk * 4, up, s));
    CgState plain = make_state(M.w, (int)k, 0.0, 0);
    plain.done = nullptr;
    launch_spmm<float>(M, M.val32, M.w.P32, M.w.AP32, M.w.part, M.w.lpart, 0, plain, s);
    HIP_TRY(hipGetLastError());
    const hipMemcpyKind down = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    HIP_TRY(hipMemcpyAsync(y, M.w.AP32, Lk * 4, down, s));
    HIP_TRY(hipStreamSynchronize(s));
    return HDFE_STATUS_OK;
}

int hdfe_stream_create(int64_t L, int64_t Gi, int64_t M, int D, const int *igstart,
                       const int *icodes, int64_t ld, const int64_t *offs, const double *cell_n,
                       const double *dinv, void **out) {
    ARG_CHECK(out && igstart && icodes && cell_n && dinv && offs && L >= 1 && Gi >= 0 &&
                  M >= 0 && D >= 1 && D <= kMfMaxD && ld >= M,
              "hdfe_stream_create: bad argument");
    ARG_CHECK(M * D < INT32_MAX && L < INT32_MAX, "hdfe_stream_create: too many cells");
    for (int d = 0; d < D; ++d)
        ARG_CHECK(offs[d] >= 0 && offs[d] < L, "hdfe_stream_create: level offset out of range");
    return matfree_create(L, Gi, M, D, igstart, icodes, ld, offs, cell_n, dinv, out);
}

// A matrix-free operator over new cell arrays of the same structure (the
// IRLS reweighting: the same cells, new weights): when the new igstart and
// icodes have the operator's sizes and fingerprint, it takes the new
// pointers, weights and inverse diagonal and keeps its plan (*rebound = 1);
// otherwise nothing changes (*rebound = 0).
int hdfe_stream_rebind(void *h, int64_t L, int64_t Gi, int64_t M, int D, const int *igstart,
                       const int *icodes, int64_t ld, const double *cell_n, const double *dinv,
                       int *rebound) {
    ARG_CHECK(h && igstart && icodes && cell_n && dinv && rebound && ld >= M,
              "hdfe_stream_rebind: bad argument");
    Matrix &Mx = *(Matrix *)h;
    ARG_CHECK(Mx.mf != nullptr, "hdfe_stream_rebind: not a matrix-free operator");
    std::lock_guard<std::mutex> lock(Mx.mu);
    HIP_TRY(hipSetDevice(Mx.device));
    MatFree &F = *Mx.mf;
    *rebound = 0;
    if (L != Mx.L || Gi != F.Gi || M != F.Mc || D != F.dm.D) return HDFE_STATUS_OK;
    unsigned long long fp = 0;
    TRY(cells_fingerprint(Gi, M, D, igstart, icodes, ld, Mx.stream, &fp));
    if (fp != F.fp) return HDFE_STATUS_OK;
    Mx.drop_graphs();  // their kernels hold the old pointers
    F.igstart = igstart;
    F.icodes = icodes;
    F.ld = ld;
    F.cn = cell_n;
    hipStream_t ms = Mx.stream;
    if (L) HIP_TRY(hipMemcpyAsync(Mx.dinv64, dinv, L * 8, hipMemcpyDeviceToDevice, ms));
    hipLaunchKernelGGL(k_cast_f64_to_f32, dim3(grid_for(std::max<int64_t>(L, 1), 1024, 1, 4096)),
                       dim3(kThreads), 0, ms, L, Mx.dinv64, Mx.dinv32);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(ms));
    *rebound = 1;
    return HDFE_STATUS_OK;
}

void hdfe_destroy(void *h) {
    if (!h) return;
    Matrix *M = (Matrix *)h;
    (void)hipSetDevice(M->device);
    if (M->stream) (void)hipStreamSynchronize(M->stream);
    M->drop_graphs();
    M->w.release(M->stream);
    void *ps[] = {M->indptr, M->indptr64, M->chunk_beg64, M->indices, M->val64, M->val32,
                  M->dinv64, M->dinv32,
                  M->tbp, M->off16, M->bidx16,
                  M->long_rows, M->long_cptr, M->chunk_row, M->chunk_beg, M->col_lo, M->bptr,
                  M->bidx, M->bval64, M->bval32, M->dl_slot, M->dl_cptr, M->dt64, M->dt32,
                  M->ll64, M->ll32, M->dw64, M->dw32};
    for (void *p : ps) pfree(p, M->stream);
    if (M->mf) {
        MatFree *F = M->mf;
        void *fs[] = {F->lcell, F->ch_beg, F->ch_end, F->lch, F->v, F->cpart, F->slev,
                      F->lptr, F->cgroup, F->rng, F->rbase, F->rsoff, F->dpoff, F->dlev,
                      F->dcptr, F->dcrow, F->mg, F->dpart, F->offs_d};
        for (void *p : fs) pfree(p, M->stream);
        delete F;
        M->mf = nullptr;
    }
    if (M->stream) (void)hipStreamSynchronize(M->stream);
    if (M->ev) (void)hipEventDestroy(M->ev);
    stream_put(M->stream, M->device);
    delete M;
}

}  // extern "C"
