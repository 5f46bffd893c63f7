// Row passes of the leave-out operator on the device, for gfx950: the
// Rademacher draws, the reduction of row vectors onto the levels (the reduced
// right-hand side) and the projection back to rows, fused with the JLA
// leverage moments.
//
// Reference (hdfe_stream/):
//   _nb_rademacher            core/kernels_groups.py:52-77
//   project_rows / _reduce    inverse.py:162-180, 240-309
//   _nb_center_rows           kernels_inverse.py:34-71
//   _nb_project_rows          kernels_inverse.py:74-108
//   the jla_leverages sink    leaveout.py:181-190
//   _combine_moments          leaveout.py:209-282
//
// The draws are never stored: entry (row ordinal i, draw j) is a counter-based
// hash, regenerated in both passes.  A block of k draws is processed with a
// wavefront per group (or per chunk of a level's rows), lanes over the draws
// (draw j = lane + 64 t), so a row's k values are k contiguous lanes and the
// level-indexed vectors (k per level) are read with coalesced loads.
#include <algorithm>
#include <memory>
#include <vector>

#include "devbuf.h"

namespace hdfe {

constexpr int kRowChunk = 1024;  // rows of a level per wave (a year holds millions)
constexpr int kMaxDraws = 1024;  // draws per block
constexpr int kDPL = kMaxDraws / 64;
constexpr int kMaxCov = 64;       // covariates in a leave-out design
constexpr int kCovBlocks = 512;   // row blocks of the covariate reduction

// The splitmix64-finalised counter hash of (seed, row ordinal, draw); the top
// bit gives the sign (core/kernels_groups.py:71-77).
__device__ __forceinline__ double rademacher(uint64_t seed, uint64_t row, uint64_t draw) {
    uint64_t x = seed ^ (row * 0x9E3779B97F4A7C15ull) ^ (draw * 0xBF58476D1CE4E5B9ull);
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    x = x ^ (x >> 31);
    return (x >> 63) ? -1.0 : 1.0;
}

// out[i, j] = draw (ordinal + i, first + j), row major (n, k).
__global__ void k_rademacher(int64_t ordinal, int64_t n, int k, uint64_t seed, int64_t first,
                             double *__restrict__ out) {
    const int64_t nk = n * k;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nk;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e / k;
        const int j = (int)(e - i * k);
        out[e] = rademacher(seed, (uint64_t)(ordinal + i), (uint64_t)(first + j));
    }
}

// The row vector entry the passes work on: the draw, in the square-root
// weight metric divided by sqrt(w) (inverse.py:172-174, 299-305).
__device__ __forceinline__ double row_value(uint64_t seed, int64_t i, int64_t draw,
                                            const double *w, int sqrtw) {
    const double r = rademacher(seed, (uint64_t)i, (uint64_t)draw);
    return (sqrtw && w) ? r / sqrt(w[i]) : r;
}

// Group means of the row vectors: Rbar[g, j] = sum w r / sum w.  One wave per
// group (kernels_inverse.py:52-60).
__global__ void k_group_rbar(int64_t G, int k, const int *__restrict__ gstart,
                             const double *__restrict__ w, int sqrtw, uint64_t seed,
                             int64_t first, double *__restrict__ Rbar) {
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x / 64);
    for (int64_t g = blockIdx.x * (int64_t)(blockDim.x / 64) + threadIdx.x / 64; g < G;
         g += waves) {
        const int s = gstart[g], e = gstart[g + 1];
        double sw = 0.0;
        for (int i = s; i < e; ++i) sw += w ? w[i] : 1.0;
        for (int t = 0; 64 * t < k; ++t) {
            const int j = lane + 64 * t;
            if (j >= k) break;
            double acc = 0.0;
            for (int i = s; i < e; ++i)
                acc += (w ? w[i] : 1.0) * row_value(seed, i, first + j, w, sqrtw);
            Rbar[g * k + j] = sw > 0.0 ? acc / sw : 0.0;
        }
    }
}

// Reduction onto the levels: part[c, j] = sum over chunk c's rows i of
// w_i (r_ij - Rbar[g(i), j]) (kernels_inverse.py:61-71), rows of each level
// from the level-to-rows CSR of its dimension.
__global__ void k_center_levels(int64_t nchunk, int k, const int *__restrict__ chunk_lvl,
                                const int *__restrict__ chunk_beg, const int *__restrict__ rptr,
                                const int *__restrict__ rrow, const int *__restrict__ g,
                                const double *__restrict__ w, int sqrtw, uint64_t seed,
                                int64_t first, const double *__restrict__ Rbar,
                                double *__restrict__ part) {
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x / 64);
    for (int64_t c = blockIdx.x * (int64_t)(blockDim.x / 64) + threadIdx.x / 64; c < nchunk;
         c += waves) {
        const int a = chunk_lvl[c];
        const int b = chunk_beg[c], e = min(b + kRowChunk, rptr[a + 1]);
        for (int t = 0; 64 * t < k; ++t) {
            const int j = lane + 64 * t;
            if (j >= k) break;
            double acc = 0.0;
            for (int p = b; p < e; ++p) {
                const int i = rrow[p];
                const double wi = w ? w[i] : 1.0;
                acc += wi * (row_value(seed, i, first + j, w, sqrtw) - Rbar[(int64_t)g[i] * k + j]);
            }
            part[c * k + j] = acc;
        }
    }
}

__global__ void k_center_sum(int64_t L, int k, const int *__restrict__ cptr,
                             const double *__restrict__ part, double *__restrict__ rhs) {
    const int64_t n = L * k;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t a = e / k;
        const int j = (int)(e - a * k);
        double s = 0.0;
        for (int c = cptr[a]; c < cptr[a + 1]; ++c) s += part[(int64_t)c * k + j];
        rhs[e] = s;
    }
}

// The covariate block of the reduced right-hand side of a block of draws
// (inverse.py:162-180 with a border): part[b, c, j] = sum over row block b
// of w_i Xd[i, c] r_ij, Xd the covariates less their weighted group means,
// so that the sum is X'W M_0 r; then a fixed-order sum over the blocks.
__global__ __launch_bounds__(256) void k_cov_center(int64_t n, int kc, int k,
                                                    const double *__restrict__ Xd,
                                                    const double *__restrict__ w, int sqrtw,
                                                    uint64_t seed, int64_t first,
                                                    double *__restrict__ part) {
    const int64_t per = (n + gridDim.x - 1) / gridDim.x;
    const int64_t r0 = min(n, (int64_t)blockIdx.x * per), r1 = min(n, r0 + per);
    for (int j = threadIdx.x; j < k; j += blockDim.x) {
        for (int c0 = 0; c0 < kc; c0 += 8) {
            const int cn = min(8, kc - c0);
            double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            for (int64_t i = r0; i < r1; ++i) {
                const double v = (w ? w[i] : 1.0) * row_value(seed, i, first + j, w, sqrtw);
                const double *x = Xd + i * kc + c0;
#pragma unroll
                for (int c = 0; c < 8; ++c)
                    if (c < cn) acc[c] += v * x[c];
            }
            for (int c = 0; c < cn; ++c)
                part[((int64_t)blockIdx.x * kc + c0 + c) * k + j] = acc[c];
        }
    }
}

__global__ void k_cov_finish(int nb, int64_t m, const double *__restrict__ part,
                             double *__restrict__ out) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < m;
         e += (int64_t)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int b = 0; b < nb; ++b) s += part[(int64_t)b * m + e];
        out[e] = s;
    }
}

// Projection back to rows fused with the JLA moments (kernels_inverse.py:
// 74-108, leaveout.py:181-190): per group, h_ij = sum_d U[level_d(i), j],
// u0_j = weighted mean over the group of (r_ij - h_ij), proj = h + u0 (times
// sqrt(w_i) in the square-root metric); with covariates in the design, plus
// Xd_i' C_j (Xd the covariates less their weighted group means, C (kc, k)
// the covariate block of the solution), which is x_i'C with its group mean
// taken out as u0 takes it.  Then for each row, over the block's draws, the
// sums of P = proj^2, M = (r - proj)^2, P^2, M^2 and P M are added to
// moments[0..4][i].  One wave per group.
__global__ void k_project_moments(int64_t G, int D, int k, const int *__restrict__ gstart,
                                  const int *__restrict__ codes, int64_t ld,
                                  const int64_t *__restrict__ offs, const double *__restrict__ w,
                                  int sqrtw, uint64_t seed, int64_t first,
                                  const double *__restrict__ U, int kc,
                                  const double *__restrict__ Xd, const double *__restrict__ Cb,
                                  double *__restrict__ moments, int64_t n) {
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x / 64);
    for (int64_t g = blockIdx.x * (int64_t)(blockDim.x / 64) + threadIdx.x / 64; g < G;
         g += waves) {
        const int s = gstart[g], e = gstart[g + 1];
        double sw = 0.0;
        for (int i = s; i < e; ++i) sw += w ? w[i] : 1.0;
        double u0[kDPL];
        for (int t = 0; t < kDPL; ++t) u0[t] = 0.0;
        // the group means of r - h
        for (int i = s; i < e; ++i) {
            const double wi = w ? w[i] : 1.0;
            for (int t = 0; 64 * t < k && t < kDPL; ++t) {
                const int j = lane + 64 * t;
                if (j >= k) break;
                double h = 0.0;
                for (int d = 0; d < D; ++d) h += U[(offs[d] + codes[d * ld + i]) * k + j];
                u0[t] += wi * (row_value(seed, i, first + j, w, sqrtw) - h);
            }
        }
        for (int t = 0; t < kDPL; ++t) u0[t] = sw > 0.0 ? u0[t] / sw : 0.0;
        for (int i = s; i < e; ++i) {
            const double root = (sqrtw && w) ? sqrt(w[i]) : 1.0;
            double m0 = 0, m1 = 0, m2 = 0, m3 = 0, m4 = 0;
            for (int t = 0; 64 * t < k && t < kDPL; ++t) {
                const int j = lane + 64 * t;
                if (j >= k) break;
                double h = 0.0;
                for (int d = 0; d < D; ++d) h += U[(offs[d] + codes[d * ld + i]) * k + j];
                for (int c = 0; c < kc; ++c) h += Xd[(int64_t)i * kc + c] * Cb[c * k + j];
                const double proj = (h + u0[t]) * root;
                const double r = rademacher(seed, (uint64_t)i, (uint64_t)(first + j));
                const double P = proj * proj, dm = r - proj, M = dm * dm;
                m0 += P;
                m1 += M;
                m2 += P * P;
                m3 += M * M;
                m4 += P * M;
            }
            m0 = wave_sum(m0);
            m1 = wave_sum(m1);
            m2 = wave_sum(m2);
            m3 = wave_sum(m3);
            m4 = wave_sum(m4);
            if (lane == 0) {
                moments[i] += m0;
                moments[n + i] += m1;
                moments[2 * n + i] += m2;
                moments[3 * n + i] += m3;
                moments[4 * n + i] += m4;
            }
        }
    }
}

// The 2021 note's algebra per row (leaveout.py:209-282): in place, P -> P_bar,
// M -> M_bar = 1 - P_bar, mPP -> the non-linearity factor; per-block partials
// of (min P_bar, max P_bar, sum P_bar, max |P + M - 1|).
__global__ void k_combine_moments(int64_t n, double n_draws, double *__restrict__ mom,
                                  double *__restrict__ part) {
    __shared__ double lds[256 / 64];
    double lo = 1e300, hi = -1e300, sum = 0.0, worst = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double P = mom[i] / n_draws, M = mom[n + i] / n_draws;
        const double mPP = mom[2 * n + i] / n_draws, mMM = mom[3 * n + i] / n_draws;
        const double mPM = mom[4 * n + i] / n_draws;
        const double total = P + M;
        const double Pb = total > 0.0 ? P / total : 0.0;
        const double Mb = 1.0 - Pb;
        const double V = (Mb * Mb * mPP + Pb * Pb * mMM - 2.0 * Pb * Mb * mPM) / n_draws;
        const double B = (Mb * mPP - Pb * mMM + (Mb - Pb) * mPM) / n_draws;
        const double f = Mb > 0.0 ? 1.0 - V / (Mb * Mb) + B / Mb : 1.0;
        mom[i] = Pb;
        mom[n + i] = Mb;
        mom[2 * n + i] = f;
        lo = fmin(lo, Pb);
        hi = fmax(hi, Pb);
        sum += Pb;
        worst = fmax(worst, fabs(total - 1.0));
    }
    // fixed-shape reductions (min/max are order-free; the sum is a tree)
    for (int o = 32; o > 0; o >>= 1) {
        lo = fmin(lo, __shfl_xor(lo, o, 64));
        hi = fmax(hi, __shfl_xor(hi, o, 64));
        worst = fmax(worst, __shfl_xor(worst, o, 64));
    }
    __shared__ double s_lo[4], s_hi[4], s_w[4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        s_lo[wv] = lo;
        s_hi[wv] = hi;
        s_w[wv] = worst;
    }
    sum = block_sum<double>(sum, lds);
    if (threadIdx.x == 0) {
        for (int q = 1; q < (int)(blockDim.x / 64); ++q) {
            lo = fmin(lo, s_lo[q]);
            hi = fmax(hi, s_hi[q]);
            worst = fmax(worst, s_w[q]);
        }
        lo = fmin(lo, s_lo[0]);
        hi = fmax(hi, s_hi[0]);
        worst = fmax(worst, s_w[0]);
        part[blockIdx.x * 4] = lo;
        part[blockIdx.x * 4 + 1] = hi;
        part[blockIdx.x * 4 + 2] = sum;
        part[blockIdx.x * 4 + 3] = worst;
    }
}

// Level-to-rows CSR helpers: keys (level << rbits | row) per dimension.
__global__ void k_level_row_keys(int64_t n, const int *__restrict__ code, int64_t off, int rbits,
                                 uint64_t *__restrict__ key) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        key[i] = ((uint64_t)(off + code[i]) << rbits) | (uint64_t)i;
}

__global__ void k_key_rows(int64_t n, const uint64_t *__restrict__ key, int rbits,
                           int *__restrict__ row) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        row[i] = (int)(key[i] & ((1ull << rbits) - 1));
}

// rptr[a] = first position of level a among the sorted keys of its
// dimension's block [base, base + n), for the levels [lo, hi) of that dimension.
__global__ void k_key_ptr(int64_t lo, int64_t hi, int64_t base, int64_t n,
                          const uint64_t *__restrict__ key, int rbits, int *__restrict__ rptr) {
    for (int64_t a = lo + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; a <= hi;
         a += (int64_t)gridDim.x * blockDim.x) {
        int64_t l = 0, h = n;
        const uint64_t target = (uint64_t)a << rbits;
        while (l < h) {
            const int64_t mid = (l + h) >> 1;
            if (key[mid] < target) l = mid + 1;
            else h = mid;
        }
        rptr[a] = (int)(base + l);
    }
}

__global__ void k_chunk_count(int64_t L, const int *__restrict__ rptr, int *__restrict__ nch) {
    for (int64_t a = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; a <= L;
         a += (int64_t)gridDim.x * blockDim.x)
        nch[a] = a < L ? max(1, (rptr[a + 1] - rptr[a] + kRowChunk - 1) / kRowChunk) : 0;
}

__global__ void k_chunk_fill(int64_t L, const int *__restrict__ rptr, const int *__restrict__ cptr,
                             int *__restrict__ clvl, int *__restrict__ cbeg) {
    for (int64_t a = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; a < L;
         a += (int64_t)gridDim.x * blockDim.x)
        for (int c = cptr[a], b = rptr[a]; c < cptr[a + 1]; ++c, b += kRowChunk) {
            clvl[c] = (int)a;
            cbeg[c] = b;
        }
}

}  // namespace hdfe

using namespace hdfe;

// The row-pass state kept per panel: the level-to-rows CSR of every
// dimension and its chunks, and scratch for the group means.
struct RowPasses {
    hipStream_t s = nullptr;
    int dev = 0;  // the device of s
    int64_t n = 0, G = 0, L = 0;
    int D = 0;
    DevBuf<int> rptr, rrow, cptr, clvl, cbeg;
    int64_t nchunk = 0;
    DevBuf<double> Rbar, part, cpart;
    int64_t kcap = 0, ccap = 0;
    Temp tmp;
};

extern "C" {

// The entry points below that launch on the null stream synchronise the
// device before their first launch and after their last, so that work the
// caller queued on any stream (torch's current one included) is done before
// they read or write its buffers, and theirs is done when they return.

// A draw table on the device: out (n, k) = draws (ordinal + i, first + j).
int hdfe_rademacher(int64_t ordinal, int64_t n, int k, uint64_t seed, int64_t first_draw,
                    double *out) {
    ARG_CHECK(out && n >= 0 && k >= 1, "hdfe_rademacher: bad argument");
    if (n == 0) return HDFE_STATUS_OK;
    HIP_TRY(hipDeviceSynchronize());
    hipLaunchKernelGGL(k_rademacher, dim3(grid1(n * k)), dim3(256), 0, 0, ordinal, n, k, seed,
                       first_draw, out);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    return HDFE_STATUS_OK;
}

// Row-pass state for sorted rows (device arrays of a panel): gstart (G+1),
// codes (D, n) row-major by dimension, offs (D, host), n_levels (D, host).
int hdfe_rows_create(int64_t n, int64_t G, int D, const int *codes, const int64_t *offs_h,
                     const int64_t *nlev_h, void **out) {
    ARG_CHECK(out && codes && offs_h && nlev_h && D >= 1, "hdfe_rows_create: bad argument");
    std::unique_ptr<RowPasses> R(new RowPasses());
    HIP_TRY(hipGetDevice(&R->dev));
    HIP_TRY(stream_get(&R->s));
    hipStream_t s = R->s;
    R->n = n;
    R->G = G;
    R->D = D;
    int64_t L = 0;
    for (int d = 0; d < D; ++d) L += nlev_h[d];
    R->L = L;
    const int rbits = std::max(1, bits_for((uint64_t)n));
    const int lbits = std::max(1, bits_for((uint64_t)L));
    ARG_CHECK(rbits + lbits <= 64, "hdfe_rows_create: keys need more than 64 bits");
    DEV_ALLOC(R->rptr, L + 1, s);
    DEV_ALLOC(R->rrow, n * D, s);
    DevBuf<uint64_t> k1, k2;
    DevBuf<int> dummy, dummy2;
    DEV_ALLOC(k1, n, s);
    DEV_ALLOC(k2, n, s);
    for (int d = 0; d < D; ++d) {
        hipLaunchKernelGGL(k_level_row_keys, dim3(grid1(n)), dim3(256), 0, s, n, codes + d * n,
                           offs_h[d], rbits, k1.p);
        size_t bytes = 0;
        HIP_TRY(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, k1.p, k2.p, (int)n, 0,
                                                  rbits + lbits, s));
        HIP_TRY(R->tmp.ensure(bytes, s));
        bytes = (size_t)R->tmp.buf.n;
        HIP_TRY(hipcub::DeviceRadixSort::SortKeys(R->tmp.buf.p, bytes, k1.p, k2.p, (int)n, 0,
                                                  rbits + lbits, s));
        hipLaunchKernelGGL(k_key_rows, dim3(grid1(n)), dim3(256), 0, s, n, k2.p, rbits,
                           R->rrow.p + d * n);
        const int64_t lo = offs_h[d], hi = offs_h[d] + nlev_h[d];
        hipLaunchKernelGGL(k_key_ptr, dim3(grid1(nlev_h[d] + 1)), dim3(256), 0, s, lo, hi,
                           (int64_t)d * n, n, k2.p, rbits, R->rptr.p);
    }
    // the last entry: every dimension's block ends at its own base + n, and
    // the next dimension's first pointer overwrote this one's end; fix the
    // global end
    {
        const int end = (int)(n * D);
        HIP_TRY(hipMemcpyAsync(R->rptr.p + L, &end, 4, hipMemcpyHostToDevice, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    DevBuf<int> nch;
    DEV_ALLOC(nch, L + 1, s);
    DEV_ALLOC(R->cptr, L + 1, s);
    hipLaunchKernelGGL(k_chunk_count, dim3(grid1(L + 1)), dim3(256), 0, s, L, R->rptr.p, nch.p);
    HIP_TRY(exclusive_sum(R->tmp, nch.p, R->cptr.p, L + 1, s));
    int nchunk = 0;
    HIP_TRY(scalar_to_host(&nchunk, R->cptr.p + L, s));
    R->nchunk = nchunk;
    DEV_ALLOC(R->clvl, nchunk, s);
    DEV_ALLOC(R->cbeg, nchunk, s);
    hipLaunchKernelGGL(k_chunk_fill, dim3(grid1(L)), dim3(256), 0, s, L, R->rptr.p, R->cptr.p,
                       R->clvl.p, R->cbeg.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    *out = R.release();
    return HDFE_STATUS_OK;
}

static int ensure_k(RowPasses &R, int k) {
    if (R.kcap >= k) return HDFE_STATUS_OK;
    DEV_ALLOC(R.Rbar, R.G * (int64_t)k, R.s);
    DEV_ALLOC(R.part, R.nchunk * (int64_t)k, R.s);
    R.kcap = k;
    return HDFE_STATUS_OK;
}

// The reduced right-hand side of a block of k draws (inverse.py:162-180 with
// the Rademacher vectors of leaveout.py:179): rhs (L, k) device output.
// g (n) are the rows' group codes, gstart (G+1), w (n) or NULL.
int hdfe_rows_center(void *h, const int *g, const int *gstart, const double *w, int sqrtw,
                     uint64_t seed, int64_t first_draw, int k, double *rhs) {
    ARG_CHECK(h && g && gstart && rhs, "hdfe_rows_center: null argument");
    ARG_CHECK(k >= 1 && k <= kMaxDraws, "hdfe_rows_center: k must be in [1, 1024]");
    RowPasses &R = *(RowPasses *)h;
    int st = ensure_k(R, k);
    if (st) return st;
    hipStream_t s = R.s;
    hipLaunchKernelGGL(k_group_rbar, dim3(grid1(R.G * 64)), dim3(256), 0, s, R.G, k, gstart, w,
                       sqrtw, seed, first_draw, R.Rbar.p);
    hipLaunchKernelGGL(k_center_levels, dim3(grid1(R.nchunk * 64)), dim3(256), 0, s, R.nchunk, k,
                       R.clvl.p, R.cbeg.p, R.rptr.p, R.rrow.p, g, w, sqrtw, seed, first_draw,
                       R.Rbar.p, R.part.p);
    hipLaunchKernelGGL(k_center_sum, dim3(grid1(R.L * k)), dim3(256), 0, s, R.L, k, R.cptr.p,
                       R.part.p, rhs);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    return HDFE_STATUS_OK;
}

// The covariate block of a block of draws' reduced right-hand side: out
// (kc, k) device = X'W M_0 r from Xd (n, kc), the covariates less their
// weighted group means.
int hdfe_rows_cov_center(void *h, int kc, const double *Xd, const double *w, int sqrtw,
                         uint64_t seed, int64_t first_draw, int k, double *out) {
    ARG_CHECK(h && Xd && out, "hdfe_rows_cov_center: null argument");
    ARG_CHECK(kc >= 1 && kc <= kMaxCov && k >= 1 && k <= kMaxDraws,
              "hdfe_rows_cov_center: kc must be in [1, 64] and k in [1, 1024]");
    RowPasses &R = *(RowPasses *)h;
    const int64_t m = (int64_t)kc * k;
    const int nb = (int)std::max<int64_t>(
        1, std::min<int64_t>({(int64_t)kCovBlocks, (R.n + 255) / 256, (int64_t)(64 << 20) / (8 * m)}));
    if (R.ccap < nb * m) {
        DEV_ALLOC(R.cpart, nb * m, R.s);
        R.ccap = nb * m;
    }
    hipLaunchKernelGGL(k_cov_center, dim3(nb), dim3(256), 0, R.s, R.n, kc, k, Xd, w, sqrtw, seed,
                       first_draw, R.cpart.p);
    hipLaunchKernelGGL(k_cov_finish, dim3(grid1(m)), dim3(256), 0, R.s, nb, m, R.cpart.p, out);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(R.s));
    return HDFE_STATUS_OK;
}

// Back to rows and the JLA moments of the block: moments (5, n) device,
// accumulated; U (L, k) device, the reduced solution; with covariates in the
// design, Xd (n, kc) and the solution's covariate block Cb (kc, k).
int hdfe_rows_project_moments(void *h, const int *gstart, const int *codes, const double *w,
                              int sqrtw, uint64_t seed, int64_t first_draw, int k,
                              const int64_t *offs_d, const double *U, int kc, const double *Xd,
                              const double *Cb, double *moments) {
    ARG_CHECK(h && gstart && codes && U && moments && offs_d, "hdfe_rows_project_moments: null");
    ARG_CHECK(k >= 1 && k <= kMaxDraws, "hdfe_rows_project_moments: k must be in [1, 1024]");
    ARG_CHECK(kc >= 0 && kc <= kMaxCov && (kc == 0 || (Xd && Cb)),
              "hdfe_rows_project_moments: kc must be in [0, 64] with Xd and Cb");
    RowPasses &R = *(RowPasses *)h;
    hipStream_t s = R.s;
    hipLaunchKernelGGL(k_project_moments, dim3(grid1(R.G * 64)), dim3(256), 0, s, R.G, R.D, k,
                       gstart, codes, R.n, offs_d, w, sqrtw, seed, first_draw, U, kc, Xd, Cb,
                       moments, R.n);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    return HDFE_STATUS_OK;
}

// Leverages from the moments, in place; stats (host, 4): min, max, mean of
// P_bar and the largest |P + M - 1| of the raw estimates.
int hdfe_jla_combine(int64_t n, int64_t n_draws, double *moments, double *stats) {
    ARG_CHECK(moments && stats && n >= 1 && n_draws >= 1, "hdfe_jla_combine: bad argument");
    HIP_TRY(hipDeviceSynchronize());
    const int nb = 256;
    DevBuf<double> part;
    DEV_ALLOC(part, nb * 4, nullptr);
    hipLaunchKernelGGL(k_combine_moments, dim3(nb), dim3(256), 0, 0, n, (double)n_draws, moments,
                       part.p);
    HIP_TRY(hipGetLastError());
    std::vector<double> h;
    HIP_TRY(to_host(h, part.p, nb * 4, 0));
    double lo = 1e300, hi = -1e300, sum = 0.0, worst = 0.0;
    for (int b = 0; b < nb; ++b) {
        lo = std::min(lo, h[b * 4]);
        hi = std::max(hi, h[b * 4 + 1]);
        sum += h[b * 4 + 2];
        worst = std::max(worst, h[b * 4 + 3]);
    }
    stats[0] = lo;
    stats[1] = hi;
    stats[2] = sum / (double)n;
    stats[3] = worst;
    return HDFE_STATUS_OK;
}

// The level-to-rows CSR and its chunks (device pointers), for kss.hip.
int hdfe_rows_csr(void *h, const int **rptr, const int **rrow, const int **cptr, const int **clvl,
                  const int **cbeg, int64_t *nchunk) {
    ARG_CHECK(h, "hdfe_rows_csr: null handle");
    RowPasses &R = *(RowPasses *)h;
    *rptr = R.rptr.p;
    *rrow = R.rrow.p;
    *cptr = R.cptr.p;
    *clvl = R.clvl.p;
    *cbeg = R.cbeg.p;
    *nchunk = R.nchunk;
    return HDFE_STATUS_OK;
}

// out (L, k) = per level, the sum of its chunks' partials (nchunk, k) in
// chunk order.
int hdfe_rows_chunk_sum(void *h, int k, const double *part, double *out) {
    ARG_CHECK(h && part && out, "hdfe_rows_chunk_sum: null argument");
    RowPasses &R = *(RowPasses *)h;
    HIP_TRY(hipDeviceSynchronize());
    hipLaunchKernelGGL(k_center_sum, dim3(grid1(R.L * k)), dim3(256), 0, 0, R.L, k, R.cptr.p,
                       part, out);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    return HDFE_STATUS_OK;
}

void hdfe_rows_destroy(void *h) {
    if (!h) return;
    RowPasses *R = (RowPasses *)h;
    // buffers are freed on the passes' stream, which must outlive them
    (void)hipDeviceSynchronize();
    hipStream_t s = R->s;
    const int dev = R->dev;
    delete R;
    stream_put(s, dev);
}

}  // extern "C"
