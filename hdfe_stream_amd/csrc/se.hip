// Standard errors of the KSS variance components on the device, for gfx950:
// the row-level passes around the two coefficient-space solves of applying the
// estimator's kernel C, the per-row diagonals B_ii, and the smoother of sigma2.
//
// Reference (hdfe_stream/):
//   bii_rows                      leaveout_se.py:228-327
//   _nb_bii_moments               kernels_se.py:75-98
//   _apply_q, _nb_cov_cross       leaveout_se.py:339-385, kernels_se.py:44-72
//   _apply_c / _apply_c_passes    leaveout_se.py:426-596
//   component_variances, _cc_trace  leaveout_se.py:599-784, 891-937
//   fit_smoother / smooth_eval    leaveout_se.py:982-1054
//
// Everything is in the square-root-weight metric X~ = W^(1/2) A.  Blocks of s
// vectors are row-major (rows, s).  Reductions to scalars are per-workgroup
// partials over fixed row ranges, added on the host in workgroup order, and
// per-cell sums are over rows sorted by cell, one wavefront per cell: every
// result is independent of scheduling.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "devbuf.h"

namespace hdfe {

constexpr int kSeBlocks = 1024;  // workgroups of the reducing passes

// The stacked inputs of the first solve (leaveout_se.py:522-530): with a
// shared input v (n, s): [v, b_0 v, b_1 v, b_2 v]; per component (z != null):
// v_k = z sqrt(max(s_k, 0)) -- zero for a component not in the bit mask --
// and [v_0, v_1, v_2, b_0 v_0, b_1 v_1, b_2 v_2].  b and s3 are (3, n).
__global__ void k_se_stack(int64_t n, int s, const double *__restrict__ v,
                           const double *__restrict__ z, const double *__restrict__ s3,
                           const double *__restrict__ b, int mask, double *__restrict__ out) {
    const int nin = z ? 3 : 1, wide = nin + 3;
    const int64_t total = n * s;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e / s;
        const int j = (int)(e - i * s);
        double vk[3];
        if (z) {
            const double zz = z[e];
#pragma unroll
            for (int k = 0; k < 3; ++k)
                vk[k] = ((mask >> k) & 1) ? zz * sqrt(fmax(s3[k * n + i], 0.0)) : 0.0;
        } else {
            vk[0] = vk[1] = vk[2] = v[e];
        }
        double *o = out + i * (int64_t)wide * s;
        for (int k = 0; k < nin; ++k) o[k * s + j] = vk[k];
#pragma unroll
        for (int k = 0; k < 3; ++k) o[(nin + k) * s + j] = vk[k] * b[k * n + i];
    }
}

// C_k v assembled at row level (leaveout_se.py:575-592) and reduced:
//   M v    = v_k - sqrt(w) base[in(k)],    B v = sqrt(w) thq[k]
//   M(b v) = b_k v_k - sqrt(w) base[b(k)]
//   C v    = B v - (b_k M v + M(b v)) / 2  [+ c (u_k'v 1 + (1'v) u_k)]
// With x1 (3, n) and defl = (lambda (3), x1_k'v_k (3, s)), B v loses
// lambda_k x1_k (x1_k'v_k): C2 of the deflated B2 (leaveout_se.py:579-580).
// mode 0 (C y, s = 1): part[w] = (4 s_k (Cy)_k^2, yt (Cy)_k) for k < 3;
// mode 1 (trace): part[w] = (s_k sum_j (C v)_kj^2) for k < 3;
// mode 2 (C2 y, s = 1, s floored at 0): (2 x1_k s_k (C2y)_k, 4 s_k (C2y)_k^2,
//   yt (C2y)_k) for k < 3; mode 3 (trace, s floored at 0): as mode 1.
// base (n, wide s), thq (n, 3 s); u (3, n), udot/odot (3, s) or null.
__global__ __launch_bounds__(256) void k_se_assemble(
    int64_t n, int s, int mode, const double *__restrict__ v, const double *__restrict__ z,
    const double *__restrict__ s3, const double *__restrict__ b, const double *__restrict__ w,
    const double *__restrict__ base, const double *__restrict__ thq, const double *__restrict__ u,
    double c2, const double *__restrict__ udot, const double *__restrict__ odot,
    const double *__restrict__ yt, const double *__restrict__ x1,
    const double *__restrict__ defl, int mask, double *__restrict__ part) {
    __shared__ double lds[256 / 64];
    const int nin = z ? 3 : 1, wide = nin + 3;
    const int64_t r0 = part_begin(n, blockIdx.x, gridDim.x), r1 = part_begin(n, blockIdx.x + 1, gridDim.x);
    double acc[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const bool floor0 = mode >= 2;
    const int64_t total = (r1 - r0) * s;
    for (int64_t e = threadIdx.x; e < total; e += blockDim.x) {
        const int64_t i = r0 + e / s;
        const int j = (int)(e % s);
        const double root = w ? sqrt(w[i]) : 1.0;
        const double *bs = base + i * (int64_t)wide * s;
        const double *tq = thq + i * (int64_t)3 * s;
        const double zz = z ? z[i * s + j] : 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double sr = s3[k * n + i];
            const double sk = floor0 ? fmax(sr, 0.0) : sr;
            const bool on = (mask >> k) & 1;
            const double vk = z ? (on ? zz * sqrt(fmax(sr, 0.0)) : 0.0) : v[i * s + j];
            const double bk = b[k * n + i];
            const double Mv = vk - root * bs[(z ? k : 0) * s + j];
            double Bv = root * tq[k * s + j];
            if (x1) Bv -= defl[k] * x1[k * n + i] * defl[3 + k * s + j];
            const double Mbv = vk * bk - root * bs[(nin + k) * s + j];
            double cv = Bv - 0.5 * (bk * Mv + Mbv);
            if (u) cv += c2 * (udot[k * s + j] + u[k * n + i] * odot[k * s + j]);
            if (mode == 0) {
                acc[k] += 4.0 * sk * cv * cv;
                acc[3 + k] += yt[i] * cv;
            } else if (mode == 2) {
                acc[k] += 2.0 * x1[k * n + i] * sk * cv;
                acc[3 + k] += 4.0 * sk * cv * cv;
                acc[6 + k] += yt[i] * cv;
            } else if (on) {
                acc[k] += sk * cv * cv;
            }
        }
    }
    const int nout = mode == 0 ? 6 : mode == 2 ? 9 : 3;
    for (int t = 0; t < nout; ++t) {
        const double r = block_sum<double>(acc[t], lds);
        if (threadIdx.x == 0) part[(int64_t)blockIdx.x * nout + t] = r;
    }
}

// B_ii moments of one block (kernels_se.py:75-98): with r1, r2 the first and
// second s columns of rows (n, 2 s), x = sqrt(w) r1, y = sqrt(w) r2:
// out[0] += sum x^2, out[1] += sum y^2, out[2] += sum x y (out (3, n)).
__global__ void k_se_bii(int64_t n, int s, const double *__restrict__ rows,
                         const double *__restrict__ w, double *__restrict__ out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double root2 = w ? w[i] : 1.0;
        const double *r = rows + i * (int64_t)2 * s;
        double a = 0.0, bb = 0.0, c = 0.0;
        for (int j = 0; j < s; ++j) {
            const double x = r[j], y = r[s + j];
            a += x * x;
            bb += y * y;
            c += x * y;
        }
        out[i] += root2 * a;
        out[n + i] += root2 * bb;
        out[2 * n + i] += root2 * c;
    }
}

// The covariance's cross sum over the groups (kernels_se.py:44-72):
// out[g, j] = sum over g's rows of w_i Ul[psi_off + psi_code_i, j].
__global__ void k_se_cross_groups(int64_t G, int s, const int *__restrict__ gstart,
                                  const int *__restrict__ psi_codes, int64_t psi_off,
                                  const double *__restrict__ w, const double *__restrict__ Ul,
                                  double *__restrict__ out) {
    const int64_t total = G * s;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t g = e / s;
        const int j = (int)(e - g * s);
        double t = 0.0;
        for (int i = gstart[g]; i < gstart[g + 1]; ++i)
            t += (w ? w[i] : 1.0) * Ul[(psi_off + psi_codes[i]) * s + j];
        out[e] = t;
    }
}

// np.interp(x, knots, linspace(0, 1, nk)) as numpy evaluates it (the same
// operations, none fused).
__device__ __forceinline__ double se_rank(const double *__restrict__ kn, int nk, double x) {
#pragma clang fp contract(off)
    const double step = 1.0 / (double)(nk - 1);
    if (x < kn[0]) return 0.0;
    if (x >= kn[nk - 1]) return 1.0;
    int lo = 0, hi = nk - 1;  // kn[lo] <= x < kn[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (kn[mid] <= x) lo = mid;
        else hi = mid;
    }
    const double f0 = lo * step, f1 = (lo + 1 == nk - 1) ? 1.0 : (lo + 1) * step;
    if (kn[lo] == x) return f0;
    const double slope = (f1 - f0) / (kn[lo + 1] - kn[lo]);
    return slope * (x - kn[lo]) + f0;
}

// The smoother's cell of each row (leaveout_se.py:1024-1029); without bii,
// the leverage's alone.
__global__ void k_se_cells(int64_t n, const double *__restrict__ lev,
                           const double *__restrict__ bii, const double *__restrict__ kp,
                           const double *__restrict__ kb, int nb, uint32_t *__restrict__ cell,
                           int *__restrict__ idx) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t cp = min((int64_t)(se_rank(kp, nb + 1, lev[i]) * nb), (int64_t)nb - 1);
        if (bii) {
            const int64_t cb = min((int64_t)(se_rank(kb, nb + 1, bii[i]) * nb), (int64_t)nb - 1);
            cell[i] = (uint32_t)(cp * nb + cb);
        } else {
            cell[i] = (uint32_t)cp;
        }
        idx[i] = (int)i;
    }
}

// Per cell, the sum of sigma2 over its rows and their count: rows sorted by
// cell (idx the rows in that order), a wavefront per cell.
__global__ void k_se_cell_sums(int64_t n, int ncell, const uint32_t *__restrict__ scell,
                               const int *__restrict__ idx, const double *__restrict__ sigma2,
                               double *__restrict__ tot, double *__restrict__ cnt) {
    const int lane = threadIdx.x & 63;
    const int64_t nwave = (int64_t)gridDim.x * (blockDim.x / 64);
    for (int64_t c = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / 64; c < ncell; c += nwave) {
        // [lo, hi): the rows of cell c
        int64_t lo = 0, hi = n;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (scell[mid] < (uint32_t)c) lo = mid + 1;
            else hi = mid;
        }
        int64_t e = lo, top = n;
        while (e < top) {
            const int64_t mid = (e + top) >> 1;
            if (scell[mid] <= (uint32_t)c) e = mid + 1;
            else top = mid;
        }
        double t = 0.0;
        for (int64_t r = lo + lane; r < e; r += 64) t += sigma2[idx[r]];
        t = wave_sum(t);
        if (lane == 0) {
            tot[c] = t;
            cnt[c] = (double)(e - lo);
        }
    }
}

// The smoothed sigma2, bilinear between cell centres (leaveout_se.py:1038-1054);
// without bii, linear along the leverage.
__global__ void k_se_smooth(int64_t n, const double *__restrict__ lev,
                            const double *__restrict__ bii, const double *__restrict__ kp,
                            const double *__restrict__ kb, const double *__restrict__ means,
                            int nb, double *__restrict__ out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double fp = se_rank(kp, nb + 1, lev[i]) * nb - 0.5;
        const int64_t i0 = min(max((int64_t)floor(fp), (int64_t)0), (int64_t)nb - 2);
        const double a = fmin(fmax(fp - i0, 0.0), 1.0);
        if (!bii) {
            out[i] = (1.0 - a) * means[i0] + a * means[i0 + 1];
            continue;
        }
        const double fb = se_rank(kb, nb + 1, bii[i]) * nb - 0.5;
        const int64_t j0 = min(max((int64_t)floor(fb), (int64_t)0), (int64_t)nb - 2);
        const double c = fmin(fmax(fb - j0, 0.0), 1.0);
        out[i] = (1.0 - a) * (1.0 - c) * means[i0 * nb + j0] + a * (1.0 - c) * means[(i0 + 1) * nb + j0] +
                 (1.0 - a) * c * means[i0 * nb + j0 + 1] + a * c * means[(i0 + 1) * nb + j0 + 1];
    }
}

}  // namespace hdfe

using namespace hdfe;

// A' diag(w) B for tall A (K, a) and B (K, b), rows strided by lda and ldb (a b small,
// K in the millions: the shapes of x1'v, u'v and the Krylov projections, for
// which a library GEMM without a split over K runs one tile down all K rows).
// Each workgroup sums its row range for every pair (p, q) in a fixed order;
// the finish adds the workgroups' partials in order.
__global__ __launch_bounds__(256) void k_tall_dot(int64_t K, int a, int b,
                                                  const double *__restrict__ A, int64_t lda,
                                                  const double *__restrict__ B, int64_t ldb,
                                                  const double *__restrict__ w,
                                                  double *__restrict__ part) {
    __shared__ double lds[256];
    const int P = a * b;
    // with few pairs, nrl row lanes per pair, added in lane order
    const int nrl = P < (int)blockDim.x ? (int)blockDim.x / P : 1;
    const int rl = (int)threadIdx.x / (P < (int)blockDim.x ? P : (int)blockDim.x);
    const int64_t r0 = part_begin(K, blockIdx.x, gridDim.x), r1 = part_begin(K, blockIdx.x + 1, gridDim.x);
    for (int t0 = 0; t0 < P; t0 += (nrl > 1 ? P : (int)blockDim.x)) {
        const int t = t0 + (nrl > 1 ? (int)threadIdx.x % P : (int)threadIdx.x);
        const bool act = rl < nrl && t < P;
        double acc = 0.0;
        if (act) {
            const int p = t / b, q = t - p * b;
            if (w)
                for (int64_t i = r0 + rl; i < r1; i += nrl)
                    acc += w[i] * A[i * lda + p] * B[i * ldb + q];
            else
                for (int64_t i = r0 + rl; i < r1; i += nrl) acc += A[i * lda + p] * B[i * ldb + q];
        }
        if (nrl > 1) {
            lds[threadIdx.x] = act ? acc : 0.0;
            __syncthreads();
            if ((int)threadIdx.x < P) {
                double v = 0.0;
                for (int r = 0; r < nrl; ++r) v += lds[r * P + threadIdx.x];
                part[(int64_t)blockIdx.x * P + threadIdx.x] = v;
            }
            __syncthreads();
        } else if (act) {
            part[(int64_t)blockIdx.x * P + t] = acc;
        }
    }
}

__global__ void k_tall_dot_finish(int nb, int P, const double *__restrict__ part,
                                  double *__restrict__ out) {
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < P; t += gridDim.x * blockDim.x) {
        double v = 0.0;
        for (int w = 0; w < nb; ++w) v += part[(int64_t)w * P + t];
        out[t] = v;
    }
}

// The allocation-free variant's partials: a file-scope device array, one copy
// per GPU, held until the process ends.  Host code never passes `scratch`
// itself to a kernel (on the host the name is a stand-in, not a device
// address): it uses the address hipGetSymbolAddress returns, kept per device.
constexpr int64_t kNoAllocScratch = 1 << 20;  // doubles (8 MiB)
static __device__ __align__(256) double scratch[kNoAllocScratch];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[kMaxDevices];

// This device's address of `scratch`, looked up once per device with that
// device current.
static int scratch_ptr(double **out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    ARG_CHECK(dev >= 0 && dev < kMaxDevices, "hdfe_tall_dot_noalloc: device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        ARG_CHECK(p != nullptr, "hdfe_tall_dot_noalloc: no device address for the scratch array");
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return HDFE_STATUS_OK;
}

extern "C" {

// out (a, b) device, row major = A' diag(w) B for A (K, a), B (K, b) device
// with row strides lda, ldb (elements), w (K) or NULL; fixed-order sums.
int hdfe_tall_dot(int64_t K, int a, int b, const double *A, int64_t lda, const double *B,
                  int64_t ldb, const double *w, double *out) {
    ARG_CHECK(K >= 0 && a >= 1 && b >= 1 && (int64_t)a * b <= (1 << 16) && A && B && out &&
                  lda >= a && ldb >= b,
              "hdfe_tall_dot: bad argument");
    const int P = a * b;
    const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(kSeBlocks, (K + 4095) / 4096));
    DevBuf<double> part;
    DEV_ALLOC(part, (int64_t)nb * P, nullptr);
    hipLaunchKernelGGL(k_tall_dot, dim3(nb), dim3(256), 0, 0, K, a, b, A, lda, B, ldb, w, part.p);
    hipLaunchKernelGGL(k_tall_dot_finish, dim3(grid1(P)), dim3(256), 0, 0, nb, P, part.p, out);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    return HDFE_STATUS_OK;
}

// hdfe_tall_dot without any allocation: the partials go to the device's
// `scratch` array (nb * a * b doubles, at most kNoAllocScratch), the launches
// to `stream` (NULL: the null stream), and nothing waits for them.  Same
// partition and order of sums as hdfe_tall_dot, so the same result bit for
// bit.  Calls that overlap on one GPU share its array and race.
int hdfe_tall_dot_noalloc(int64_t K, int a, int b, const double *A, int64_t lda, const double *B,
                          int64_t ldb, const double *w, double *out, void *stream) {
    ARG_CHECK(K >= 0 && a >= 1 && b >= 1 && (int64_t)a * b <= (1 << 16) && A && B && out &&
                  lda >= a && ldb >= b,
              "hdfe_tall_dot_noalloc: bad argument");
    const int P = a * b;
    const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(kSeBlocks, (K + 4095) / 4096));
    if ((int64_t)nb * P > kNoAllocScratch) {
        set_error("hdfe_tall_dot_noalloc: " + std::to_string(nb) + " partials of " +
                  std::to_string(P) + " pairs exceed the " + std::to_string(kNoAllocScratch) +
                  "-double scratch array (use hdfe_tall_dot)");
        return HDFE_STATUS_ARG;
    }
    double *part = nullptr;
    TRY(scratch_ptr(&part));
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_tall_dot, dim3(nb), dim3(256), 0, st, K, a, b, A, lda, B, ldb, w, part);
    hipLaunchKernelGGL(k_tall_dot_finish, dim3(grid1(P)), dim3(256), 0, st, nb, P, part, out);
    HIP_TRY(hipGetLastError());
    return HDFE_STATUS_OK;
}

// The stacked first-solve input (see k_se_stack); out (n, (nin + 3) s); mask
// the components with draws (bits 0-2).
int hdfe_se_stack(int64_t n, int s, const double *v, const double *z, const double *s3,
                  const double *b, int mask, double *out) {
    ARG_CHECK(n >= 1 && s >= 1 && b && out && (v || (z && s3)), "hdfe_se_stack: bad argument");
    hipLaunchKernelGGL(k_se_stack, dim3(grid1(n * s)), dim3(256), 0, 0, n, s, v, z, s3, b, mask,
                       out);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    return HDFE_STATUS_OK;
}

// C v reduced (see k_se_assemble): out (host) = mode 0: first (3), theta (3);
// mode 1: sum_i s_k,i sum_j (C v)_kij^2 (3); mode 2: s12 (3), first (3),
// quad (3); mode 3: as mode 1 with s floored, components in mask.
int hdfe_se_assemble(int64_t n, int s, int mode, const double *v, const double *z,
                     const double *s3, const double *b, const double *w, const double *base,
                     const double *thq, const double *u, double c2, const double *udot,
                     const double *odot, const double *yt, const double *x1, const double *defl,
                     int mask, double *out) {
    ARG_CHECK(n >= 1 && s >= 1 && mode >= 0 && mode <= 3 && s3 && b && base && thq && out &&
                  (v || z) && ((mode & 1) || yt) && (!u || (udot && odot)) &&
                  (!x1 == !defl) && (mode != 2 || x1),
              "hdfe_se_assemble: bad argument");
    const int nb = (int)std::min<int64_t>(kSeBlocks, n), nout = mode == 0 ? 6 : mode == 2 ? 9 : 3;
    DevBuf<double> part;
    DEV_ALLOC(part, (int64_t)nb * nout, nullptr);
    hipLaunchKernelGGL(k_se_assemble, dim3(nb), dim3(256), 0, 0, n, s, mode, v, z, s3, b, w,
                       base, thq, u, c2, udot, odot, yt, x1, defl, mask, part.p);
    HIP_TRY(hipGetLastError());
    std::vector<double> h;
    HIP_TRY(to_host(h, part.p, (int64_t)nb * nout, nullptr));
    for (int t = 0; t < nout; ++t) {
        double a = 0.0;
        for (int q = 0; q < nb; ++q) a += h[(size_t)q * nout + t];
        out[t] = a;
    }
    return HDFE_STATUS_OK;
}

// out (3, n, device) += the B_ii moments of rows (n, 2 s) (see k_se_bii).
int hdfe_se_bii(int64_t n, int s, const double *rows, const double *w, double *out) {
    ARG_CHECK(n >= 1 && s >= 1 && rows && out, "hdfe_se_bii: bad argument");
    hipLaunchKernelGGL(k_se_bii, dim3(grid1(n)), dim3(256), 0, 0, n, s, rows, w, out);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    return HDFE_STATUS_OK;
}

// out (G, s) = per group, sum of w_i Ul[psi_off + psi_code_i] (see k_se_cross_groups).
int hdfe_se_cross_groups(int64_t G, int s, const int *gstart, const int *psi_codes,
                         int64_t psi_off, const double *w, const double *Ul, double *out) {
    ARG_CHECK(G >= 1 && s >= 1 && gstart && psi_codes && Ul && out,
              "hdfe_se_cross_groups: bad argument");
    hipLaunchKernelGGL(k_se_cross_groups, dim3(grid1(G * s)), dim3(256), 0, 0, G, s, gstart,
                       psi_codes, psi_off, w, Ul, out);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    return HDFE_STATUS_OK;
}

// The smoother's cell sums: per cell of the nb x nb grid of (P_ii, B_ii)
// ranks through the knots kp, kb (device, nb + 1 each), the sum of sigma2
// and the row count; tot, cnt host (nb * nb).  With bii and kb null, the nb
// cells of the leverage's ranks alone.
int hdfe_se_cell_sums(int64_t n, const double *lev, const double *bii, const double *sigma2,
                      const double *kp, const double *kb, int nb, double *tot, double *cnt) {
    ARG_CHECK(n >= 1 && n < INT32_MAX && nb >= 2 && nb <= 500 && lev && sigma2 && kp &&
                  (!bii == !kb) && tot && cnt, "hdfe_se_cell_sums: bad argument");
    hipStream_t s = nullptr;
    const int ncell = bii ? nb * nb : nb;
    Temp tmp;
    DevBuf<uint32_t> c1, c2;
    DevBuf<int> i1, i2;
    DevBuf<double> t_d, n_d;
    DEV_ALLOC(c1, n, s);
    DEV_ALLOC(c2, n, s);
    DEV_ALLOC(i1, n, s);
    DEV_ALLOC(i2, n, s);
    DEV_ALLOC(t_d, ncell, s);
    DEV_ALLOC(n_d, ncell, s);
    hipLaunchKernelGGL(k_se_cells, dim3(grid1(n)), dim3(256), 0, s, n, lev, bii, kp, kb, nb,
                       c1.p, i1.p);
    HIP_TRY(sort_pairs(tmp, c1.p, c2.p, i1.p, i2.p, n, 0, bits_for((uint64_t)ncell), s));
    hipLaunchKernelGGL(k_se_cell_sums, dim3(grid1((int64_t)ncell * 64)), dim3(256), 0, s, n,
                       ncell, c2.p, i2.p, sigma2, t_d.p, n_d.p);
    HIP_TRY(hipGetLastError());
    std::vector<double> ht, hn;
    HIP_TRY(to_host(ht, t_d.p, ncell, s));
    HIP_TRY(to_host(hn, n_d.p, ncell, s));
    std::copy(ht.begin(), ht.end(), tot);
    std::copy(hn.begin(), hn.end(), cnt);
    return HDFE_STATUS_OK;
}

// The smoothed sigma2 of every row (see k_se_smooth); means (device, nb x nb,
// or nb with bii and kb null).
int hdfe_se_smooth(int64_t n, const double *lev, const double *bii, const double *kp,
                   const double *kb, const double *means, int nb, double *out) {
    ARG_CHECK(n >= 1 && nb >= 2 && lev && kp && (!bii == !kb) && means && out,
              "hdfe_se_smooth: bad argument");
    hipLaunchKernelGGL(k_se_smooth, dim3(grid1(n)), dim3(256), 0, 0, n, lev, bii, kp, kb, means,
                       nb, out);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    return HDFE_STATUS_OK;
}

}  // extern "C"
