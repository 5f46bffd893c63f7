// Poisson regression with fixed effects by IRLS on the device, many models
// at once, for gfx950.
//
// Reference (hdfe_stream/):
//   _row_terms, _nb_glm_rows       core/kernels_glm.py:99-211 (one fused row pass)
//   out.W += V.T @ V               glm.py:990 (the dense X'WX)
//   _wls, _update                  core/irls.py:100-144
//   _nb_group_effects              core/kernels_glm.py:11-33
//   _irls                          glm.py:743-838 (the loop, step-halving)
//
// Several models are one stacked panel: their fixed-effect levels are kept
// apart, so the fixed-effect Gram matrix is block diagonal and one block CG
// solves every model's [z, X] columns at once.  Rows of a model are
// contiguous in the panel's sorted order; `rs` (models + 1) marks them.
//
// Each step's covariate Hessian X~'W X~ (X~ = X partialled on the fixed
// effects) is a split-BF16 product on the matrix cores: s = sqrt(w) x~ in
// FP32 is split s = hi + lo (two bf16), and s's' ~ hi hi' + hi lo' + lo hi'
// (three v_mfma_f32_32x32x16_bf16, FP32 accumulation).  The coefficient
// update is the Newton step beta += H^-1 X~'W (z~ - X~ beta) with that
// Hessian and the FP64 score, so the converged beta solves the FP64 normal
// equations whatever the Hessian's rounding (only the path differs from the
// reference's, which solves A_xx beta = A_xz exactly each step).
#include <cmath>
#include <vector>

#include "devbuf.h"

namespace hdfe {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kGlmRowChunk = 1024;   // level chunks of the row CSR (rows.hip)
constexpr int kSyrkThreads = 512;    // 8 waves
constexpr int kSyrkMaxBlk = 8;       // 32x32 output blocks per wave
constexpr int kGlmRowsPerWave = 4;   // the IRLS row pass's rows in flight per wave

// ------------------------------------------------------------ row pass

// One IRLS row pass (core/kernels_glm.py:99-211): per row its linear
// predictor at (beta_m, fe, a) -- or the start: log((y + ybar_m) / 2) for
// Poisson, ybar_m for Gaussian -- then, Poisson (family 0): mu = exp(eta),
// W = mu f, u = y / mu - 1, z = eta + u, and the deviance and
// log-likelihood less its constant; Gaussian (family 1, identity link, the
// OLS batch): W = f, z = y, deviance the weighted sum of squared residuals;
// family 2: the linear predictor alone, written to w.  A wave per four rows
// (lanes over the covariates for the dots, then a lane per row); a workgroup
// per chunk of one model's rows; part[c] =
// (dev, sum f (y eta - mu), f y, f) of chunk c (the fifth slot is
// k_glm_lgamma's).
__global__ __launch_bounds__(256) void k_glm_rows(
    int nchunk, const int *__restrict__ ch_model, const int64_t *__restrict__ ch_begin,
    const int64_t *__restrict__ ch_end, int init, int family, const double *__restrict__ ybar,
    const double *__restrict__ y, const float *__restrict__ X, int kx,
    const int *__restrict__ g, const int *__restrict__ codes, int64_t ld, int D,
    const int64_t *__restrict__ offs, const double *__restrict__ fe,
    const double *__restrict__ a, const double *__restrict__ beta,
    const double *__restrict__ f, double *__restrict__ w, double *__restrict__ z,
    double *__restrict__ part) {
    __shared__ double lds[4][4];
    const int c = blockIdx.x;
    const int m = ch_model[c];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t b0 = ch_begin[c], e0 = ch_end[c];
    const double *bm = init ? nullptr : beta + (int64_t)m * kx;
    double acc[4] = {0, 0, 0, 0};
    // a wave takes kGlmRowsPerWave rows at a time: their X rows' loads in
    // flight together, one dot per row across the lanes, then lane u does
    // row u's link, weight and deviance
    for (int64_t i0 = b0 + (int64_t)wv * kGlmRowsPerWave; i0 < e0; i0 += 4 * kGlmRowsPerWave) {
        double s[kGlmRowsPerWave];
#pragma unroll
        for (int u = 0; u < kGlmRowsPerWave; ++u) s[u] = 0.0;
        if (!init) {
            if ((kx & 3) == 0) {
                // 16-byte loads: four covariates per lane (rows start 16-byte aligned)
                const int kq = kx >> 2;
                for (int jq = lane; jq < kq; jq += 64) {
                    const double b0 = bm[4 * jq], b1 = bm[4 * jq + 1], b2 = bm[4 * jq + 2],
                                 b3 = bm[4 * jq + 3];
#pragma unroll
                    for (int u = 0; u < kGlmRowsPerWave; ++u)
                        if (i0 + u < e0) {
                            const float4 xv = *(const float4 *)(X + (i0 + u) * kx + 4 * jq);
                            s[u] += (double)xv.x * b0 + (double)xv.y * b1 + (double)xv.z * b2 +
                                    (double)xv.w * b3;
                        }
                }
            } else {
                for (int j = lane; j < kx; j += 64) {
                    const double bj = bm[j];
#pragma unroll
                    for (int u = 0; u < kGlmRowsPerWave; ++u)
                        if (i0 + u < e0) s[u] += (double)X[(i0 + u) * kx + j] * bj;
                }
            }
#pragma unroll
            for (int u = 0; u < kGlmRowsPerWave; ++u) s[u] = wave_sum(s[u]);
        }
        double su = s[0];
#pragma unroll
        for (int u = 1; u < kGlmRowsPerWave; ++u)
            if (lane == u) su = s[u];
        const int64_t i = i0 + lane;
        if (lane >= kGlmRowsPerWave || i >= e0) continue;
        const double yi = y[i];
        double eta;
        if (init) {
            eta = family == 1 ? ybar[m] : log(0.5 * (yi + ybar[m]));
        } else {
            eta = su + a[g[i]];
            for (int d = 0; d < D; ++d) eta += fe[offs[d] + codes[d * ld + i]];
        }
        if (family == 2) {
            // the linear predictor alone (w = eta), for the fitted values
            w[i] = eta;
            z[i] = yi;
        } else if (family == 1) {
            // Gaussian, identity link: W = f, z = y, deviance = f (y - eta)^2
            const double fi = f ? f[i] : 1.0;
            const double r = yi - eta;
            acc[0] += fi * r * r;
            acc[2] += fi * yi;
            acc[3] += fi;
            w[i] = fi;
            z[i] = yi;
        } else {
            const double mu = exp(eta);
            const double fi = f ? f[i] : 1.0;
            const double u = mu > 0.0 ? yi / mu - 1.0 : 0.0;
            const double xlogy = yi != 0.0 ? yi * log(yi) : 0.0;
            acc[0] += fi * 2.0 * (xlogy - yi * eta - yi + mu);
            acc[1] += fi * (yi * eta - mu);
            acc[2] += fi * yi;
            acc[3] += fi;
            w[i] = mu * fi;
            z[i] = eta + u;
        }
    }
    // the chunk's sums: the wave's lanes (a fixed tree), then the waves in order
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double v = wave_sum(acc[q]);
        if (lane == 0) lds[q][wv] = v;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        const int q = threadIdx.x;
        part[(int64_t)c * 5 + q] = lds[q][0] + lds[q][1] + lds[q][2] + lds[q][3];
    }
}

// part[c * 5 + 4] = sum of f lgamma(y + 1) over chunk c's rows (the
// log-likelihood's constant, computed once).
__global__ __launch_bounds__(256) void k_glm_lgamma(const int64_t *__restrict__ ch_begin,
                                                    const int64_t *__restrict__ ch_end,
                                                    const double *__restrict__ y,
                                                    const double *__restrict__ f,
                                                    double *__restrict__ part) {
    __shared__ double lds[4];
    const int c = blockIdx.x;
    double s = 0.0;
    for (int64_t i = ch_begin[c] + threadIdx.x; i < ch_end[c]; i += blockDim.x)
        s += (f ? f[i] : 1.0) * lgamma(y[i] + 1.0);
    s = block_sum<double>(s, lds);
    if (threadIdx.x == 0) part[(int64_t)c * 5 + 4] = s;
}

// ----------------------------------------------- right-hand sides [z, X]

// The row kernels below walk a group's or a level chunk's rows eight at a
// time with every load's address valid (rows past the end repeat the last,
// weight 0; columns past C and column 0's X read clamped ones), so the rows'
// loads carry no branches and fly together; with per-row guarded loads the
// compiler waited on each row's loads in turn.  Sums keep the rows' order.
constexpr int kGlmBatch = 8;

// Weighted group means of v = [z, x] (G, C): a wave per group, lanes over
// the C = kx + 1 columns (J a lane).
template <int J>
__global__ __launch_bounds__(256) void k_glm_group_means(
    int64_t G, const int *__restrict__ gstart, const double *__restrict__ z,
    const float *__restrict__ X, int kx, const double *__restrict__ w,
    double *__restrict__ vbar) {
    const int C = kx + 1;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t waves = (int64_t)gridDim.x * 4;
    for (int64_t gg = blockIdx.x * 4 + wv; gg < G; gg += waves) {
        const int s = gstart[gg], e = gstart[gg + 1];
        double W = 0.0, acc[J];
#pragma unroll
        for (int t = 0; t < J; ++t) acc[t] = 0.0;
        for (int i0 = s; i0 < e; i0 += kGlmBatch) {
            double wr[kGlmBatch], v[kGlmBatch][J];
#pragma unroll
            for (int u = 0; u < kGlmBatch; ++u) {
                const int i = min(i0 + u, e - 1);
                const double wi = w[i];
                wr[u] = i0 + u < e ? wi : 0.0;
                const double zi = z[i];
                const float *xr = X + (int64_t)i * kx;
#pragma unroll
                for (int t = 0; t < J; ++t) {
                    const int j = lane + 64 * t;
                    const float xv = xr[max(min(j, C - 1) - 1, 0)];
                    v[u][t] = j == 0 ? zi : (double)xv;
                }
            }
#pragma unroll
            for (int u = 0; u < kGlmBatch; ++u) {
                W += wr[u];
#pragma unroll
                for (int t = 0; t < J; ++t) acc[t] += wr[u] * v[u][t];
            }
        }
        const double inv = W > 0.0 ? 1.0 / W : 0.0;
#pragma unroll
        for (int t = 0; t < J; ++t) {
            const int j = lane + 64 * t;
            if (j < C) vbar[gg * C + j] = acc[t] * inv;
        }
    }
}

// part[chunk, j] = sum over the chunk's rows i (of one level, the level-to-
// rows CSR of rows.hip) of w_i (v_ij - vbar[g_i, j]): D_o' W M_0 v
// (kernels_inverse.py:34-71 applied to the data), a wave per chunk.
template <int J>
__global__ __launch_bounds__(256) void k_glm_center(
    int64_t nchunk, const int *__restrict__ chunk_lvl, const int *__restrict__ chunk_beg,
    const int *__restrict__ rptr, const int *__restrict__ rrow, const int *__restrict__ g,
    const double *__restrict__ z, const float *__restrict__ X, int kx,
    const double *__restrict__ w, const double *__restrict__ vbar, double *__restrict__ part) {
    const int C = kx + 1;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t waves = (int64_t)gridDim.x * 4;
    for (int64_t c = blockIdx.x * 4 + wv; c < nchunk; c += waves) {
        const int lv = chunk_lvl[c];
        const int b = chunk_beg[c], e = min(b + kGlmRowChunk, rptr[lv + 1]);
        double acc[J];
#pragma unroll
        for (int t = 0; t < J; ++t) acc[t] = 0.0;
        for (int p0 = b; p0 < e; p0 += kGlmBatch) {
            double wr[kGlmBatch], v[kGlmBatch][J];
#pragma unroll
            for (int u = 0; u < kGlmBatch; ++u) {
                const int i = rrow[min(p0 + u, e - 1)];
                const double wi = w[i];
                wr[u] = p0 + u < e ? wi : 0.0;
                const double zi = z[i];
                const double *vb = vbar + (int64_t)g[i] * C;
                const float *xr = X + (int64_t)i * kx;
#pragma unroll
                for (int t = 0; t < J; ++t) {
                    const int j = lane + 64 * t;
                    const int jj = min(j, C - 1);
                    const float xv = xr[max(jj - 1, 0)];
                    v[u][t] = (j == 0 ? zi : (double)xv) - vb[jj];
                }
            }
#pragma unroll
            for (int u = 0; u < kGlmBatch; ++u)
#pragma unroll
                for (int t = 0; t < J; ++t) acc[t] += wr[u] * v[u][t];
        }
#pragma unroll
        for (int t = 0; t < J; ++t) {
            const int j = lane + 64 * t;
            if (j < C) part[c * C + j] = acc[t];
        }
    }
}

// ------------------------------------------------------ partialled rows

// For the groups of one model per workgroup (gb..ge): x~ = x less the levels'
// effects in Gm (L, C) less its weighted group mean, likewise z~; the score
// partial gpart[wg, j] = sum w_i x~_ij (z~_i - x~_i'beta_m); and s = sqrt(w)
// x~ split into bf16 hi + lo, row-major: element (p, j) of the padded rows at
// p * Cp + j (a row's lanes store one contiguous stretch; the rows-of-eight
// layout the matrix cores read is made while k_glm_syrk stages them).
// A wave per group, lanes over columns, J = ceil(C / 64) columns a lane.  A
// group of at most kGlmStage rows (a worker's years) is read once: its rows'
// x - sum_d Gm[level] stay in registers for the mean and then the values,
// every row's loads in flight together; a longer group takes two sweeps of its
// rows (the means, then the values), gathering the effects twice.
constexpr int kGlmStage = 8;
// Row i's x - sum_d Gm[level_d] (and z likewise in column 0) for the lane's
// columns j = lane + 64 t: every load's address is valid (columns past C and
// column 0's X read clamped ones, the values selected afterwards), so a row's
// loads carry no branches and the rows' loads of a staged group all fly
// together.  i is wave-uniform.
template <int J, int D>
__device__ __forceinline__ void glm_partial_row(int lane, int C, int i, const double *__restrict__ z,
                                                const float *__restrict__ X, int kx,
                                                const int *__restrict__ codes, int64_t ld,
                                                const int64_t *__restrict__ offs,
                                                const double *__restrict__ Gm, double (&u)[J]) {
    int64_t lv[D > 0 ? D : 1];
#pragma unroll
    for (int d = 0; d < D; ++d) lv[d] = (offs[d] + codes[d * ld + i]) * C;
    const double zi = z[i];
    const float *xr = X + (int64_t)i * kx;
#pragma unroll
    for (int t = 0; t < J; ++t) {
        const int j = lane + 64 * t;
        const int jj = min(j, C - 1);
        const float xv = xr[max(jj - 1, 0)];   // loaded by every lane
        double v = j == 0 ? zi : (double)xv;
#pragma unroll
        for (int d = 0; d < D; ++d) v -= Gm[lv[d] + jj];
        u[t] = j < C ? v : 0.0;
    }
}

template <int J>
__device__ __forceinline__ void glm_partial_out(int lane, int C, int64_t base, double wi,
                                                const double (&v)[J], const double *__restrict__ bm,
                                                double (&gacc)[J], __bf16 *__restrict__ hi,
                                                __bf16 *__restrict__ lo) {
    double dot = 0.0, zt = 0.0;
#pragma unroll
    for (int t = 0; t < J; ++t) {
        const int j = lane + 64 * t;
        const double b = bm[max(min(j, C - 1) - 1, 0)];
        if (j == 0) zt = v[t];
        else dot += v[t] * b;   // v is 0 past C
    }
    dot = wave_sum(dot);
    zt = __shfl(zt, 0, 64);
    const double r = zt - dot;
    const double sw = sqrt(wi);
#pragma unroll
    for (int t = 0; t < J; ++t) {
        const int j = lane + 64 * t;
        const float sv = (float)(sw * v[t]);
        const __bf16 h = (__bf16)sv;
        const __bf16 l = (__bf16)(sv - (float)h);
        if (j >= 1 && j < C) {
            gacc[t] += wi * v[t] * r;
            hi[base + j - 1] = h;
            lo[base + j - 1] = l;
        }
    }
}

template <int J, int D>
__global__ __launch_bounds__(256) void k_glm_partial(
    const int *__restrict__ wg_model, const int64_t *__restrict__ wg_gb,
    const int64_t *__restrict__ wg_ge, const int *__restrict__ gstart,
    const int64_t *__restrict__ rs, const int64_t *__restrict__ pstart, int Cp,
    const double *__restrict__ z, const float *__restrict__ X, int kx,
    const double *__restrict__ w, const int *__restrict__ codes, int64_t ld,
    const int64_t *__restrict__ offs, const double *__restrict__ Gm,
    const double *__restrict__ beta, __bf16 *__restrict__ hi, __bf16 *__restrict__ lo,
    double *__restrict__ gpart) {
    extern __shared__ double glds[];  // [4][kx]
    const int C = kx + 1;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int m = wg_model[blockIdx.x];
    const double *bm = beta + (int64_t)m * kx;
    const int64_t prow = pstart[m] - rs[m];   // padded row of panel row i: prow + i
    double gacc[J];
#pragma unroll
    for (int t = 0; t < J; ++t) gacc[t] = 0.0;
    for (int64_t gg = wg_gb[blockIdx.x] + wv; gg < wg_ge[blockIdx.x]; gg += 4) {
        const int s = gstart[gg], e = gstart[gg + 1];
        double mean[J], W = 0.0;
#pragma unroll
        for (int t = 0; t < J; ++t) mean[t] = 0.0;
        if (e - s <= kGlmStage) {
            // rows past the group's end repeat its last row, weight 0
            double u[kGlmStage][J], wr[kGlmStage];
#pragma unroll
            for (int r = 0; r < kGlmStage; ++r) {
                const int i = s + min(r, e - s - 1);
                const double wi = w[i];
                wr[r] = r < e - s ? wi : 0.0;
                glm_partial_row<J, D>(lane, C, i, z, X, kx, codes, ld, offs, Gm, u[r]);
            }
#pragma unroll
            for (int r = 0; r < kGlmStage; ++r) {
                W += wr[r];
#pragma unroll
                for (int t = 0; t < J; ++t) mean[t] += wr[r] * u[r][t];
            }
            const double inv = W > 0.0 ? 1.0 / W : 0.0;
#pragma unroll
            for (int r = 0; r < kGlmStage; ++r) {
                if (r >= e - s) break;
#pragma unroll
                for (int t = 0; t < J; ++t) u[r][t] -= mean[t] * inv;
                glm_partial_out<J>(lane, C, (prow + s + r) * (int64_t)Cp, wr[r], u[r], bm, gacc,
                                   hi, lo);
            }
        } else {
            for (int i = s; i < e; ++i) {
                const double wi = w[i];
                double u[J];
                glm_partial_row<J, D>(lane, C, i, z, X, kx, codes, ld, offs, Gm, u);
                W += wi;
#pragma unroll
                for (int t = 0; t < J; ++t) mean[t] += wi * u[t];
            }
            const double inv = W > 0.0 ? 1.0 / W : 0.0;
            for (int i = s; i < e; ++i) {
                double u[J];
                glm_partial_row<J, D>(lane, C, i, z, X, kx, codes, ld, offs, Gm, u);
#pragma unroll
                for (int t = 0; t < J; ++t) u[t] -= mean[t] * inv;
                glm_partial_out<J>(lane, C, (prow + i) * (int64_t)Cp, w[i], u, bm, gacc, hi, lo);
            }
        }
    }
    // the workgroup's score partial: waves in order
#pragma unroll
    for (int t = 0; t < J; ++t) {
        const int j = lane + 64 * t;
        if (j >= 1 && j < C) glds[wv * kx + j - 1] = gacc[t];
    }
    __syncthreads();
    for (int j = threadIdx.x; j < kx; j += blockDim.x)
        gpart[(int64_t)blockIdx.x * kx + j] =
            glds[j] + glds[kx + j] + glds[2 * kx + j] + glds[3 * kx + j];
}

// ------------------------------------------------ split-BF16 MFMA SYRK

// Upper 32x32 blocks (bi <= bj) of s's over one model's padded rows
// [kb, ke) (multiples of 16) for the blocks q0 + wave + 8 t of this launch:
// the workgroup stages 16 rows of every column (hi and lo, row-major in
// memory) in LDS as rows of eight (a column's eight rows in 16 bytes: one
// MFMA operand per lane), each wave
// runs three MFMAs per block and step.  out[wg, q, lane * 16 + r]: the MFMA
// accumulator as it lies in the lanes.
__global__ __launch_bounds__(kSyrkThreads) void k_glm_syrk(
    const int64_t *__restrict__ wg_kb, const int64_t *__restrict__ wg_ke, int Cp, int nb,
    int q0, int nq, const __bf16 *__restrict__ hi, const __bf16 *__restrict__ lo,
    float *__restrict__ out) {
    extern __shared__ __bf16 slds[];  // [2: hi, lo][2 row groups][Cp][8]
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t slice = (int64_t)2 * Cp * 8;  // bf16 per 16 rows of one array
    __bf16 *sh = slds, *sl = slds + slice;
    int bi[kSyrkMaxBlk], bj[kSyrkMaxBlk];
    int nmine = 0;
#pragma unroll
    for (int t = 0; t < kSyrkMaxBlk; ++t) {
        const int q = q0 + wv + 8 * t;
        bi[t] = bj[t] = 0;
        if (wv + 8 * t < nq) {
            // q -> (bi, bj), bi <= bj, row-major over the upper triangle
            int r = 0, qq = q;
            while (qq >= nb - r) {
                qq -= nb - r;
                ++r;
            }
            bi[t] = r;
            bj[t] = r + qq;
            nmine = t + 1;
        }
    }
    f32x16 acc[kSyrkMaxBlk];
#pragma unroll
    for (int t = 0; t < kSyrkMaxBlk; ++t)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[t][r] = 0.f;
    const int64_t kb = wg_kb[blockIdx.x], ke = wg_ke[blockIdx.x];
    const int cpair = Cp / 2;  // column pairs per eight rows
    for (int64_t k0 = kb; k0 < ke; k0 += 16) {
        // rows k0..k0+15 (row-major, 8 columns per vector) into the LDS
        // rows of eight: element (r, c) at ((r / 8) * Cp + c) * 8 + r % 8
        // a thread per two columns of eight rows: eight 4-byte loads per
        // array (a row's pairs are contiguous across the threads), two
        // columns' 16 bytes out
        __syncthreads();
        for (int tile = threadIdx.x; tile < Cp; tile += blockDim.x) {
            const int rg = tile / cpair, c0 = (tile - rg * cpair) * 2;
            bf16x8 h0, h1, l0, l1;
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                const int64_t off = (k0 + rg * 8 + r) * (int64_t)Cp + c0;
                const bf16x2 a = *(const bf16x2 *)(hi + off), b = *(const bf16x2 *)(lo + off);
                h0[r] = a[0];
                h1[r] = a[1];
                l0[r] = b[0];
                l1[r] = b[1];
            }
            *(bf16x8 *)(sh + (rg * Cp + c0) * 8) = h0;
            *(bf16x8 *)(sh + (rg * Cp + c0 + 1) * 8) = h1;
            *(bf16x8 *)(sl + (rg * Cp + c0) * 8) = l0;
            *(bf16x8 *)(sl + (rg * Cp + c0 + 1) * 8) = l1;
        }
        __syncthreads();
        const int rg = lane >> 5, col = lane & 31;
#pragma unroll
        for (int t = 0; t < kSyrkMaxBlk; ++t) {
            if (t >= nmine) continue;
            const int64_t oa = ((int64_t)rg * Cp + bi[t] * 32 + col) * 8;
            const int64_t ob = ((int64_t)rg * Cp + bj[t] * 32 + col) * 8;
            const bf16x8 ah = *(const bf16x8 *)(sh + oa);
            const bf16x8 al = *(const bf16x8 *)(sl + oa);
            const bf16x8 bh = *(const bf16x8 *)(sh + ob);
            const bf16x8 bl = *(const bf16x8 *)(sl + ob);
            acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc[t], 0, 0, 0);
            acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc[t], 0, 0, 0);
            acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < kSyrkMaxBlk; ++t) {
        if (t >= nmine) continue;
        float *o = out + ((int64_t)blockIdx.x * nq + wv + 8 * t) * 1024 + lane * 16;
#pragma unroll
        for (int r = 0; r < 16; ++r) o[r] = acc[t][r];
    }
}

// A (nb 32, nb 32) of model m: the sum over its workgroups (wb..we) of the
// blocks' accumulators, in workgroup order, in FP64; element (row, col) of
// block (bi, bj) from lane col + 32 (row / 4 % 2), register 4 (row / 8) +
// row % 4, and the transposed block below the diagonal.
__global__ void k_glm_syrk_reduce(int wb, int we, int nq_launch, int q0, int nq, int nb,
                                  const float *__restrict__ part, double *__restrict__ A) {
    const int Cp = nb * 32;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < (int64_t)nq * 1024;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int ql = (int)(e / 1024);
        const int idx = (int)(e % 1024);
        const int lane = idx / 16, r = idx % 16;
        double s = 0.0;
        for (int wg = wb; wg < we; ++wg) s += (double)part[((int64_t)wg * nq_launch + ql) * 1024 + idx];
        int q = q0 + ql, bi = 0;
        while (q >= nb - bi) {
            q -= nb - bi;
            ++bi;
        }
        const int bj = bi + q;
        const int row = 8 * (r / 4) + 4 * (lane / 32) + (r % 4);
        const int col = lane % 32;
        // a diagonal block holds (row, col) and (col, row) from different
        // accumulation orders: its upper half alone is written, both ways
        if (bi == bj && row > col) continue;
        const int R = bi * 32 + row, Cc = bj * 32 + col;
        A[(int64_t)R * Cp + Cc] = s;
        A[(int64_t)Cc * Cp + R] = s;
    }
}

// ------------------------------------------------------- coefficients

// fe[l] = Gm[l, 0] - Gm[l, 1:] beta_{model(l)} (core/irls.py:135).
__global__ void k_glm_fe(int64_t L, int kx, const double *__restrict__ Gm,
                         const int *__restrict__ lmodel, const double *__restrict__ beta,
                         double *__restrict__ fe) {
    const int C = kx + 1;
    for (int64_t l = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; l < L;
         l += (int64_t)gridDim.x * blockDim.x) {
        const double *bm = beta + (int64_t)lmodel[l] * kx;
        double v = Gm[l * C];
        for (int j = 0; j < kx; ++j) v -= Gm[l * C + 1 + j] * bm[j];
        fe[l] = v;
    }
}

// a[g] = the weighted group mean of z - x'beta_m - sum_d fe (the group's own
// effect, core/kernels_glm.py:11-33): a wave per group, lanes over the kx
// slopes (J a lane), its rows kGlmBatch at a time as in the row kernels above.
template <int J, int D>
__global__ __launch_bounds__(256) void k_glm_group_effects(
    int64_t G, const int *__restrict__ gstart, const int *__restrict__ gmodel,
    const double *__restrict__ z, const float *__restrict__ X, int kx,
    const double *__restrict__ w, const int *__restrict__ codes, int64_t ld,
    const int64_t *__restrict__ offs, const double *__restrict__ fe,
    const double *__restrict__ beta, double *__restrict__ a) {
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t waves = (int64_t)gridDim.x * 4;
    for (int64_t gg = blockIdx.x * 4 + wv; gg < G; gg += waves) {
        const double *bm = beta + (int64_t)gmodel[gg] * kx;
        double bl[J > 0 ? J : 1];
#pragma unroll
        for (int t = 0; t < J; ++t) {
            const int j = lane + 64 * t;
            bl[t] = j < kx ? bm[min(j, kx - 1)] : 0.0;
        }
        const int s = gstart[gg], e = gstart[gg + 1];
        double W = 0.0, S = 0.0;
        for (int i0 = s; i0 < e; i0 += kGlmBatch) {
            double sr[kGlmBatch], ur[kGlmBatch], wr[kGlmBatch];
#pragma unroll
            for (int u = 0; u < kGlmBatch; ++u) {
                const int i = min(i0 + u, e - 1);
                const float *xr = X + (int64_t)i * kx;
                double x = 0.0;
#pragma unroll
                for (int t = 0; t < J; ++t) x += (double)xr[min(lane + 64 * t, kx - 1)] * bl[t];
                sr[u] = x;
                double v = z[i];
#pragma unroll
                for (int d = 0; d < D; ++d) v -= fe[offs[d] + codes[d * ld + i]];
                ur[u] = v;
                const double wi = w[i];
                wr[u] = i0 + u < e ? wi : 0.0;
            }
#pragma unroll
            for (int u = 0; u < kGlmBatch; ++u) {
                const double xb = wave_sum(sr[u]);
                if (i0 + u < e) {
                    W += wr[u];
                    S += wr[u] * (ur[u] - xb);
                }
            }
        }
        if (lane == 0) a[gg] = W > 0.0 ? S / W : 0.0;
    }
}

// x = base + t_model (target - base), per entry's model (step-halving,
// glm.py:810-823): fe by level, a by group, beta by model (kx each).
__global__ void k_glm_toward(int64_t n, const int *__restrict__ model, int per,
                             const double *__restrict__ t, const double *__restrict__ base,
                             const double *__restrict__ target, double *__restrict__ out) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e / per;
        const double tm = t[model ? model[i] : i];
        out[e] = base[e] + tm * (target[e] - base[e]);
    }
}

}  // namespace hdfe

using namespace hdfe;

extern "C" {

int hdfe_rows_csr(void *h, const int **rptr, const int **rrow, const int **cptr, const int **clvl,
                  const int **cbeg, int64_t *nchunk);
int hdfe_rows_chunk_sum(void *h, int k, const double *part, double *out);

// One IRLS row pass over nchunk chunks (model, rows [begin, end)), host
// chunk lists; device in/out.  part (host, nchunk x 5): deviance, sum f (y
// eta - mu), f y, f, and with init the sum of f lgamma(y + 1).
int hdfe_glm_rows(int nchunk, const int *ch_model, const int64_t *ch_begin,
                  const int64_t *ch_end, int init, int family, const double *ybar, const double *y,
                  const float *X, int kx, const int *g, const int *codes, int64_t ld, int D,
                  const int64_t *offs, const double *fe, const double *a, const double *beta,
                  const double *f, double *w, double *z, double *part_h) {
    ARG_CHECK(nchunk >= 1 && y && w && z && part_h && (init || (X || kx == 0)) &&
                  family >= 0 && family <= 2,
              "hdfe_glm_rows: bad argument");
    DevBuf<int> cm;
    DevBuf<int64_t> cb, ce;
    DevBuf<double> part;
    DEV_ALLOC(cm, nchunk, nullptr);
    DEV_ALLOC(cb, nchunk, nullptr);
    DEV_ALLOC(ce, nchunk, nullptr);
    DEV_ALLOC(part, (int64_t)nchunk * 5, nullptr);
    HIP_TRY(hipMemcpy(cm.p, ch_model, nchunk * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(cb.p, ch_begin, nchunk * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ce.p, ch_end, nchunk * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(part.p, 0, (int64_t)nchunk * 5 * 8));
    hipLaunchKernelGGL(k_glm_rows, dim3(nchunk), dim3(256), 0, 0, nchunk, cm.p, cb.p, ce.p, init,
                       family, ybar, y, X, kx, g, codes, ld, D, offs, fe, a, beta, f, w, z,
                       part.p);
    if (init && family == 0)
        hipLaunchKernelGGL(k_glm_lgamma, dim3(nchunk), dim3(256), 0, 0, cb.p, ce.p, y, f, part.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(part_h, part.p, (int64_t)nchunk * 5 * 8, hipMemcpyDeviceToHost));
    return HDFE_STATUS_OK;
}

// rhs (L, kx + 1) = D_o' W M_0 [z, X] (device), with vbar (G, kx + 1) scratch.
int hdfe_glm_rhs(void *rows_h, int64_t G, const int *gstart, const int *g, const double *z,
                 const float *X, int kx, const double *w, double *vbar, double *rhs) {
    ARG_CHECK(rows_h && gstart && g && z && w && vbar && rhs && kx >= 0 && kx + 1 <= 512,
              "hdfe_glm_rhs: bad argument");
    const int C = kx + 1;
    const int *rptr, *rrow, *cptr, *clvl, *cbeg;
    int64_t nchunk;
    int st = hdfe_rows_csr(rows_h, &rptr, &rrow, &cptr, &clvl, &cbeg, &nchunk);
    if (st) return st;
    DevBuf<double> part;
    DEV_ALLOC(part, nchunk * (int64_t)C, nullptr);
    // without covariates the kernels' clamped X loads read z's first bytes
    // (their values unused)
    const float *Xs = kx > 0 ? X : (const float *)z;
#define HDFE_RHS(JJ)                                                                              \
    hipLaunchKernelGGL((k_glm_group_means<JJ>), dim3(grid1(G * 64)), dim3(256), 0, 0, G, gstart, \
                       z, Xs, kx, w, vbar);                                                       \
    hipLaunchKernelGGL((k_glm_center<JJ>), dim3(grid1(nchunk * 64)), dim3(256), 0, 0, nchunk,    \
                       clvl, cbeg, rptr, rrow, g, z, Xs, kx, w, vbar, part.p)
    switch ((C + 63) / 64) {
        case 1: HDFE_RHS(1); break;
        case 2: HDFE_RHS(2); break;
        case 3: HDFE_RHS(3); break;
        case 4: HDFE_RHS(4); break;
        case 5: HDFE_RHS(5); break;
        case 6: HDFE_RHS(6); break;
        case 7: HDFE_RHS(7); break;
        default: HDFE_RHS(8); break;
    }
#undef HDFE_RHS
    HIP_TRY(hipGetLastError());
    st = hdfe_rows_chunk_sum(rows_h, C, part.p, rhs);
    if (st) return st;
    HIP_TRY(hipDeviceSynchronize());
    return HDFE_STATUS_OK;
}

// The partialled rows (split bf16, hi and lo (npad, Cp) row-major, device)
// and the score partials of nwg workgroups (host lists: model, group range);
// gpart (host, nwg x kx).  D <= 4 fixed-effect dimensions besides the groups.
int hdfe_glm_partial(int nwg, const int *wg_model, const int64_t *wg_gb, const int64_t *wg_ge,
                     const int *gstart, const int64_t *rs_h, const int64_t *pstart_h, int nmodels,
                     int Cp, const double *z, const float *X, int kx, const double *w,
                     const int *codes, int64_t ld, int D, const int64_t *offs, const double *Gm,
                     const double *beta, void *hi, void *lo, double *gpart_h) {
    ARG_CHECK(nwg >= 1 && kx >= 1 && kx + 1 <= 512 && Cp >= kx && Cp % 32 == 0 && D >= 0 &&
                  D <= 4,
              "hdfe_glm_partial: bad argument");
    DevBuf<int> wm;
    DevBuf<int64_t> gb, ge, rs, ps;
    DevBuf<double> gpart;
    DEV_ALLOC(wm, nwg, nullptr);
    DEV_ALLOC(gb, nwg, nullptr);
    DEV_ALLOC(ge, nwg, nullptr);
    DEV_ALLOC(rs, nmodels + 1, nullptr);
    DEV_ALLOC(ps, nmodels + 1, nullptr);
    DEV_ALLOC(gpart, (int64_t)nwg * kx, nullptr);
    HIP_TRY(hipMemcpy(wm.p, wg_model, nwg * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(gb.p, wg_gb, nwg * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ge.p, wg_ge, nwg * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(rs.p, rs_h, (nmodels + 1) * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ps.p, pstart_h, (nmodels + 1) * 8, hipMemcpyHostToDevice));
#define HDFE_GP(JJ, DD)                                                                          \
    hipLaunchKernelGGL((k_glm_partial<JJ, DD>), dim3(nwg), dim3(256), 4 * kx * 8, 0, wm.p, gb.p,   \
                       ge.p, gstart, rs.p, ps.p, Cp, z, X, kx, w, codes, ld, offs, Gm, beta,      \
                       (__bf16 *)hi, (__bf16 *)lo, gpart.p)
#define HDFE_GPD(JJ)                      \
    switch (D) {                          \
        case 0: HDFE_GP(JJ, 0); break;    \
        case 1: HDFE_GP(JJ, 1); break;    \
        case 2: HDFE_GP(JJ, 2); break;    \
        case 3: HDFE_GP(JJ, 3); break;    \
        default: HDFE_GP(JJ, 4); break;   \
    }
    switch ((kx + 1 + 63) / 64) {
        case 1: HDFE_GPD(1); break;
        case 2: HDFE_GPD(2); break;
        case 3: HDFE_GPD(3); break;
        case 4: HDFE_GPD(4); break;
        case 5: HDFE_GPD(5); break;
        case 6: HDFE_GPD(6); break;
        case 7: HDFE_GPD(7); break;
        default: HDFE_GPD(8); break;
    }
#undef HDFE_GPD
#undef HDFE_GP
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(gpart_h, gpart.p, (int64_t)nwg * kx * 8, hipMemcpyDeviceToHost));
    return HDFE_STATUS_OK;
}

// A (nmodels, Cp, Cp, device FP64) = s's per model from the split rows:
// model m's padded rows [pstart[m], pstart[m + 1]) (multiples of 16) in
// workgroups of at most `rows_per_wg` rows.
int hdfe_glm_syrk(int nmodels, const int64_t *pstart_h, int Cp, const void *hi, const void *lo,
                  int64_t rows_per_wg, double *A) {
    ARG_CHECK(nmodels >= 1 && Cp % 32 == 0 && Cp >= 32 && rows_per_wg >= 16 &&
                  rows_per_wg % 16 == 0 && A,
              "hdfe_glm_syrk: bad argument");
    const int nb = Cp / 32;
    const int nblk = nb * (nb + 1) / 2;
    std::vector<int64_t> kb, ke;
    std::vector<int> wb(nmodels + 1, 0);
    for (int m = 0; m < nmodels; ++m) {
        ARG_CHECK(pstart_h[m] % 16 == 0 && pstart_h[m + 1] % 16 == 0,
                  "hdfe_glm_syrk: model rows must start at multiples of 16");
        for (int64_t k = pstart_h[m]; k < pstart_h[m + 1]; k += rows_per_wg) {
            kb.push_back(k);
            ke.push_back(std::min(pstart_h[m + 1], k + rows_per_wg));
        }
        wb[m + 1] = (int)kb.size();
    }
    const int nwg = (int)kb.size();
    HIP_TRY(hipMemset(A, 0, (int64_t)nmodels * Cp * Cp * 8));
    if (nwg == 0) return HDFE_STATUS_OK;
    DevBuf<int64_t> dkb, dke;
    DEV_ALLOC(dkb, nwg, nullptr);
    DEV_ALLOC(dke, nwg, nullptr);
    HIP_TRY(hipMemcpy(dkb.p, kb.data(), nwg * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dke.p, ke.data(), nwg * 8, hipMemcpyHostToDevice));
    const int per_launch = 8 * kSyrkMaxBlk;
    DevBuf<float> part;
    DEV_ALLOC(part, (int64_t)nwg * per_launch * 1024, nullptr);
    const size_t lds = (size_t)2 * 2 * Cp * 8 * sizeof(__bf16);
    for (int q0 = 0; q0 < nblk; q0 += per_launch) {
        const int nq = std::min(per_launch, nblk - q0);
        hipLaunchKernelGGL(k_glm_syrk, dim3(nwg), dim3(kSyrkThreads), lds, 0, dkb.p, dke.p, Cp, nb,
                           q0, nq, (const __bf16 *)hi, (const __bf16 *)lo, part.p);
        HIP_TRY(hipGetLastError());
        for (int m = 0; m < nmodels; ++m) {
            if (wb[m + 1] == wb[m]) continue;
            hipLaunchKernelGGL(k_glm_syrk_reduce, dim3(grid1((int64_t)nq * 1024)), dim3(256), 0, 0,
                               wb[m], wb[m + 1], nq, q0, nq, nb, part.p,
                               A + (int64_t)m * Cp * Cp);
        }
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipDeviceSynchronize());
    return HDFE_STATUS_OK;
}

// fe (L) and the group effects a (G) of new slopes beta (nmodels, kx, device).
int hdfe_glm_coefficients(int64_t L, int64_t G, int kx, const double *Gm, const int *lmodel,
                          const int *gmodel, const int *gstart, const double *z, const float *X,
                          const double *w, const int *codes, int64_t ld, int D,
                          const int64_t *offs, const double *beta, double *fe, double *a) {
    ARG_CHECK(kx >= 0 && kx <= 512 && D >= 0 && D <= 4, "hdfe_glm_coefficients: bad argument");
    hipLaunchKernelGGL(k_glm_fe, dim3(grid1(L)), dim3(256), 0, 0, L, kx, Gm, lmodel, beta, fe);
#define HDFE_GE(JJ, DD)                                                                       \
    hipLaunchKernelGGL((k_glm_group_effects<JJ, DD>), dim3(grid1(G * 64)), dim3(256), 0, 0, G, \
                       gstart, gmodel, z, X, kx, w, codes, ld, offs, fe, beta, a)
#define HDFE_GED(JJ)                    \
    switch (D) {                        \
        case 0: HDFE_GE(JJ, 0); break;  \
        case 1: HDFE_GE(JJ, 1); break;  \
        case 2: HDFE_GE(JJ, 2); break;  \
        case 3: HDFE_GE(JJ, 3); break;  \
        default: HDFE_GE(JJ, 4); break; \
    }
    switch ((kx + 63) / 64) {
        case 0: HDFE_GED(0); break;   // no slopes: X is not read
        case 1: HDFE_GED(1); break;
        case 2: HDFE_GED(2); break;
        case 3: HDFE_GED(3); break;
        case 4: HDFE_GED(4); break;
        case 5: HDFE_GED(5); break;
        case 6: HDFE_GED(6); break;
        case 7: HDFE_GED(7); break;
        default: HDFE_GED(8); break;
    }
#undef HDFE_GED
#undef HDFE_GE
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    return HDFE_STATUS_OK;
}

// out = base + t[model] (target - base) over n entries of `per` values each
// (model NULL: entry i's own t[i]).
int hdfe_glm_toward(int64_t n, const int *model, int per, const double *t, const double *base,
                    const double *target, double *out) {
    hipLaunchKernelGGL(k_glm_toward, dim3(grid1(n * per)), dim3(256), 0, 0, n * per, model, per,
                       t, base, target, out);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    return HDFE_STATUS_OK;
}

}  // extern "C"
