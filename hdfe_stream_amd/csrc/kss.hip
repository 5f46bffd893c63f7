// The KSS bias term on the device: the coefficient-space inverse
// (apply_inverse) and Hutchinson's trace moments, for gfx950.
//
// Reference (hdfe_stream/):
//   apply_inverse                 inverse.py:311-374
//   _null_space / _drop_null_...  inverse.py:376-460
//   _nb_reduce_rows               kernels_inverse.py:111-130
//   _nb_spread_groups             kernels_inverse.py:133-152
//   _nb_project_coef              kernels_inverse.py:156-188
//   _nb_group_sums                core/kernels_groups.py:16-27
//   _nb_trace_moments             kernels_inverse.py:191-231
//   hutchinson_trace / _trace_block  leaveout.py:321-448
//
// Blocks of k vectors are row-major (rows, k) arrays: per row, k contiguous
// values.  Each pass is a wavefront per group (rows of a group are
// contiguous) or per chunk of a level's rows (the level-to-rows CSR of
// rows.hip), lanes over the k vectors, partial sums added in a fixed order.
#include <algorithm>
#include <cmath>
#include <vector>

#include "devbuf.h"

namespace hdfe {

constexpr int kRowChunk2 = 1024;

// Group weights W_g = sum of w over the group's rows.
__global__ void k_group_weight(int64_t G, const int *__restrict__ gstart,
                               const double *__restrict__ w, double *__restrict__ Wg) {
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < G;
         g += (int64_t)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int i = gstart[g]; i < gstart[g + 1]; ++i) s += w ? w[i] : 1.0;
        Wg[g] = s;
    }
}

// part[c, j] = sum over chunk c's rows i of w_i * x_i * v_ij, where v is
// either a row block R (n, k) or a group block V (G, k) divided by W_g (the
// spread of a streamed-block vector, kernels_inverse.py:133-152), and x an
// optional per-row factor (sigma2).  kernels_inverse.py:111-130.
__global__ void k_reduce_levels(int64_t nchunk, int k, const int *__restrict__ chunk_lvl,
                                const int *__restrict__ chunk_beg, const int *__restrict__ rptr,
                                const int *__restrict__ rrow, const int *__restrict__ g,
                                const double *__restrict__ w, const double *__restrict__ x,
                                const double *__restrict__ R, const double *__restrict__ V,
                                const double *__restrict__ Wg, double *__restrict__ part) {
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x / 64);
    for (int64_t c = blockIdx.x * (int64_t)(blockDim.x / 64) + threadIdx.x / 64; c < nchunk;
         c += waves) {
        const int a = chunk_lvl[c];
        const int b = chunk_beg[c], e = min(b + kRowChunk2, rptr[a + 1]);
        for (int t = 0; 64 * t < k; ++t) {
            const int j = lane + 64 * t;
            if (j >= k) break;
            double acc = 0.0;
            for (int p = b; p < e; ++p) {
                const int i = rrow[p];
                double f = (w ? w[i] : 1.0) * (x ? x[i] : 1.0);
                double v;
                if (R) {
                    v = R[(int64_t)i * k + j];
                } else {
                    const double W = Wg[g[i]];
                    v = W > 0.0 ? V[(int64_t)g[i] * k + j] / W : 0.0;
                }
                acc += f * v;
            }
            part[c * k + j] = acc;
        }
    }
}

// Row values and streamed block of a coefficient-space solution
// (kernels_inverse.py:156-188): h_i = sum_d U[level_d(i)], u0_g = weighted
// mean over the group of (rho_i - h_i) with rho_i = V[g] / W_g, rows = h + u0.
__global__ void k_project_coef(int64_t G, int D, int k, const int *__restrict__ gstart,
                               const int *__restrict__ codes, int64_t ld,
                               const int64_t *__restrict__ offs, const double *__restrict__ w,
                               const double *__restrict__ Wg, const double *__restrict__ U,
                               const double *__restrict__ V, double *__restrict__ rows,
                               double *__restrict__ Ug) {
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x / 64);
    for (int64_t g = blockIdx.x * (int64_t)(blockDim.x / 64) + threadIdx.x / 64; g < G;
         g += waves) {
        const int s = gstart[g], e = gstart[g + 1];
        const double W = Wg[g];
        for (int t = 0; 64 * t < k; ++t) {
            const int j = lane + 64 * t;
            if (j >= k) break;
            const double rho = W > 0.0 ? V[g * k + j] / W : 0.0;
            double tot = 0.0;
            for (int i = s; i < e; ++i) {
                double h = 0.0;
                for (int d = 0; d < D; ++d) h += U[(offs[d] + codes[d * ld + i]) * k + j];
                if (rows) rows[(int64_t)i * k + j] = h;
                tot += (w ? w[i] : 1.0) * (rho - h);
            }
            const double mean = W > 0.0 ? tot / W : 0.0;
            Ug[g * k + j] = mean;
            if (rows)
                for (int i = s; i < e; ++i) rows[(int64_t)i * k + j] += mean;
        }
    }
}

// out[g, j] = sum over the group's rows of w_i x_i R_ij (core/kernels_groups.py:16-27).
__global__ void k_group_sums(int64_t G, int k, const int *__restrict__ gstart,
                             const double *__restrict__ w, const double *__restrict__ x,
                             const double *__restrict__ R, double *__restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x / 64);
    for (int64_t g = blockIdx.x * (int64_t)(blockDim.x / 64) + threadIdx.x / 64; g < G;
         g += waves)
        for (int t = 0; 64 * t < k; ++t) {
            const int j = lane + 64 * t;
            if (j >= k) break;
            double s = 0.0;
            for (int i = gstart[g]; i < gstart[g + 1]; ++i)
                s += (w ? w[i] : 1.0) * (x ? x[i] : 1.0) * R[(int64_t)i * k + j];
            out[g * k + j] = s;
        }
}

// Hutchinson's cross-moments per draw j (kernels_inverse.py:191-231), per
// workgroup partials part[b, j, 0..7]:
//   0 sum w, 1 sum w a1 a2, 2 sum w b1 b2, 3 sum w (a1 b2 + b1 a2),
//   4 sum w a1, 5 sum w a2, 6 sum w b1, 7 sum w b2,
// a = the psi dimension's value (Z / U at the row's psi level), b = the
// streamed value (Z / U at the row's group).  Thread per draw, rows of the
// workgroup's groups in order.
__global__ void k_trace_moments(int64_t G, int k, const int *__restrict__ gstart,
                                const int *__restrict__ psi_codes, int64_t psi_off,
                                const double *__restrict__ w, const double *__restrict__ Zl,
                                const double *__restrict__ Zg, const double *__restrict__ Ul,
                                const double *__restrict__ Ug, double *__restrict__ part) {
    const int64_t g0 = part_begin(G, blockIdx.x, gridDim.x), g1 = part_begin(G, blockIdx.x + 1, gridDim.x);
    for (int j = threadIdx.x; j < k; j += blockDim.x) {
        double m[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int64_t g = g0; g < g1; ++g) {
            const double b1 = Zg[g * k + j], b2 = Ug[g * k + j];
            for (int i = gstart[g]; i < gstart[g + 1]; ++i) {
                const int64_t lv = psi_off + psi_codes[i];
                const double a1 = Zl[lv * k + j], a2 = Ul[lv * k + j];
                const double wi = w ? w[i] : 1.0;
                m[0] += wi;
                m[1] += wi * a1 * a2;
                m[2] += wi * b1 * b2;
                m[3] += wi * (a1 * b2 + b1 * a2);
                m[4] += wi * a1;
                m[5] += wi * a2;
                m[6] += wi * b1;
                m[7] += wi * b2;
            }
        }
        for (int q = 0; q < 8; ++q) part[((int64_t)blockIdx.x * k + j) * 8 + q] = m[q];
    }
}

// Null-space overlaps (inverse.py:376-460): per column j, the sums of the
// streamed block over each component's groups, of the first dimension's
// levels over each component, and of each other dimension's levels:
// out[(c or D-1 extra) * k + j].  Thread per (block, column), fixed order.
__global__ void k_null_sums(int64_t G, int64_t n1, int D, int64_t ncomp, int k,
                            const int *__restrict__ gcomp, const int *__restrict__ lcomp,
                            const int64_t *__restrict__ offs_h, const int64_t *__restrict__ nlev,
                            const double *__restrict__ Vg, const double *__restrict__ Vl,
                            double *__restrict__ sg, double *__restrict__ s1,
                            double *__restrict__ so) {
    // sg[c, j], s1[c, j]: a thread per (component, j) would walk all groups;
    // instead a thread per j accumulates every component (components are few
    // after pruning), and the other dimensions' sums likewise.
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < k; j += gridDim.x * blockDim.x) {
        for (int64_t g = 0; g < G; ++g) sg[gcomp[g] * (int64_t)k + j] += Vg[g * k + j];
        for (int64_t a = 0; a < n1; ++a) s1[lcomp[a] * (int64_t)k + j] += Vl[a * k + j];
        for (int d = 1; d < D; ++d) {
            double s = 0.0;
            for (int64_t a = offs_h[d]; a < offs_h[d] + nlev[d]; ++a) s += Vl[a * k + j];
            so[(d - 1) * (int64_t)k + j] = s;
        }
    }
}

// Column sums of rows [lo, hi) of a (rows, k) block in fixed row blocks:
// part[b * k + j].
__global__ void k_rows_colsum(int64_t lo, int64_t hi, int k, const double *__restrict__ X,
                              double *__restrict__ part) {
    const int64_t n = hi - lo;
    const int64_t r0 = lo + part_begin(n, blockIdx.x, gridDim.x);
    const int64_t r1 = lo + part_begin(n, blockIdx.x + 1, gridDim.x);
    for (int j = threadIdx.x; j < k; j += blockDim.x) {
        double s = 0.0;
        for (int64_t r = r0; r < r1; ++r) s += X[r * k + j];
        part[(int64_t)blockIdx.x * k + j] = s;
    }
}

// Subtract the null-space part given the coefficients: component c's
// groups get -cg[c], its first-dimension levels -(c1[c] + sum_t co[t]), and
// dimension t's levels +co[t] (the basis vectors are (+1 groups, -1 first
// dim) per component and (+1 first dim, -1 dim t) per other dimension).
__global__ void k_null_apply(int64_t G, int64_t L, int64_t n1, int D, int k,
                             const int *__restrict__ gcomp, const int *__restrict__ lcomp,
                             const int *__restrict__ ldim, const double *__restrict__ cc,
                             const double *__restrict__ co, double *__restrict__ Vg,
                             double *__restrict__ Vl) {
    const int64_t ng = G * k, nl = L * k;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < ng + nl;
         e += (int64_t)gridDim.x * blockDim.x) {
        if (e < ng) {
            const int64_t g = e / k;
            const int j = (int)(e - g * k);
            Vg[e] -= cc[gcomp[g] * (int64_t)k + j];
        } else {
            const int64_t f = e - ng;
            const int64_t a = f / k;
            const int j = (int)(f - a * k);
            const int d = ldim[a];
            if (d == 0) {
                double s = -cc[lcomp[a] * (int64_t)k + j];
                for (int t = 1; t < D; ++t) s += co[(t - 1) * (int64_t)k + j];
                Vl[f] -= s;
            } else {
                Vl[f] += co[(d - 1) * (int64_t)k + j];
            }
        }
    }
}

// Component of each group: that of the first dimension's level of its first row.
__global__ void k_group_comp(int64_t G, const int *__restrict__ gstart,
                             const int *__restrict__ codes0, const int *__restrict__ lcomp,
                             int *__restrict__ gcomp) {
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < G;
         g += (int64_t)gridDim.x * blockDim.x)
        gcomp[g] = lcomp[codes0[gstart[g]]];
}

// Partial sums for the outcome mean: part[b] = (sum w, sum w y).
__global__ void k_wmean_part(int64_t n, int m, const double *__restrict__ V,
                             const double *__restrict__ w, double *__restrict__ part) {
    __shared__ double lds[4];
    const int64_t r0 = part_begin(n, blockIdx.x, gridDim.x), r1 = part_begin(n, blockIdx.x + 1, gridDim.x);
    double sw = 0.0, swy = 0.0;
    for (int64_t i = r0 + threadIdx.x; i < r1; i += blockDim.x) {
        const double wi = w ? w[i] : 1.0;
        sw += wi;
        swy += wi * V[i * m];
    }
    sw = block_sum<double>(sw, lds);
    swy = block_sum<double>(swy, lds);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sw;
        part[2 * blockIdx.x + 1] = swy;
    }
}

// sigma2_i (leaveout.py:949-1016), every row its own estimate: centering
// "weighted" (mode 0) w_i (y_i - ybar_w) e_i / M_ii * factor_i; "reference"
// (mode 1) (sqrt(w_i) y_i - the mean of sqrt(w) y) sqrt(w_i) e_i / M_ii *
// factor_i, `ybar` the respective mean.
__global__ void k_sigma2(int64_t n, int m, const double *__restrict__ V,
                         const double *__restrict__ w, double ybar, int mode,
                         const double *__restrict__ e, const double *__restrict__ Mb,
                         const double *__restrict__ F, double *__restrict__ out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double wi = w ? w[i] : 1.0;
        if (mode == 1) {
            const double r = sqrt(wi);
            out[i] = (r * V[i * m] - ybar) * r * e[i] / Mb[i] * F[i];
        } else {
            out[i] = wi * (V[i * m] - ybar) * e[i] / Mb[i] * F[i];
        }
    }
}

// Partial sums for the reference centring's mean: part[b] = (count, sum
// sqrt(w) y).
__global__ void k_rootmean_part(int64_t n, int m, const double *__restrict__ V,
                                const double *__restrict__ w, double *__restrict__ part) {
    __shared__ double lds[4];
    const int64_t r0 = part_begin(n, blockIdx.x, gridDim.x), r1 = part_begin(n, blockIdx.x + 1, gridDim.x);
    double c = 0.0, s = 0.0;
    for (int64_t i = r0 + threadIdx.x; i < r1; i += blockDim.x) {
        c += 1.0;
        s += sqrt(w ? w[i] : 1.0) * V[i * m];
    }
    c = block_sum<double>(c, lds);
    s = block_sum<double>(s, lds);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = c;
        part[2 * blockIdx.x + 1] = s;
    }
}

// The plug-in moments' sums over rows (leaveout.py:807-832): per block,
// (sum w, sum w a, sum w b, sum w a^2, sum w b^2, sum w a b) with a the row's
// psi effect and b its streamed effect.
__global__ void k_plugin_part(int64_t G, const int *__restrict__ gstart,
                              const int *__restrict__ psi_codes, int64_t psi_off,
                              const double *__restrict__ gam, const double *__restrict__ geff,
                              const double *__restrict__ w, double *__restrict__ part) {
    __shared__ double lds[4];
    const int64_t g0 = part_begin(G, blockIdx.x, gridDim.x), g1 = part_begin(G, blockIdx.x + 1, gridDim.x);
    double m[6] = {0, 0, 0, 0, 0, 0};
    for (int64_t g = g0 + threadIdx.x; g < g1; g += blockDim.x) {
        const double b = geff[g];
        for (int i = gstart[g]; i < gstart[g + 1]; ++i) {
            const double a = gam[psi_off + psi_codes[i]];
            const double wi = w ? w[i] : 1.0;
            m[0] += wi;
            m[1] += wi * a;
            m[2] += wi * b;
            m[3] += wi * a * a;
            m[4] += wi * b * b;
            m[5] += wi * a * b;
        }
    }
    for (int q = 0; q < 6; ++q) {
        const double v = block_sum<double>(m[q], lds);
        if (threadIdx.x == 0) part[blockIdx.x * 6 + q] = v;
    }
}

}  // namespace hdfe

using namespace hdfe;

extern "C" {

// The level-to-rows CSR of a row-pass state, and the per-level sum of chunk
// partials (rows.hip).
int hdfe_rows_csr(void *h, const int **rptr, const int **rrow, const int **cptr, const int **clvl,
                  const int **cbeg, int64_t *nchunk);
int hdfe_rows_chunk_sum(void *h, int k, const double *part, double *out);

// Group weights W_g (G, device out).
int hdfe_kss_group_weight(int64_t G, const int *gstart, const double *w, double *Wg) {
    hipLaunchKernelGGL(k_group_weight, dim3(grid1(G)), dim3(256), 0, 0, G, gstart, w, Wg);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    return HDFE_STATUS_OK;
}

// out (L, k) = D_o' W diag(x) R (R: rows (n, k)), or D_o' W rho with rho the
// spread of V (G, k) (V / W_g on each row of the group) when R is NULL.
// kernels_inverse.py:111-152.
int hdfe_kss_reduce(void *rows_h, int k, const int *g, const double *w, const double *x,
                    const double *R, const double *V, const double *Wg, double *out) {
    ARG_CHECK(rows_h && g && out && (R || (V && Wg)), "hdfe_kss_reduce: bad argument");
    const int *rptr, *rrow, *cptr, *clvl, *cbeg;
    int64_t nchunk;
    int st = hdfe_rows_csr(rows_h, &rptr, &rrow, &cptr, &clvl, &cbeg, &nchunk);
    if (st) return st;
    // the chunk sums, then per level in chunk order (rows.hip's k_center_sum)
    DevBuf<double> part;
    DEV_ALLOC(part, nchunk * (int64_t)k, nullptr);
    hipLaunchKernelGGL(k_reduce_levels, dim3(grid1(nchunk * 64)), dim3(256), 0, 0, nchunk, k, clvl,
                       cbeg, rptr, rrow, g, w, x, R, V, Wg, part.p);
    HIP_TRY(hipGetLastError());
    st = hdfe_rows_chunk_sum(rows_h, k, part.p, out);
    if (st) return st;
    HIP_TRY(hipDeviceSynchronize());
    return HDFE_STATUS_OK;
}

// Row values (n, k; skipped when rows is NULL) and streamed block Ug (G, k)
// of a coefficient-space solution U (L, k) with streamed right-hand side V (G, k).
int hdfe_kss_project(int64_t G, int D, int k, const int *gstart, const int *codes, int64_t n,
                     const int64_t *offs_d, const double *w, const double *Wg, const double *U,
                     const double *V, double *rows, double *Ug) {
    hipLaunchKernelGGL(k_project_coef, dim3(grid1(G * 64)), dim3(256), 0, 0, G, D, k, gstart,
                       codes, n, offs_d, w, Wg, U, V, rows, Ug);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    return HDFE_STATUS_OK;
}

// out (G, k) = per group sum of w x R.
int hdfe_kss_group_sums(int64_t G, int k, const int *gstart, const double *w, const double *x,
                        const double *R, double *out) {
    hipLaunchKernelGGL(k_group_sums, dim3(grid1(G * 64)), dim3(256), 0, 0, G, k, gstart, w, x, R,
                       out);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    return HDFE_STATUS_OK;
}

// Hutchinson's 8 moments summed over rows and draws (host out, 8).
int hdfe_kss_trace_moments(int64_t G, int k, const int *gstart, const int *psi_codes,
                           int64_t psi_off, const double *w, const double *Zl, const double *Zg,
                           const double *Ul, const double *Ug, double *out) {
    const int nb = 512;
    DevBuf<double> part;
    DEV_ALLOC(part, (int64_t)nb * k * 8, nullptr);
    hipLaunchKernelGGL(k_trace_moments, dim3(nb), dim3(256), 0, 0, G, k, gstart, psi_codes,
                       psi_off, w, Zl, Zg, Ul, Ug, part.p);
    HIP_TRY(hipGetLastError());
    std::vector<double> h;
    HIP_TRY(to_host(h, part.p, (int64_t)nb * k * 8, 0));
    for (int q = 0; q < 8; ++q) out[q] = 0.0;
    for (int b = 0; b < nb; ++b)
        for (int j = 0; j < k; ++j)
            for (int q = 0; q < 8; ++q) out[q] += h[((size_t)b * k + j) * 8 + q];
    return HDFE_STATUS_OK;
}

// Remove the null-space part of a coefficient-space block (Vg (G, k), Vl
// (L, k), device, in place): the basis of inverse.py:376-433, one vector per
// component and one per further dimension, the projection solved on the
// host from the block sums (a (ncomp + D - 1)-square Gram system).
// lcomp: components of the first dimension's levels (n1, device); codes0:
// the rows' first-dimension codes; offs / nlev host.
int hdfe_kss_drop_null(int64_t G, int64_t L, int D, int64_t ncomp, int k, const int *gstart,
                       const int *codes0, const int *lcomp, const int64_t *offs_h,
                       const int64_t *nlev_h, double *Vg, double *Vl) {
    ARG_CHECK(D >= 1 && ncomp >= 1 && k >= 1, "hdfe_kss_drop_null: bad argument");
    const int64_t n1 = nlev_h[0];
    DevBuf<int> gcomp, ldim;
    DEV_ALLOC(gcomp, G, nullptr);
    DEV_ALLOC(ldim, L, nullptr);
    hipLaunchKernelGGL(k_group_comp, dim3(grid1(G)), dim3(256), 0, 0, G, gstart, codes0, lcomp,
                       gcomp.p);
    std::vector<int> hd(L);
    for (int d = 0; d < D; ++d)
        for (int64_t a = offs_h[d]; a < offs_h[d] + nlev_h[d]; ++a) hd[a] = d;
    HIP_TRY(hipMemcpy(ldim.p, hd.data(), L * 4, hipMemcpyHostToDevice));
    DevBuf<int64_t> offs_d, nlev_d;
    DEV_ALLOC(offs_d, D, nullptr);
    DEV_ALLOC(nlev_d, D, nullptr);
    HIP_TRY(hipMemcpy(offs_d.p, offs_h, D * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(nlev_d.p, nlev_h, D * 8, hipMemcpyHostToDevice));
    const int q = (int)ncomp + D - 1;
    DevBuf<double> sg, s1, so;
    DEV_ALLOC(sg, ncomp * k, nullptr);
    DEV_ALLOC(s1, ncomp * k, nullptr);
    DEV_ALLOC(so, std::max(D - 1, 1) * (int64_t)k, nullptr);
    HIP_TRY(hipMemset(sg.p, 0, ncomp * k * 8));
    HIP_TRY(hipMemset(s1.p, 0, ncomp * k * 8));
    std::vector<double> hsg, hs1, hso;
    if (ncomp == 1) {
        // one component (a leave-one-out connected panel): block sums
        const int nb = 256;
        DevBuf<double> part;
        DEV_ALLOC(part, (int64_t)nb * k, nullptr);
        std::vector<double> hp;
        auto colsum = [&](const double *X, int64_t lo, int64_t hi, std::vector<double> &dst,
                          size_t at) -> int {
            hipLaunchKernelGGL(k_rows_colsum, dim3(nb), dim3(256), 0, 0, lo, hi, k, X, part.p);
            HIP_TRY(hipGetLastError());
            HIP_TRY(to_host(hp, part.p, (int64_t)nb * k, 0));
            for (int j = 0; j < k; ++j) {
                double t = 0.0;
                for (int b = 0; b < nb; ++b) t += hp[(size_t)b * k + j];
                dst[at + j] = t;
            }
            return HDFE_STATUS_OK;
        };
        hsg.assign(k, 0.0);
        hs1.assign(k, 0.0);
        hso.assign((size_t)std::max(D - 1, 1) * k, 0.0);
        int st2 = colsum(Vg, 0, G, hsg, 0);
        if (st2) return st2;
        st2 = colsum(Vl, 0, n1, hs1, 0);
        if (st2) return st2;
        for (int d = 1; d < D; ++d) {
            st2 = colsum(Vl, offs_h[d], offs_h[d] + nlev_h[d], hso, (size_t)(d - 1) * k);
            if (st2) return st2;
        }
    } else {
        hipLaunchKernelGGL(k_null_sums, dim3(grid1(k, 64)), dim3(64), 0, 0, G, n1, D, ncomp, k,
                           gcomp.p, lcomp, offs_d.p, nlev_d.p, Vg, Vl, sg.p, s1.p, so.p);
        HIP_TRY(hipGetLastError());
        HIP_TRY(to_host(hsg, sg.p, ncomp * k, 0));
        HIP_TRY(to_host(hs1, s1.p, ncomp * k, 0));
        HIP_TRY(to_host(hso, so.p, std::max(D - 1, 1) * (int64_t)k, 0));
    }
    // component sizes (groups, first-dimension levels)
    std::vector<int> hgc, hlc;
    HIP_TRY(to_host(hgc, gcomp.p, G, 0));
    HIP_TRY(to_host(hlc, lcomp, n1, 0));
    std::vector<double> Gc(ncomp, 0.0), Lc(ncomp, 0.0);
    for (int64_t g = 0; g < G; ++g) Gc[hgc[g]] += 1.0;
    for (int64_t a = 0; a < n1; ++a) Lc[hlc[a]] += 1.0;
    // Gram matrix N'N of the basis (n_c; m_t) and the right-hand sides N'v
    std::vector<double> A((size_t)q * q, 0.0), B((size_t)q * k, 0.0);
    for (int c = 0; c < ncomp; ++c) {
        A[(size_t)c * q + c] = Gc[c] + Lc[c];
        for (int t = 1; t < D; ++t) {
            A[(size_t)c * q + ncomp + t - 1] = -Lc[c];
            A[(size_t)(ncomp + t - 1) * q + c] = -Lc[c];
        }
        for (int j = 0; j < k; ++j) B[(size_t)c * k + j] = hsg[(size_t)c * k + j] - hs1[(size_t)c * k + j];
    }
    for (int t = 1; t < D; ++t)
        for (int u = 1; u < D; ++u)
            A[(size_t)(ncomp + t - 1) * q + ncomp + u - 1] = (double)n1 + (t == u ? (double)nlev_h[t] : 0.0);
    for (int t = 1; t < D; ++t)
        for (int j = 0; j < k; ++j) {
            double s = 0.0;
            for (int c = 0; c < ncomp; ++c) s += hs1[(size_t)c * k + j];
            B[(size_t)(ncomp + t - 1) * k + j] = s - hso[(size_t)(t - 1) * k + j];
        }
    // solve A X = B by Gaussian elimination with partial pivoting
    for (int col = 0; col < q; ++col) {
        int piv = col;
        for (int r = col + 1; r < q; ++r)
            if (std::fabs(A[(size_t)r * q + col]) > std::fabs(A[(size_t)piv * q + col])) piv = r;
        ARG_CHECK(std::fabs(A[(size_t)piv * q + col]) > 0.0, "hdfe_kss_drop_null: singular basis");
        if (piv != col) {
            for (int c2 = 0; c2 < q; ++c2) std::swap(A[(size_t)piv * q + c2], A[(size_t)col * q + c2]);
            for (int j = 0; j < k; ++j) std::swap(B[(size_t)piv * k + j], B[(size_t)col * k + j]);
        }
        for (int r = col + 1; r < q; ++r) {
            const double f = A[(size_t)r * q + col] / A[(size_t)col * q + col];
            if (f == 0.0) continue;
            for (int c2 = col; c2 < q; ++c2) A[(size_t)r * q + c2] -= f * A[(size_t)col * q + c2];
            for (int j = 0; j < k; ++j) B[(size_t)r * k + j] -= f * B[(size_t)col * k + j];
        }
    }
    for (int r = q - 1; r >= 0; --r)
        for (int j = 0; j < k; ++j) {
            double s = B[(size_t)r * k + j];
            for (int c2 = r + 1; c2 < q; ++c2) s -= A[(size_t)r * q + c2] * B[(size_t)c2 * k + j];
            B[(size_t)r * k + j] = s / A[(size_t)r * q + r];
        }
    DevBuf<double> cc, co;
    DEV_ALLOC(cc, ncomp * k, nullptr);
    DEV_ALLOC(co, std::max(D - 1, 1) * (int64_t)k, nullptr);
    HIP_TRY(hipMemcpy(cc.p, B.data(), ncomp * k * 8, hipMemcpyHostToDevice));
    if (D > 1)
        HIP_TRY(hipMemcpy(co.p, B.data() + (size_t)ncomp * k, (D - 1) * (int64_t)k * 8,
                          hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_null_apply, dim3(grid1((G + L) * k)), dim3(256), 0, 0, G, L, n1, D, k,
                       gcomp.p, lcomp, ldim.p, cc.p, co.p, Vg, Vl);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    return HDFE_STATUS_OK;
}

// sigma2 (n, device out) from the outcome (column 0 of V (n, m)), the
// residuals, the leverage complements and factors; centering 0 "weighted",
// 1 "reference"; *ybar_out: the mean the outcome was centred at.
int hdfe_kss_sigma2(int64_t n, int m, const double *V, const double *w, int centering,
                    const double *e, const double *Mb, const double *F, double *out,
                    double *ybar_out) {
    ARG_CHECK(V && e && Mb && F && out && (centering == 0 || centering == 1),
              "hdfe_kss_sigma2: bad argument");
    const int nb = 256;
    DevBuf<double> part;
    DEV_ALLOC(part, 2 * nb, nullptr);
    if (centering == 1)
        hipLaunchKernelGGL(k_rootmean_part, dim3(nb), dim3(256), 0, 0, n, m, V, w, part.p);
    else
        hipLaunchKernelGGL(k_wmean_part, dim3(nb), dim3(256), 0, 0, n, m, V, w, part.p);
    std::vector<double> h;
    HIP_TRY(to_host(h, part.p, 2 * nb, 0));
    double sw = 0.0, swy = 0.0;
    for (int b = 0; b < nb; ++b) {
        sw += h[2 * b];
        swy += h[2 * b + 1];
    }
    const double ybar = swy / sw;
    hipLaunchKernelGGL(k_sigma2, dim3(grid1(n)), dim3(256), 0, 0, n, m, V, w, ybar, centering, e,
                       Mb, F, out);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    if (ybar_out) *ybar_out = ybar;
    return HDFE_STATUS_OK;
}

// The six weighted sums of the plug-in moments (host out, 6).
int hdfe_kss_plugin(int64_t G, const int *gstart, const int *psi_codes, int64_t psi_off,
                    const double *gam, const double *geff, const double *w, double *out) {
    const int nb = 256;
    DevBuf<double> part;
    DEV_ALLOC(part, 6 * nb, nullptr);
    hipLaunchKernelGGL(k_plugin_part, dim3(nb), dim3(256), 0, 0, G, gstart, psi_codes, psi_off, gam,
                       geff, w, part.p);
    HIP_TRY(hipGetLastError());
    std::vector<double> h;
    HIP_TRY(to_host(h, part.p, 6 * nb, 0));
    for (int q = 0; q < 6; ++q) {
        double v = 0.0;
        for (int b = 0; b < nb; ++b) v += h[b * 6 + q];
        out[q] = v;
    }
    return HDFE_STATUS_OK;
}

}  // extern "C"
