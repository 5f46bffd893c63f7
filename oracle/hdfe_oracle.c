/*
 * hdfe_oracle.c -- CPU oracle for the hdfe-stream solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, in plain C (OpenMP for the
 * loops the reference runs with numba prange), the reference's numba kernels
 * on the solve path.  It is the checker the HIP library is compared against
 * and the timed CPU baseline.  Nothing in the product links or calls it: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Each function names the reference function and file:line it follows.  The
 * reference is FP64 throughout; so is this file.
 *
 * Layouts (the reference's, core/kernels_base.py:11-17):
 *   starts (G+1) int64   group boundaries over the identifying cells
 *   codes  (M, D) uint32 level codes of the non-streamed dimensions, row major
 *   n      (M)   double  cell weights
 *   sums   (M, m) double centred cell sums, row major
 *   offs   (D)   int64   start of each dimension's block in the stacked vector
 * Dense blocks (L, k) are row major (C order), as in the reference.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_CG_BLOCKS 256 /* core/solve.py:49 _CG_BLOCKS */

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* b += D_o' M_0 V for the columns of V given as cell sums.
 * core/kernels_base.py:64-83 (_nb_rhs). */
void orc_rhs(int64_t G, const int64_t *starts, int32_t D, const uint32_t *codes,
             const double *n, int32_t m, const double *sums, const int64_t *offs,
             double *b) {
    double *vbar = (double *)malloc(sizeof(double) * (m > 0 ? m : 1));
    for (int64_t g = 0; g < G; ++g) {
        int64_t s = starts[g], e = starts[g + 1];
        double Ng = 0.0;
        for (int j = 0; j < m; ++j) vbar[j] = 0.0;
        for (int64_t i = s; i < e; ++i) {
            Ng += n[i];
            for (int j = 0; j < m; ++j) vbar[j] += sums[i * m + j];
        }
        for (int j = 0; j < m; ++j) vbar[j] /= Ng;
        for (int64_t i = s; i < e; ++i)
            for (int d = 0; d < D; ++d) {
                int64_t a = offs[d] + codes[i * D + d];
                for (int j = 0; j < m; ++j) b[a * m + j] += sums[i * m + j] - n[i] * vbar[j];
            }
    }
    free(vbar);
}

/* Exact diagonal of S = D_o' M_0 D_o.  core/kernels_base.py:86-110 (_nb_diag). */
void orc_diag(int64_t G, const int64_t *starts, int32_t D, const uint32_t *codes,
              const double *n, const int64_t *offs, double *diag) {
    int64_t cap = 64;
    int64_t *lev = (int64_t *)malloc(sizeof(int64_t) * cap);
    double *u = (double *)malloc(sizeof(double) * cap);
    for (int64_t g = 0; g < G; ++g) {
        int64_t s = starts[g], e = starts[g + 1];
        if ((e - s) * D > cap) {
            cap = (e - s) * D;
            lev = (int64_t *)realloc(lev, sizeof(int64_t) * cap);
            u = (double *)realloc(u, sizeof(double) * cap);
        }
        double Ng = 0.0;
        int64_t q = 0;
        for (int64_t i = s; i < e; ++i) {
            Ng += n[i];
            for (int d = 0; d < D; ++d) {
                int64_t a = offs[d] + codes[i * D + d];
                diag[a] += n[i];
                int64_t r = 0;
                while (r < q && lev[r] != a) ++r;
                if (r == q) { lev[q] = a; u[q] = 0.0; ++q; }
                u[r] += n[i];
            }
        }
        for (int64_t r = 0; r < q; ++r) diag[lev[r]] -= u[r] * u[r] / Ng;
    }
    free(lev);
    free(u);
}

/* For each level, the groups that contain it, ascending.
 * core/kernels_base.py:200-228 (_nb_level_groups).  ptr has L+1 entries;
 * returns a malloc'd grp array. */
static int32_t *level_groups(int64_t G, const int64_t *starts, int32_t D, const uint32_t *codes,
                             const int64_t *offs, int64_t L, int64_t *ptr) {
    int64_t *mark = (int64_t *)malloc(sizeof(int64_t) * (L > 0 ? L : 1));
    for (int64_t a = 0; a < L; ++a) mark[a] = -1;
    memset(ptr, 0, sizeof(int64_t) * (L + 1));
    for (int64_t g = 0; g < G; ++g)
        for (int64_t i = starts[g]; i < starts[g + 1]; ++i)
            for (int d = 0; d < D; ++d) {
                int64_t a = offs[d] + codes[i * D + d];
                if (mark[a] != g) { mark[a] = g; ptr[a + 1] += 1; }
            }
    for (int64_t a = 0; a < L; ++a) ptr[a + 1] += ptr[a];
    int32_t *grp = (int32_t *)malloc(sizeof(int32_t) * (ptr[L] > 0 ? ptr[L] : 1));
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (L > 0 ? L : 1));
    for (int64_t a = 0; a < L; ++a) { fill[a] = ptr[a]; mark[a] = -1; }
    for (int64_t g = 0; g < G; ++g)
        for (int64_t i = starts[g]; i < starts[g + 1]; ++i)
            for (int d = 0; d < D; ++d) {
                int64_t a = offs[d] + codes[i * D + d];
                if (mark[a] != g) { mark[a] = g; grp[fill[a]++] = (int32_t)g; }
            }
    free(mark);
    free(fill);
    return grp;
}

static int dim_of(int32_t D, const int64_t *offs, int64_t r) {
    int dr = 0;
    while (dr + 1 < D && offs[dr + 1] <= r) ++dr;
    return dr;
}

/* Row nonzeros of S: the distinct levels of every group a level is in.
 * core/kernels_base.py:242-276 (_nb_row_nnz).  Returns nnz. */
int64_t orc_s_count(int64_t G, const int64_t *starts, int32_t D, const uint32_t *codes,
                    const int64_t *offs, int64_t L, int64_t *rnnz) {
    int64_t *ptr = (int64_t *)malloc(sizeof(int64_t) * (L + 1));
    int32_t *grp = level_groups(G, starts, D, codes, offs, L, ptr);
    int64_t total = 0;
#pragma omp parallel reduction(+ : total)
    {
        int64_t *mark = (int64_t *)malloc(sizeof(int64_t) * (L > 0 ? L : 1));
        for (int64_t a = 0; a < L; ++a) mark[a] = -1;
#pragma omp for schedule(dynamic, 64)
        for (int64_t r = 0; r < L; ++r) {
            int64_t cnt = 0;
            for (int64_t j = ptr[r]; j < ptr[r + 1]; ++j) {
                int64_t g = grp[j];
                for (int64_t i = starts[g]; i < starts[g + 1]; ++i)
                    for (int d = 0; d < D; ++d) {
                        int64_t c = offs[d] + codes[i * D + d];
                        if (mark[c] != r) { mark[c] = r; ++cnt; }
                    }
            }
            rnnz[r] = cnt;
            total += cnt;
        }
        free(mark);
    }
    free(ptr);
    free(grp);
    return total;
}

static int cmp_i32(const void *a, const void *b) {
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

/* The rows of S, sorted columns.  Row r sums, over the groups g containing
 * level r, the cells' n a' minus u_r U_g' / N_g.
 * core/kernels_base.py:192-198 and 279-358 (_nb_row_fill, the plain case:
 * the correction divides u_r by N_g, then multiplies by each cell's n). */
void orc_s_fill(int64_t G, const int64_t *starts, int32_t D, const uint32_t *codes,
                const double *n, const int64_t *offs, int64_t L, const int64_t *indptr,
                int32_t *indices, double *data) {
    int64_t *ptr = (int64_t *)malloc(sizeof(int64_t) * (L + 1));
    int32_t *grp = level_groups(G, starts, D, codes, offs, L, ptr);
    double *Ng = (double *)malloc(sizeof(double) * (G > 0 ? G : 1));
    for (int64_t g = 0; g < G; ++g) {
        double t = 0.0;
        for (int64_t i = starts[g]; i < starts[g + 1]; ++i) t += n[i];
        Ng[g] = t;
    }
#pragma omp parallel
    {
        double *acc = (double *)calloc((size_t)(L > 0 ? L : 1), sizeof(double));
        int64_t *mark = (int64_t *)malloc(sizeof(int64_t) * (L > 0 ? L : 1));
        for (int64_t a = 0; a < L; ++a) mark[a] = -1;
#pragma omp for schedule(dynamic, 64)
        for (int64_t r = 0; r < L; ++r) {
            int dr = dim_of(D, offs, r);
            uint32_t lr = (uint32_t)(r - offs[dr]);
            int64_t p0 = indptr[r], q = 0;
            for (int64_t j = ptr[r]; j < ptr[r + 1]; ++j) {
                int64_t g = grp[j];
                int64_t s = starts[g], e = starts[g + 1];
                double ur = 0.0;
                for (int64_t i = s; i < e; ++i)
                    if (codes[i * D + dr] == lr) ur += n[i];
                double w = ur / Ng[g];
                for (int64_t i = s; i < e; ++i) {
                    double v = -n[i] * w;
                    if (codes[i * D + dr] == lr) v += n[i];
                    for (int d = 0; d < D; ++d) {
                        int64_t c = offs[d] + codes[i * D + d];
                        if (mark[c] != r) { mark[c] = r; indices[p0 + q++] = (int32_t)c; acc[c] = 0.0; }
                        acc[c] += v;
                    }
                }
            }
            qsort(indices + p0, (size_t)q, sizeof(int32_t), cmp_i32);
            for (int64_t t = 0; t < q; ++t) data[p0 + t] = acc[indices[p0 + t]];
        }
        free(acc);
        free(mark);
    }
    free(ptr);
    free(grp);
    free(Ng);
}

/* out = S X for CSR S, dense X (L, k).  core/kernels_base.py:361-371. */
void orc_csr_matmat(int64_t L, const int64_t *indptr, const int32_t *indices,
                    const double *data, int64_t k, const double *X, double *out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < L; ++i) {
        for (int64_t j = 0; j < k; ++j) out[i * k + j] = 0.0;
        for (int64_t p = indptr[i]; p < indptr[i + 1]; ++p) {
            int64_t c = indices[p];
            double v = data[p];
            for (int64_t j = 0; j < k; ++j) out[i * k + j] += v * X[c * k + j];
        }
    }
}

/* parts[blk, j] = sum over the block's rows of A * B.  core/kernels_base.py:431-441. */
static void cg_dot(int64_t L, int64_t k, int64_t nblk, const double *A, const double *B,
                   double *parts) {
#pragma omp parallel for schedule(static)
    for (int64_t b = 0; b < nblk; ++b) {
        for (int64_t j = 0; j < k; ++j) parts[b * k + j] = 0.0;
        for (int64_t i = b * L / nblk; i < (b + 1) * L / nblk; ++i)
            for (int64_t j = 0; j < k; ++j) parts[b * k + j] += A[i * k + j] * B[i * k + j];
    }
}

/* X += P alpha, R -= AP alpha, with block sums of r^2 and r^2 dinv.
 * core/kernels_base.py:444-462. parts is (nblk, 2, k). */
static void cg_step(int64_t L, int64_t k, int64_t nblk, double *X, double *R, const double *P,
                    const double *AP, const double *alpha, const double *dinv, double *parts) {
#pragma omp parallel for schedule(static)
    for (int64_t b = 0; b < nblk; ++b) {
        double *p0 = parts + b * 2 * k, *p1 = p0 + k;
        for (int64_t j = 0; j < k; ++j) { p0[j] = 0.0; p1[j] = 0.0; }
        for (int64_t i = b * L / nblk; i < (b + 1) * L / nblk; ++i)
            for (int64_t j = 0; j < k; ++j) {
                X[i * k + j] += alpha[j] * P[i * k + j];
                double r = R[i * k + j] - alpha[j] * AP[i * k + j];
                R[i * k + j] = r;
                p0[j] += r * r;
                p1[j] += r * r * dinv[i];
            }
    }
}

/* P = dinv R + P beta.  core/kernels_base.py:465-474 (Jacobi form). */
static void cg_direction(int64_t L, int64_t k, double *P, const double *R, const double *dinv,
                         const double *beta) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < L; ++i)
        for (int64_t j = 0; j < k; ++j) P[i * k + j] = R[i * k + j] * dinv[i] + beta[j] * P[i * k + j];
}

/* Block Jacobi-PCG, one independent CG per column in lockstep, with the
 * reference's zero-column skip, warm start and per-column convergence.
 * core/solve.py:66-140 (_SolveMixin._pcg, the Jacobi branch).
 * zero: (k) uint8 or NULL; x0: (L, k) or NULL.  Returns 0. */
int orc_pcg(int64_t L, const int64_t *indptr, const int32_t *indices, const double *data,
            const double *dinv, int64_t k, const double *b, const uint8_t *zero,
            const double *x0, double tol, int64_t maxiter, double *X, int64_t *it_out,
            int *conv_out, double *rel_out) {
    size_t Lk = (size_t)L * (size_t)k;
    double *bnorm = (double *)calloc((size_t)k, sizeof(double));
    double *R = (double *)malloc(sizeof(double) * (Lk ? Lk : 1));
    double *P = (double *)malloc(sizeof(double) * (Lk ? Lk : 1));
    double *AP = (double *)malloc(sizeof(double) * (Lk ? Lk : 1));
    uint8_t *active = (uint8_t *)malloc((size_t)k + 1);
    double *rel = (double *)calloc((size_t)k + 1, sizeof(double));
    double *rz = (double *)calloc((size_t)k + 1, sizeof(double));
    double *alpha = (double *)calloc((size_t)k + 1, sizeof(double));
    double *beta = (double *)calloc((size_t)k + 1, sizeof(double));
    for (int64_t i = 0; i < L; ++i)
        for (int64_t j = 0; j < k; ++j) bnorm[j] += b[i * k + j] * b[i * k + j];
    for (int64_t j = 0; j < k; ++j) {
        bnorm[j] = sqrt(bnorm[j]);
        if (bnorm[j] == 0.0) bnorm[j] = 1.0;
    }
    if (x0 == NULL) {
        memset(X, 0, sizeof(double) * Lk);
        memcpy(R, b, sizeof(double) * Lk);
    } else {
        memcpy(X, x0, sizeof(double) * Lk);
        orc_csr_matmat(L, indptr, indices, data, k, X, AP);
        for (size_t t = 0; t < Lk; ++t) R[t] = b[t] - AP[t];
    }
    int any_zero = 0, all_zero = 1;
    for (int64_t j = 0; j < k; ++j) {
        int z = zero ? zero[j] != 0 : 0;
        any_zero |= z;
        all_zero &= z;
        active[j] = (uint8_t)!z;
    }
    if (any_zero) {
        for (int64_t i = 0; i < L; ++i)
            for (int64_t j = 0; j < k; ++j)
                if (zero[j]) { X[i * k + j] = 0.0; R[i * k + j] = 0.0; }
    }
    int64_t it = 0;
    int conv = 1;
    double relmax = 0.0;
    if (any_zero && all_zero) goto done;
    if (x0 != NULL) {
        int any = 0;
        for (int64_t j = 0; j < k; ++j) {
            double s = 0.0;
            for (int64_t i = 0; i < L; ++i) s += R[i * k + j] * R[i * k + j];
            rel[j] = sqrt(s) / bnorm[j];
            active[j] = active[j] && rel[j] > tol;
            any |= active[j];
            if (rel[j] > relmax) relmax = rel[j];
        }
        if (!any) goto done;
    }
    {
        int64_t nblk = L < ORC_CG_BLOCKS ? (L > 0 ? L : 1) : ORC_CG_BLOCKS;
        double *parts = (double *)calloc((size_t)(nblk * 2 * k), sizeof(double));
        double *dots = (double *)calloc((size_t)(nblk * k), sizeof(double));
        for (int64_t i = 0; i < L; ++i)
            for (int64_t j = 0; j < k; ++j) P[i * k + j] = R[i * k + j] * dinv[i];
        cg_dot(L, k, nblk, R, P, dots);
        for (int64_t j = 0; j < k; ++j) {
            double s = 0.0;
            for (int64_t q = 0; q < nblk; ++q) s += dots[q * k + j];
            rz[j] = s;
        }
        int any = 1;
        for (it = 1; it <= maxiter; ++it) {
            orc_csr_matmat(L, indptr, indices, data, k, P, AP);
            cg_dot(L, k, nblk, P, AP, dots);
            for (int64_t j = 0; j < k; ++j) {
                double pAp = 0.0;
                for (int64_t q = 0; q < nblk; ++q) pAp += dots[q * k + j];
                alpha[j] = (active[j] && pAp > 0) ? rz[j] / pAp : 0.0;
            }
            cg_step(L, k, nblk, X, R, P, AP, alpha, dinv, parts);
            any = 0;
            relmax = 0.0;
            for (int64_t j = 0; j < k; ++j) {
                double s0 = 0.0, s1 = 0.0;
                for (int64_t q = 0; q < nblk; ++q) {
                    s0 += parts[(q * 2) * k + j];
                    s1 += parts[(q * 2 + 1) * k + j];
                }
                rel[j] = sqrt(s0) / bnorm[j];
                active[j] = rel[j] > tol;
                any |= active[j];
                if (rel[j] > relmax) relmax = rel[j];
                beta[j] = s1; /* rz_new, held until the test below */
            }
            if (!any) break;
            for (int64_t j = 0; j < k; ++j) {
                double rz_new = beta[j];
                beta[j] = active[j] ? rz_new / (rz[j] != 0 ? rz[j] : 1.0) : 0.0;
                rz[j] = rz_new;
            }
            cg_direction(L, k, P, R, dinv, beta);
        }
        if (it > maxiter) it = maxiter;
        conv = !any;
        free(parts);
        free(dots);
    }
done:
    *it_out = it;
    *conv_out = conv;
    *rel_out = relmax;
    free(bnorm); free(R); free(P); free(AP); free(active); free(rel); free(rz);
    free(alpha); free(beta);
    return 0;
}

/* (D_o' M_0 D_o) P without forming S.  core/kernels_base.py:477-507.
 * out (L, k) is overwritten.  Serial: the oracle needs the value only. */
void orc_stream_matvec(int64_t G, const int64_t *starts, int32_t D, const uint32_t *codes,
                       const double *n, const int64_t *offs, int64_t L, int64_t k,
                       const double *P, double *out) {
    memset(out, 0, sizeof(double) * (size_t)L * (size_t)k);
    double *mg = (double *)malloc(sizeof(double) * (k > 0 ? k : 1));
    for (int64_t g = 0; g < G; ++g) {
        int64_t s = starts[g], e = starts[g + 1];
        double Ng = 0.0;
        for (int64_t j = 0; j < k; ++j) mg[j] = 0.0;
        for (int64_t i = s; i < e; ++i) {
            Ng += n[i];
            for (int64_t j = 0; j < k; ++j) {
                double tv = 0.0;
                for (int d = 0; d < D; ++d) tv += P[(offs[d] + codes[i * D + d]) * k + j];
                mg[j] += n[i] * tv;
            }
        }
        for (int64_t j = 0; j < k; ++j) mg[j] /= Ng;
        for (int64_t i = s; i < e; ++i)
            for (int64_t j = 0; j < k; ++j) {
                double tv = 0.0;
                for (int d = 0; d < D; ++d) tv += P[(offs[d] + codes[i * D + d]) * k + j];
                double v = n[i] * (tv - mg[j]);
                for (int d = 0; d < D; ++d) out[(offs[d] + codes[i * D + d]) * k + j] += v;
            }
    }
    free(mg);
}

/* V' M_D V from the cells (between-cell part; the caller adds the
 * within-cell scatter).  core/kernels_base.py:510-540 (_nb_assemble). */
void orc_assemble(int64_t G, const int64_t *starts, int32_t D, const uint32_t *codes,
                  const double *n, int32_t m, const double *sums, const int64_t *offs,
                  const double *Gamma, double *acc) {
    double *q = (double *)malloc(sizeof(double) * (m > 0 ? m : 1));
    double *qg = (double *)malloc(sizeof(double) * (m > 0 ? m : 1));
    memset(acc, 0, sizeof(double) * (size_t)m * (size_t)m);
    for (int64_t g = 0; g < G; ++g) {
        int64_t s = starts[g], e = starts[g + 1];
        double Ng = 0.0;
        for (int j = 0; j < m; ++j) qg[j] = 0.0;
        for (int64_t i = s; i < e; ++i) {
            Ng += n[i];
            for (int j = 0; j < m; ++j) {
                double v = sums[i * m + j] / n[i];
                for (int d = 0; d < D; ++d) v -= Gamma[(offs[d] + codes[i * D + d]) * m + j];
                qg[j] += n[i] * v;
            }
        }
        for (int j = 0; j < m; ++j) qg[j] /= Ng;
        for (int64_t i = s; i < e; ++i) {
            for (int j = 0; j < m; ++j) {
                double v = sums[i * m + j] / n[i];
                for (int d = 0; d < D; ++d) v -= Gamma[(offs[d] + codes[i * D + d]) * m + j];
                q[j] = v - qg[j];
            }
            for (int a = 0; a < m; ++a)
                for (int c = 0; c < m; ++c) acc[a * m + c] += n[i] * q[a] * q[c];
        }
    }
    free(q);
    free(qg);
}

/* Union-find with path halving.  core/kernels_base.py:147-152. */
static int64_t uf_find(int64_t *parent, int64_t x) {
    while (parent[x] != x) {
        parent[x] = parent[parent[x]];
        x = parent[x];
    }
    return x;
}

/* Levels of dimension `col` joined by shared groups; root = smallest level.
 * core/kernels_base.py:155-174 (_nb_components). */
void orc_components(int64_t G, const int64_t *starts, int32_t D, const uint32_t *codes,
                    int32_t col, int64_t n_levels, int64_t *parent) {
    for (int64_t x = 0; x < n_levels; ++x) parent[x] = x;
    for (int64_t g = 0; g < G; ++g) {
        int64_t s = starts[g], e = starts[g + 1];
        if (e <= s) continue;
        int64_t r0 = uf_find(parent, codes[s * D + col]);
        for (int64_t i = s + 1; i < e; ++i) {
            int64_t r = uf_find(parent, codes[i * D + col]);
            if (r != r0) {
                if (r < r0) { parent[r0] = r; r0 = r; }
                else parent[r] = r0;
            }
        }
    }
    for (int64_t x = 0; x < n_levels; ++x) parent[x] = uf_find(parent, x);
}

/* Rademacher +/-1 by a counter-based splitmix64 finaliser of
 * (seed, row ordinal, draw).  core/kernels_groups.py:52-77 (_nb_rademacher).
 * out is (n_rows, draws) row major. */
void orc_rademacher(int64_t ordinal, int64_t n_rows, int64_t draws, uint64_t seed,
                    int64_t first_draw, double *out) {
    for (int64_t i = 0; i < n_rows; ++i) {
        uint64_t row = (uint64_t)(ordinal + i);
        for (int64_t j = 0; j < draws; ++j) {
            uint64_t d = (uint64_t)(first_draw + j);
            uint64_t x = seed ^ (row * 0x9E3779B97F4A7C15ULL) ^ (d * 0xBF58476D1CE4E5B9ULL);
            x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
            x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
            x = x ^ (x >> 31);
            out[i * draws + j] = (x >> 63) == 0 ? 1.0 : -1.0;
        }
    }
}

/* acc += D_o' W M_0 R and acc_x += X' W M_0 R over rows grouped by the
 * streamed dimension.  inverse/kernels_inverse.py:34-71 (_nb_center_rows).
 * acc (L, m), acc_x (kc, m). */
void orc_center_rows(int64_t G, const int64_t *starts, int32_t D, const int64_t *codes,
                     const int64_t *offs, const double *w, int32_t kc, const double *X,
                     int32_t m, const double *R, double *acc, double *acc_x) {
    for (int64_t g = 0; g < G; ++g) {
        int64_t s = starts[g], e = starts[g + 1];
        double sw = 0.0;
        for (int64_t i = s; i < e; ++i) sw += w[i];
        if (sw <= 0.0) continue;
        for (int j = 0; j < m; ++j) {
            double total = 0.0;
            for (int64_t i = s; i < e; ++i) total += w[i] * R[i * m + j];
            double mean = total / sw;
            for (int64_t i = s; i < e; ++i) {
                double v = w[i] * (R[i * m + j] - mean);
                for (int d = 0; d < D; ++d) acc[(offs[d] + codes[i * D + d]) * m + j] += v;
                for (int c = 0; c < kc; ++c) acc_x[c * m + j] += X[i * kc + c] * v;
            }
        }
    }
}

/* out[i, j] = (A u_j)_i from the level block U and covariate block C.
 * kernels_inverse.py:74-108 (_nb_project_rows). */
void orc_project_rows(int64_t G, const int64_t *starts, int32_t D, const int64_t *codes,
                      const int64_t *offs, const double *w, int32_t kc, const double *X,
                      const double *C, int32_t m, const double *R, const double *U,
                      double *out) {
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < G; ++g) {
        int64_t s = starts[g], e = starts[g + 1];
        double sw = 0.0;
        for (int64_t i = s; i < e; ++i) sw += w[i];
        for (int j = 0; j < m; ++j) {
            double total = 0.0;
            for (int64_t i = s; i < e; ++i) {
                double h = 0.0;
                for (int d = 0; d < D; ++d) h += U[(offs[d] + codes[i * D + d]) * m + j];
                for (int c = 0; c < kc; ++c) h += X[i * kc + c] * C[c * m + j];
                out[i * m + j] = h;
                total += w[i] * (R[i * m + j] - h);
            }
            double mean = sw > 0.0 ? total / sw : 0.0;
            for (int64_t i = s; i < e; ++i) out[i * m + j] += mean;
        }
    }
}

/* JLA per-row running moments for one block of draws, over rows in their
 * ordinal order: P = (proj)^2, M = (r - proj)^2 and their squares and
 * product, summed over the block's draws.  leaveout.py:181-190 (the sink of
 * jla_leverages).  moments is (5, n_rows). */
void orc_jla_moments(int64_t n_rows, int32_t k, const double *r, const double *proj,
                     double *moments) {
    for (int64_t i = 0; i < n_rows; ++i) {
        double sP = 0, sM = 0, sPP = 0, sMM = 0, sPM = 0;
        for (int j = 0; j < k; ++j) {
            double p = proj[i * k + j];
            double P = p * p, d = r[i * k + j] - p, M = d * d;
            sP += P; sM += M; sPP += P * P; sMM += M * M; sPM += P * M;
        }
        moments[0 * n_rows + i] += sP;
        moments[1 * n_rows + i] += sM;
        moments[2 * n_rows + i] += sPP;
        moments[3 * n_rows + i] += sMM;
        moments[4 * n_rows + i] += sPM;
    }
}

/* Worker effects and residuals: per group the weighted mean of
 * y - X beta - sum_d gamma[level], then e = that less the mean.
 * core/kernels_base.py:543-619 (_nb_pass2, the residual and g_eff outputs). */
void orc_pass2(int64_t G, const int64_t *starts, int32_t D, const int64_t *codes,
               const int64_t *offs, const double *w, const double *y, int32_t k,
               const double *X, const double *beta, const double *gam_y, double *g_eff,
               double *e_out) {
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < G; ++g) {
        int64_t s = starts[g], e = starts[g + 1];
        double Wg = 0.0, mu = 0.0;
        for (int64_t i = s; i < e; ++i) {
            double u = y[i];
            for (int j = 0; j < k; ++j) u -= X[i * k + j] * beta[j];
            for (int d = 0; d < D; ++d) u -= gam_y[offs[d] + codes[i * D + d]];
            Wg += w[i];
            mu += w[i] * u;
            e_out[i] = u;
        }
        mu /= Wg;
        g_eff[g] = mu;
        for (int64_t i = s; i < e; ++i) e_out[i] -= mu;
    }
}

/* Articulation points of the alive subgraph by iterative Tarjan.
 * core/kernels_graph.py:22-86 (_nb_articulation_points). */
void orc_articulation_points(int64_t nv, const int64_t *indptr, const int64_t *indices,
                             uint8_t *is_cut, const uint8_t *alive) {
    int64_t *disc = (int64_t *)malloc(sizeof(int64_t) * (nv ? nv : 1));
    int64_t *low = (int64_t *)malloc(sizeof(int64_t) * (nv ? nv : 1));
    int64_t *parent = (int64_t *)malloc(sizeof(int64_t) * (nv ? nv : 1));
    int64_t *sv = (int64_t *)malloc(sizeof(int64_t) * (nv ? nv : 1));
    int64_t *sp = (int64_t *)malloc(sizeof(int64_t) * (nv ? nv : 1));
    for (int64_t v = 0; v < nv; ++v) { disc[v] = -1; low[v] = 0; parent[v] = -1; }
    int64_t timer = 0;
    for (int64_t start = 0; start < nv; ++start) {
        if (disc[start] != -1 || !alive[start]) continue;
        disc[start] = low[start] = timer++;
        sv[0] = start;
        sp[0] = indptr[start];
        int64_t top = 0, root_children = 0;
        while (top >= 0) {
            int64_t u = sv[top];
            if (sp[top] < indptr[u + 1]) {
                int64_t v = indices[sp[top]++];
                if (v == parent[u] || !alive[v]) continue;
                if (disc[v] == -1) {
                    parent[v] = u;
                    disc[v] = low[v] = timer++;
                    ++top;
                    sv[top] = v;
                    sp[top] = indptr[v];
                    if (u == start) ++root_children;
                } else if (disc[v] < low[u]) {
                    low[u] = disc[v];
                }
            } else {
                --top;
                if (top >= 0) {
                    int64_t p = sv[top];
                    if (low[u] < low[p]) low[p] = low[u];
                    if (p != start && low[u] >= disc[p]) is_cut[p] = 1;
                }
            }
        }
        if (root_children > 1) is_cut[start] = 1;
    }
    free(disc); free(low); free(parent); free(sv); free(sp);
}

/* Components of the alive subgraph, labelled in the order of their smallest
 * vertex; dead vertices get -1.  core/kernels_graph.py:89-116. */
int64_t orc_bfs_components(int64_t nv, const int64_t *indptr, const int64_t *indices,
                           const uint8_t *alive, int64_t *label) {
    int64_t *queue = (int64_t *)malloc(sizeof(int64_t) * (nv ? nv : 1));
    for (int64_t v = 0; v < nv; ++v) label[v] = -1;
    int64_t comp = 0;
    for (int64_t s = 0; s < nv; ++s) {
        if (!alive[s] || label[s] != -1) continue;
        label[s] = comp;
        queue[0] = s;
        int64_t head = 0, tail = 1;
        while (head < tail) {
            int64_t u = queue[head++];
            for (int64_t q = indptr[u]; q < indptr[u + 1]; ++q) {
                int64_t v = indices[q];
                if (alive[v] && label[v] == -1) { label[v] = comp; queue[tail++] = v; }
            }
        }
        ++comp;
    }
    free(queue);
    return comp;
}

/* The splitmix64 finaliser alone (the mixing step of orc_rademacher), for
 * checking its constants against the generator's published outputs. */
uint64_t orc_mix64(uint64_t x) {
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}
