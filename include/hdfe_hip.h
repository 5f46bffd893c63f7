/*
 * hdfe_hip.h -- C ABI of libhdfe_hip.so, the MI355X (gfx950) solve path of
 * hdfe-stream.
 *
 * Plain pointers and sizes only.  Every function returns a status:
 *   0 ok, 1 invalid argument, 2 HIP error, 3 out of device memory, 4 no device;
 * hdfe_last_error() gives the text of the calling thread's last failure.
 * Host bindings map 1 -> ValueError, 3 -> MemoryError, 4 -> RuntimeError,
 * a missing library -> ImportError (the reference's contract for its device
 * backend, hdfe_stream/core/gpu.py:33-51, tests/test_gpu.py:113-136).
 *
 * Arrays marked "host or device" are host pointers when on_device == 0 and
 * device pointers when on_device == 1.  (L, k) blocks are row major.
 */
#ifndef HDFE_HIP_H
#define HDFE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------- library */

/* Text of this thread's last error. */
const char *hdfe_last_error(void);

/* ABI version (1). */
int hdfe_version(void);

/* Name of the current device and its free/total memory in bytes.
 * Replaces the device probe of cuda_backend(), core/gpu.py:41-51. */
int hdfe_device_info(char *name, int len, int64_t *free_bytes, int64_t *total_bytes);

/* ------------------------------------------------------ explicit solver */

/* Copy S (CSR, L x L, sorted column indices) and the Jacobi inverse diagonal
 * to the device once; *out receives the handle.  indptr int64 (L+1), indices
 * int32 (nnz), data float64 (nnz), dinv float64 (L); host or device.  L must
 * be below 2^31; with 2^31 nonzeros or more the device keeps 64-bit row
 * pointers (the reference's int64 index arrays, core/solve.py:169, 188-189).
 * Replaces explicit_solver's upload, core/gpu.py:186-192 (and the caller at
 * core/solve.py:221-229). */
int hdfe_csr_create(int64_t L, int64_t nnz, const int64_t *indptr, const int32_t *indices,
                    const double *data, const double *dinv, int on_device, void **out);

/* L, nnz and the number of rows split across workgroups (long rows). */
int hdfe_csr_info(void *h, int64_t *L, int64_t *nnz, int *nlong);

/* Block Jacobi-PCG for k right-hand sides b (L, k), host or device:
 * zero (k, uint8, host, may be NULL) marks columns solved by 0; x0 (L, k) or
 * NULL is the starting guess; columns stop at ||r|| <= tol ||b||.
 * precision 0: FP64 throughout.  precision 1: FP32 CG (relative inner
 * tolerance inner_tol) inside FP64 iterative refinement on the true FP64
 * residual.  Outputs: X (L, k), iterations (total inner iterations in mode 1),
 * converged, max_rel (the largest relative residual), outer (refinement
 * steps; may be NULL).  stream: the caller's hipStream_t for on_device == 1
 * (work is ordered after it and X is ready on it).
 * Replaces solve(b, zero, x0) / pcg, core/gpu.py:194-253, and the CPU
 * _SolveMixin._pcg, core/solve.py:66-140. */
int hdfe_pcg_solve(void *h, int64_t k, const double *b, const uint8_t *zero, const double *x0,
                   double tol, int64_t maxiter, int precision, double inner_tol, double *X,
                   int64_t *iterations, int *converged, double *max_rel, int64_t *outer,
                   int on_device, void *stream);

/* A matrix-free S over a panel's identifying cells, for the same solves as
 * an explicit S (hdfe_pcg_solve, hdfe_csr_matmat, hdfe_destroy): the
 * reference's stream_cg fallback (`_nb_stream_matvec`,
 * core/kernels_base.py:477-507, chosen at core/solve.py:205-218, 251-271
 * when S does not fit).  Device arrays: igstart (Gi + 1) cells per group,
 * icodes (D, ld) the cells' level codes per non-streamed dimension, cell_n
 * (M) the cells' weights, dinv (L); offs (host, D) each dimension's first
 * level.  The arrays must outlive the handle (they stay the panel's). */
int hdfe_stream_create(int64_t L, int64_t Gi, int64_t M, int D, const int *igstart,
                       const int *icodes, int64_t ld, const int64_t *offs, const double *cell_n,
                       const double *dinv, void **out);

/* The operator of hdfe_stream_create over new cell arrays of the same
 * structure -- the IRLS reweighting keeps the cells and changes their
 * weights: when igstart and icodes have the operator's sizes and content
 * (a 64-bit fingerprint of both), it takes the new pointers, weights and
 * inverse diagonal and keeps its plan (*rebound = 1); otherwise it is left
 * as it was (*rebound = 0) and the caller creates a new one.  No
 * counterpart in the reference, which rebuilds nothing for stream_cg. */
int hdfe_stream_rebind(void *h, int64_t L, int64_t Gi, int64_t M, int D, const int *igstart,
                       const int *icodes, int64_t ld, const double *cell_n, const double *dinv,
                       int *rebound);

/* y = S x (k columns, FP64), host or device.  The operator of
 * _nb_csr_matmat, core/kernels_base.py:361-371. */
int hdfe_csr_matmat(void *h, int64_t k, const double *x, double *y, int on_device);

/* y = S x in FP32 (FP32 values and vectors: the operator of the mixed
 * precision inner solve), host or device float32 (L, k) blocks. */
int hdfe_csr_matmat_f32(void *h, int64_t k, const float *x, float *y, int on_device);

/* Copy S as the device holds it to host arrays (indptr int64 (L+1),
 * indices int32, data float64, dinv float64 (L)); any may be NULL.  For
 * checks of the device-built S against the reference's (_build_explicit,
 * core/solve.py:148-194). */
int hdfe_csr_export(void *h, int64_t *indptr, int32_t *indices, double *data, double *dinv);

/* HIP-event timing of the CG's SpMM launches of later solves on this handle
 * (enable resets the totals) and the totals (milliseconds, launches).  The
 * reference has no counterpart; its timings are the solver-info seconds
 * (core/solve.py:370-381). */
int hdfe_csr_timing(void *h, int enable);
int hdfe_csr_timing_read(void *h, double *ms, int64_t *launches);

/* The blocked small-k SpMM's row parts per XCD for k (1..8) columns in one
 * precision (0 FP64, 1 FP32), as chosen by timing at the matrix's first use of
 * that k (0: untuned, 128); set >= 1 overrides it.  Only the row partition
 * depends on it, so results are bit-identical for every value.  No
 * counterpart in the reference (its SpMM is scipy's / cuSPARSE's); for the
 * launch autotuner's report and tests. */
int hdfe_csr_blk_parts(void *h, int k, int precision, int set, int *parts);

/* The wide SpMM's form for k > 8 columns at `precision` (0 FP64, 1 FP32), as
 * the launch autotuner chose it at the matrix's first use with that k (-1: not
 * yet): 0 a wave per row over all columns, 1 and 2 column slabs per XCD with
 * 64 or 128 workgroups per XCD; set >= 0 overrides it.  No counterpart in the
 * reference; for the autotuner's report and tests. */
int hdfe_csr_wide_mode(void *h, int k, int precision, int set, int *mode);

/* The column-blocked small-k product's form for k <= 8 columns (no
 * counterpart in the reference, whose cuSPARSE SpMM has one form): 0 the
 * default (tiled where it applies), 1 groups of column blocks with their
 * partial products and a combine, 2 tiled (a workgroup's rows streamed block
 * by block through LDS, no partial products; 16-bit block-local column
 * indices and row offsets where they fit), 3 tiled with 32-bit indices (the
 * same sums as 2); `set` >= 0 overrides it.  Each form is deterministic;
 * 1 and 2 differ in rounding. */
int hdfe_csr_small_form(void *h, int k, int precision, int set, int *form);

/* The row pointers' width on the device, 32 or 64 (*bits); set = 64 widens
 * 32-bit ones (the matrix then runs the plain CSR products only, as one of
 * 2^31 nonzeros or more does), set = -1 reads.  No counterpart in the
 * reference (scipy picks the index type by size); for tests of the 64-bit
 * path at sizes that fit a test. */
int hdfe_csr_row_pointer_bits(void *h, int set, int *bits);

/* The number of CG batches held as HIP graphs on this matrix.  Batches are
 * captured only on systems small enough that launching costs more than the
 * device time of an iteration.  No counterpart in the reference (it has no
 * graphs); for tests. */
int hdfe_csr_graphs(void *h, int64_t *count);

/* Average duration in microseconds of the solve's SpMM kernel on this
 * matrix (k right-hand sides, fused pAp sums): reps back-to-back launches
 * between two HIP events on the solver's stream; precision 0 FP64, 1 FP32.
 * No counterpart in the reference (measurement only). */
int hdfe_csr_spmm_time(void *h, int64_t k, int reps, int precision, double *avg_us);

/* Column norms sqrt(sum_i X_ij^2) of a device (L, k) block, to the host,
 * in a fixed order: the zero-column test of core/solve.py:354-355. */
int hdfe_col_norms(const double *X, int64_t L, int64_t k, double *out);

/* Returns the library's cached free device blocks to the driver (*freed:
 * their bytes; the slabs of blocks under 256 KB stay), after synchronising
 * the device.  The library keeps freed blocks for reuse and otherwise
 * returns them only when one of its own allocations fails, so a caller that
 * allocates elsewhere (torch) calls this after a large phase or before
 * retrying an allocation that failed.  No counterpart in the reference (its
 * device memory is CuPy's pool, released by the backend's release(),
 * core/gpu.py:33-51). */
int hdfe_cache_release(int64_t *freed);

/* dst <- src; kind 0 device to host, 1 host to device, 2 device to device. */
int hdfe_memcpy(void *dst, const void *src, int64_t bytes, int kind);

/* Free the handle and its device memory (the host S release of
 * core/solve.py:222-223 has no counterpart: nothing is kept on the host). */
void hdfe_destroy(void *h);

/* --------------------------------------------------- panel (passes 0-2) */

/* Passes 0 and 1 on the device for a panel resident in HBM: fe (nfe device
 * pointers to int64 id columns), vars (m device pointers to float64 columns,
 * the outcome first), weights (NULL, or a float32 (weights_f32 = 1) or
 * float64 device column).  Drops singletons when fixef_rm_singleton (dimensions
 * in turn, core/passes.py:859-904), streams the dimension with the most levels
 * unless stream_dim >= 0 (core/estimator.py:371-385), codes every other
 * dimension in sorted level order and sorts the rows by (group, codes[,
 * variables if sort_vars]) (core/passes.py:311-394), then collapses cells and
 * keeps the identifying ones (core/passes.py:397-589).  *out: the panel. */
int hdfe_panel_create(int64_t n_rows, int nfe, const int64_t *const *fe, int m,
                      const double *const *vars, const void *weights, int weights_f32,
                      int fixef_rm_singleton, int stream_dim, int sort_vars, void **out);

/* info[16] = n_in, n, G, C, M, Gi, L, D, gdim, singletons, singleton rounds,
 * incidences, nnz(S), components, m, weighted; nlev[D], dims[D] (may be NULL). */
int hdfe_panel_info(void *h, int64_t *info, int64_t *nlev, int *dims);

/* Device pointer and length of a panel array by name ("g", "codes", "V", "w",
 * "perm", "gstart", "cstart", "glevels", "levels<t>", "means", "raw_ss", "W",
 * "obs", "cnt", "icodes", "in", "isums", "igstart", "igroup", "diag", "rhs",
 * "comp", "offs"). */
int hdfe_panel_array(void *h, const char *name, void **ptr, int64_t *count);

/* Copies `bytes` of device memory at src (a panel array) into page-locked
 * host memory at dst on a stream of the panel's own, after the work queued on
 * the panel so far, so that it overlaps the panel's later kernels;
 * hdfe_panel_copy_wait waits for every such copy.  No counterpart in the
 * reference (its level values are host arrays throughout); the fit starts
 * the streamed dimension's level values this way after pass 0. */
int hdfe_panel_copy_async(void *h, const void *src, void *dst, int64_t bytes);
int hdfe_panel_copy_wait(void *h);

/* Step 2's system on the device: diag (core/kernels_base.py:86-110), the
 * right-hand side (:64-83) and the explicit S (core/solve.py:148-194,
 * core/kernels_base.py:192-358) with its solver handle in *matrix (owned by the
 * panel: use it with hdfe_pcg_solve, never hdfe_destroy).  *seconds: wall time. */
int hdfe_panel_system(void *h, int with_rhs, void **matrix, double *seconds);

/* The explicit S build's budget in bytes (set >= 0; -1: no limit; < -1:
 * unchanged): over it hdfe_panel_system builds diag and rhs but no S
 * (*matrix NULL), the reference's _STooLarge test before allocating S
 * (core/solve.py:148-194); *skipped is the estimate of the S not built (0:
 * built).  Either output may be NULL. */
int hdfe_panel_budget(void *h, int64_t set, int64_t *budget, int64_t *skipped);

/* Components of the first non-streamed dimension through the identifying
 * groups, ranked by rows with ties to the smallest level
 * (core/kernels_base.py:155-174, core/passes.py:933-979). */
int hdfe_panel_components(void *h, int64_t *ncomp);

/* Connected components of two fixed-effect dimensions' levels linked by the
 * rows (core/passes.py:964-979 with the streamed dimension, :525-527 for two
 * others; the `pair_components` that fe_dof="exact" takes the largest of,
 * core/estimator.py:446-464).  a < b index the dimensions in fit order: 0
 * the streamed one, t + 1 the non-streamed t. */
int hdfe_panel_pair_components(void *h, int a, int b, int64_t *count);

/* The coefficients' scores u_i = w_i e_i x~_i for the covariates cols (host,
 * 1 <= kc <= 8; x~ partialled on the fixed effects from the reduced
 * solutions Gm, e the residuals in sorted row order) summed in one pass in a
 * fixed order, as the reference's residual pass collects its meats
 * (core/inference.py:430-668): out (host, 1 + 2 kc^2) = [sum w e^2,
 * sum_i u_i u_i', sum_g s_g s_g' with s_g the scores' sum over group g of
 * the streamed dimension]; U (n, kc, device) receives u when not NULL. */
int hdfe_panel_scores(void *h, const double *Gm, const int *cols, int kc, const double *e,
                      double *U, double *out);

/* Normalise Gamma (L, k), device, in place: obs-weighted mean zero per
 * component for the first dimension, overall for the others
 * (core/solve.py:383-399).  stream may be NULL. */
int hdfe_panel_normalize(void *h, double *Gamma, int k, void *stream);

/* A = V' M_D V (m x m, host) from the cells: the within-cell scatter of pass 1
 * plus the between-cell part (core/inference.py:109-115,
 * core/kernels_base.py:510-540).  m <= 8. */
int hdfe_panel_assemble(void *h, const double *Gamma, double *A);

/* Pass 2 (core/kernels_base.py:543-619): gam_y = Gamma[:, 0] - Gamma[:, xcols]
 * beta (L), the streamed effects geff (G) and residuals (n, sorted row
 * order), all device outputs; beta and xcols (kx) on the host. */
int hdfe_panel_pass2(void *h, const double *Gamma, const double *beta, const int *xcols, int kx,
                     double *gam_y, double *geff, double *resid);

/* New variables V (n, m) and weights w (n), device, in the panel's sorted row
 * order, and pass 1 again over the same rows and levels; the caller then
 * rebuilds the system.  One weighted least-squares step of IRLS
 * (core/irls.py:92-123: _wls after each row pass). */
int hdfe_panel_reweight(void *h, const double *V, const double *w);

/* out (n, kc, device) = the columns cols (host, kc) of the panel's variables
 * partialled on the fixed effects (M_F x), from their reduced solutions Gm
 * (L, m): per group, x less the levels' effects, less its weighted mean.
 * The covariate block of the adjoint; no counterpart in the reference, whose
 * adjoint hook is solve_normal_equations (normal_equations.py:115-231). */
int hdfe_panel_partial(void *h, const double *Gm, const int *cols, int kc, double *out);

/* Free the panel, its device arrays and its matrix. */
void hdfe_panel_destroy(void *h);

/* Measurement only (no counterpart in the reference): copies `bytes` (a
 * multiple of 16) between two fresh device buffers `reps` times with a plain
 * streaming kernel (k_copy_calibrate) and returns the mean launch time in
 * microseconds.  Its known traffic calibrates the HBM byte counters. */
int hdfe_copy_calibrate(int64_t bytes, int reps, double *avg_us);

/* ------------------------------------------------ leave-out row passes */

/* Rademacher +/-1 draws (n, k) on the device: entry (i, j) from the
 * counter-based splitmix64 hash of (seed, ordinal + i, first_draw + j)
 * (core/kernels_groups.py:52-77, _nb_rademacher). */
int hdfe_rademacher(int64_t ordinal, int64_t n, int k, uint64_t seed, int64_t first_draw,
                    double *out);

/* Row-pass state for a panel's sorted rows: the level-to-rows CSR of each of
 * the D dimensions (codes (D, n) device; offs, n_levels host).  The reference
 * iterates row files in whole groups instead (inverse.py:141-160). */
int hdfe_rows_create(int64_t n, int64_t G, int D, const int *codes, const int64_t *offs_h,
                     const int64_t *nlev_h, void **out);

/* The reduced right-hand side of a block of k Rademacher vectors over the
 * rows: rhs (L, k) = D_o' W M_0 R, the draws keyed on the sorted row ordinal,
 * in the square-root weight metric when sqrtw (inverse.py:162-180,
 * kernels_inverse.py:34-71, leaveout.py:179). */
int hdfe_rows_center(void *h, const int *g, const int *gstart, const double *w, int sqrtw,
                     uint64_t seed, int64_t first_draw, int k, double *rhs);

/* The projection back to rows (A S^- A' r)_i from the reduced solution U
 * (L, k), fused with the JLA sink: moments (5, n) += the block's sums of P,
 * M, P^2, M^2, P M per row (kernels_inverse.py:74-108, leaveout.py:181-190).
 * With covariates in the design (kc >= 1): Xd (n, kc) the covariates less
 * their weighted group means and Cb (kc, k) the solution's covariate block
 * (the x_i'C term of _nb_project_rows). */
int hdfe_rows_project_moments(void *h, const int *gstart, const int *codes, const double *w,
                              int sqrtw, uint64_t seed, int64_t first_draw, int k,
                              const int64_t *offs_d, const double *U, int kc, const double *Xd,
                              const double *Cb, double *moments);

/* The covariate block of a block of draws' reduced right-hand side, out
 * (kc, k) = X'W M_0 r = sum_i w_i Xd_i r_i' (the acc_x of
 * _nb_center_rows, kernels_inverse.py:34-71, inverse.py:162-180); the draws
 * as hdfe_rows_center's. */
int hdfe_rows_cov_center(void *h, int kc, const double *Xd, const double *w, int sqrtw,
                         uint64_t seed, int64_t first_draw, int k, double *out);

/* Leverages from the moments, in place: rows 0, 1, 2 become P_bar, M_bar and
 * the non-linearity factor (leaveout.py:209-282); stats (host, 4) = min, max,
 * mean of P_bar, largest |P + M - 1|. */
int hdfe_jla_combine(int64_t n, int64_t n_draws, double *moments, double *stats);

/* The level-to-rows CSR of a row-pass state (device pointers): rptr (L+1),
 * rrow (n D), the chunks of at most 1024 rows (cptr (L+1), clvl, cbeg) and
 * their count. */
int hdfe_rows_csr(void *h, const int **rptr, const int **rrow, const int **cptr, const int **clvl,
                  const int **cbeg, int64_t *nchunk);

/* out (L, k) = per level, the sum in chunk order of part (nchunk, k). */
int hdfe_rows_chunk_sum(void *h, int k, const double *part, double *out);

void hdfe_rows_destroy(void *h);

/* ------------------------------------------------- leave-out (KSS) */

/* Group weights W_g (G, device) = sum of w over the group's rows (w NULL:
 * ones); the denominators of _nb_spread_groups (kernels_inverse.py:133-152). */
int hdfe_kss_group_weight(int64_t G, const int *gstart, const double *w, double *Wg);

/* out (L, k) = D_o' W diag(x) R for a row block R (n, k), or D_o' W rho with
 * rho the streamed block V (G, k) spread over the rows (V / W_g) when R is
 * NULL (_nb_reduce_rows + _nb_spread_groups, kernels_inverse.py:111-152).
 * x (n) may be NULL (ones). */
int hdfe_kss_reduce(void *rows_h, int k, const int *g, const double *w, const double *x,
                    const double *R, const double *V, const double *Wg, double *out);

/* The rows (A u)_i (n, k; skipped when rows is NULL) and streamed block Ug
 * (G, k) of a coefficient-space solution U (L, k) with streamed right-hand
 * side V (G, k) (_nb_project_coef, kernels_inverse.py:156-188). */
int hdfe_kss_project(int64_t G, int D, int k, const int *gstart, const int *codes, int64_t n,
                     const int64_t *offs_d, const double *w, const double *Wg, const double *U,
                     const double *V, double *rows, double *Ug);

/* out (G, k) = per group sum of w x R (_nb_group_sums, core/kernels_groups.py:
 * 16-27, with the sigma2 factor x of _trace_block's sink, leaveout.py:421-428). */
int hdfe_kss_group_sums(int64_t G, int k, const int *gstart, const double *w, const double *x,
                        const double *R, double *out);

/* Hutchinson's eight cross-moments summed over rows and the k draws (host
 * out, 8): sum w, w a1 a2, w b1 b2, w (a1 b2 + b1 a2), w a1, w a2, w b1, w b2
 * with a the psi dimension's level value and b the group value of Z (1) and
 * U (2) (_nb_trace_moments, kernels_inverse.py:191-231). */
int hdfe_kss_trace_moments(int64_t G, int k, const int *gstart, const int *psi_codes,
                           int64_t psi_off, const double *w, const double *Zl, const double *Zg,
                           const double *Ul, const double *Ug, double *out);

/* Remove, in place, the null-space part of a coefficient-space block (Vg
 * (G, k), Vl (L, k), device): one basis vector per connected component and
 * one per further dimension (_null_space / _drop_null_component,
 * inverse.py:376-460).  lcomp: component of each first-dimension level
 * (device); codes0: the rows' first-dimension codes; offs / nlev host. */
int hdfe_kss_drop_null(int64_t G, int64_t L, int D, int64_t ncomp, int k, const int *gstart,
                       const int *codes0, const int *lcomp, const int64_t *offs_h,
                       const int64_t *nlev_h, double *Vg, double *Vl);

/* sigma2_i from y (column 0 of V (n, m)), the residuals e, the leverage
 * complements Mb and factors F, device in/out (_leave_one_out_sigma2,
 * leaveout.py:949-1016): centering 0 "weighted", w_i (y_i - ybar_w) e_i /
 * Mb_i F_i; centering 1 "reference", (sqrt(w_i) y_i - mean sqrt(w) y)
 * sqrt(w_i) e_i / Mb_i F_i; *ybar_out the mean used. */
int hdfe_kss_sigma2(int64_t n, int m, const double *V, const double *w, int centering,
                    const double *e, const double *Mb, const double *F, double *out,
                    double *ybar_out);

/* The plug-in moments' sums (host out, 6): sum w, w a, w b, w a^2, w b^2,
 * w a b with a the row's psi effect gam[psi_off + code] and b its group
 * effect geff[g] (_plug_in_moments, leaveout.py:807-832). */
int hdfe_kss_plugin(int64_t G, const int *gstart, const int *psi_codes, int64_t psi_off,
                    const double *gam, const double *geff, const double *w, double *out);

/* ------------------------------------------------- connected sets */

/* The leave-one-out connected set of the rows (a[i], b[i]), a the workers
 * (device int64 id columns): keep (n, device uint8) marks the rows whose worker
 * and firm both survive; info (host, 8) = components before, workers with one
 * row, cut workers, rounds, workers and firms left, distinct pairs, vertices.
 * Replaces _find_components(leave_one_out=True) / _leave_one_out
 * (core/connected.py:378-562) and its graph kernels (core/kernels_graph.py). */
int hdfe_loo_prune(int64_t n, const int64_t *a_ids, const int64_t *b_ids, uint8_t *keep,
                   int64_t *info);

/* Standalone connected components of the graph whose vertices are the levels
 * of two id columns (a_ids, b_ids: device int64, n rows) and whose edges are
 * the rows.  min_deg0/min_deg1 >= 1 first prune to the (k0, k1)-core of the
 * distinct pairs; leave_one_out then keeps the leave-one-out connected set;
 * rank_by -1 ranks components by rows, 0 or 1 by that dimension's levels.
 * Levels are coded in sorted order; components are numbered by their
 * smallest first-dimension level and ranked by size, stably.  *out is a
 * handle read by the three calls below and freed by hdfe_components_destroy.
 * Replaces connected_components (core/connected.py:211-338),
 * _find_components (:378-493) and _nb_peel / _nb_union_keys /
 * _nb_dense_labels / _nb_pair_rows (core/kernels_graph.py:187-284). */
int hdfe_components(int64_t n, const int64_t *a_ids, const int64_t *b_ids, int min_deg0,
                    int min_deg1, int leave_one_out, int rank_by, void **out);
/* info (host, 16): levels of each dimension, components, rows, distinct
 * pairs; min_degree's rounds, removed levels of each dimension and removed
 * rows; the leave-one-out step's components before, workers with one row,
 * cut workers, rounds, removed levels of each dimension and removed rows
 * (the reference's diagnostics, core/connected.py:418-456). */
int hdfe_components_info(void *h, int64_t *info);
/* Dimension dim's level ids, sorted, and their components (-1: removed):
 * host arrays of info[dim] entries (ConnectedComponents.levels,
 * core/connected.py:474-480). */
int hdfe_components_levels(void *h, int dim, int64_t *levels, int64_t *comp);
/* Sizes (host, info[2] x 3), largest first: rows and the levels of each
 * dimension (ConnectedComponents.sizes, core/connected.py:481-483). */
int hdfe_components_sizes(void *h, int64_t *sizes);
void hdfe_components_destroy(void *h);

/* ------------------------------------------ KSS standard errors (se.hip)
 * Row passes of the components' standard errors, in the square-root-weight
 * metric; blocks of s vectors are row-major (rows, s); b and s3 are (3, n). */

/* out (a, b) device = A' diag(w) B for tall A (K, a) and B (K, b), device,
 * row strides lda and ldb, w (K) or NULL; fixed-order sums.  No reference counterpart: it replaces the
 * numpy products x1'v, u'v and the Krylov projections of leaveout_se.py and
 * leaveout_weakid.py (e.g. leaveout_se.py:536-540, leaveout_weakid.py:505-512)
 * on the device. */
int hdfe_tall_dot(int64_t K, int a, int b, const double *A, int64_t lda, const double *B,
                  int64_t ldb, const double *w, double *out);

/* hdfe_tall_dot that allocates nothing.  No counterpart in the reference: it
 * is for callers that cannot call hipMalloc (code under HIP graph capture,
 * tight serving loops, harnesses that forbid allocation).  The per-workgroup
 * partials live in a file-scope device array of 1,048,576 doubles (8 MiB) that
 * each GPU the library runs on holds until the process ends; a call whose
 * partials (min(1024, ceil(K / 4096)) * a * b doubles) exceed it returns
 * HDFE_STATUS_ARG.  Launches go to `stream` (NULL: the null stream) and the
 * call does not wait for them; the result is the same as hdfe_tall_dot's, bit
 * for bit.  Every call on one GPU shares that GPU's array: calls that overlap
 * on one GPU race. */
int hdfe_tall_dot_noalloc(int64_t K, int a, int b, const double *A, int64_t lda,
                          const double *B, int64_t ldb, const double *w, double *out,
                          void *stream);

/* The first solve's stacked input of applying C: [v, b_k v] for a shared v
 * (n, s), or v_k = z sqrt(max(s_k, 0)) and [v_k, b_k v_k] per component from
 * draws z (n, s); out (n, (1 or 3) s + 3 s).  Replaces the stacking of
 * _apply_c_passes (leaveout_se.py:515-530). */
int hdfe_se_stack(int64_t n, int s, const double *v, const double *z, const double *s3,
                  const double *b, int mask, double *out);
/* C v at row level from the two solves' read-backs base and thq, reduced in
 * a fixed order (host out): mode 0, v = yt (s = 1): 4 sum s_k (C y)_k^2 and
 * yt'(C y)_k; mode 1, per-component draws: sum_i s_k,i sum_j (C v)_kij^2;
 * modes 2 and 3 the same for the deflated C2 of the q = 1 interval, s floored
 * at zero (mode 2 adds 2 sum x1_k s_k (C2 y)_k).  u, c2, udot, odot: the
 * rank-two term of the reference centring; x1 (3, n) and defl (lambda (3),
 * x1_k'v_k (3, s)): the deflation; mask: the components with draws.
 * Replaces _apply_c_passes' pass 3 with collect_y / _cc_trace's collect and
 * _weak_moments' collect (leaveout_se.py:565-592, 722-731, 843-853, 911-925). */
int hdfe_se_assemble(int64_t n, int s, int mode, const double *v, const double *z,
                     const double *s3, const double *b, const double *w, const double *base,
                     const double *thq, const double *u, double c2, const double *udot,
                     const double *odot, const double *yt, const double *x1, const double *defl,
                     int mask, double *out);
/* out (3, n) += w_i (sum r1^2, sum r2^2, sum r1 r2) of rows (n, 2 s).
 * Replaces _nb_bii_moments (kernels_se.py:75-98). */
int hdfe_se_bii(int64_t n, int s, const double *rows, const double *w, double *out);
/* out (G, s) = per group, sum of w_i Ul[psi_off + psi_code_i]: the
 * covariance's cross sum over the groups.  Replaces the out_groups half of
 * _nb_cov_cross (kernels_se.py:44-72). */
int hdfe_se_cross_groups(int64_t G, int s, const int *gstart, const int *psi_codes,
                         int64_t psi_off, const double *w, const double *Ul, double *out);
/* The smoother's cells: per cell of the nb x nb grid of (P_ii, B_ii)
 * quantile ranks, the sum of sigma2 and the rows (host tot, cnt); with bii
 * and kb null, the nb cells of the leverage alone.  Replaces fit_smoother's
 * binning (leaveout_se.py:1024-1033). */
int hdfe_se_cell_sums(int64_t n, const double *lev, const double *bii, const double *sigma2,
                      const double *kp, const double *kb, int nb, double *tot, double *cnt);
/* The smoothed sigma2, bilinear between cell centres (linear along the
 * leverage with bii and kb null).  Replaces smooth_eval
 * (leaveout_se.py:1038-1054). */
int hdfe_se_smooth(int64_t n, const double *lev, const double *bii, const double *kp,
                   const double *kb, const double *means, int nb, double *out);

/* ------------------------------------------------- adjoint (autograd)
 * Row passes of the backward of a fitted regression: z_i'u for the adjoint
 * solution u = (Z'WZ)^- g of a loss gradient g, and from it dL/dy_i =
 * w_i z_i'u and dL/dw_i = e_i z_i'u.  The reference has no autograd; its
 * adjoint hook is solve_normal_equations / predict (normal_equations.py:
 * 65-231), whose z'u and w_i z_i'u these are. */

/* Per row (device, n each): r_psi, r_alpha = the plug-in components'
 * derivative in the row's psi and alpha effects over w_i, and direct = their
 * derivative in w_i at fixed effects (leaveout.py:807-832).  coef (host, 10)
 * = weights of var(psi), var(alpha), cov in the loss; the weighted means of
 * psi and alpha; the divisor W (R-1)/R; W; the three components. */
int hdfe_adj_component_rows(int64_t G, const int *gstart, const int *psi_codes, int64_t psi_off,
                            const double *gam, const double *geff, const double *coef,
                            double *r_psi, double *r_alpha, double *direct);

/* out (host, kx) = X'W R over the covariate columns xcols (host) of V (n, m)
 * and a row vector R (device). */
int hdfe_adj_xwr(int64_t n, int m, const double *V, const int *xcols, int kx, const double *w,
                 const double *R, double *out);

/* gy, gw (n, device): zu_i = rows_i + Xt_i'u_X (rows may be NULL),
 * gy = w zu, gw = e zu + direct (direct may be NULL); u_X host, kx <= 16. */
int hdfe_adj_gradients(int64_t n, int kx, const double *rows, const double *Xt, const double *uX,
                       const double *w, const double *e, const double *direct, double *gy,
                       double *gw);

/* ------------------------------------------------- Poisson IRLS (GLM)
 * Many models as one stacked panel (levels kept apart per model), rows of a
 * model contiguous in sorted order.  X: covariates (n, kx) float32, row
 * major, sorted order. */

/* One IRLS row pass (_nb_glm_rows / _row_terms, core/kernels_glm.py:99-211):
 * per row the linear predictor at (beta_m, fe, a), or the start with init
 * (Poisson log((y + ybar_m) / 2), Gaussian ybar_m); family 0 (Poisson):
 * w = mu f, z = eta + y / mu - 1; family 1 (Gaussian, identity link: the
 * OLS batch): w = f, z = y; family 2: w = eta alone.  Chunks (host lists:
 * model, row range) of one
 * model each; part (host, nchunk x 5) = deviance (Gaussian: the weighted
 * residual sum of squares), sum f (y eta - mu) (Poisson), f y, f, and with
 * init (Poisson) sum f lgamma(y + 1). */
int hdfe_glm_rows(int nchunk, const int *ch_model, const int64_t *ch_begin,
                  const int64_t *ch_end, int init, int family, const double *ybar, const double *y,
                  const float *X, int kx, const int *g, const int *codes, int64_t ld, int D,
                  const int64_t *offs, const double *fe, const double *a, const double *beta,
                  const double *f, double *w, double *z, double *part_h);

/* rhs (L, kx + 1) = D_o' W M_0 [z, X] (the cell sums' reduced right-hand
 * side of _wls, core/irls.py:100-123), vbar (G, kx + 1) scratch. */
int hdfe_glm_rhs(void *rows_h, int64_t G, const int *gstart, const int *g, const double *z,
                 const float *X, int kx, const double *w, double *vbar, double *rhs);

/* The covariates and working response partialled on the fixed effects
 * (from Gm (L, kx + 1)); sqrt(w) X~ split into bf16 hi + lo, row-major
 * ((padded rows, Cp), model m's padded rows from pstart[m]); and per workgroup
 * (host lists: model, group range) the score X~'W (z~ - X~ beta_m), gpart
 * (host, nwg x kx).  At most four fixed-effect dimensions besides the
 * groups (D <= 4). */
int hdfe_glm_partial(int nwg, const int *wg_model, const int64_t *wg_gb, const int64_t *wg_ge,
                     const int *gstart, const int64_t *rs_h, const int64_t *pstart_h, int nmodels,
                     int Cp, const double *z, const float *X, int kx, const double *w,
                     const int *codes, int64_t ld, int D, const int64_t *offs, const double *Gm,
                     const double *beta, void *hi, void *lo, double *gpart_h);

/* A (nmodels, Cp, Cp, FP64, device) = (hi + lo)'(hi + lo) per model over its
 * padded rows (hi, lo row-major (padded rows, Cp) bf16), as hi'hi + hi'lo +
 * lo'hi on the matrix cores (split-BF16, FP32
 * accumulation per workgroup of rows_per_wg rows, FP64 across them): the
 * dense X'WX of glm.py:990. */
int hdfe_glm_syrk(int nmodels, const int64_t *pstart_h, int Cp, const void *hi, const void *lo,
                  int64_t rows_per_wg, double *A);

/* fe (L) = Gm[:, 0] - Gm[:, 1:] beta_model and the group effects a (G)
 * (_update, core/irls.py:127-144; _nb_group_effects, kernels_glm.py:11-33). */
int hdfe_glm_coefficients(int64_t L, int64_t G, int kx, const double *Gm, const int *lmodel,
                          const int *gmodel, const int *gstart, const double *z, const float *X,
                          const double *w, const int *codes, int64_t ld, int D,
                          const int64_t *offs, const double *beta, double *fe, double *a);

/* out = base + t[model] (target - base) over n entries of `per` values
 * (model NULL: t[i]): step-halving, _State.toward (core/irls.py:44-48). */
int hdfe_glm_toward(int64_t n, const int *model, int per, const double *t, const double *base,
                    const double *target, double *out);

/* ------------------------------------------ ingest (ingest.hip)
 * Host columns to HBM in batches of rows through page-locked slots: the
 * batch loop of the reference's Parquet passes (iter_group_chunks,
 * core/utils.py:26-60, and the chunk loop of _pass1_cells,
 * core/passes.py:397-531), whose batches this build moves to the device
 * instead of working on them on the CPU. */

/* A stager for ncols columns of elem_bytes[c] bytes per element: `depth`
 * page-locked slots of batch_rows rows, `threads` host threads packing them. */
int hdfe_stage_create(int ncols, const int *elem_bytes, int64_t batch_rows, int depth,
                      int threads, void **out);

/* Rows [0, n_rows) of host_cols[c] (ordinary host memory) into dev_cols[c]
 * from row dev_row on, batch by batch (the pack of one batch overlapping the
 * DMA of the previous).  The host columns are free again on return; with
 * `wait` the copies have also landed, else hdfe_stage_wait waits for them.
 * stats (or NULL): seconds in all, seconds packing, bytes, batches. */
int hdfe_stage_columns(void *h, int64_t n_rows, const void *const *host_cols,
                       void *const *dev_cols, int64_t dev_row, int wait, double *stats);

/* Waits for every copy the stager has enqueued. */
int hdfe_stage_wait(void *h);

void hdfe_stage_destroy(void *h);

#ifdef __cplusplus
}
#endif

#endif /* HDFE_HIP_H */
