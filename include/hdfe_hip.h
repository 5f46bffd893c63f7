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
 * int32 (nnz), data float64 (nnz), dinv float64 (L); host or device.
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

/* y = S x (k columns, FP64), host or device.  The operator of
 * _nb_csr_matmat, core/kernels_base.py:361-371. */
int hdfe_csr_matmat(void *h, int64_t k, const double *x, double *y, int on_device);

/* Free the handle and its device memory (the host S release of
 * core/solve.py:222-223 has no counterpart: nothing is kept on the host). */
void hdfe_destroy(void *h);

#ifdef __cplusplus
}
#endif

#endif /* HDFE_HIP_H */
