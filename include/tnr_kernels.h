/*
 * tnr_kernels.h - kernel-level entry points of libtnr_hip.so.
 *
 * These expose the engine's building blocks on plain host buffers (each call
 * copies in, runs on the device, copies out) for unit tests, autotuning and
 * microbenchmarks.  They have no counterpart in the reference: TNRKit.jl
 * reaches the same operations through TensorKit / MatrixAlgebraKit / BLAS
 * (`@tensor` contractions, `svd_trunc`, `eigh_trunc!`).
 */
#ifndef TNR_KERNELS_H
#define TNR_KERNELS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* C = alpha op(A) op(B) + beta C, column-major FP64; split32 != 0 uses the
 * FP32-MFMA hi/lo split kernel.  repeat >= 1 launches the product that many
 * times on device-resident copies; *ms_out (optional) receives the mean
 * device time per launch from HIP events. */
int tnr_kernel_gemm(int M, int N, int K, const double* A, int lda, int ta,
                    const double* B, int ldb, int tb, double* C, int ldc,
                    double alpha, double beta, int split32, int repeat,
                    double* ms_out);

/* Top-k singular triplets of the m x n matrix M (column-major, ld = m):
 * U (m x k), s (k), V (n x k), as the engine's truncated SVD computes them.
 * *iters receives the block-iteration count (0 = direct solve). */
int tnr_kernel_svd(int m, int n, const double* M, int k, double* U,
                   double* s, double* V, int* iters);

/* Largest-|lambda| k eigenpairs of the symmetrised n x n matrix M:
 * U (n x k), lam (k). */
int tnr_kernel_eigh(int n, const double* M, int k, double* U, double* lam,
                    int* iters);

/* Eigen-decomposition of `count` symmetric n x n matrices (n <= 96) by the
 * grouped parallel-order Jacobi kernel: W (n x n each), lam (n each), by
 * |lambda| descending; *sweeps receives the largest sweep count. */
int tnr_kernel_jacobi(int count, int n, const double* G, double* W,
                      double* lam, int* sweeps);

#ifdef __cplusplus
}
#endif
#endif /* TNR_KERNELS_H */
