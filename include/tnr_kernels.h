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

/* C = alpha op(A) op(B) + beta C, column-major FP64.  mode & 7: 0 = FP64
 * MFMA with the tile shape the engine would choose, 1 = FP32-MFMA hi/lo
 * split kernel, 2 / 3 / 4 / 5 = FP64 MFMA with 64x64 / 128x64 / 128x128 /
 * 64x64 (32-deep k-tiles) tiles;
 * mode | 8 never splits K over workgroups, mode | 16 always does (for K of
 * two or more 256-deep chunks); the default decides by launch size;
 * mode | 32 loads the 16-deep tiles through registers instead of LDS-DMA
 * (the same bits).
 * repeat >= 1 launches the product that many times on device-resident
 * copies; *ms_out (optional) receives the device time of the last launch
 * from HIP events. */
int tnr_kernel_gemm(int M, int N, int K, const double* A, int lda, int ta,
                    const double* B, int ldb, int tb, double* C, int ldc,
                    double alpha, double beta, int mode, int repeat,
                    double* ms_out);

/* Top-k singular triplets of the m x n matrix M (column-major, ld = m):
 * U (m x k), s (k), V (n x k), as the engine's truncated SVD computes them.
 * *iters receives the block-iteration count (0 = direct solve). */
int tnr_kernel_svd(int m, int n, const double* M, int k, double* U,
                   double* s, double* V, int* iters);

/* Pivoted-row gauge of k kept left singular vectors U (m x k, column-major,
 * ld = m) with singular values s (descending): Q (k x k, column-major) such
 * that U Q (and V Q on the right side) is the basis the impurity TRG's SVDs
 * use.  Singular values within 1e-9 * s[0] of their neighbour form a
 * cluster, values at or below 1e-14 * s[0] are null (Q = I there); inside a
 * cluster column t of U Q is the normalised projection of the coordinate
 * vector e_r onto the part of the span orthogonal to columns 0..t-1, r the
 * first row whose residual norm is within 1e-10 of the largest.  No
 * counterpart in the reference, which keeps LAPACK's basis; k <= 96. */
int tnr_kernel_pivoted_gauge(int m, int k, const double* U, const double* s, double* Q);

/* All eigenvalues of `count` real n x n matrices A (column-major, ld = n,
 * stored one after another; n <= 512): wr/wi (count x n) receive real and
 * imaginary parts, conjugate pairs adjacent, in the solver's deflation order.
 * Balancing, Householder reduction to Hessenberg form and Francis
 * double-shift QR, one wavefront per matrix; the engine of cft_data and
 * ground_state_degeneracy (eig_full in src/utility/cft.jl:16). */
int tnr_kernel_eig(int count, int n, const double* A, double* wr, double* wi);

/* Largest-|lambda| k eigenpairs of the symmetrised n x n matrix M:
 * U (n x k), lam (k). */
int tnr_kernel_eigh(int n, const double* M, int k, double* U, double* lam,
                    int* iters);

/* Eigen-decomposition of `count` symmetric n x n matrices (n <= 96) by the
 * grouped parallel-order Jacobi kernel: W (n x n each), lam (n each), by
 * |lambda| descending; sweeps (count ints, optional) receives each sweep
 * count and *ms_out (optional) the device time of the launch. */
int tnr_kernel_jacobi(int count, int n, const double* G, double* W,
                      double* lam, int* sweeps, double* ms_out);

/* The eigensolver of tnr_kernel_jacobi on device buffers, allocating
 * nothing: no counterpart in the reference; for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  G, W (count x n x n, each matrix
 * column-major, ld = n, one after another) and lam (count x n) are device
 * pointers; sweeps (count ints, optional) is host memory; stream is a
 * hipStream_t (NULL = the default stream), synchronised before return.  The
 * results equal tnr_kernel_jacobi's.  The solver descriptors and sweep
 * counts live in a 1 MiB array each GPU holds for the life of the process:
 * count is limited to what it holds (TNR_ERR_ARG beyond, naming the limit:
 * 13797 problems at 76 bytes each), and calls that overlap on one GPU race. */
int tnr_kernel_jacobi_noalloc(int count, int n, const double* G, double* W,
                              double* lam, int* sweeps, void* stream);
/* Device-time accounting per kernel class.  A launch of one kernel is timed
 * by its dispatch's start and end timestamps (the events of
 * hipExtLaunchKernelGGL, no marker packets between launches); a class call
 * of two kernels by events recorded around both.  tnr_profile_enable(1)
 * synchronises the device, clears the totals and starts recording (no host
 * synchronisation while recording), (2) also reads back after each Jacobi
 * launch the sweeps it took (its flops), (0) stops it.  Classes: 0 FP64
 * GEMM, 1 split-FP32 GEMM, 2 split-K reduce, 3 Jacobi, 4 permute/copy,
 * 5 column norms, 6 gather, 7 trace/normalise, 8 random fill, 9 other.
 * tnr_profile_read: *ms sums the launch durations, *busy_ms is the union of
 * the launch intervals (launches of one class on different streams
 * overlap, so busy_ms <= ms and <= the wall time).  *flops / *bytes are the
 * algorithmic work: GEMM flops 2MNK are counted on the device per output
 * tile, for the problems that actually ran (a problem the solver has
 * finished skips its tiles), *flops_planned is what the launches were asked
 * to do; copies count bytes read + written. */
int tnr_profile_enable(int on);
/* GEMM tile-shape autotuner: on = 1 (TNR_AUTOTUNE=1 in the environment
 * starts it on) times each tile shape on the first launches of every launch
 * signature and then keeps the fastest; on = 0 (the default) uses the
 * static rule.  Every call forgets what was learned.  Tile shapes never change
 * results: each output element is summed in the same order under all. */
int tnr_autotune_enable(int on);
/* Solver limits and thresholds of the truncated SVD / eigensolver: the
 * defaults, then the TNR_SOLVER_CFG environment variable, then `spec`
 * ("key=value,..."; NULL or "" = defaults and environment only).  Keys:
 * max_iters (block iterations, default 80), stall_from (8: from that
 * iteration on, a block iteration of a matrix whose smaller side is at most
 * 384 stops as soon as even twice its current convergence rate could not
 * reach conv_target by max_iters; 0 = iterate to max_iters; such a stall,
 * like one at max_iters, is factorised whole), sweeps / sweeps_iter (Jacobi
 * sweep caps of final and in-loop solves, 40), tol_iter (in-loop Jacobi
 * threshold, 1e-10), tol_first (the first iteration's, 1e-4),
 * conv_target (1e-14), block_mult (2), block_extra (2),
 * cholqr_y (1), cholqr_z (1), cholqr_eigh (0) (block iteration steps by
 * pivoted Cholesky QR instead of Rayleigh-Ritz), rv_passes (2), rf_passes (1)
 * (Jacobi passes of the final extraction).  A solve of a larger matrix that
 * reaches max_iters, or one that reaches a sweep cap, unconverged fails its
 * step with TNR_ERR_NOCONV.  Not thread-safe: call it
 * while no other entry point runs. */
int tnr_solver_config(const char* spec);
/* hipDeviceSynchronize() on the current device (timing brackets). */
int tnr_device_sync(void);
int tnr_profile_read(int kind, int64_t* launches, double* ms, double* flops,
                     double* bytes, double* busy_ms, double* flops_planned);

#ifdef __cplusplus
}
#endif
#endif /* TNR_KERNELS_H */
