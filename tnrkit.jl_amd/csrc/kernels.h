// Descriptors and launch wrappers of the HIP kernels (kernels.hip).
#pragma once

#include "common.h"

namespace tnr {

// Sub-block copy / scale job used by the permute kernel.  Jobs of many
// tensors can share one launch (each carries its own pointers).
struct CopyDesc {
    const double* src;
    double* dst;
    const double* w[2];           // weight vectors (already offset), or null
    int64_t sstr[MAXLEG];         // src stride per dst-ordered dim
    int64_t dstr[MAXLEG];         // dst stride per dst-ordered dim
    int32_t dim[MAXLEG];          // extents (dst order)
    int32_t nd;                   // number of dims
    int32_t wleg[2];              // dims (dst order) carrying weights, -1 = none
    int64_t start;                // prefix sum of element counts
};

// Matrix product job: C = alpha * op(A) * op(B) + beta * C, column-major.
struct GemmDesc {
    const double* A;
    const double* B;
    double* C;
    int32_t M, N, K;
    int32_t lda, ldb, ldc;
    int32_t ta, tb;               // 1 = operand stored transposed
    double alpha, beta;
    int32_t tiles_m, tiles_n;
    int32_t ksplit;               // >1: partial products into `work`
    int64_t tile_start;
    double* work;                 // ksplit partials, each M*N (ld = M)
};

// Weighted partial trace sum_{x1,x2} T[x1*s1 + x2*s2] w1[x1] w2[x2] over one
// sub-block; problem `prob` owns descriptors [first, first + count).
struct TraceDesc {
    const double* T;
    const double* w1;             // already offset, or null
    const double* w2;
    int64_t s1, s2;
    int32_t d1, d2;
    int64_t start;                // prefix within its problem
};
struct TraceProb {
    int32_t first, count;
    int64_t total;                // elements of the problem
    double* T;                    // whole tensor (for the normalise pass)
    int64_t size;
    double* partial;              // TRACE_PARTS partial sums
    double* norm_slot;
};
constexpr int TRACE_PARTS = 64;

// Symmetric n x n eigenproblem (n <= LMAX).
struct JacobiDesc {
    const double* G;
    double* W;
    double* lam;
    int* sweeps;
    int32_t n, ldg, ldw;
};

// Column-norm / normalise job (also the Frobenius-norm job).
struct ColDesc {
    const double* Y;
    double* U;
    double* sigma;
    int32_t nrow, ncol, ldy, ldu;
};

// Deterministic uniform(-1,1) fill of a rows x cols column-major matrix.
struct FillDesc {
    double* p;
    int32_t rows, cols, ld;
    uint64_t seed;
    int64_t start;
};

// out = (trans ? in^T : in) scaled by row weights rw and column weights cw.
struct GatherDesc {
    const double* in;
    double* out;
    const double* rw;
    const double* cw;
    int32_t rows, cols, ldi, ldo, trans;
    int64_t start;
};

void launch_copy(hipStream_t st, const CopyDesc* d_descs, int nd, int64_t total);
void launch_gemm(hipStream_t st, const GemmDesc* d_descs, int nd, int64_t tiles, bool split32);
void launch_gemm_reduce(hipStream_t st, const GemmDesc* d_descs, int nd, const int64_t* d_starts,
                        int64_t total);
void launch_trace_normalize(hipStream_t st, const TraceDesc* d_descs, const TraceProb* d_probs,
                            int nprob, int64_t max_size);
size_t jacobi_lds_bytes(int n);
void launch_jacobi(hipStream_t st, const JacobiDesc* d_descs, int nd, int nmax, double tol,
                   int max_sweeps);
void launch_colnorm(hipStream_t st, const ColDesc* d_descs, int nd, double rtol);
void launch_symmetrize(hipStream_t st, const ColDesc* d_descs, int nd, int nmax);
void launch_frob(hipStream_t st, const ColDesc* d_descs, int nd);
void launch_randfill(hipStream_t st, const FillDesc* d_descs, int nd, int64_t total);
void launch_gather(hipStream_t st, const GatherDesc* d_descs, int nd, int64_t total);

// GEMM tile geometry shared with the host planner.
constexpr int GEMM_BM = 64, GEMM_BN = 64, GEMM_BK = 16;

}  // namespace tnr
