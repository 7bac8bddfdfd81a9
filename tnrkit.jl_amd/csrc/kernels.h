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
    int64_t start;                // prefix sum of 4096-element chunks
};

// Tiled transpose job: a sub-block whose source-contiguous dim (extent n0,
// source stride 1, destination stride ds0) differs from its
// destination-contiguous dim (extent n1, destination stride 1, source
// stride ss1); up to TR_BATCH further batch dims.  Handled in 64 x 64 LDS
// tiles.
constexpr int TR_BATCH = MAXLEG - 2;
struct TransDesc {
    const double* src;
    double* dst;
    int64_t ds0, ss1;
    int64_t bs[TR_BATCH], bd[TR_BATCH];   // batch strides (src, dst)
    int32_t bdim[TR_BATCH];
    int32_t n0, n1, nb;           // nb = product of batch extents
    int32_t tx, ty, sub;          // sub-tile extents (powers of two <= 64), sub-tiles per tile
    int32_t t0, t1;               // tiles along n0, n1
    int64_t tile_start;
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
    // K is summed in chunks of kchunk (a function of K alone): each chunk
    // is accumulated from zero and the chunk sums are added in order, so
    // ksplit = 1 (one workgroup runs all chunks) and ksplit = #chunks (one
    // chunk per workgroup, partials added in order by k_gemm_reduce) give
    // identical bits
    int32_t kchunk;
    int32_t ksplit;               // 1 or #chunks: partial products into `work`
    int64_t tile_start;
    double* work;                 // ksplit partials, each M*N (ld = M)
    const int* skip;              // optional: problem finished, do nothing
};

// Weighted partial trace sum_{x1,x2} T[x1*s1 + x2*s2] w1[x1] w2[x2] over one
// sub-block; problem `prob` owns descriptors [first, first + count).
// sum_{x1, x2, x3} T[x1 s1 + x2 s2 + x3 s3] w1[x1] w2[x2] over one sub-block
// (d3 = 1 for the 2D traces, 3 index pairs for the 3D one)
struct TraceDesc {
    const double* T;
    const double* w1;             // already offset, or null
    const double* w2;
    int64_t s1, s2, s3;
    int32_t d1, d2, d3;
    int64_t start;                // prefix within its problem
};
struct TraceProb {
    int32_t first, count;
    int64_t total;                // elements of the problem
    double* T;                    // whole tensor (for the normalise pass)
    int64_t size;
    double* partial;              // TRACE_PARTS partial sums
    double* norm_slot;
    const double* divide_by;      // null: T /= |trace|; else T /= *divide_by (written earlier)
};

// out[e] = sum_i coef[i] * in[i][e], added left to right (in-place allowed)
struct LinDesc {
    double* out;
    const double* in[4];
    double coef[4];
    int32_t nin;
    int64_t n;
    int64_t start;
};
void launch_lincomb(hipStream_t st, const LinDesc* d_descs, int nd, int64_t total);
constexpr int TRACE_PARTS = 64;

// Symmetric n x n eigenproblem (n <= LMAX).
struct JacobiDesc {
    const double* G;
    double* W;
    double* lam;
    int* sweeps;
    int32_t n, ldg, ldw;
    const int* skip;              // optional
    double* fail = nullptr;       // optional: set to 1 when the sweep cap is
                                  // reached with a rotation still due
    int32_t natural = 0;          // 1: leave eigenpair i in position i (the
                                  // accumulated rotation, close to the
                                  // identity for a nearly diagonal G) instead
                                  // of ordering by |lambda| descending
};

// Column-norm / normalise job (also the Frobenius-norm job).
struct ColDesc {
    const double* Y;
    double* U;
    double* sigma;
    int32_t nrow, ncol, ldy, ldu;
    const int* skip;              // optional
};

// Pivoted-row gauge of one truncated SVD problem (k_canon): Q (k x k,
// column-major, ld k) rotates the kept left vectors U (m x k) inside each
// cluster of equal singular values, cluster c spanning columns
// [cl[c], cl[c+1]) with cl[0] = 0 and cl[nc] = knn; columns from knn on are
// null and get Q = I.
struct CanonDesc {
    const double* U;
    double* Q;
    const int* cl;
    int32_t m, ldu, k, knn, nc;
};

// Grouped dot product: *out = sum_i x[i] y[i] over n elements.  DOT_PARTS
// workgroups per problem each sum one fixed contiguous chunk into part[],
// then the parts are added in order (the result does not depend on timing).
constexpr int DOT_PARTS = 256;
struct DotDesc {
    const double* x;
    const double* y;
    double* out;
    double* part;      // DOT_PARTS doubles of scratch
    int64_t n;
};

// All eigenvalues of a real n x n matrix, k_eig_hqr (n <= EIG_NMAX, the
// matrix in LDS) or k_eig_hqr_g (n <= EIG_NMAX_G, in `work`,
// eig_work_doubles(n) doubles of device memory): wr/wi receive the
// real/imaginary parts (conjugate pairs adjacent), *info 0 or 1 when the QR
// iteration did not converge.
constexpr int EIG_NMAX = 128;
constexpr int EIG_NMAX_G = 512;
inline int64_t eig_work_doubles(int n) { return (int64_t)n * (n + 1) + n; }
struct EigDesc {
    const double* A;
    double* wr;
    double* wi;
    int* info;
    int32_t n, lda;
    double* work;      // n > EIG_NMAX only
};

// Device-side stopping rule of the block iteration (one thread per problem):
// with sigma the Ritz values of the current basis (descending), kk the
// number of the first kmax above null_rtol * sigma_0 and
// r = min(1, sigma_{l-1} / sigma_{kk-1})^(squared ? 2 : 1), iteration `it`
// multiplies prod by r once it >= min_iters and sets done when
// prod < target or it >= max_iters.
struct ConvDesc {
    const double* sigma;
    int32_t l, kmax, squared;
    int* it;
    double* prod;
    int* done;
    int32_t whole_ok;             // a stall is refactorised whole: give up early
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

// start = prefix of 4096-element chunks; every job holds < 2^31 elements
void launch_copy(hipStream_t st, const CopyDesc* d_descs, int nd, int64_t chunks, int64_t elems);
void launch_transpose(hipStream_t st, const TransDesc* d_descs, int nd, int64_t tiles,
                      int64_t elems);
// tile shapes of the FP64 GEMM kernel (the split-FP32 kernel is 64 x 64)
// FP64 tile shapes: workgroup tile BM x BN, waves, k-tile depth
//   GEMM_64x64       64x64,  4 waves of 32x32, k 16 (LDS-DMA, 3 stages)
//   GEMM_128x64     128x64,  8 waves of 32x32, k 16 (LDS-DMA, 3 stages)
//   GEMM_128x128   128x128,  8 waves of 32x64, k 16 (LDS-DMA, 3 stages)
//   GEMM_64x64_K32   64x64,  4 waves of 32x32, k 32 (register-staged loads)
enum GemmShape { GEMM_64x64 = 0, GEMM_128x64 = 1, GEMM_128x128 = 2, GEMM_64x64_K32 = 3 };
constexpr int GEMM_NSHAPES = 4;
inline int gemm_shape_bm(int s) { return (s == GEMM_64x64 || s == GEMM_64x64_K32) ? 64 : 128; }
inline int gemm_shape_bn(int s) { return s == GEMM_128x128 ? 128 : 64; }
// dma: the 16-deep shapes load through LDS-DMA (k_gemm_dl, same bits); needs
// every descriptor's per-tile offsets below 2^31 bytes (gemm_dma_ok)
void launch_gemm(hipStream_t st, const GemmDesc* d_descs, int nd, int64_t tiles, bool split32,
                 double flops, int shape, bool dma);
inline bool gemm_dma_ok(int ta, int64_t lda, int tb, int64_t ldb) {
    const int64_t lim = (int64_t)1 << 31;
    const int64_t la = ta ? 8 * (16 + 128 * lda) : 8 * (128 + 16 * lda);
    const int64_t lb = tb ? 8 * (128 + 16 * ldb) : 8 * (16 + 128 * ldb);
    return la < lim && lb < lim;
}
inline bool gemm_dma_ok(const GemmDesc& g) { return gemm_dma_ok(g.ta, g.lda, g.tb, g.ldb); }
void launch_gemm_reduce(hipStream_t st, const GemmDesc* d_descs, int nd, const int64_t* d_starts,
                        int64_t total);
void launch_trace_normalize(hipStream_t st, const TraceDesc* d_descs, const TraceProb* d_probs,
                            int nprob, int64_t max_size);
size_t jacobi_lds_bytes(int n);
// pivoted Cholesky QR of Gram matrices (JacobiDesc: G in, W = P R^{-1} out,
// lam = R_jj); columns below null_rel2 * max diag are dropped
void launch_cholqr(hipStream_t st, const JacobiDesc* d_descs, int nd, int nmin, int nmax,
                   double null_rel2);
void launch_jacobi(hipStream_t st, const JacobiDesc* d_descs, int nd, int nmax, double tol,
                   int max_sweeps);
void launch_colnorm(hipStream_t st, const ColDesc* d_descs, int nd, int ncol_max, double rtol);
void launch_symmetrize(hipStream_t st, const ColDesc* d_descs, int nd, int nmax);
// stall_from > 0: from that iteration on, a problem with whole_ok set stops
// (unconverged) once prod * r^(2 (max_iters - it)) >= target, i.e. even at twice
// its current rate it could not reach the target in the iterations left
void launch_conv(hipStream_t st, const ConvDesc* d_descs, int nd, double null_rtol, double target,
                 int min_iters, int max_iters, int stall_from);
// sigma[0] of each descriptor = squared Frobenius norm of its nrow x ncol
// matrix Y (ld ldy), summed in a fixed order over FROB_PARTS column ranges;
// partial: nd * FROB_PARTS doubles of scratch
constexpr int FROB_PARTS = 64;
void launch_frob(hipStream_t st, const ColDesc* d_descs, int nd, double* partial);
// kmax: the largest k of the problems (sizes the LDS); tie: relative pivot tie
void launch_canon(hipStream_t st, const CanonDesc* d_descs, int nd, int kmax, double tie);
void launch_dot(hipStream_t st, const DotDesc* d_descs, int nd);
// nmax: the largest n of the problems (sizes the LDS)
void launch_eig(hipStream_t st, const EigDesc* d_descs, int nd, int nmin, int nmax);
void launch_randfill(hipStream_t st, const FillDesc* d_descs, int nd, int64_t total);
void launch_gather(hipStream_t st, const GatherDesc* d_descs, int nd, int64_t total);

// ------------------------------------------------------------- profiling
// Optional per-kernel-class device timing (HIP events on the launching
// stream) with the algorithmic flops / bytes each launch was asked to do.
enum ProfKind {
    PK_GEMM_F64 = 0, PK_GEMM_SPLIT32, PK_GEMM_REDUCE, PK_JACOBI, PK_COPY, PK_COLNORM,
    PK_GATHER, PK_TRACE, PK_FILL, PK_MISC, PK_COUNT
};
// level 0 off; 1 device time per launch, no host synchronisation; 2 also
// the work of launches whose work is only known on the device (the Jacobi
// sweeps), read back after each such launch
void prof_enable(int level);
bool prof_enabled();
int prof_level();
struct ProfScope {
    ProfScope(hipStream_t st, int kind, double flops, double bytes);
    ~ProfScope();
    void bracket();                   // time the kernels launched from here on by events
                                      // recorded around them (a scope of several kernels)
    void end();                       // record the closing event of a bracket now
    hipStream_t st;
    int kind;
    bool active;
    bool ended = false;
    int mode = 0;                     // 0 nothing launched, 1 bracket, 2 dispatch timestamps
    double flops, bytes;
    hipEvent_t ea = nullptr, eb = nullptr;
};

// Launch one kernel inside a profiling scope.  While profiling, a lone kernel
// is timed by its dispatch's own start and end timestamps (the events of
// hipExtLaunchKernelGGL), so no marker packets go between launches; a
// bracketed scope launches plainly.
#define TNR_LAUNCH(ps, kernel, grid, block, shm, st, ...)                                   \
    do {                                                                                    \
        if ((ps).active && (ps).mode == 0) {                                                \
            hipExtLaunchKernelGGL(kernel, grid, block, shm, st, (ps).ea, (ps).eb, 0u,       \
                                  __VA_ARGS__);                                             \
            (ps).mode = 2;                                                                  \
            (ps).ended = true;                                                              \
        } else {                                                                            \
            hipLaunchKernelGGL(kernel, grid, block, shm, st, __VA_ARGS__);                  \
        }                                                                                   \
    } while (0)
// ms: summed launch durations; busy_ms: the union of the launch intervals
// (launches on different streams overlap); flops: GEMM flops of the problems
// that ran (device-counted), flops_planned: what the launches were asked to do
struct ProfTotals {
    int64_t launches = 0;
    double ms = 0, busy_ms = 0, flops = 0, flops_planned = 0, bytes = 0;
};
// device counter of the GEMM flops that run while profiling is on (FP64 or
// split), nullptr when it is off
unsigned long long* prof_flop_counter(bool split32);
ProfTotals prof_read(int kind);   // synchronises the device

// GEMM k-tile depth shared with the host planner.
constexpr int GEMM_BK = 16;

}  // namespace tnr
