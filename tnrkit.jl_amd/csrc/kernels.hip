// HIP kernels for gfx950 (CDNA4): grouped sub-block permutes, grouped FP64
// MFMA GEMM (and FP32-MFMA hi/lo split variant), trace/normalise, leg
// scaling, grouped parallel-order Jacobi eigensolver for small symmetric
// matrices, column norms, deterministic pseudo-random fill.
//
// Every kernel here is launched with its grid checked on the host against
// the descriptors it reads (see kernels.h); no kernel uses atomics, so every
// reduction has a fixed order and results are bitwise reproducible run to
// run and alone vs batched.
#include "kernels.h"

namespace tnr {

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ int find_desc_copy(const CopyDesc* d, int n, int64_t e) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (d[mid].start <= e) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// ------------------------------------------------------------------ permute
// dst[sum i_k dstr_k] = src[sum i_k sstr_k] * w0[i_wleg0] * w1[i_wleg1]
__global__ __launch_bounds__(256) void k_copy(const CopyDesc* __restrict__ descs, int nd,
                                              int64_t total) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const CopyDesc& D = descs[find_desc_copy(descs, nd, e)];
        int64_t r = e - D.start;
        int64_t so = 0, dof = 0;
        int idx[MAXLEG];
#pragma unroll
        for (int k = 0; k < MAXLEG; ++k) {
            if (k < D.nd) {
                int64_t ik = r % D.dim[k];
                r /= D.dim[k];
                idx[k] = (int)ik;
                so += ik * D.sstr[k];
                dof += ik * D.dstr[k];
            } else {
                idx[k] = 0;
            }
        }
        double v = D.src[so];
        if (D.wleg[0] >= 0) v *= D.w[0][idx[D.wleg[0]]];
        if (D.wleg[1] >= 0) v *= D.w[1][idx[D.wleg[1]]];
        D.dst[dof] = v;
    }
}

// --------------------------------------------------------------- FP64 GEMM
// 64x64 output tile per 256-thread workgroup; 4 waves as 2x2 of 32x32; each
// wave issues 2x2 v_mfma_f64_16x16x4_f64 per k-step of 4.  Operand lane
// layout of the 16x16x4 MFMAs (measured on gfx950): A[i=l%16][k=l/16],
// B[k=l/16][j=l%16]; result register r of lane l holds
//   f64: D[i = l/16 + 4r][j = l%16]   (rows interleaved)
//   f32: D[i = 4(l/16) + r][j = l%16] (rows contiguous).
constexpr int BM = 64, BN = 64, BK = 16;

__device__ __forceinline__ int find_desc_gemm(const GemmDesc* d, int n, int64_t t) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (d[mid].tile_start <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}

struct TileCoord {
    int m0, n0, kbeg, kend, kz;
};

__device__ __forceinline__ TileCoord tile_coord(const GemmDesc& g, int64_t local) {
    TileCoord c;
    int tm = (int)(local % g.tiles_m);
    int64_t rest = local / g.tiles_m;
    int tn = (int)(rest % g.tiles_n);
    int kz = (int)(rest / g.tiles_n);
    c.m0 = tm * BM;
    c.n0 = tn * BN;
    int ktiles = (g.K + BK - 1) / BK;
    int per = (ktiles + g.ksplit - 1) / g.ksplit;
    c.kbeg = kz * per * BK;
    c.kend = min(g.K, (kz + 1) * per * BK);
    c.kz = kz;
    return c;
}

__device__ __forceinline__ void load_a(const GemmDesc& g, int m0, int k0, int t, double (&ra)[4]) {
    if (!g.ta) {  // A[i + k*lda], M x K
        int i = m0 + (t & 63), kk = t >> 6;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int k = k0 + kk + 4 * r;
            ra[r] = (i < g.M && k < g.K) ? g.A[(int64_t)i + (int64_t)k * g.lda] : 0.0;
        }
    } else {      // stored K x M: A[k + i*lda]
        int kk = t & 15, i0 = t >> 4;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int i = m0 + i0 + 16 * r, k = k0 + kk;
            ra[r] = (i < g.M && k < g.K) ? g.A[(int64_t)k + (int64_t)i * g.lda] : 0.0;
        }
    }
}

__device__ __forceinline__ void store_a(const GemmDesc& g, int t, const double (&ra)[4],
                                        double (*As)[BM + 1]) {
    if (!g.ta) {
        int i = t & 63, kk = t >> 6;
#pragma unroll
        for (int r = 0; r < 4; ++r) As[kk + 4 * r][i] = ra[r];
    } else {
        int kk = t & 15, i0 = t >> 4;
#pragma unroll
        for (int r = 0; r < 4; ++r) As[kk][i0 + 16 * r] = ra[r];
    }
}

__device__ __forceinline__ void load_b(const GemmDesc& g, int n0, int k0, int t, double (&rb)[4]) {
    if (!g.tb) {  // B[k + j*ldb], K x N
        int kk = t & 15, j0 = t >> 4;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int j = n0 + j0 + 16 * r, k = k0 + kk;
            rb[r] = (j < g.N && k < g.K) ? g.B[(int64_t)k + (int64_t)j * g.ldb] : 0.0;
        }
    } else {      // stored N x K: B[j + k*ldb]
        int j = n0 + (t & 63), kk = t >> 6;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int k = k0 + kk + 4 * r;
            rb[r] = (j < g.N && k < g.K) ? g.B[(int64_t)j + (int64_t)k * g.ldb] : 0.0;
        }
    }
}

__device__ __forceinline__ void store_b(const GemmDesc& g, int t, const double (&rb)[4],
                                        double (*Bs)[BN + 1]) {
    if (!g.tb) {
        int kk = t & 15, j0 = t >> 4;
#pragma unroll
        for (int r = 0; r < 4; ++r) Bs[kk][j0 + 16 * r] = rb[r];
    } else {
        int j = t & 63, kk = t >> 6;
#pragma unroll
        for (int r = 0; r < 4; ++r) Bs[kk + 4 * r][j] = rb[r];
    }
}

template <bool F64_LAYOUT>
__device__ __forceinline__ void write_tile(const GemmDesc& g, const TileCoord& tc, int lane,
                                           int wm, int wn, const dbl4 (&acc)[2][2]) {
    const bool part = g.ksplit > 1;
    double* out = part ? g.work + (int64_t)tc.kz * g.M * g.N : g.C;
    const int64_t ld = part ? g.M : g.ldc;
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni) {
            int col = tc.n0 + wn * 32 + ni * 16 + (lane & 15);
            if (col >= g.N) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int ri = F64_LAYOUT ? (lane >> 4) + 4 * r : (lane >> 4) * 4 + r;
                int row = tc.m0 + wm * 32 + mi * 16 + ri;
                if (row >= g.M) continue;
                double v = acc[mi][ni][r];
                double* p = out + (int64_t)row + (int64_t)col * ld;
                if (part) *p = v;
                else *p = (g.beta == 0.0) ? g.alpha * v : g.alpha * v + g.beta * *p;
            }
        }
}

__global__ __launch_bounds__(256) void k_gemm_f64(const GemmDesc* __restrict__ descs, int nd) {
    __shared__ double As[BK][BM + 1];
    __shared__ double Bs[BK][BN + 1];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, wm = w & 1, wn = w >> 1;
    const GemmDesc g = descs[find_desc_gemm(descs, nd, blockIdx.x)];
    const TileCoord tc = tile_coord(g, (int64_t)blockIdx.x - g.tile_start);
    dbl4 acc[2][2];
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni) acc[mi][ni] = dbl4{0.0, 0.0, 0.0, 0.0};
    double ra[4], rb[4];
    if (tc.kbeg < tc.kend) {
        load_a(g, tc.m0, tc.kbeg, t, ra);
        load_b(g, tc.n0, tc.kbeg, t, rb);
    }
    for (int k0 = tc.kbeg; k0 < tc.kend; k0 += BK) {
        store_a(g, t, ra, As);
        store_b(g, t, rb, Bs);
        __syncthreads();
        if (k0 + BK < tc.kend) {
            load_a(g, tc.m0, k0 + BK, t, ra);
            load_b(g, tc.n0, k0 + BK, t, rb);
        }
#pragma unroll
        for (int ks = 0; ks < BK / 4; ++ks) {
            const int k = ks * 4 + (lane >> 4);
            double a0 = As[k][wm * 32 + (lane & 15)];
            double a1 = As[k][wm * 32 + 16 + (lane & 15)];
            double b0 = Bs[k][wn * 32 + (lane & 15)];
            double b1 = Bs[k][wn * 32 + 16 + (lane & 15)];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
        }
        __syncthreads();
    }
    write_tile<true>(g, tc, lane, wm, wn, acc);
}

// ------------------------------------------------ FP32 MFMA, hi/lo split
// x = hi + lo with hi = fp32(x), lo = fp32(x - hi).  Per k-tile the three
// significant products hi*hi, hi*lo, lo*hi are accumulated on FP32 MFMA
// (v_mfma_f32_16x16x4_f32) in separate accumulators, then flushed into FP64
// registers, so FP32 rounding only acts within one 16-deep k-tile.
__global__ __launch_bounds__(256) void k_gemm_split32(const GemmDesc* __restrict__ descs, int nd) {
    __shared__ float Ah[BK][BM + 1], Al[BK][BM + 1];
    __shared__ float Bh[BK][BN + 1], Bl[BK][BN + 1];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, wm = w & 1, wn = w >> 1;
    const GemmDesc g = descs[find_desc_gemm(descs, nd, blockIdx.x)];
    const TileCoord tc = tile_coord(g, (int64_t)blockIdx.x - g.tile_start);
    dbl4 acc[2][2];
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni) acc[mi][ni] = dbl4{0.0, 0.0, 0.0, 0.0};
    double ra[4], rb[4];
    for (int k0 = tc.kbeg; k0 < tc.kend; k0 += BK) {
        load_a(g, tc.m0, k0, t, ra);
        load_b(g, tc.n0, k0, t, rb);
        // split and stage
        {
            float ah[4], al[4], bh[4], bl[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                ah[r] = (float)ra[r]; al[r] = (float)(ra[r] - (double)ah[r]);
                bh[r] = (float)rb[r]; bl[r] = (float)(rb[r] - (double)bh[r]);
            }
            if (!g.ta) {
                int i = t & 63, kk = t >> 6;
#pragma unroll
                for (int r = 0; r < 4; ++r) { Ah[kk + 4 * r][i] = ah[r]; Al[kk + 4 * r][i] = al[r]; }
            } else {
                int kk = t & 15, i0 = t >> 4;
#pragma unroll
                for (int r = 0; r < 4; ++r) { Ah[kk][i0 + 16 * r] = ah[r]; Al[kk][i0 + 16 * r] = al[r]; }
            }
            if (!g.tb) {
                int kk = t & 15, j0 = t >> 4;
#pragma unroll
                for (int r = 0; r < 4; ++r) { Bh[kk][j0 + 16 * r] = bh[r]; Bl[kk][j0 + 16 * r] = bl[r]; }
            } else {
                int j = t & 63, kk = t >> 6;
#pragma unroll
                for (int r = 0; r < 4; ++r) { Bh[kk + 4 * r][j] = bh[r]; Bl[kk + 4 * r][j] = bl[r]; }
            }
        }
        __syncthreads();
        flt4 hh[2][2], cr[2][2];
#pragma unroll
        for (int mi = 0; mi < 2; ++mi)
#pragma unroll
            for (int ni = 0; ni < 2; ++ni) { hh[mi][ni] = flt4{0, 0, 0, 0}; cr[mi][ni] = flt4{0, 0, 0, 0}; }
#pragma unroll
        for (int ks = 0; ks < BK / 4; ++ks) {
            const int k = ks * 4 + (lane >> 4);
            float a0h = Ah[k][wm * 32 + (lane & 15)], a1h = Ah[k][wm * 32 + 16 + (lane & 15)];
            float a0l = Al[k][wm * 32 + (lane & 15)], a1l = Al[k][wm * 32 + 16 + (lane & 15)];
            float b0h = Bh[k][wn * 32 + (lane & 15)], b1h = Bh[k][wn * 32 + 16 + (lane & 15)];
            float b0l = Bl[k][wn * 32 + (lane & 15)], b1l = Bl[k][wn * 32 + 16 + (lane & 15)];
            float ah_[2] = {a0h, a1h}, al_[2] = {a0l, a1l}, bh_[2] = {b0h, b1h}, bl_[2] = {b0l, b1l};
#pragma unroll
            for (int mi = 0; mi < 2; ++mi)
#pragma unroll
                for (int ni = 0; ni < 2; ++ni) {
                    hh[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah_[mi], bh_[ni], hh[mi][ni], 0, 0, 0);
                    cr[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah_[mi], bl_[ni], cr[mi][ni], 0, 0, 0);
                    cr[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x4f32(al_[mi], bh_[ni], cr[mi][ni], 0, 0, 0);
                }
        }
#pragma unroll
        for (int mi = 0; mi < 2; ++mi)
#pragma unroll
            for (int ni = 0; ni < 2; ++ni)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    acc[mi][ni][r] += (double)hh[mi][ni][r] + (double)cr[mi][ni][r];
        __syncthreads();
    }
    write_tile<false>(g, tc, lane, wm, wn, acc);
}

// sum ksplit partials in fixed order, apply alpha/beta
__global__ __launch_bounds__(256) void k_gemm_reduce(const GemmDesc* __restrict__ descs, int nd,
                                                     const int64_t* __restrict__ starts,
                                                     int64_t total) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        int lo = 0, hi = nd - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (starts[mid] <= e) lo = mid; else hi = mid - 1;
        }
        const GemmDesc& g = descs[lo];
        int64_t r = e - starts[lo];
        int64_t mn = (int64_t)g.M * g.N;
        double s = 0.0;
        for (int z = 0; z < g.ksplit; ++z) s += g.work[z * mn + r];
        int row = (int)(r % g.M), col = (int)(r / g.M);
        double* p = g.C + row + (int64_t)col * g.ldc;
        *p = (g.beta == 0.0) ? g.alpha * s : g.alpha * s + g.beta * *p;
    }
}

// --------------------------------------------------------- trace / normalise
__device__ __forceinline__ double block_sum_fixed(double v, double* red) {
    // deterministic tree over 256 threads
    const int t = threadIdx.x;
    red[t] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) red[t] += red[t + s];
        __syncthreads();
    }
    double r = red[0];
    __syncthreads();
    return r;
}

// grid (TRACE_PARTS, nprob): partial sums of each problem's weighted trace
__global__ __launch_bounds__(256) void k_trace_partial(const TraceDesc* __restrict__ descs,
                                                       const TraceProb* __restrict__ probs) {
    __shared__ double red[256];
    const TraceProb P = probs[blockIdx.y];
    const TraceDesc* ds = descs + P.first;
    double acc = 0.0;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < P.total;
         e += (int64_t)gridDim.x * blockDim.x) {
        int lo = 0, hi = P.count - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (ds[mid].start <= e) lo = mid; else hi = mid - 1;
        }
        const TraceDesc& D = ds[lo];
        int64_t r = e - D.start;
        int x1 = (int)(r % D.d1), x2 = (int)(r / D.d1);
        double v = D.T[x1 * D.s1 + x2 * D.s2];
        if (D.w1) v *= D.w1[x1];
        if (D.w2) v *= D.w2[x2];
        acc += v;
    }
    double s = block_sum_fixed(acc, red);
    if (threadIdx.x == 0) P.partial[blockIdx.x] = s;
}

// grid (X, nprob): every block re-derives its problem's norm from the
// partials in the same order, block 0 stores it, all scale T by 1/|n|.
__global__ __launch_bounds__(256) void k_normalize(const TraceProb* __restrict__ probs) {
    __shared__ double red[256];
    const TraceProb P = probs[blockIdx.y];
    double v = (threadIdx.x < TRACE_PARTS) ? P.partial[threadIdx.x] : 0.0;
    double s = fabs(block_sum_fixed(v, red));
    if (blockIdx.x == 0 && threadIdx.x == 0) *P.norm_slot = s;
    const double inv = 1.0 / s;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < P.size;
         e += (int64_t)gridDim.x * blockDim.x)
        P.T[e] *= inv;
}

// ------------------------------------------------------------- Jacobi eigen
// One workgroup per symmetric n x n matrix (n <= LMAX).  Parallel cyclic
// Jacobi with the circle-method tournament ordering: each round rotates n/2
// disjoint pairs (p,q) simultaneously, G <- J^T G J, W <- W J.  A pair is
// skipped when |g_pq| <= tol sqrt(|g_pp g_qq|) (relative criterion, which
// keeps small eigenvalues of graded Gram matrices relatively accurate).  The
// rotated 2x2 block is written from the closed form (g_pp - t g_pq,
// g_qq + t g_pq, 0).  The input is symmetrised on load ((G + G^T)/2).
// Output: eigenvalues by |lambda| descending, eigenvectors as columns of W
// in the same order.
__global__ __launch_bounds__(256) void k_jacobi(const JacobiDesc* __restrict__ descs, int nd,
                                                double tol, int max_sweeps) {
    extern __shared__ double lds[];
    const JacobiDesc D = descs[blockIdx.x];
    const int n = D.n;
    const int np = n + (n & 1);
    double* G = lds;                 // np x np, column-major (G[i + j*np])
    double* W = lds + np * np;
    double* cs = W + np * np;        // c, s, t per pair (3 * np/2)
    double* g0 = cs + 3 * (LMAX / 2);   // gpp, gqq, gpq per pair
    int* act = (int*)(g0 + 3 * (LMAX / 2));
    int* anyrot = act + LMAX / 2;
    const int t = threadIdx.x, nt = blockDim.x;
    for (int e = t; e < np * np; e += nt) {
        int i = e % np, j = e / np;
        G[e] = (i < n && j < n)
                   ? 0.5 * (D.G[i + (int64_t)j * D.ldg] + D.G[j + (int64_t)i * D.ldg])
                   : 0.0;
        W[e] = (i == j) ? 1.0 : 0.0;
    }
    __syncthreads();
    const int half = np / 2, m = np - 1;
    int sweeps = 0;
    for (int sw = 0; sw < max_sweeps; ++sw) {
        if (t == 0) *anyrot = 0;
        __syncthreads();
        for (int rnd = 0; rnd < m; ++rnd) {
            if (t < half) {
                int a = (t == 0) ? 0 : 1 + ((t - 1 + rnd) % m);
                int x = np - 1 - t;
                int b = 1 + ((x - 1 + rnd) % m);
                int p = min(a, b), q = max(a, b);
                double gpp = G[p + p * np], gqq = G[q + q * np], gpq = G[p + q * np];
                int on = (gpq != 0.0) && (fabs(gpq) > tol * sqrt(fabs(gpp * gqq)));
                double c = 1.0, s = 0.0, tt = 0.0;
                if (on) {
                    double tau = (gqq - gpp) / (2.0 * gpq);
                    tt = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                    c = 1.0 / sqrt(1.0 + tt * tt);
                    s = tt * c;
                    *anyrot = 1;
                }
                act[t] = on;
                cs[3 * t] = c; cs[3 * t + 1] = s; cs[3 * t + 2] = tt;
                g0[3 * t] = gpp; g0[3 * t + 1] = gqq; g0[3 * t + 2] = gpq;
            }
            __syncthreads();
            // rows p, q of G (all columns)
            for (int e = t; e < half * np; e += nt) {
                int i = e / np, j = e % np;
                if (!act[i]) continue;
                int a = (i == 0) ? 0 : 1 + ((i - 1 + rnd) % m);
                int b = 1 + ((np - 1 - i - 1 + rnd) % m);
                int p = min(a, b), q = max(a, b);
                double c = cs[3 * i], s = cs[3 * i + 1];
                double gp = G[p + j * np], gq = G[q + j * np];
                G[p + j * np] = c * gp - s * gq;
                G[q + j * np] = s * gp + c * gq;
            }
            __syncthreads();
            // columns p, q of G and W (all rows); closed form on the 2x2 block
            for (int e = t; e < half * np; e += nt) {
                int i = e / np, j = e % np;
                if (!act[i]) continue;
                int a = (i == 0) ? 0 : 1 + ((i - 1 + rnd) % m);
                int b = 1 + ((np - 1 - i - 1 + rnd) % m);
                int p = min(a, b), q = max(a, b);
                double c = cs[3 * i], s = cs[3 * i + 1];
                if (j == p || j == q) {
                    double tt = cs[3 * i + 2];
                    double gpp = g0[3 * i], gqq = g0[3 * i + 1], gpq = g0[3 * i + 2];
                    if (j == p) { G[p + p * np] = gpp - tt * gpq; G[q + p * np] = 0.0; }
                    else        { G[q + q * np] = gqq + tt * gpq; G[p + q * np] = 0.0; }
                } else {
                    double gp = G[j + p * np], gq = G[j + q * np];
                    G[j + p * np] = c * gp - s * gq;
                    G[j + q * np] = s * gp + c * gq;
                }
                double wp = W[j + p * np], wq = W[j + q * np];
                W[j + p * np] = c * wp - s * wq;
                W[j + q * np] = s * wp + c * wq;
            }
            __syncthreads();
        }
        ++sweeps;
        int done = (*anyrot == 0);
        __syncthreads();
        if (done) break;
    }
    // sort by |lambda| descending (stable by index) and write out
    for (int i = t; i < n; i += nt) {
        double li = G[i + i * np], ai = fabs(li);
        int pos = 0;
        for (int j = 0; j < n; ++j) {
            double aj = fabs(G[j + j * np]);
            pos += (aj > ai) || (aj == ai && j < i);
        }
        D.lam[pos] = li;
        for (int r = 0; r < n; ++r) D.W[r + (int64_t)pos * D.ldw] = W[r + i * np];
    }
    if (t == 0 && D.sweeps) *D.sweeps = sweeps;
}

// ------------------------------------------------- column norms / normalise
// sigma_j = ||Y[:, j]||; U[:, j] = Y[:, j] / sigma_j, or 0 when
// sigma_j <= rtol * max_j sigma_j (numerical null direction).  One
// workgroup per problem; wave w owns columns j = w (mod 4) and reduces with a
// fixed butterfly, so the result does not depend on scheduling.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__global__ __launch_bounds__(256) void k_colnorm(const ColDesc* __restrict__ descs, int nd,
                                                 double rtol) {
    __shared__ double sig[LMAX];
    const ColDesc D = descs[blockIdx.x];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    for (int j = w; j < D.ncol; j += 4) {
        const double* y = D.Y + (int64_t)j * D.ldy;
        double acc = 0.0;
        for (int i = lane; i < D.nrow; i += 64) acc += y[i] * y[i];
        acc = wave_sum(acc);
        if (lane == 0) sig[j] = sqrt(acc);
    }
    __syncthreads();
    double smax = 0.0;
    for (int j = 0; j < D.ncol; ++j) smax = fmax(smax, sig[j]);
    for (int j = w; j < D.ncol; j += 4) {
        const double sj = sig[j];
        const bool live = sj > rtol * smax && sj > 0.0;
        const double inv = live ? 1.0 / sj : 0.0;
        const double* y = D.Y + (int64_t)j * D.ldy;
        double* u = D.U + (int64_t)j * D.ldu;
        for (int i = lane; i < D.nrow; i += 64) u[i] = y[i] * inv;
        if (lane == 0) D.sigma[j] = live ? sj : 0.0;
    }
}

// in-place (M + M^T)/2 of square blocks (project_hermitian)
__global__ __launch_bounds__(256) void k_symmetrize(const ColDesc* __restrict__ descs, int nd) {
    const ColDesc D = descs[blockIdx.y];
    const int64_t n = D.nrow;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n * n;
         e += (int64_t)gridDim.x * blockDim.x) {
        int64_t i = e % n, j = e / n;
        if (i >= j) continue;
        double* a = const_cast<double*>(D.Y);
        double x = a[i + j * D.ldy], y = a[j + i * D.ldy];
        double m = 0.5 * (x + y);
        a[i + j * D.ldy] = m;
        a[j + i * D.ldy] = m;
    }
}

// ------------------------------------------------------ deterministic fill
__device__ __forceinline__ uint64_t splitmix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__global__ __launch_bounds__(256) void k_randfill(const FillDesc* __restrict__ descs, int nd,
                                                  int64_t total) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        int lo = 0, hi = nd - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (descs[mid].start <= e) lo = mid; else hi = mid - 1;
        }
        const FillDesc& D = descs[lo];
        int64_t r = e - D.start;
        int i = (int)(r % D.rows), j = (int)(r / D.rows);
        uint64_t h = splitmix(D.seed ^ splitmix((uint64_t)i * 0x100000001B3ull + (uint64_t)j));
        D.p[i + (int64_t)j * D.ld] = ((double)(h >> 11) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0;
    }
}

// --------------------------------------------------- strided matrix gather
// out[i + j*ldo] = in[(trans ? j + i*ldi : i + j*ldi)] * rw[i] * cw[j]
__global__ __launch_bounds__(256) void k_gather(const GatherDesc* __restrict__ descs, int nd,
                                                int64_t total) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        int lo = 0, hi = nd - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (descs[mid].start <= e) lo = mid; else hi = mid - 1;
        }
        const GatherDesc& D = descs[lo];
        int64_t r = e - D.start;
        int i = (int)(r % D.rows), j = (int)(r / D.rows);
        double v = D.trans ? D.in[(int64_t)j + (int64_t)i * D.ldi] : D.in[(int64_t)i + (int64_t)j * D.ldi];
        if (D.rw) v *= D.rw[i];
        if (D.cw) v *= D.cw[j];
        D.out[(int64_t)i + (int64_t)j * D.ldo] = v;
    }
}

// ------------------------------------------------ Frobenius norm of blocks
__global__ __launch_bounds__(256) void k_frob(const ColDesc* __restrict__ descs, int nd) {
    __shared__ double red[256];
    const ColDesc D = descs[blockIdx.x];
    double acc = 0.0;
    for (int j = 0; j < D.ncol; ++j) {
        const double* y = D.Y + (int64_t)j * D.ldy;
        for (int i = threadIdx.x; i < D.nrow; i += 256) acc += y[i] * y[i];
    }
    double s = block_sum_fixed(acc, red);
    if (threadIdx.x == 0) D.sigma[0] = s;
}

// =================================================================== launch
static int grid_for(int64_t total, int cap = 4096) {
    int64_t g = (total + 255) / 256;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

void launch_copy(hipStream_t st, const CopyDesc* d_descs, int nd, int64_t total) {
    if (total <= 0 || nd <= 0) return;
    hipLaunchKernelGGL(k_copy, dim3(grid_for(total, 8192)), dim3(256), 0, st, d_descs, nd, total);
    TNR_HIP(hipGetLastError());
}

void launch_gemm(hipStream_t st, const GemmDesc* d_descs, int nd, int64_t tiles, bool split32) {
    if (tiles <= 0 || nd <= 0) return;
    if (split32)
        hipLaunchKernelGGL(k_gemm_split32, dim3((unsigned)tiles), dim3(256), 0, st, d_descs, nd);
    else
        hipLaunchKernelGGL(k_gemm_f64, dim3((unsigned)tiles), dim3(256), 0, st, d_descs, nd);
    TNR_HIP(hipGetLastError());
}

void launch_gemm_reduce(hipStream_t st, const GemmDesc* d_descs, int nd, const int64_t* d_starts,
                        int64_t total) {
    if (total <= 0 || nd <= 0) return;
    hipLaunchKernelGGL(k_gemm_reduce, dim3(grid_for(total, 8192)), dim3(256), 0, st, d_descs, nd,
                       d_starts, total);
    TNR_HIP(hipGetLastError());
}

void launch_trace_normalize(hipStream_t st, const TraceDesc* d_descs, const TraceProb* d_probs,
                            int nprob, int64_t max_size) {
    if (nprob <= 0) return;
    hipLaunchKernelGGL(k_trace_partial, dim3(TRACE_PARTS, nprob), dim3(256), 0, st, d_descs,
                       d_probs);
    TNR_HIP(hipGetLastError());
    int gx = grid_for(max_size, 1024);
    hipLaunchKernelGGL(k_normalize, dim3(gx, nprob), dim3(256), 0, st, d_probs);
    TNR_HIP(hipGetLastError());
}

size_t jacobi_lds_bytes(int n) {
    int np = n + (n & 1);
    return (size_t)2 * np * np * sizeof(double) + 6 * (LMAX / 2) * sizeof(double) +
           (LMAX / 2 + 4) * sizeof(int);
}

void launch_jacobi(hipStream_t st, const JacobiDesc* d_descs, int nd, int nmax, double tol,
                   int max_sweeps) {
    if (nd <= 0) return;
    static bool attr = false;
    if (!attr) {
        TNR_HIP(hipFuncSetAttribute((const void*)k_jacobi,
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)jacobi_lds_bytes(LMAX)));
        attr = true;
    }
    hipLaunchKernelGGL(k_jacobi, dim3(nd), dim3(256), jacobi_lds_bytes(nmax), st, d_descs, nd, tol,
                       max_sweeps);
    TNR_HIP(hipGetLastError());
}

void launch_colnorm(hipStream_t st, const ColDesc* d_descs, int nd, double rtol) {
    if (nd <= 0) return;
    hipLaunchKernelGGL(k_colnorm, dim3(nd), dim3(256), 0, st, d_descs, nd, rtol);
    TNR_HIP(hipGetLastError());
}

void launch_symmetrize(hipStream_t st, const ColDesc* d_descs, int nd, int nmax) {
    if (nd <= 0) return;
    int64_t tot = (int64_t)nmax * nmax;
    hipLaunchKernelGGL(k_symmetrize, dim3(grid_for(tot, 1024), nd), dim3(256), 0, st, d_descs, nd);
    TNR_HIP(hipGetLastError());
}

void launch_frob(hipStream_t st, const ColDesc* d_descs, int nd) {
    if (nd <= 0) return;
    hipLaunchKernelGGL(k_frob, dim3(nd), dim3(256), 0, st, d_descs, nd);
    TNR_HIP(hipGetLastError());
}

void launch_randfill(hipStream_t st, const FillDesc* d_descs, int nd, int64_t total) {
    if (total <= 0 || nd <= 0) return;
    hipLaunchKernelGGL(k_randfill, dim3(grid_for(total, 4096)), dim3(256), 0, st, d_descs, nd, total);
    TNR_HIP(hipGetLastError());
}

void launch_gather(hipStream_t st, const GatherDesc* d_descs, int nd, int64_t total) {
    if (total <= 0 || nd <= 0) return;
    hipLaunchKernelGGL(k_gather, dim3(grid_for(total, 8192)), dim3(256), 0, st, d_descs, nd, total);
    TNR_HIP(hipGetLastError());
}

}  // namespace tnr
