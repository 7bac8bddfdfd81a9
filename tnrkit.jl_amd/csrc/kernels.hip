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

#include <hip/hip_ext.h>

#include <condition_variable>
#include <map>
#include <mutex>

#include <algorithm>
#include <vector>

namespace tnr {

// Pointers taken from descriptor arrays are generic to the compiler, so plain
// reads through them become flat loads, which also count in lgkmcnt (an LDS
// wait would then also wait for them).  A global-address-space view makes
// them global loads, tracked by vmcnt alone.
using gconst_double = const __attribute__((address_space(1))) double;
using gmut_double = __attribute__((address_space(1))) double;
__device__ __forceinline__ gconst_double* as_global(const double* p) { return (gconst_double*)p; }
__device__ __forceinline__ gmut_double* as_global(double* p) { return (gmut_double*)p; }

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
// One chunk of COPY_CHUNK consecutive elements (destination order) of one
// job per workgroup iteration: the job is looked up once per chunk, and the
// index decomposition runs in 32-bit arithmetic (a job holds < 2^31
// elements; checked on the host).  Each thread decodes COPY_GROUP elements,
// issues their loads together, then scales and stores them, so that many
// loads are in flight per thread.
constexpr int COPY_CHUNK = 4096;
constexpr int COPY_GROUP = 8;

__global__ __launch_bounds__(256) void k_copy(const CopyDesc* __restrict__ descs, int nd,
                                              int64_t chunks) {
    const int t = threadIdx.x;
    for (int64_t c = blockIdx.x; c < chunks; c += gridDim.x) {
        int lo = 0, hi = nd - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (descs[mid].start <= c) lo = mid; else hi = mid - 1;
        }
        const CopyDesc& D = descs[lo];
        const int ndim = D.nd;
        uint32_t dim[MAXLEG];
        uint32_t cnt = 1;
#pragma unroll
        for (int q = 0; q < MAXLEG; ++q) {
            dim[q] = q < ndim ? (uint32_t)D.dim[q] : 1u;
            cnt *= dim[q];
        }
        const uint32_t base = (uint32_t)(c - D.start) * COPY_CHUNK;
        const int wl0 = D.wleg[0], wl1 = D.wleg[1];
        gconst_double* src = as_global(D.src);
        gmut_double* dst = as_global(D.dst);
        for (int u0 = 0; u0 < COPY_CHUNK / 256; u0 += COPY_GROUP) {
            if (base + (uint32_t)(256 * u0) >= cnt) break;
            int64_t so[COPY_GROUP], dof[COPY_GROUP];
            uint32_t iw0[COPY_GROUP], iw1[COPY_GROUP];
            bool ok[COPY_GROUP];
#pragma unroll
            for (int g = 0; g < COPY_GROUP; ++g) {
                uint32_t r = base + (uint32_t)(t + 256 * (u0 + g));
                ok[g] = r < cnt;
                if (!ok[g]) r = 0;          // element 0: a valid address, not stored
                so[g] = 0;
                dof[g] = 0;
                iw0[g] = 0;
                iw1[g] = 0;
#pragma unroll
                for (int q = 0; q < MAXLEG; ++q) {
                    if (q < ndim) {
                        const uint32_t nq = r / dim[q];
                        const uint32_t iq = r - nq * dim[q];
                        r = nq;
                        so[g] += (int64_t)iq * D.sstr[q];
                        dof[g] += (int64_t)iq * D.dstr[q];
                        if (q == wl0) iw0[g] = iq;
                        if (q == wl1) iw1[g] = iq;
                    }
                }
            }
            double v[COPY_GROUP];
#pragma unroll
            for (int g = 0; g < COPY_GROUP; ++g) v[g] = src[so[g]];
#pragma unroll
            for (int g = 0; g < COPY_GROUP; ++g) {
                double x = v[g];
                if (wl0 >= 0) x *= D.w[0][iw0[g]];
                if (wl1 >= 0) x *= D.w[1][iw1[g]];
                if (ok[g]) dst[dof[g]] = x;
            }
        }
    }
}

// ------------------------------------------------------- tiled transpose
// One 64 x 64 tile per workgroup iteration: 256 threads read 4 rows of 64
// consecutive source elements per pass (coalesced), and write 4 rows of 64
// consecutive destination elements per pass (coalesced) through LDS.
__global__ __launch_bounds__(256) void k_transpose(const TransDesc* __restrict__ descs, int nd,
                                                   int64_t tiles) {
    // 4096 doubles of LDS hold `sub` sub-tiles of tx (along n0) by ty (along
    // n1), one per consecutive batch index; rows padded by one element
    __shared__ double buf[4096 + 64 * 16];
    __shared__ int64_t soff[64], doff[64];           // batch offsets per sub-tile (-1: none)
    const int t = threadIdx.x;
    int lo = 0;
    for (int64_t tid = blockIdx.x; tid < tiles; tid += gridDim.x) {
        // tiles come in increasing order: continue the search from the last
        // job (usually the same one or the next)
        if (descs[lo].tile_start > tid) lo = 0;
        if (lo + 1 < nd && descs[lo + 1].tile_start <= tid) {
            int hi = nd - 1;
            while (lo < hi) {
                int mid = (lo + hi + 1) >> 1;
                if (descs[mid].tile_start <= tid) lo = mid; else hi = mid - 1;
            }
        }
        const TransDesc& D = descs[lo];
        const int tx = D.tx, ty = D.ty, sub = D.sub, pitch = tx + 1;
        const int lx = __ffs(tx) - 1, ly = __ffs(ty) - 1, la = lx + ly;   // log2 extents
        // a job has fewer than 2^31 tiles and batch entries: 32-bit decode
        int r = (int)(tid - D.tile_start);
        const int i0 = (r % D.t0) * tx;
        r /= D.t0;
        const int i1 = (r % D.t1) * ty;
        r /= D.t1;
        const int b0 = r * sub;                     // first batch index of this tile
        __syncthreads();                            // previous tile done with buf / offsets
        if (t < sub) {
            const int bi = b0 + t;
            int64_t so = -1, dof = -1;
            if (bi < D.nb) {
                int q = bi;
                so = 0;
                dof = 0;
                for (int kk = 0; kk < TR_BATCH; ++kk) {
                    const int ik = q % D.bdim[kk];
                    q /= D.bdim[kk];
                    so += (int64_t)ik * D.bs[kk];
                    dof += (int64_t)ik * D.bd[kk];
                }
            }
            soff[t] = so;
            doff[t] = dof;
        }
        __syncthreads();
        gconst_double* src = as_global(D.src);
        gmut_double* dst = as_global(D.dst);
        // all 16 loads of this thread are issued before the first LDS store
        double v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {              // x along n0 (source-contiguous)
            const int e = t + 256 * u;
            const int s = e >> la, rem = e & ((1 << la) - 1), j = rem >> lx, x = rem & (tx - 1);
            const int a = i0 + x, c = i1 + j;
            const int64_t so = soff[s];
            v[u] = (so >= 0 && a < D.n0 && c < D.n1) ? src[so + a + (int64_t)c * D.ss1] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int e = t + 256 * u;
            const int s = e >> la, rem = e & ((1 << la) - 1), j = rem >> lx, x = rem & (tx - 1);
            buf[s * ty * pitch + j * pitch + x] = v[u];
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 16; ++u) {              // x along n1 (destination-contiguous)
            const int e = t + 256 * u;
            const int s = e >> la, rem = e & ((1 << la) - 1), j = rem >> ly, x = rem & (ty - 1);
            v[u] = buf[s * ty * pitch + x * pitch + j];
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int e = t + 256 * u;
            const int s = e >> la, rem = e & ((1 << la) - 1), j = rem >> ly, x = rem & (ty - 1);
            const int a = i0 + j, c = i1 + x;
            const int64_t dof = doff[s];
            if (dof >= 0 && a < D.n0 && c < D.n1) dst[dof + c + (int64_t)a * D.ds0] = v[u];
        }
    }
}

// --------------------------------------------------------------- FP64 GEMM
// Output tile BM x BN per 256-thread workgroup; 4 waves as 2 x 2, each wave
// owning (BM/2) x (BN/2) = MI x NI blocks of v_mfma_f64_16x16x4_f64 per
// k-step of 4.  The k order of every output element is the same for all
// tile shapes, so results do not depend on the shape chosen.  Operand lane
// layout of the 16x16x4 MFMAs (measured on gfx950): A[i=l%16][k=l/16],
// B[k=l/16][j=l%16]; result register r of lane l holds
//   f64: D[i = l/16 + 4r][j = l%16]   (rows interleaved)
//   f32: D[i = 4(l/16) + r][j = l%16] (rows contiguous).
constexpr int BK = 16;

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

// Profiling: the algorithmic flops of one output tile (its rows and columns
// inside M x N, its K range) added to a device counter, so only problems
// that actually ran are counted (a finished problem's tiles return before
// this).  The counts are integers, so the sum does not depend on order.
template <int BM, int BN>
__device__ __forceinline__ void count_tile_flops(unsigned long long* ctr, const GemmDesc& g,
                                                 const TileCoord& tc) {
    if (!ctr || threadIdx.x != 0) return;
    const long long rows = min(BM, g.M - tc.m0), cols = min(BN, g.N - tc.n0);
    const long long kk = tc.kend - tc.kbeg;
    if (rows > 0 && cols > 0 && kk > 0) atomicAdd(ctr, (unsigned long long)(2 * rows * cols * kk));
}

template <int BM, int BN>
__device__ __forceinline__ TileCoord tile_coord(const GemmDesc& g, int64_t local) {
    TileCoord c;
    // Workgroups go round-robin over the 8 XCDs (each with its own L2): the
    // physical ids of one residue class mod 8 get one contiguous range of
    // the tile order, which walks GROUP_M tile rows at a time (then the next
    // tile column), so the tiles an XCD runs together share A and B slices.
    // A bijection on each K chunk's tiles; every tile's arithmetic is the
    // same wherever it runs.
    constexpr int NXCD = 8, GROUP_M = 8;
    const int64_t per_k = (int64_t)g.tiles_m * g.tiles_n;
    const int kz = (int)(local / per_k);
    int64_t r = local - (int64_t)kz * per_k;
    const int64_t F = per_k / NXCD;
    if (r < NXCD * F) r = (r % NXCD) * F + r / NXCD;
    const int64_t width = (int64_t)GROUP_M * g.tiles_n;
    const int64_t grp = r / width, inner = r - grp * width;
    const int first_m = (int)(grp * GROUP_M);
    const int gsz = min(g.tiles_m - first_m, GROUP_M);
    const int tm = first_m + (int)(inner % gsz);
    const int tn = (int)(inner / gsz);
    c.m0 = tm * BM;
    c.n0 = tn * BN;
    if (g.ksplit > 1) {           // one K chunk per workgroup
        c.kbeg = kz * g.kchunk;
        c.kend = min(g.K, (kz + 1) * g.kchunk);
    } else {
        c.kbeg = 0;
        c.kend = g.K;
    }
    c.kz = kz;
    return c;
}

// global -> registers: the tile's BM x KB slab of op(A), KB BM / NT values
// per thread of an NT-thread workgroup
template <int BM, int NT, int KB>
__device__ __forceinline__ void load_a(const GemmDesc& g, int m0, int k0, int t,
                                       double (&ra)[KB * BM / NT]) {
    constexpr int R = KB * BM / NT;
    gconst_double* A = as_global(g.A);
    if (!g.ta) {  // A[i + k*lda], M x K: thread -> (i, k-group)
        const int i = m0 + (t % BM), kk = t / BM;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int k = k0 + kk + (NT / BM) * r;
            ra[r] = (i < g.M && k < g.K) ? A[(int64_t)i + (int64_t)k * g.lda] : 0.0;
        }
    } else {      // stored K x M: A[k + i*lda]
        const int kk = t % KB, i0 = t / KB;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = m0 + i0 + (NT / KB) * r, k = k0 + kk;
            ra[r] = (i < g.M && k < g.K) ? A[(int64_t)k + (int64_t)i * g.lda] : 0.0;
        }
    }
}

template <int BM, int NT, int KB>
__device__ __forceinline__ void store_a(const GemmDesc& g, int t, const double (&ra)[KB * BM / NT],
                                        double (*As)[BM + 1]) {
    constexpr int R = KB * BM / NT;
    if (!g.ta) {
        const int i = t % BM, kk = t / BM;
#pragma unroll
        for (int r = 0; r < R; ++r) As[kk + (NT / BM) * r][i] = ra[r];
    } else {
        const int kk = t % KB, i0 = t / KB;
#pragma unroll
        for (int r = 0; r < R; ++r) As[kk][i0 + (NT / KB) * r] = ra[r];
    }
}

template <int BN, int NT, int KB>
__device__ __forceinline__ void load_b(const GemmDesc& g, int n0, int k0, int t,
                                       double (&rb)[KB * BN / NT]) {
    constexpr int R = KB * BN / NT;
    gconst_double* B = as_global(g.B);
    if (!g.tb) {  // B[k + j*ldb], K x N
        const int kk = t % KB, j0 = t / KB;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int j = n0 + j0 + (NT / KB) * r, k = k0 + kk;
            rb[r] = (j < g.N && k < g.K) ? B[(int64_t)k + (int64_t)j * g.ldb] : 0.0;
        }
    } else {      // stored N x K: B[j + k*ldb]
        const int j = n0 + (t % BN), kk = t / BN;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int k = k0 + kk + (NT / BN) * r;
            rb[r] = (j < g.N && k < g.K) ? B[(int64_t)j + (int64_t)k * g.ldb] : 0.0;
        }
    }
}

template <int BN, int NT, int KB>
__device__ __forceinline__ void store_b(const GemmDesc& g, int t, const double (&rb)[KB * BN / NT],
                                        double (*Bs)[BN + 1]) {
    constexpr int R = KB * BN / NT;
    if (!g.tb) {
        const int kk = t % KB, j0 = t / KB;
#pragma unroll
        for (int r = 0; r < R; ++r) Bs[kk][j0 + (NT / KB) * r] = rb[r];
    } else {
        const int j = t % BN, kk = t / BN;
#pragma unroll
        for (int r = 0; r < R; ++r) Bs[kk + (NT / BN) * r][j] = rb[r];
    }
}

template <bool F64_LAYOUT, int MI, int NI>
__device__ __forceinline__ void write_tile(const GemmDesc& g, const TileCoord& tc, int lane,
                                           int wm0, int wn0, const dbl4 (&acc)[MI][NI]) {
    const bool part = g.ksplit > 1;
    gmut_double* out = as_global(part ? g.work + (int64_t)tc.kz * g.M * g.N : g.C);
    const int64_t ld = part ? g.M : g.ldc;
#pragma unroll
    for (int mi = 0; mi < MI; ++mi)
#pragma unroll
        for (int ni = 0; ni < NI; ++ni) {
            const int col = tc.n0 + wn0 + ni * 16 + (lane & 15);
            if (col >= g.N) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int ri = F64_LAYOUT ? (lane >> 4) + 4 * r : (lane >> 4) * 4 + r;
                const int row = tc.m0 + wm0 + mi * 16 + ri;
                if (row >= g.M) continue;
                const double v = acc[mi][ni][r];
                gmut_double* p = out + (int64_t)row + (int64_t)col * ld;
                if (part) *p = v;
                else *p = (g.beta == 0.0) ? g.alpha * v : g.alpha * v + g.beta * *p;
            }
        }
}

// Output through LDS: a register of the FP64 MFMA result holds 4 rows x 16
// columns, which a direct store writes as 16 pieces of 32 bytes.  Instead
// each wave parks 32 x 16 blocks of its result in its own LDS region `ep`
// (EP_REGION doubles, column stride 34: conflict-free both ways) and stores
// whole 256-byte column runs.  Same values and formula as write_tile.
constexpr int EP_LD = 34, EP_REGION = 16 * EP_LD;

template <int MI, int NI>
__device__ __forceinline__ void write_tile_lds(const GemmDesc& g, const TileCoord& tc, int lane,
                                               int wm0, int wn0, const dbl4 (&acc)[MI][NI],
                                               double* ep) {
    static_assert(MI == 2, "32-row wave blocks");
    const bool part = g.ksplit > 1;
    gmut_double* out = as_global(part ? g.work + (int64_t)tc.kz * g.M * g.N : g.C);
    const int64_t ldo = part ? g.M : g.ldc;
    const int row = lane & 31, cpar = lane >> 5;
    const int grow = tc.m0 + wm0 + row;
#pragma unroll
    for (int ni = 0; ni < NI; ++ni) {
#pragma unroll
        for (int mi = 0; mi < MI; ++mi)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                ep[(lane & 15) * EP_LD + mi * 16 + (lane >> 4) + 4 * r] = acc[mi][ni][r];
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int col = 2 * c + cpar, gcol = tc.n0 + wn0 + ni * 16 + col;
            const double v = ep[col * EP_LD + row];
            if (grow < g.M && gcol < g.N) {
                gmut_double* p = out + (int64_t)grow + (int64_t)gcol * ldo;
                if (part) *p = v;
                else *p = (g.beta == 0.0) ? g.alpha * v : g.alpha * v + g.beta * *p;
            }
        }
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    }
}

// WM x WN waves, each owning a (BM/WM) x (BN/WN) block of the tile; k-tiles
// of KB.  Every shape runs the MFMAs over k in the same 4-wide steps.
template <int BM, int BN, int WM, int WN, int KB>
__global__ __launch_bounds__(64 * WM * WN) void k_gemm_f64(const GemmDesc* __restrict__ descs,
                                                           int nd, unsigned long long* flop_ctr) {
    constexpr int NT = 64 * WM * WN;
    constexpr int MI = BM / WM / 16, NI = BN / WN / 16;
    // two LDS stages: the next k-tile is stored while the current one is
    // consumed, one barrier per k-tile
    __shared__ double As[2][KB][BM + 1];
    __shared__ double Bs[2][KB][BN + 1];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm0 = (w % WM) * (BM / WM), wn0 = (w / WM) * (BN / WN);
    const GemmDesc g = descs[find_desc_gemm(descs, nd, blockIdx.x)];
    if (g.skip && *g.skip) return;
    const TileCoord tc = tile_coord<BM, BN>(g, (int64_t)blockIdx.x - g.tile_start);
    count_tile_flops<BM, BN>(flop_ctr, g, tc);
    dbl4 acc[MI][NI], tot[MI][NI];
#pragma unroll
    for (int mi = 0; mi < MI; ++mi)
#pragma unroll
        for (int ni = 0; ni < NI; ++ni) {
            acc[mi][ni] = dbl4{0.0, 0.0, 0.0, 0.0};
            tot[mi][ni] = dbl4{0.0, 0.0, 0.0, 0.0};
        }
    double ra[KB * BM / NT], rb[KB * BN / NT];
    if (tc.kbeg < tc.kend) {
        load_a<BM, NT, KB>(g, tc.m0, tc.kbeg, t, ra);
        load_b<BN, NT, KB>(g, tc.n0, tc.kbeg, t, rb);
        store_a<BM, NT, KB>(g, t, ra, As[0]);
        store_b<BN, NT, KB>(g, t, rb, Bs[0]);
    }
    __syncthreads();
    int buf = 0;
    for (int k0 = tc.kbeg; k0 < tc.kend; k0 += KB) {
        const bool more = k0 + KB < tc.kend;
        if (more) {
            load_a<BM, NT, KB>(g, tc.m0, k0 + KB, t, ra);
            load_b<BN, NT, KB>(g, tc.n0, k0 + KB, t, rb);
        }
        // operands of k-step ks + 1 are read from LDS while the MFMAs of
        // k-step ks run (two register sets)
        double a[2][MI], b[2][NI];
#pragma unroll
        for (int mi = 0; mi < MI; ++mi) a[0][mi] = As[buf][lane >> 4][wm0 + mi * 16 + (lane & 15)];
#pragma unroll
        for (int ni = 0; ni < NI; ++ni) b[0][ni] = Bs[buf][lane >> 4][wn0 + ni * 16 + (lane & 15)];
#pragma unroll
        for (int ks = 0; ks < KB / 4; ++ks) {
            const int cur = ks & 1;
            if (ks + 1 < KB / 4) {
                const int k = (ks + 1) * 4 + (lane >> 4);
#pragma unroll
                for (int mi = 0; mi < MI; ++mi)
                    a[cur ^ 1][mi] = As[buf][k][wm0 + mi * 16 + (lane & 15)];
#pragma unroll
                for (int ni = 0; ni < NI; ++ni)
                    b[cur ^ 1][ni] = Bs[buf][k][wn0 + ni * 16 + (lane & 15)];
            }
#pragma unroll
            for (int mi = 0; mi < MI; ++mi)
#pragma unroll
                for (int ni = 0; ni < NI; ++ni)
                    acc[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[cur][mi], b[cur][ni],
                                                                       acc[mi][ni], 0, 0, 0);
        }
        // end of a K chunk: add its sum to the running total (chunk order)
        if ((k0 + KB) % g.kchunk == 0 || !more) {
#pragma unroll
            for (int mi = 0; mi < MI; ++mi)
#pragma unroll
                for (int ni = 0; ni < NI; ++ni) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) tot[mi][ni][r] += acc[mi][ni][r];
                    acc[mi][ni] = dbl4{0.0, 0.0, 0.0, 0.0};
                }
        }
        if (more) {
            store_a<BM, NT, KB>(g, t, ra, As[buf ^ 1]);
            store_b<BN, NT, KB>(g, t, rb, Bs[buf ^ 1]);
        }
        __syncthreads();
        buf ^= 1;
    }
    // the loop's last barrier ends every read of As and Bs
    if constexpr (MI == 2 && 2 * KB * (BM + 1) >= (NT / 64) * EP_REGION)
        write_tile_lds<MI, NI>(g, tc, lane, wm0, wn0, tot, &As[0][0][0] + w * EP_REGION);
    else
        write_tile<true, MI, NI>(g, tc, lane, wm0, wn0, tot);
}

// ------------------------------------------------ FP64 MFMA, LDS-DMA pipeline
// The same products as k_gemm_f64 (same MFMA sequence over k, same K-chunk
// totals, so the same bits), with the operands going from global memory
// straight into LDS (buffer_load_dwordx4 ... lds: lane l of a wave writes its
// 16 bytes at the instruction's LDS base + 16 l) through NS LDS stages (three
// in every launch): the loads of k-tile t + NS - 1 are in flight while
// k-tile t is multiplied, and no registers hold tiles in transit.
//
// Each operand's stage region takes one of two layouts, by storage order:
//  * row layout (operand contiguous along its M / N index: A as stored, B
//    stored transposed): k-row kk holds the D values of that k; even rows
//    are packed from offset 0, odd rows from 8 D + 16, so rows kk and kk + 1
//    sit in opposite halves of the 64 LDS banks;
//  * column layout (contiguous along k): 16-byte slots (i, kp) hold
//    k = 2 kp, 2 kp + 1 of index i, slot i * 8 + (kp ^ ((i >> 1) & 7)): one
//    DMA instruction fetches 8 whole 128-byte lines, and the MFMA operand
//    reads of 16 consecutive i at one k fall in distinct banks.
// Loads go through buffer descriptors rebased per k-tile whose extent ends
// at the operand's last element: reads past it return zero and never fault.
// Rows past M (columns past N) may hold other elements of the operand; they
// only reach output rows (columns) that are not written.  k past K is zeroed
// in LDS before the last k-tile is used, so padded products are 0 * 0 as in
// k_gemm_f64.
// (KB, the k-tile depth, is 16 or 32: KB / 2 k-pairs per index; the column
// layout's XOR mask is (i >> 1) & 7 for 8 pairs and i & 15 for 16, either
// way distinct over 16 consecutive i.  Every launch uses 16: a 32-deep
// 128 x 128 tile on two stages needs more than 256 VGPRs and spills.)
template <int KB>
__host__ __device__ constexpr int dl_region(int d) { return KB * d + 16; }

template <int KB>
__device__ __forceinline__ int dl_mask(int i) { return KB == 16 ? ((i >> 1) & 7) : (i & 15); }

// LDS offset (doubles) of element (i, kk) of a D-wide operand region
template <int KB>
__device__ __forceinline__ int dl_off(bool rowl, int D, int i, int kk) {
    if (rowl) return ((kk & 1) ? (KB / 2) * D + 16 : 0) + (kk >> 1) * D + i;
    return (i * (KB / 2) + ((kk >> 1) ^ dl_mask<KB>(i))) * 2 + (kk & 1);
}

// What DMA instruction u (of D KB / 128) of a D-wide operand region fetches
// for this lane: its LDS base (doubles, uniform), the lane's element (i, k pair)
template <int KB>
__device__ __forceinline__ void dl_slot(bool rowl, int D, int u, int lane, int& base, int& i,
                                        int& kk0, int& kk1) {
    if (rowl) {
        const int half = (KB / 2) * D / 128;   // instructions per row parity
        const int odd = u >= half, uu = odd ? u - half : u;
        const int e = 128 * uu + 2 * lane;
        base = (odd ? (KB / 2) * D + 16 : 0) + 128 * uu;
        i = e % D;
        kk0 = kk1 = 2 * (e / D) + odd;
    } else {
        base = 128 * u;
        i = (128 / KB) * u + lane / (KB / 2);
        const int kp = (lane % (KB / 2)) ^ dl_mask<KB>(i);
        kk0 = 2 * kp;
        kk1 = 2 * kp + 1;
    }
}

template <int BM, int BN, int NW, int KB>
struct DlPlan {
    static constexpr int SA = dl_region<KB>(BM), SB = dl_region<KB>(BN), ST = SA + SB;
    // DMA instructions (1 KB each) per stage: for A, for both, per wave
    static constexpr int QA = BM * KB / 128, Q = (BM + BN) * KB / 128, PER = Q / NW;
    static_assert(Q % NW == 0, "DMA instructions must divide over the waves");
};

// LDS-DMA for one k-tile: this wave's PER instructions into `stage`
template <int BM, int BN, int NW, int KB>
__device__ __forceinline__ void dl_issue(double* stage, __amdgpu_buffer_rsrc_t ra,
                                         __amdgpu_buffer_rsrc_t rb,
                                         const int (&lbase)[DlPlan<BM, BN, NW, KB>::PER],
                                         const int (&vofs)[DlPlan<BM, BN, NW, KB>::PER], int w) {
    using P = DlPlan<BM, BN, NW, KB>;
#pragma unroll
    for (int r = 0; r < P::PER; ++r) {
        const int q = w + NW * r;
        __builtin_amdgcn_raw_ptr_buffer_load_lds(q < P::QA ? ra : rb,
                                                 (__attribute__((address_space(3))) void*)(stage + lbase[r]),
                                                 16, vofs[r], 0, 0, 0);
    }
}

// buffer descriptor over [base, base + n) doubles (n clamped to the 32-bit
// byte range)
__device__ __forceinline__ __amdgpu_buffer_rsrc_t dl_rsrc(const double* base, int64_t n) {
    const int64_t bytes = n > 0 ? n * 8 : 0;
    const int nb = (int)(bytes > 0x7fffffffLL ? 0x7fffffffLL : bytes);
    return __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, nb, 0x00020000);
}

// NS LDS stages (3 or 4, each its own array so that the compiler's wait
// for a stage's loads does not include the other stages'); OCC waves per
// SIMD: 4 lets two 128x64 workgroups share a CU (three stages: 73 KB of LDS
// each), 3 fits three 64x64 ones (49 KB each)
template <int BM, int BN, int WM, int WN, int OCC, int NS, int KB>
__global__ __launch_bounds__(64 * WM * WN) __attribute__((amdgpu_waves_per_eu(OCC)))
void k_gemm_dl(const GemmDesc* __restrict__ descs, int nd, unsigned long long* flop_ctr) {
    static_assert(NS >= 2 && NS <= 4, "two to four LDS stages");
    static_assert(KB == 16 || KB == 32, "16- or 32-deep k-tiles");
    constexpr int NW = WM * WN;
    using P = DlPlan<BM, BN, NW, KB>;
    constexpr int MI = BM / WM / 16, NI = BN / WN / 16;
    __shared__ __align__(16) double S0[P::ST];
    __shared__ __align__(16) double S1[P::ST];
    __shared__ __align__(16) double S2[NS > 2 ? P::ST : 2];
    __shared__ __align__(16) double S3[NS > 3 ? P::ST : 2];
    double* const SS[4] = {S0, S1, S2, S3};
    // the wave index is uniform: keep it (and the LDS bases derived from it)
    // in scalar registers, as the DMA's LDS address must be
    const int t = threadIdx.x, lane = t & 63, w = __builtin_amdgcn_readfirstlane(t >> 6);
    const int wm0 = (w % WM) * (BM / WM), wn0 = (w / WM) * (BN / WN);
    const GemmDesc g = descs[find_desc_gemm(descs, nd, blockIdx.x)];
    if (g.skip && *g.skip) return;
    const TileCoord tc = tile_coord<BM, BN>(g, (int64_t)blockIdx.x - g.tile_start);
    count_tile_flops<BM, BN>(flop_ctr, g, tc);
    if (flop_ctr) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    const bool arow = !g.ta, brow = g.tb != 0;

    // per-lane DMA offsets (bytes, relative to the k-tile's descriptors) and
    // LDS bases of this wave's instructions
    int lbase[P::PER], vofs[P::PER];
#pragma unroll
    for (int r = 0; r < P::PER; ++r) {
        const int q = w + NW * r;
        int base, i, kk0, kk1;
        if (q < P::QA) {
            dl_slot<KB>(arow, BM, q, lane, base, i, kk0, kk1);
            lbase[r] = __builtin_amdgcn_readfirstlane(base);
            vofs[r] = 8 * (arow ? i + kk0 * g.lda : kk0 + i * g.lda);
        } else {
            dl_slot<KB>(brow, BN, q - P::QA, lane, base, i, kk0, kk1);
            lbase[r] = __builtin_amdgcn_readfirstlane(P::SA + base);
            vofs[r] = 8 * (brow ? i + kk0 * g.ldb : kk0 + i * g.ldb);
        }
    }
    // the operands' last elements bound the descriptors
    const int64_t extA = arow ? (int64_t)g.lda * (g.K - 1) + g.M : (int64_t)g.lda * (g.M - 1) + g.K;
    const int64_t extB = brow ? (int64_t)g.ldb * (g.K - 1) + g.N : (int64_t)g.ldb * (g.N - 1) + g.K;
    auto rsrc_a = [&](int k0) {
        const int64_t o = arow ? (int64_t)tc.m0 + (int64_t)k0 * g.lda : (int64_t)k0 + (int64_t)tc.m0 * g.lda;
        return dl_rsrc(g.A + o, extA - o);
    };
    auto rsrc_b = [&](int k0) {
        const int64_t o = brow ? (int64_t)tc.n0 + (int64_t)k0 * g.ldb : (int64_t)k0 + (int64_t)tc.n0 * g.ldb;
        return dl_rsrc(g.B + o, extB - o);
    };
    // MFMA operand offsets for each 4-deep k-step (row i / column j at mi,
    // ni = 0; +16 rows is +16 doubles in the row layout, +16 KB doubles in
    // the column one)
    int aoff[KB / 4], boff[KB / 4];
#pragma unroll
    for (int ks = 0; ks < KB / 4; ++ks) {
        const int kk = ks * 4 + (lane >> 4);
        aoff[ks] = dl_off<KB>(arow, BM, wm0 + (lane & 15), kk);
        boff[ks] = P::SA + dl_off<KB>(brow, BN, wn0 + (lane & 15), kk);
    }
    const int astep = arow ? 16 : 16 * KB, bstep = brow ? 16 : 16 * KB;

    dbl4 acc[MI][NI], tot[MI][NI];
#pragma unroll
    for (int mi = 0; mi < MI; ++mi)
#pragma unroll
        for (int ni = 0; ni < NI; ++ni) {
            acc[mi][ni] = dbl4{0.0, 0.0, 0.0, 0.0};
            tot[mi][ni] = dbl4{0.0, 0.0, 0.0, 0.0};
        }
    const int nt = (tc.kend - tc.kbeg + KB - 1) / KB;
#pragma unroll
    for (int s = 0; s < NS - 1; ++s)
        if (s < nt)
            dl_issue<BM, BN, NW, KB>(SS[s], rsrc_a(tc.kbeg + s * KB), rsrc_b(tc.kbeg + s * KB),
                                 lbase, vofs, w);

    // one k-tile: wait for this wave's loads of tile it, zero k >= K on the
    // last tile, barrier (every wave's loads of tile it have landed and every
    // wave is done with the stage tile it + NS - 1 goes into), issue tile
    // it + NS - 1, multiply
    auto step = [&](double* cur, double* nxt, int it) {
        const int k0 = tc.kbeg + it * KB;
        const int ahead = min(NS - 2, nt - 1 - it);   // tiles issued after this one
        if (NS > 3 && ahead >= 2) asm volatile("s_waitcnt vmcnt(%0)" ::"n"(2 * P::PER) : "memory");
        else if (ahead >= 1) asm volatile("s_waitcnt vmcnt(%0)" ::"n"(P::PER) : "memory");
        else asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        if (k0 + KB > g.K) {
            const int kv = g.K - k0;
#pragma unroll
            for (int r = 0; r < P::PER; ++r) {
                const int q = w + NW * r;
                int base, i, kk0, kk1;
                if (q < P::QA) dl_slot<KB>(arow, BM, q, lane, base, i, kk0, kk1);
                else dl_slot<KB>(brow, BN, q - P::QA, lane, base, i, kk0, kk1);
                double* d = cur + lbase[r] + 2 * lane;
                if (kk0 >= kv) d[0] = 0.0;
                if (kk1 >= kv) d[1] = 0.0;
            }
        }
        asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
        if (it + NS - 1 < nt)
            dl_issue<BM, BN, NW, KB>(nxt, rsrc_a(k0 + (NS - 1) * KB), rsrc_b(k0 + (NS - 1) * KB),
                                 lbase, vofs, w);
        __builtin_amdgcn_sched_barrier(0);
        double a[2][MI], b[2][NI];
#pragma unroll
        for (int mi = 0; mi < MI; ++mi) a[0][mi] = cur[aoff[0] + mi * astep];
#pragma unroll
        for (int ni = 0; ni < NI; ++ni) b[0][ni] = cur[boff[0] + ni * bstep];
#pragma unroll
        for (int ks = 0; ks < KB / 4; ++ks) {
            const int c = ks & 1;
            if (ks + 1 < KB / 4) {
#pragma unroll
                for (int mi = 0; mi < MI; ++mi) a[c ^ 1][mi] = cur[aoff[ks + 1] + mi * astep];
#pragma unroll
                for (int ni = 0; ni < NI; ++ni) b[c ^ 1][ni] = cur[boff[ks + 1] + ni * bstep];
            }
#pragma unroll
            for (int mi = 0; mi < MI; ++mi)
#pragma unroll
                for (int ni = 0; ni < NI; ++ni)
                    acc[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[c][mi], b[c][ni],
                                                                       acc[mi][ni], 0, 0, 0);
        }
        if ((k0 + KB) % g.kchunk == 0 || it + 1 == nt) {
#pragma unroll
            for (int mi = 0; mi < MI; ++mi)
#pragma unroll
                for (int ni = 0; ni < NI; ++ni) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) tot[mi][ni][r] += acc[mi][ni][r];
                    acc[mi][ni] = dbl4{0.0, 0.0, 0.0, 0.0};
                }
        }
        // the stage's reads are complete before the next barrier lets another
        // wave's DMA overwrite it
        __builtin_amdgcn_sched_barrier(0);
    };
    for (int it = 0; it < nt; it += NS) {
#pragma unroll
        for (int s = 0; s < NS; ++s)
            if (it + s < nt) step(SS[s], SS[(s + NS - 1) % NS], it + s);
    }
    // the stages are free: the last k-tile waited for every load, and the
    // barrier below for every wave's reads; wave w parks its blocks in slot
    // w / NS of stage w % NS
    static_assert(((NW + NS - 1) / NS) * EP_REGION <= P::ST, "epilogue fits the stages");
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    const int es = w % NS;
    write_tile_lds<MI, NI>(g, tc, lane, wm0, wn0, tot,
                           (es == 0 ? S0 : es == 1 ? S1 : es == 2 ? S2 : S3) + (w / NS) * EP_REGION);
}

// ------------------------------------------------ FP32 MFMA, hi/lo split
// x = hi + lo with hi = fp32(x), lo = fp32(x - hi).  Per k-tile the three
// significant products hi*hi, hi*lo, lo*hi are accumulated on FP32 MFMA
// (v_mfma_f32_16x16x4_f32) in separate accumulators, then flushed into FP64
// registers, so FP32 rounding only acts within one 16-deep k-tile.
__global__ __launch_bounds__(256) void k_gemm_split32(const GemmDesc* __restrict__ descs, int nd,
                                                      unsigned long long* flop_ctr) {
    __shared__ float Ah[BK][65], Al[BK][65];
    __shared__ float Bh[BK][65], Bl[BK][65];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, wm = w & 1, wn = w >> 1;
    const GemmDesc g = descs[find_desc_gemm(descs, nd, blockIdx.x)];
    if (g.skip && *g.skip) return;
    const TileCoord tc = tile_coord<64, 64>(g, (int64_t)blockIdx.x - g.tile_start);
    count_tile_flops<64, 64>(flop_ctr, g, tc);
    dbl4 acc[2][2], tot[2][2];
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni) {
            acc[mi][ni] = dbl4{0.0, 0.0, 0.0, 0.0};
            tot[mi][ni] = dbl4{0.0, 0.0, 0.0, 0.0};
        }
    double ra[4], rb[4];
    for (int k0 = tc.kbeg; k0 < tc.kend; k0 += BK) {
        load_a<64, 256, BK>(g, tc.m0, k0, t, ra);
        load_b<64, 256, BK>(g, tc.n0, k0, t, rb);
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
        if ((k0 + BK) % g.kchunk == 0 || k0 + BK >= tc.kend) {
#pragma unroll
            for (int mi = 0; mi < 2; ++mi)
#pragma unroll
                for (int ni = 0; ni < 2; ++ni) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) tot[mi][ni][r] += acc[mi][ni][r];
                    acc[mi][ni] = dbl4{0.0, 0.0, 0.0, 0.0};
                }
        }
        __syncthreads();
    }
    write_tile<false, 2, 2>(g, tc, lane, wm * 32, wn * 32, tot);
}

// sum ksplit partials in fixed order, apply alpha/beta
__global__ __launch_bounds__(256) void k_gemm_reduce(const GemmDesc* __restrict__ descs, int nd,
                                                     const int64_t* __restrict__ starts,
                                                     int64_t total) {
    // A thread's elements increase, so its job index only moves forward: the
    // search continues from the last job (most steps stay in it or move to
    // the next), and that job's fields stay in registers until it changes.
    int lo = 0, cur = -1;
    const double* work = nullptr;
    double* C = nullptr;
    int64_t mn = 0, base = 0;
    int M = 1, ldc = 1, ks = 1;
    double alpha = 1.0, beta = 0.0;
    bool skip = false;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        if (lo + 1 < nd && starts[lo + 1] <= e) {
            int hi = nd - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (starts[mid] <= e) lo = mid; else hi = mid - 1;
            }
        }
        if (lo != cur) {
            cur = lo;
            const GemmDesc& g = descs[lo];
            skip = g.skip && *g.skip;
            work = g.work;
            C = g.C;
            M = g.M;
            ldc = g.ldc;
            ks = g.ksplit;
            alpha = g.alpha;
            beta = g.beta;
            mn = (int64_t)g.M * g.N;
            base = starts[lo];
        }
        if (skip) continue;
        const int64_t r = e - base;
        double s = 0.0;
        for (int z = 0; z < ks; ++z) s += work[z * mn + r];
        const int row = (int)(r % M), col = (int)(r / M);
        double* p = C + row + (int64_t)col * ldc;
        *p = (beta == 0.0) ? alpha * s : alpha * s + beta * *p;
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
        const int x1 = (int)(r % D.d1);
        r /= D.d1;
        const int x2 = (int)(r % D.d2), x3 = (int)(r / D.d2);
        double v = D.T[x1 * D.s1 + x2 * D.s2 + x3 * D.s3];
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
    const double inv = 1.0 / (P.divide_by ? *P.divide_by : s);
    // four elements a grid-stride apart per pass, loaded before they are
    // stored, so each thread keeps four loads in flight
    gmut_double* T = as_global(P.T);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; e + 3 * stride < P.size; e += 4 * stride) {
        const double x0 = T[e], x1 = T[e + stride], x2 = T[e + 2 * stride], x3 = T[e + 3 * stride];
        T[e] = x0 * inv;
        T[e + stride] = x1 * inv;
        T[e + 2 * stride] = x2 * inv;
        T[e + 3 * stride] = x3 * inv;
    }
    for (; e < P.size; e += stride) T[e] *= inv;
}

// ---------------------------------------------------- linear combinations
__global__ __launch_bounds__(256) void k_lincomb(const LinDesc* __restrict__ descs, int nd,
                                                 int64_t total) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        int lo = 0, hi = nd - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (descs[mid].start <= e) lo = mid; else hi = mid - 1;
        }
        const LinDesc& D = descs[lo];
        const int64_t r = e - D.start;
        double v = D.coef[0] * D.in[0][r];
        for (int i = 1; i < D.nin; ++i) v += D.coef[i] * D.in[i][r];
        D.out[r] = v;
    }
}

// ------------------------------------------------------------- Jacobi eigen
// One 1024-thread workgroup per symmetric n x n matrix (n <= LMAX).
// Parallel cyclic Jacobi with the circle-method tournament ordering: each
// round rotates the n/2 disjoint pairs (p_i, q_i) simultaneously,
// G <- J^T G J, W <- W J.  A pair is rotated only while
// |g_pq| > tol sqrt(|g_pp g_qq|) (relative criterion, which keeps the small
// eigenvalues of graded Gram matrices relatively accurate).  A round is two
// phases: (1) one thread per pair computes its rotation (the tournament
// positions advance in registers); (2) every 2x2 block
// B_ij = G[{p_i,q_i},{p_j,q_j}] becomes J_i^T B_ij J_j (diagonal blocks from
// the closed form g_pp - t g_pq, g_qq + t g_pq, 0) and every row of W is
// rotated.  A sweep starts with a parallel scan of all off-diagonal pairs and
// the solve stops when none violates the criterion.  The input is
// symmetrised on load ((G + G^T)/2).  Output: eigenvalues by |lambda|
// descending, eigenvectors as columns of W in the same order.
constexpr int JT = 1024;

// Pair (p < q) that tournament slot `slot` (0 <= slot < np / 2) rotates in
// round r of a sweep: slot 0 keeps position 0, every other position moves
// one step round the circle 1..np-1 per round.
__device__ __forceinline__ void jacobi_pair(int slot, int r, int np, int& p, int& q) {
    const int m = np - 1;
    int a = 0;
    if (slot > 0) {
        a = slot - 1 + r;
        if (a >= m) a -= m;
        a += 1;
    }
    int b = np - 2 - slot + r;
    if (b >= m) b -= m;
    b += 1;
    p = min(a, b);
    q = max(a, b);
}

// A pair is rotated while g_pq^2 > tol^2 |g_pp g_qq| (the relative
// criterion keeps the small eigenvalues of graded Gram matrices relatively
// accurate); the sweep scan and the rounds use this same predicate.
__device__ __forceinline__ int jacobi_violates(double gpp, double gqq, double gpq, double tol2) {
    return gpq != 0.0 && gpq * gpq > tol2 * fabs(gpp * gqq);
}

// Rotation zeroing g_pq: with a = g_qq - g_pp, b = 2 g_pq and
// h = sqrt(a^2 + b^2), t = sign(a) b / (|a| + h) is the tangent of the
// smaller angle, c = 1 / sqrt(1 + t^2), s = t c.  The square roots and the
// reciprocal come from the hardware estimates (v_rsq_f64, v_rcp_f64)
// refined by two Newton steps each, a short dependent chain instead of
// IEEE division and square-root sequences; c^2 + s^2 = 1 to a few ulp.
__device__ __forceinline__ double rsqrt_refined(double x) {
    double y = __builtin_amdgcn_rsq(x);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const double e = fma(-x * y, y, 1.0);       // 1 - x y^2
        y = fma(0.5 * y, e, y);
    }
    return y;
}

__device__ __forceinline__ double rcp_refined(double d) {
    double y = __builtin_amdgcn_rcp(d);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const double e = fma(-d, y, 1.0);
        y = fma(y, e, y);
    }
    return y;
}

__device__ __forceinline__ void jacobi_rotation(double gpp, double gqq, double gpq, double& t,
                                                double& c, double& s) {
    const double a = gqq - gpp, b = 2.0 * gpq;
    const double h2 = fma(a, a, b * b);
    const double h = h2 * rsqrt_refined(h2);
    t = (a >= 0.0 ? b : -b) * rcp_refined(fabs(a) + h);
    c = rsqrt_refined(fma(t, t, 1.0));
    s = t * c;
}

__global__ __launch_bounds__(JT) void k_jacobi(const JacobiDesc* __restrict__ descs, int nd,
                                               double tol, int max_sweeps) {
    extern __shared__ double lds[];
    const JacobiDesc D = descs[blockIdx.x];
    if (D.skip && *D.skip) return;
    const int n = D.n;
    const int np = n + (n & 1);
    const int ld = np + 1;              // padded column stride (bank spread)
    double* G = lds;                    // np x np, G[i + j*ld]
    double* W = G + np * ld;
    double4* rot = (double4*)(W + np * ld);      // (c, s, t, active) per pair
    int* pos = (int*)(rot + LMAX / 2);           // output order (LMAX ints)
    const int t = threadIdx.x;
    const float inv_np = 1.0f / (float)np;
    // G is symmetric and only its upper triangle (row <= column) is read
    // or written after the load: gx(r, c) addresses entry (r, c) there
    auto gx = [ld](int r, int c) { return r <= c ? r + c * ld : c + r * ld; };
    for (int e = t; e < np * np; e += JT) {
        const int j = (int)(((float)e + 0.5f) * inv_np), i = e - j * np;
        G[i + j * ld] = (i < n && j < n)
                            ? 0.5 * (D.G[i + (int64_t)j * D.ldg] + D.G[j + (int64_t)i * D.ldg])
                            : 0.0;
        W[i + j * ld] = (i == j) ? 1.0 : 0.0;
    }
    const int half = np / 2, m = np - 1;
    const int ntri = half * (half + 1) / 2;      // pair blocks (i <= j), <= 2 * JT
    const double tol2 = tol * tol;
    // this thread's fixed tasks: pair blocks e = t, t + JT of the upper
    // triangle, and (pair, row) entries e = t + u JT of W
    int gi[2] = {-1, -1}, gj[2] = {-1, -1};
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int e = t + u * JT;
        if (e < ntri) {
            int j = (int)((sqrtf(8.0f * (float)e + 1.0f) - 1.0f) * 0.5f);
            while (j * (j + 1) / 2 > e) --j;
            while ((j + 1) * (j + 2) / 2 <= e) ++j;
            gi[u] = e - j * (j + 1) / 2;
            gj[u] = j;
        }
    }
    constexpr int WT = (LMAX / 2) * LMAX / JT + 1;   // W tasks per thread, max
    int wslot[WT], wrow[WT];
    {
        int j = t / np, r = t - (t / np) * np;
        const int wdj = JT / np, wdr = JT - (JT / np) * np;
#pragma unroll
        for (int u = 0; u < WT; ++u) {
            wslot[u] = j < half ? j : -1;
            wrow[u] = r;
            r += wdr;
            j += wdj;
            if (r >= np) { r -= np; ++j; }
        }
    }
    // any off-diagonal pair p < q that still violates the criterion
    auto violated = [&]() {
        int viol = 0;
        for (int e = t; e < np * np; e += JT) {
            const int q = (int)(((float)e + 0.5f) * inv_np), p = e - q * np;
            if (p < q)
                viol |= jacobi_violates(G[p + p * ld], G[q + q * ld], G[p + q * ld], tol2);
        }
        return __syncthreads_or(viol);
    };
    int sweeps = 0;
    bool converged = false;
    for (int sw = 0; sw < max_sweeps; ++sw) {
        if (!violated()) {
            converged = true;
            break;
        }
        ++sweeps;
        for (int rnd = 0; rnd < m; ++rnd) {
            // The pairs of a round are fixed by the tournament schedule, so
            // this thread's G block and W entries are loaded before the
            // rotations are known (nothing writes G or W until the barrier
            // after them).
            int bp[2] = {0, 0}, bq[2] = {0, 0}, cp[2] = {0, 0}, cq[2] = {0, 0};
            double b00[2], b01[2], b10[2], b11[2];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                if (gi[u] < 0) continue;
                jacobi_pair(gi[u], rnd, np, bp[u], bq[u]);
                jacobi_pair(gj[u], rnd, np, cp[u], cq[u]);
                b00[u] = G[gx(bp[u], cp[u])];
                b01[u] = G[gx(bp[u], cq[u])];
                b10[u] = G[gx(bq[u], cp[u])];
                b11[u] = G[gx(bq[u], cq[u])];
            }
            int wa[WT], wb[WT];
            double wp[WT], wq[WT];
#pragma unroll
            for (int u = 0; u < WT; ++u) {
                if (wslot[u] < 0) continue;
                int p, q;
                jacobi_pair(wslot[u], rnd, np, p, q);
                wa[u] = wrow[u] + p * ld;
                wb[u] = wrow[u] + q * ld;
                wp[u] = W[wa[u]];
                wq[u] = W[wb[u]];
            }
            if (t < half) {
                int p, q;
                jacobi_pair(t, rnd, np, p, q);
                const double gpp = G[p + p * ld], gqq = G[q + q * ld], gpq = G[p + q * ld];
                const int on = jacobi_violates(gpp, gqq, gpq, tol2);
                double c = 1.0, sn = 0.0, tt = 0.0;
                if (on) jacobi_rotation(gpp, gqq, gpq, tt, c, sn);
                rot[t] = make_double4(c, sn, tt, on ? 1.0 : 0.0);
            }
            // every task checks its pairs' flags below, so a plain barrier
            // suffices (a workgroup OR-reduction costs two)
            __syncthreads();
            // G <- J^T G J on the upper pair blocks; the lower ones are their
            // transposes, kept only through gx
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int i = gi[u], j = gj[u];
                if (i < 0) continue;
                const double4 ri = rot[i], rj = rot[j];
                if (ri.w == 0.0 && rj.w == 0.0) continue;
                const int pi = bp[u], qi = bq[u], pj = cp[u], qj = cq[u];
                if (i == j) {
                    const double tt = ri.z, gpq = b01[u];
                    G[pi + pi * ld] = b00[u] - tt * gpq;
                    G[qi + qi * ld] = b11[u] + tt * gpq;
                    G[pi + qi * ld] = 0.0;   // pi < qi
                    continue;
                }
                const double ci = ri.x, si = ri.y, cj = rj.x, sj = rj.y;
                const double x00 = ci * b00[u] - si * b10[u], x01 = ci * b01[u] - si * b11[u];
                const double x10 = si * b00[u] + ci * b10[u], x11 = si * b01[u] + ci * b11[u];
                const double y00 = cj * x00 - sj * x01, y01 = sj * x00 + cj * x01;   // (.) J_j
                const double y10 = cj * x10 - sj * x11, y11 = sj * x10 + cj * x11;
                G[gx(pi, pj)] = y00;
                G[gx(pi, qj)] = y01;
                G[gx(qi, pj)] = y10;
                G[gx(qi, qj)] = y11;
            }
            // W <- W J: this thread's (pair, row) tasks touch distinct columns
#pragma unroll
            for (int u = 0; u < WT; ++u) {
                if (wslot[u] < 0) continue;
                const double4 rj = rot[wslot[u]];
                if (rj.w == 0.0) continue;
                W[wa[u]] = rj.x * wp[u] - rj.y * wq[u];
                W[wb[u]] = rj.y * wp[u] + rj.x * wq[u];
            }
            __syncthreads();
        }
    }
    // sweep cap reached: converged only if the last sweep left nothing to rotate
    if (!converged && violated() && t == 0 && D.fail) *D.fail = 1.0;
    // order by |lambda| descending (stable by index), or keep the diagonal
    // positions, and write out
    if (t < n) {
        const double li = G[t + t * ld], ai = fabs(li);
        int k = t;
        if (!D.natural) {
            k = 0;
            for (int j = 0; j < n; ++j) {
                const double aj = fabs(G[j + j * ld]);
                k += (aj > ai) || (aj == ai && j < t);
            }
        }
        pos[t] = k;
        D.lam[k] = li;
    }
    __syncthreads();
    const float inv_n = 1.0f / (float)n;
    for (int e = t; e < n * n; e += JT) {
        const int i = (int)(((float)e + 0.5f) * inv_n), r = e - i * n;
        D.W[r + (int64_t)pos[i] * D.ldw] = W[r + i * ld];
    }
    if (t == 0 && D.sweeps) *D.sweeps = sweeps;
}

// ------------------------------------------------ pivoted Cholesky QR
// One workgroup per n x n Gram matrix G = Y^T Y (n <= LMAX):
// diagonal-pivoted Cholesky G[p, p] = R^T R, stopped where the largest
// remaining diagonal falls to null_rel2 * max diag(G) (numerical rank r),
// then X = R^{-1} (r x r).  Output W = P X (W[perm[i], j] = X[i, j], zero
// columns from r on), so that Y W = Q has orthonormal columns spanning
// span(Y), ordered by decreasing pivot; lam[j] = R[j][j] (the norm of
// column j of Y after projecting out the earlier pivots).  Ties in the
// pivot search go to the lowest index, so the result is a function of G.
// k_cholqr64 takes n <= 64 in one wavefront, k_cholqr96 the larger sizes in
// four; both keep the working matrix in registers.
constexpr int CQ64 = 64;   // Gram matrices up to this size go to k_cholqr64

// Pivoted Cholesky QR of Gram matrices with at most CQ64 = 64 columns in
// one wavefront that keeps the working matrix in registers, lane i
// holding row i, so a pivot step has one wave-local barrier and no LDS
// round trip per element.  Pivoting does not move data: the pivot's row is
// broadcast through LDS and rows already pivoted drop out of the search.
// Step j with pivot p, r = sqrt(A_j[p][p]): R[j][i] = A_j[p][i] / r (kept
// over the original column indices), A_{j+1}[i][k] = A_j[i][k] -
// (A_j[i][p] / r^2) A_j[p][k].  X = R^{-1} (R in pivot order) is solved
// column by column, lane c holding column c in registers, by right-looking
// elimination: rows kk from the last up, each scaled by 1 / R[kk][kk] and
// eliminated from the rows above it.  Ties in the pivot search go to the
// lowest row index.
constexpr int CQLD = CQ64 + 1;   // padded stride of the LDS matrices

size_t cholqr64_lds_bytes() {
    return 2 * (size_t)CQ64 * CQLD * sizeof(double)   // R rows / staging, pivot-ordered R
           + 2 * (size_t)CQ64 * sizeof(double)       // pivot rows (two, by step parity)
           + 2 * (size_t)CQ64 * sizeof(double)       // R[j][j], 1 / R[j][j]
           + 2 * (size_t)CQ64 * sizeof(int);         // perm, inverse perm
}

// One step of a wave argmax over (v, i) through DPP lane exchange CTRL:
// larger v wins, equal v goes to the lower index.
template <int CTRL>
__device__ __forceinline__ void argmax_dpp(double& v, int& i) {
    const int lo = __double2loint(v), hi = __double2hiint(v);
    const int lo2 = __builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xF, 0xF, false);
    const int hi2 = __builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xF, 0xF, false);
    const int i2 = __builtin_amdgcn_update_dpp(i, i, CTRL, 0xF, 0xF, false);
    const double v2 = __hiloint2double(hi2, lo2);
    const bool take = (v2 > v) | ((v2 == v) & (i2 < i));
    v = take ? v2 : v;
    i = take ? i2 : i;
}

// Argmax of (v, i) over the wave, the same result in every lane: within
// each row of 16 lanes by lane swaps (xor 1, xor 2, mirror within 8, mirror
// within 16), then over the four rows' results read from lanes 0, 16, 32, 48.
__device__ __forceinline__ void argmax_wave64(double& v, int& i) {
    argmax_dpp<0xB1>(v, i);    // quad_perm [1, 0, 3, 2]
    argmax_dpp<0x4E>(v, i);    // quad_perm [2, 3, 0, 1]
    argmax_dpp<0x141>(v, i);   // row_half_mirror
    argmax_dpp<0x140>(v, i);   // row_mirror
    const int lo = __double2loint(v), hi = __double2hiint(v);
    double bv = __hiloint2double(__builtin_amdgcn_readlane(hi, 0), __builtin_amdgcn_readlane(lo, 0));
    int bi = __builtin_amdgcn_readlane(i, 0);
#pragma unroll
    for (int r = 1; r < 4; ++r) {
        const double cv = __hiloint2double(__builtin_amdgcn_readlane(hi, 16 * r),
                                           __builtin_amdgcn_readlane(lo, 16 * r));
        const int ci = __builtin_amdgcn_readlane(i, 16 * r);
        const bool take = (cv > bv) | ((cv == bv) & (ci < bi));
        bv = take ? cv : bv;
        bi = take ? ci : bi;
    }
    v = bv;
    i = bi;
}

// (one wave per SIMD: the registers hold the rows and a whole pivot row in
// flight)
__global__ __launch_bounds__(CQ64) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_cholqr64(
    const JacobiDesc* __restrict__ descs, int nd, double null_rel2) {
    extern __shared__ double lds[];
    const JacobiDesc D = descs[blockIdx.x];
    if (D.n > CQ64 || (D.skip && *D.skip)) return;
    const int n = D.n, lane = threadIdx.x;
    double* Rr = lds;                        // R[j][k] by original column: Rr[j * CQLD + k]
    double* Rp = Rr + CQ64 * CQLD;           // G staging; later R[i][kk] at Rp[kk * CQ64 + i]
    double* prow = Rp + CQ64 * CQLD;         // [2][CQ64]
    double* rdiag = prow + 2 * CQ64;         // R[j][j]
    double* rinv = rdiag + CQ64;             // 1 / R[j][j]
    int* perm = (int*)(rinv + CQ64);
    int* iperm = perm + CQ64;
    const bool live = lane < n;
    // G through LDS (coalesced column loads, zero outside n x n), then
    // symmetrised rows into registers
    {
        // every lane loads (clamped, in bounds) and masks the bits, so the
        // loads are not predicated and stay in flight together
        const int ii = min(lane, n - 1);
#pragma unroll
        for (int c0 = 0; c0 < CQ64; c0 += 16) {
            long long g[16];
#pragma unroll
            for (int u = 0; u < 16; ++u)
                g[u] = __double_as_longlong(D.G[ii + (int64_t)min(c0 + u, n - 1) * D.ldg]);
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const long long keep = (live && c0 + u < n) ? -1LL : 0LL;
                Rp[lane + (c0 + u) * CQLD] = __longlong_as_double(g[u] & keep);
            }
        }
    }
    __syncthreads();
    double a[CQ64];
#pragma unroll
    for (int k = 0; k < CQ64; ++k) a[k] = 0.5 * (Rp[lane + k * CQLD] + Rp[k + lane * CQLD]);
    double d = Rp[lane + lane * CQLD];   // = a[lane]
    double dmax = d;
    {
        int ii = lane;
        argmax_wave64(dmax, ii);
    }
    const double nfloor = null_rel2 * dmax;
    bool done = !live;
    int rank = n;
    double v = live ? d : -1.0;
    int p = live ? lane : n;
    argmax_wave64(v, p);
    for (int j = 0; j < n; ++j) {
        if (!(v > nfloor && dmax > 0.0)) {
            rank = j;
            break;
        }
        double* pr = prow + (j & 1) * CQ64;
        if (lane == p) {
#pragma unroll
            for (int k = 0; k < CQ64; k += 2) *(double2*)(pr + k) = make_double2(a[k], a[k + 1]);
            done = true;
        }
        const double r = sqrt(v), ri = 1.0 / r;
        __syncthreads();
        const double aip = pr[lane];                  // A_j[p][i] = A_j[i][p]
        Rr[j * CQLD + lane] = lane == p ? r : aip * ri;
        if (lane == 0) {
            perm[j] = p;
            rdiag[j] = r;
            rinv[j] = ri;
        }
        // rows already pivoted keep their values (zero multiplier); the next
        // pivot search needs only d and overlaps the row updates
        const double f = done ? 0.0 : (aip * ri) * ri;
        d = d - f * aip;
        v = done ? -1.0 : d;
        p = done ? n : lane;
        double2 row[CQ64 / 2];
#pragma unroll
        for (int k = 0; k < CQ64 / 2; ++k) row[k] = *(const double2*)(pr + 2 * k);
        argmax_wave64(v, p);
#pragma unroll
        for (int k = 0; k < CQ64 / 2; ++k) {
            a[2 * k] = a[2 * k] - f * row[k].x;
            a[2 * k + 1] = a[2 * k + 1] - f * row[k].y;
        }
    }
    __syncthreads();
    // R in pivot order, column kk contiguous: Rp[kk * CQ64 + i] = R[i][kk]
    iperm[lane] = -1;
    for (int kk = 0; kk < rank; ++kk) Rp[kk * CQ64 + lane] = Rr[lane * CQLD + perm[kk]];
    __syncthreads();
    if (lane < rank) iperm[perm[lane]] = lane;
    // X = R^{-1}: lane c holds column c; rows kk from the last up, each
    // scaled by 1 / R[kk][kk] and eliminated from the rows above it
    double* x = a;
#pragma unroll
    for (int i = 0; i < CQ64; ++i) x[i] = (i == lane && lane < rank) ? 1.0 : 0.0;
#pragma unroll
    for (int kk = CQ64 - 1; kk >= 0; --kk) {
        if (kk < rank) {
            x[kk] *= rinv[kk];
            const double* col = Rp + kk * CQ64;
#pragma unroll
            for (int i = 0; i + 1 < kk; i += 2) {
                const double2 rr = *(const double2*)(col + i);
                x[i] = x[i] - rr.x * x[kk];
                x[i + 1] = x[i + 1] - rr.y * x[kk];
            }
            if (kk & 1) x[kk - 1] = x[kk - 1] - col[kk - 1] * x[kk];
        }
    }
    // stage X by columns (X[i][c] at Rr[i + c * CQLD]), then W[perm[i]][c] =
    // X[i][c] for i <= c < rank and zero elsewhere, one coalesced column at a time
#pragma unroll
    for (int i = 0; i < CQ64; ++i) Rr[i + lane * CQLD] = x[i];
    __syncthreads();
    const int ip = iperm[lane];
    if (live)
        for (int c = 0; c < n; ++c)
            D.W[lane + (int64_t)c * D.ldw] =
                (ip >= 0 && ip <= c && c < rank) ? Rr[ip + c * CQLD] : 0.0;
    if (live) D.lam[lane] = lane < rank ? rdiag[lane] : 0.0;
    if (lane == 0 && D.sweeps) *D.sweeps = 0;
}

// Pivoted Cholesky QR of Gram matrices with CQ64 < n <= LMAX columns: the
// arithmetic of k_cholqr64 over four wavefronts.  Thread (w, lane) holds
// row i = lane + 64 (w & 1) of the working matrix, columns CQH (w >> 1) to
// CQH (w >> 1) + CQH - 1, in registers.  At each pivot step the best
// remaining row of each row-wave (w & 1) publishes itself in LDS (buffers
// alternate by step parity), so a step takes one workgroup barrier, after
// which every thread picks the larger of the two candidates (ties to the
// lower index) and reads its row.  X = R^{-1} is solved with thread (w,
// lane) holding column c = lane + 64 (w & 1), rows CQH (w >> 1) to
// CQH (w >> 1) + CQH - 1: the elimination runs from the last row up, so the
// waves holding rows CQH.. finish them first and hand them through LDS to
// the waves holding rows 0..CQH-1; every element takes the same updates in
// the same order as in a one-thread right-looking elimination.
constexpr int CQH = LMAX / 2;     // columns (factorisation) / rows (inverse) per thread
constexpr int CQLDM = LMAX + 1;   // padded stride of the LDS matrices
constexpr int CQ96T = 256;
static_assert(LMAX * LMAX % (4 * CQ96T) == 0 && CQH % 2 == 0 && LMAX <= 2 * 64,
              "k_cholqr96 tiling");

size_t cholqr96_lds_bytes() {
    return 2 * (size_t)LMAX * CQLDM * sizeof(double)   // G staging / pivot-ordered R; R rows / X
           + 4 * (size_t)LMAX * sizeof(double)         // candidate rows [step parity][row-wave]
           + 4 * sizeof(double)                        // candidate values
           + 2 * (size_t)LMAX * sizeof(double)         // R[j][j], 1 / R[j][j]
           + 4 * sizeof(int)                           // candidate rows' indices
           + 2 * (size_t)LMAX * sizeof(int);           // perm, inverse perm
}

__global__ __launch_bounds__(CQ96T) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_cholqr96(
    const JacobiDesc* __restrict__ descs, int nd, double null_rel2) {
    extern __shared__ double lds[];
    const JacobiDesc D = descs[blockIdx.x];
    if (D.n <= CQ64 || D.n > LMAX || (D.skip && *D.skip)) return;
    const int n = D.n, t = threadIdx.x, lane = t & 63;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6), rw = w & 1, h = w >> 1;
    const int i = lane + 64 * rw;    // row (factorisation), column (inverse)
    const int ii = min(i, n - 1);    // in-bounds stand-in for rows past n
    const bool live = i < n;
    double* Gs = lds;                          // G staging; later R[r][kk] at Gs[kk * LMAX + r]
    double* Rr = Gs + LMAX * CQLDM;            // R[j][k] by original column k; later X
    double* cand = Rr + LMAX * CQLDM;          // [2][2][LMAX]
    double* candv = cand + 4 * LMAX;           // [2][2]
    double* rdiag = candv + 4;
    double* rinv = rdiag + LMAX;
    int* candp = (int*)(rinv + LMAX);          // [2][2]
    int* perm = candp + 4;
    int* iperm = perm + LMAX;
    // G into LDS, zero outside n x n (clamped loads, masked bits, so the
    // loads are not predicated and stay in flight together)
    for (int e0 = 0; e0 < LMAX * LMAX; e0 += 4 * CQ96T) {
        long long g[4];
        int at[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = e0 + u * CQ96T + t, c = e / LMAX, r = e - c * LMAX;
            g[u] = __double_as_longlong(D.G[min(r, n - 1) + (int64_t)min(c, n - 1) * D.ldg]);
            g[u] &= (r < n && c < n) ? -1LL : 0LL;
            at[u] = r + c * CQLDM;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) Gs[at[u]] = __longlong_as_double(g[u]);
    }
    if (t < LMAX) iperm[t] = -1;
    __syncthreads();
    double a[CQH];
#pragma unroll
    for (int k = 0; k < CQH; ++k) {
        const int c = h * CQH + k;
        a[k] = 0.5 * (Gs[ii + c * CQLDM] + Gs[c + ii * CQLDM]);
    }
    double d = live ? Gs[ii + ii * CQLDM] : 0.0;
    {
        // largest diagonal over both row-waves, through the parity-1 value
        // slots (first rewritten at step 1, after the step-0 barrier)
        double dm = d;
        int di = i;
        argmax_wave64(dm, di);
        if (h == 0 && lane == 0) candv[2 + rw] = dm;
    }
    __syncthreads();
    const double dmax = candv[3] > candv[2] ? candv[3] : candv[2];
    const double nfloor = null_rel2 * dmax;
    bool done = !live;
    int rank = n;
    double v = live ? d : -1.0;
    int p = live ? i : n;
    argmax_wave64(v, p);
    for (int j = 0; j < n; ++j) {
        const int b = j & 1;
        double* cb = cand + b * 2 * LMAX;
        if (p < n && i == p) {
            double* dst = cb + rw * LMAX + h * CQH;
#pragma unroll
            for (int k = 0; k < CQH; k += 2) *(double2*)(dst + k) = make_double2(a[k], a[k + 1]);
        }
        if (h == 0 && lane == 0) {
            candv[2 * b + rw] = v;
            candp[2 * b + rw] = p;
        }
        __syncthreads();
        // buffers of parity b are next written at step j + 2, after the
        // step j + 1 barrier, which every thread reaches only once it has
        // read them here
        const double v0 = candv[2 * b], v1 = candv[2 * b + 1];
        const int p0 = candp[2 * b], p1 = candp[2 * b + 1];
        const bool s1 = (v1 > v0) | ((v1 == v0) & (p1 < p0));
        const double vb = s1 ? v1 : v0;
        const int pb = s1 ? p1 : p0;
        if (!(vb > nfloor && dmax > 0.0)) {
            rank = j;
            break;
        }
        const double* pr = cb + (s1 ? LMAX : 0);
        const double r = sqrt(vb), ri = 1.0 / r;
        const double aip = pr[ii];   // A_j[p][i] = A_j[i][p]
        if (h == 0 && live) Rr[j * CQLDM + i] = i == pb ? r : aip * ri;
        if (t == 0) {
            perm[j] = pb;
            rdiag[j] = r;
            rinv[j] = ri;
        }
        if (i == pb) done = true;
        const double f = done ? 0.0 : (aip * ri) * ri;
        d = d - f * aip;
        v = done ? -1.0 : d;
        p = done ? n : i;
        double2 row[CQH / 2];
#pragma unroll
        for (int k = 0; k < CQH / 2; ++k) row[k] = *(const double2*)(pr + h * CQH + 2 * k);
        argmax_wave64(v, p);
#pragma unroll
        for (int k = 0; k < CQH / 2; ++k) {
            a[2 * k] = a[2 * k] - f * row[k].x;
            a[2 * k + 1] = a[2 * k + 1] - f * row[k].y;
        }
    }
    __syncthreads();
    // R in pivot order, column kk contiguous: Rp[kk * LMAX + r] = R[r][kk]
    double* Rp = Gs;
    for (int e = t; e < rank * LMAX; e += CQ96T) {
        const int kk = e / LMAX, r = e - kk * LMAX;
        Rp[e] = r < rank ? Rr[r * CQLDM + perm[kk]] : 0.0;
    }
    if (t < rank) iperm[perm[t]] = t;
    __syncthreads();
    // X = R^{-1}, X[r][c] staged at X[r + c * CQLDM]
    double* X = Rr;
    const int c = i, cc = min(c, LMAX - 1);
    double x[CQH];
#pragma unroll
    for (int k = 0; k < CQH; ++k) x[k] = (h * CQH + k == c && c < rank) ? 1.0 : 0.0;
    if (h == 1) {
        // rows CQH.. : rows kk from the last up, scaled and eliminated from
        // the rows above them within this half
#pragma unroll
        for (int kk = LMAX - 1; kk >= CQH; --kk) {
            if (kk < rank) {
                const int q = kk - CQH;
                x[q] *= rinv[kk];
                const double* col = Rp + kk * LMAX + CQH;
#pragma unroll
                for (int m = 0; m + 1 < q; m += 2) {
                    const double2 rr = *(const double2*)(col + m);
                    x[m] = x[m] - rr.x * x[q];
                    x[m + 1] = x[m + 1] - rr.y * x[q];
                }
                if (q & 1) x[q - 1] = x[q - 1] - col[q - 1] * x[q];
            }
        }
        if (c < n)
#pragma unroll
            for (int k = 0; k < CQH; ++k) X[CQH + k + c * CQLDM] = x[k];
    }
    __syncthreads();
    if (h == 0) {
        // rows below CQH: first the eliminations of rows CQH.. (final now),
        // in the same descending order, then their own
#pragma unroll
        for (int kk = LMAX - 1; kk >= CQH; --kk) {
            if (kk < rank) {
                const double xk = X[kk + cc * CQLDM];
                const double* col = Rp + kk * LMAX;
#pragma unroll
                for (int m = 0; m < CQH; m += 2) {
                    const double2 rr = *(const double2*)(col + m);
                    x[m] = x[m] - rr.x * xk;
                    x[m + 1] = x[m + 1] - rr.y * xk;
                }
            }
        }
#pragma unroll
        for (int kk = CQH - 1; kk >= 0; --kk) {
            if (kk < rank) {
                x[kk] *= rinv[kk];
                const double* col = Rp + kk * LMAX;
#pragma unroll
                for (int m = 0; m + 1 < kk; m += 2) {
                    const double2 rr = *(const double2*)(col + m);
                    x[m] = x[m] - rr.x * x[kk];
                    x[m + 1] = x[m + 1] - rr.y * x[kk];
                }
                if (kk & 1) x[kk - 1] = x[kk - 1] - col[kk - 1] * x[kk];
            }
        }
        if (c < n)
#pragma unroll
            for (int k = 0; k < CQH; ++k) X[k + c * CQLDM] = x[k];
    }
    __syncthreads();
    // W[perm[r]][cj] = X[r][cj] for r <= cj < rank, zero elsewhere; one
    // coalesced column per wave at a time
    for (int cj = w; cj < n; cj += CQ96T / 64)
        for (int r = lane; r < n; r += 64) {
            const int ip = iperm[r];
            D.W[r + (int64_t)cj * D.ldw] =
                (ip >= 0 && ip <= cj && cj < rank) ? X[ip + cj * CQLDM] : 0.0;
        }
    if (t < n) D.lam[t] = t < rank ? rdiag[t] : 0.0;
    if (t == 0 && D.sweeps) *D.sweeps = 0;
}

// ------------------------------------------------- column norms / normalise
// sigma_j = ||Y[:, j]||; U[:, j] = Y[:, j] / sigma_j, or 0 when
// sigma_j <= rtol * max_j sigma_j (numerical null direction).  Two launches
// over (problem, group of CN_COLS columns) workgroups, one wave per column,
// so even one problem spreads over many CUs: k_colnorm_part writes the
// raw norms, k_colnorm_scale takes the largest of them and normalises.  A
// lane keeps four partial sums over rows 256 apart (four loads in flight),
// adds them in a fixed order and the wave reduces with a fixed butterfly,
// so the result does not depend on scheduling.
constexpr int CN_COLS = 4;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__global__ __launch_bounds__(64 * CN_COLS) void k_colnorm_part(const ColDesc* __restrict__ descs,
                                                              int nd) {
    const ColDesc D = descs[blockIdx.x];
    const int lane = threadIdx.x & 63, j = blockIdx.y * CN_COLS + (threadIdx.x >> 6);
    if (j >= D.ncol || (D.skip && *D.skip)) return;
    const double* y = D.Y + (int64_t)j * D.ldy;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int i = lane;
    for (; i + 192 < D.nrow; i += 256) {
        const double y0 = y[i], y1 = y[i + 64], y2 = y[i + 128], y3 = y[i + 192];
        a0 = fma(y0, y0, a0); a1 = fma(y1, y1, a1); a2 = fma(y2, y2, a2); a3 = fma(y3, y3, a3);
    }
    for (; i < D.nrow; i += 64) a0 = fma(y[i], y[i], a0);
    const double acc = wave_sum((a0 + a1) + (a2 + a3));
    if (lane == 0) D.sigma[j] = sqrt(acc);   // raw norm
}

// Every workgroup of a problem reads all its raw norms for the largest one;
// a column zeroed as null is never the largest, so the workgroups that
// zero theirs do not change what the others read.
__global__ __launch_bounds__(64 * CN_COLS) void k_colnorm_scale(const ColDesc* __restrict__ descs,
                                                               int nd, double rtol) {
    __shared__ double wmax[CN_COLS];
    const ColDesc D = descs[blockIdx.x];
    if (blockIdx.y * CN_COLS >= D.ncol || (D.skip && *D.skip)) return;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, j = blockIdx.y * CN_COLS + w;
    double smax = 0.0;
    for (int c = t; c < D.ncol; c += 64 * CN_COLS) smax = fmax(smax, D.sigma[c]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) smax = fmax(smax, __shfl_xor(smax, o, 64));
    if (lane == 0) wmax[w] = smax;
    __syncthreads();
    smax = wmax[0];
    for (int q = 1; q < CN_COLS; ++q) smax = fmax(smax, wmax[q]);
    if (j >= D.ncol) return;
    const double sj = D.sigma[j];
    const bool live = sj > rtol * smax && sj > 0.0;
    const double inv = live ? 1.0 / sj : 0.0;
    const double* y = D.Y + (int64_t)j * D.ldy;
    double* u = D.U + (int64_t)j * D.ldu;
    int i = lane;
    for (; i + 192 < D.nrow; i += 256) {
        const double y0 = y[i], y1 = y[i + 64], y2 = y[i + 128], y3 = y[i + 192];
        u[i] = y0 * inv; u[i + 64] = y1 * inv; u[i + 128] = y2 * inv; u[i + 192] = y3 * inv;
    }
    for (; i < D.nrow; i += 64) u[i] = y[i] * inv;
    if (lane == 0 && !live) D.sigma[j] = 0.0;
}

// ------------------------------------------------ pivoted-row SVD gauge
// One 256-thread workgroup per SVD problem.  For each cluster of columns
// [i0, i0 + d) of U (m x k) the new basis is u_t = U_c c_t with
//   rn(r) = |U_c[r, :]|^2 - sum_{s<t} (U_c[r, :] . c_s)^2   (residual row norm)
//   r_t   = first row with rn(r) >= (1 - tie) * max_r rn(r)
//   c_t   = (U_c[r_t, :]^T - sum_{s<t} c_s (U_c[r_t, :] . c_s)) / sqrt(rn(r_t)),
// i.e. the normalised projection of the coordinate vector e_{r_t} onto the
// part of span(U_c) orthogonal to u_0..u_{t-1}; the c_t are the columns of
// the cluster's diagonal block of Q (a lone vector only gets its sign,
// exactly +-1).  Every column of Q is written once.
constexpr int CNT = 256;

__device__ __forceinline__ double canon_resid(const double* __restrict__ U, int ldu, int r, int i0,
                                              int d, const double* Cq, int tt) {
    double rn = 0.0;
    for (int j = 0; j < d; ++j) {
        const double x = U[r + (int64_t)(i0 + j) * ldu];
        rn += x * x;
    }
    for (int s = 0; s < tt; ++s) {
        double pj = 0.0;
        for (int j = 0; j < d; ++j) pj += U[r + (int64_t)(i0 + j) * ldu] * Cq[j + s * d];
        rn -= pj * pj;
    }
    return rn;
}

__global__ __launch_bounds__(CNT) void k_canon(const CanonDesc* __restrict__ descs, int nd,
                                               double tie) {
    extern __shared__ double lds[];
    __shared__ double rv[CNT / 64];
    __shared__ int ri[CNT / 64];
    __shared__ double s_max, s_rn;
    __shared__ int s_row;
    const CanonDesc D = descs[blockIdx.x];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, k = D.k;
    double* Cq = lds;                   // d x d, column-major, ld d
    double* xr = Cq + (size_t)k * k;    // U_c[r_t, :]
    double* pr = xr + k;                // U_c[r_t, :] . c_s
    for (int c = 0; c < D.nc; ++c) {
        const int i0 = D.cl[c], d = D.cl[c + 1] - i0;
        for (int tt = 0; tt < d; ++tt) {
            double best = -1.0;
            for (int r = t; r < D.m; r += CNT)
                best = fmax(best, canon_resid(D.U, D.ldu, r, i0, d, Cq, tt));
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) best = fmax(best, __shfl_xor(best, o, 64));
            if (lane == 0) rv[w] = best;
            __syncthreads();
            if (t == 0) {
                double m = rv[0];
                for (int q = 1; q < CNT / 64; ++q) m = fmax(m, rv[q]);
                s_max = m;
            }
            __syncthreads();
            const double thr = (1.0 - tie) * s_max;
            int first = 0x7fffffff;
            for (int r = t; r < D.m; r += CNT)
                if (canon_resid(D.U, D.ldu, r, i0, d, Cq, tt) >= thr) {
                    first = r;
                    break;
                }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) first = min(first, __shfl_xor(first, o, 64));
            if (lane == 0) ri[w] = first;
            __syncthreads();
            if (t == 0) {
                int m = ri[0];
                for (int q = 1; q < CNT / 64; ++q) m = min(m, ri[q]);
                s_row = m;
            }
            __syncthreads();
            const int r = s_row < D.m ? s_row : 0;     // NaN input: no row qualifies
            for (int j = t; j < d; j += CNT) xr[j] = D.U[r + (int64_t)(i0 + j) * D.ldu];
            __syncthreads();
            for (int s = t; s < tt; s += CNT) {
                double pj = 0.0;
                for (int j = 0; j < d; ++j) pj += xr[j] * Cq[j + s * d];
                pr[s] = pj;
            }
            __syncthreads();
            if (t == 0) {
                double rn = 0.0;
                for (int j = 0; j < d; ++j) rn += xr[j] * xr[j];
                for (int s = 0; s < tt; ++s) rn -= pr[s] * pr[s];
                s_rn = rn;
            }
            __syncthreads();
            const double inv = s_rn > 0.0 ? 1.0 / sqrt(s_rn) : 0.0;
            if (d == 1) {                 // a lone vector: only its sign
                if (t == 0) Cq[0] = xr[0] >= 0.0 ? 1.0 : -1.0;
                __syncthreads();
                break;
            }
            for (int j = t; j < d; j += CNT) {
                double cj = xr[j];
                for (int s = 0; s < tt; ++s) cj -= Cq[j + s * d] * pr[s];
                Cq[j + tt * d] = cj * inv;
            }
            __syncthreads();
        }
        for (int e = t; e < k * d; e += CNT) {
            const int jj = e / k, i = e - jj * k;
            D.Q[i + (int64_t)(i0 + jj) * k] = (i >= i0 && i < i0 + d) ? Cq[(i - i0) + jj * d] : 0.0;
        }
        __syncthreads();
    }
    for (int e = t; e < k * (k - D.knn); e += CNT) {
        const int jj = D.knn + e / k, i = e % k;
        D.Q[i + (int64_t)jj * k] = (i == jj) ? 1.0 : 0.0;
    }
}

// ------------------------------------------------------------- dot product
// grid (DOT_PARTS, problems): workgroup b sums the chunk [b c, (b+1) c) with
// c = ceil(n / DOT_PARTS); k_dot_final adds the parts in order
__global__ __launch_bounds__(256) void k_dot_part(const DotDesc* __restrict__ descs, int nd) {
    __shared__ double wpart[4];
    const DotDesc D = descs[blockIdx.y];
    const int t = threadIdx.x;
    const int64_t c = (D.n + DOT_PARTS - 1) / DOT_PARTS;
    const int64_t i0 = (int64_t)blockIdx.x * c, i1 = min(D.n, i0 + c);
    double acc = 0.0;
    for (int64_t i = i0 + t; i < i1; i += 256) acc += D.x[i] * D.y[i];
    acc = wave_sum(acc);
    if ((t & 63) == 0) wpart[t >> 6] = acc;
    __syncthreads();
    if (t == 0) D.part[blockIdx.x] = (wpart[0] + wpart[1]) + (wpart[2] + wpart[3]);
}

__global__ __launch_bounds__(64) void k_dot_final(const DotDesc* __restrict__ descs, int nd) {
    const DotDesc D = descs[blockIdx.x];
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int b = 0; b < DOT_PARTS; ++b) s += D.part[b];
        D.out[0] = s;
    }
}

// ------------------------------------------- eigenvalues of a real matrix
// One 64-lane wavefront per matrix, the matrix in LDS (column-major, ld n+1):
//  1. balancing by powers of two (Parlett & Reinsch): row i / f, column i * f
//     while that brings the off-diagonal row and column norms closer;
//  2. Householder reduction to upper Hessenberg form;
//  3. Francis double-shift QR on the Hessenberg matrix, eigenvalues only
//     (the EISPACK hqr scheme: deflation on a negligible subdiagonal entry,
//     a 2x2 block giving a real pair or a conjugate pair, exceptional shifts
//     after 10 and 20 iterations).
// Every lane runs the same control flow on values read from LDS (the wave
// sums are xor butterflies, identical in every lane); the row and column
// updates of each step are spread over the lanes.
constexpr int EIGT = 64;
constexpr int EIG_MAXITS = 60;   // QR iterations per eigenvalue before giving up

// The solver body; `hf` holds the n x n matrix (ld n+1) and n more values:
// LDS for n <= EIG_NMAX (k_eig_hqr), the problem's workspace in global
// memory above (k_eig_hqr_g: the same code, slower memory).
__device__ __forceinline__ void eig_hqr_body(const EigDesc& D, double* hf) {
    const int nf = D.n, t = threadIdx.x, ld = nf + 1;
    double* v = hf + (size_t)nf * ld;
#define HF(i, j) hf[(i) + (j) * ld]
    for (int e = t; e < nf * nf; e += EIGT) {
        const int j = e / nf, i = e - j * nf;
        HF(i, j) = D.A[i + (int64_t)j * D.lda];
    }
    __syncthreads();
    // 0. permutation (as LAPACK's balancing 'P'): a row with no off-diagonal
    //    entry inside the active window [ilo, ihi] is swapped to ihi, a
    //    column with none to ilo; its diagonal entry is an eigenvalue and
    //    the window shrinks (exact zeros stay exact)
    int ilo = 0, ihi = nf - 1;
    auto swap_idx = [&](int a, int b) {
        if (a == b) return;
        for (int j = t; j < nf; j += EIGT) {
            const double x = HF(a, j);
            HF(a, j) = HF(b, j);
            HF(b, j) = x;
        }
        __syncthreads();
        for (int i = t; i < nf; i += EIGT) {
            const double x = HF(i, a);
            HF(i, a) = HF(i, b);
            HF(i, b) = x;
        }
        __syncthreads();
    };
    for (bool moved = true; moved && ihi >= ilo;) {
        moved = false;
        for (int j = ihi; j >= ilo && !moved; --j) {
            int nz = 0;
            for (int i = ilo + t; i <= ihi; i += EIGT)
                if (i != j && HF(j, i) != 0.0) nz = 1;
            if (!__any(nz)) {
                swap_idx(j, ihi);
                if (t == 0) {
                    D.wr[ihi] = HF(ihi, ihi);
                    D.wi[ihi] = 0.0;
                }
                --ihi;
                moved = true;
            }
        }
        for (int j = ilo; j <= ihi && !moved; ++j) {
            int nz = 0;
            for (int i = ilo + t; i <= ihi; i += EIGT)
                if (i != j && HF(i, j) != 0.0) nz = 1;
            if (!__any(nz)) {
                swap_idx(j, ilo);
                if (t == 0) {
                    D.wr[ilo] = HF(ilo, ilo);
                    D.wi[ilo] = 0.0;
                }
                ++ilo;
                moved = true;
            }
        }
        __syncthreads();
    }
#undef HF
    // the rest works on the window: n x n at (ilo, ilo), eigenvalues to wr/wi + ilo
    const int n = ihi - ilo + 1;
    double* h = hf + (size_t)ilo * (ld + 1);
    double* const wro = D.wr + ilo;
    double* const wio = D.wi + ilo;
#define HH(i, j) h[(i) + (j) * ld]
    // 1. balancing
    for (int sweep = 0; sweep < 100; ++sweep) {
        bool done = true;
        for (int i = 0; i < n; ++i) {
            double c = 0.0, r = 0.0;
            for (int j = t; j < n; j += EIGT)
                if (j != i) {
                    c += fabs(HH(j, i));
                    r += fabs(HH(i, j));
                }
            c = wave_sum(c);
            r = wave_sum(r);
            if (c != 0.0 && r != 0.0) {
                double g = r * 0.5, f = 1.0;
                const double s0 = c + r;
                while (c < g) {
                    f *= 2.0;
                    c *= 4.0;
                }
                g = r * 2.0;
                while (c > g) {
                    f *= 0.5;
                    c *= 0.25;
                }
                if ((c + r) / f < 0.95 * s0) {
                    done = false;
                    const double gi = 1.0 / f;
                    for (int j = t; j < n; j += EIGT) HH(i, j) *= gi;
                    __syncthreads();
                    for (int j = t; j < n; j += EIGT) HH(j, i) *= f;
                }
            }
            __syncthreads();
        }
        if (done) break;
    }
    // 2. Hessenberg reduction
    for (int k = 0; k + 2 < n; ++k) {
        double ss = 0.0;
        for (int i = k + 2 + t; i < n; i += EIGT) ss += HH(i, k) * HH(i, k);
        ss = wave_sum(ss);
        if (ss == 0.0) continue;
        const double x0 = HH(k + 1, k);
        const double alpha = -copysign(sqrt(x0 * x0 + ss), x0);
        const double v0 = x0 - alpha, beta = 2.0 / (v0 * v0 + ss);
        for (int i = k + 1 + t; i < n; i += EIGT) v[i] = (i == k + 1) ? v0 : HH(i, k);
        __syncthreads();
        for (int j = k + t; j < n; j += EIGT) {
            double d = 0.0;
            for (int i = k + 1; i < n; ++i) d += v[i] * HH(i, j);
            d *= beta;
            for (int i = k + 1; i < n; ++i) HH(i, j) -= d * v[i];
        }
        __syncthreads();
        for (int i = t; i < n; i += EIGT) {
            double d = 0.0;
            for (int j = k + 1; j < n; ++j) d += HH(i, j) * v[j];
            d *= beta;
            for (int j = k + 1; j < n; ++j) HH(i, j) -= d * v[j];
        }
        __syncthreads();
        if (t == 0) HH(k + 1, k) = alpha;
        for (int i = k + 2 + t; i < n; i += EIGT) HH(i, k) = 0.0;
        __syncthreads();
    }
    // 3. double-shift QR
    double anorm = 0.0;
    for (int j = t; j < n; j += EIGT)
        for (int i = 0; i <= min(j + 1, n - 1); ++i) anorm += fabs(HH(i, j));
    anorm = wave_sum(anorm);
    int nn = n - 1, info = 0;
    double tsh = 0.0;
    double p = 0.0, q = 0.0, r = 0.0, s = 0.0, w = 0.0, x = 0.0, y = 0.0, z = 0.0;
    while (nn >= 0 && !info) {
        int its = 0, l = 0;
        do {
            for (l = nn; l >= 1; --l) {
                s = fabs(HH(l - 1, l - 1)) + fabs(HH(l, l));
                if (s == 0.0) s = anorm;
                if (fabs(HH(l, l - 1)) + s == s) break;
            }
            __syncthreads();
            if (l >= 1 && t == 0) HH(l, l - 1) = 0.0;
            __syncthreads();
            x = HH(nn, nn);
            if (l == nn) {
                if (t == 0) {
                    wro[nn] = x + tsh;
                    wio[nn] = 0.0;
                }
                --nn;
            } else {
                y = HH(nn - 1, nn - 1);
                w = HH(nn, nn - 1) * HH(nn - 1, nn);
                if (l == nn - 1) {
                    p = 0.5 * (y - x);
                    q = p * p + w;
                    z = sqrt(fabs(q));
                    x += tsh;
                    if (q >= 0.0) z = p + copysign(z, p);
                    if (t == 0) {
                        if (q >= 0.0) {
                            wro[nn - 1] = wro[nn] = x + z;
                            if (z != 0.0) wro[nn] = x - w / z;
                            wio[nn - 1] = wio[nn] = 0.0;
                        } else {
                            wro[nn - 1] = wro[nn] = x + p;
                            wio[nn - 1] = z;
                            wio[nn] = -z;
                        }
                    }
                    nn -= 2;
                } else {
                    if (its == EIG_MAXITS) {
                        info = 1;
                        break;
                    }
                    if (its == 10 || its == 20) {
                        tsh += x;
                        for (int i = t; i <= nn; i += EIGT) HH(i, i) -= x;
                        __syncthreads();
                        s = fabs(HH(nn, nn - 1)) + fabs(HH(nn - 1, nn - 2));
                        y = x = 0.75 * s;
                        w = -0.4375 * s * s;
                    }
                    ++its;
                    int m = nn - 2;
                    for (; m >= l; --m) {
                        z = HH(m, m);
                        r = x - z;
                        s = y - z;
                        p = (r * s - w) / HH(m + 1, m) + HH(m, m + 1);
                        q = HH(m + 1, m + 1) - z - r - s;
                        r = HH(m + 2, m + 1);
                        s = fabs(p) + fabs(q) + fabs(r);
                        p /= s;
                        q /= s;
                        r /= s;
                        if (m == l) break;
                        const double u = fabs(HH(m, m - 1)) * (fabs(q) + fabs(r));
                        const double vv = fabs(p) * (fabs(HH(m - 1, m - 1)) + fabs(z) +
                                                     fabs(HH(m + 1, m + 1)));
                        if (u + vv == vv) break;
                    }
                    __syncthreads();
                    for (int i = m + 2 + t; i <= nn; i += EIGT) {
                        HH(i, i - 2) = 0.0;
                        if (i != m + 2) HH(i, i - 3) = 0.0;
                    }
                    __syncthreads();
                    for (int k = m; k <= nn - 1; ++k) {
                        if (k != m) {
                            p = HH(k, k - 1);
                            q = HH(k + 1, k - 1);
                            r = (k != nn - 1) ? HH(k + 2, k - 1) : 0.0;
                            x = fabs(p) + fabs(q) + fabs(r);
                            if (x != 0.0) {
                                p /= x;
                                q /= x;
                                r /= x;
                            }
                        }
                        s = copysign(sqrt(p * p + q * q + r * r), p);
                        __syncthreads();
                        if (s != 0.0) {
                            if (t == 0) {
                                if (k == m) {
                                    if (l != m) HH(k, k - 1) = -HH(k, k - 1);
                                } else {
                                    HH(k, k - 1) = -s * x;
                                }
                            }
                            p += s;
                            x = p / s;
                            y = q / s;
                            z = r / s;
                            q /= p;
                            r /= p;
                            const bool three = k != nn - 1;
                            for (int j = k + t; j <= nn; j += EIGT) {
                                double pp = HH(k, j) + q * HH(k + 1, j);
                                if (three) {
                                    pp += r * HH(k + 2, j);
                                    HH(k + 2, j) -= pp * z;
                                }
                                HH(k + 1, j) -= pp * y;
                                HH(k, j) -= pp * x;
                            }
                            __syncthreads();
                            const int mmin = nn < k + 3 ? nn : k + 3;
                            for (int i = l + t; i <= mmin; i += EIGT) {
                                double pp = x * HH(i, k) + y * HH(i, k + 1);
                                if (three) {
                                    pp += z * HH(i, k + 2);
                                    HH(i, k + 2) -= pp * r;
                                }
                                HH(i, k + 1) -= pp * q;
                                HH(i, k) -= pp;
                            }
                            __syncthreads();
                        }
                    }
                }
            }
        } while (l < nn - 1 && !info);
    }
#undef HH
    if (t == 0) D.info[0] = info;
}


__global__ __launch_bounds__(EIGT) void k_eig_hqr(const EigDesc* __restrict__ descs, int nd) {
    extern __shared__ double lds[];
    const EigDesc D = descs[blockIdx.x];
    if (D.n > EIG_NMAX) return;
    eig_hqr_body(D, lds);
}

__global__ __launch_bounds__(EIGT) void k_eig_hqr_g(const EigDesc* __restrict__ descs, int nd) {
    const EigDesc D = descs[blockIdx.x];
    if (D.n <= EIG_NMAX) return;
    eig_hqr_body(D, D.work);
}

// stopping rule of the block iteration, one thread per problem
__global__ void k_conv(const ConvDesc* __restrict__ descs, int nd, double null_rtol,
                       double target, int min_iters, int max_iters, int stall_from) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nd) return;
    const ConvDesc D = descs[p];
    if (*D.done) return;
    const int it = *D.it + 1;
    const double* sv = D.sigma;
    int kk = 0;
    for (int i = 0; i < D.kmax; ++i)
        if (sv[i] > null_rtol * sv[0]) kk = i + 1;
    double r = 0.0;
    if (kk > 0 && sv[kk - 1] > 0.0) {
        r = sv[D.l - 1] / sv[kk - 1];
        if (D.squared) r *= r;
        r = fmin(r, 1.0);
    }
    double prod = *D.prod;
    if (it >= min_iters) prod *= r;
    *D.it = it;
    *D.prod = prod;
    if ((it >= min_iters && prod < target) || it >= max_iters) *D.done = 1;
    // a stall that will be factorised whole anyway: stop iterating once even
    // twice the current rate could not reach the target in the iterations left
    else if (D.whole_ok && stall_from > 0 && it >= stall_from && it >= min_iters &&
             prod * pow(r, 2.0 * (max_iters - it)) >= target)
        *D.done = 1;
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
// Squared Frobenius norm of each problem's nrow x ncol matrix, in a fixed
// order that depends on the matrix alone: grid (FROB_PARTS, problems); part
// q sums columns [q ncol / FROB_PARTS, (q + 1) ncol / FROB_PARTS), each
// thread its rows t + 256 u into four accumulators by u mod 4 (independent
// loads in flight), then a fixed tree; k_frob_final adds the parts in order.
__global__ __launch_bounds__(256) void k_frob_part(const ColDesc* __restrict__ descs,
                                                   double* __restrict__ partial) {
    __shared__ double red[256];
    const ColDesc D = descs[blockIdx.y];
    const int q = blockIdx.x;
    const int j0 = (int)((int64_t)q * D.ncol / FROB_PARTS);
    const int j1 = (int)((int64_t)(q + 1) * D.ncol / FROB_PARTS);
    gconst_double* Y = as_global(D.Y);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int j = j0; j < j1; ++j) {
        gconst_double* y = Y + (int64_t)j * D.ldy;
        int i = threadIdx.x;
        for (; i + 768 < D.nrow; i += 1024) {
            const double a = y[i], b = y[i + 256], c = y[i + 512], d = y[i + 768];
            acc[0] += a * a;
            acc[1] += b * b;
            acc[2] += c * c;
            acc[3] += d * d;
        }
        // fewer than 1024 rows left: at most three more
        if (i < D.nrow) acc[0] += y[i] * y[i];
        if (i + 256 < D.nrow) acc[1] += y[i + 256] * y[i + 256];
        if (i + 512 < D.nrow) acc[2] += y[i + 512] * y[i + 512];
    }
    const double s = block_sum_fixed((acc[0] + acc[1]) + (acc[2] + acc[3]), red);
    if (threadIdx.x == 0) partial[(int64_t)blockIdx.y * FROB_PARTS + q] = s;
}

__global__ __launch_bounds__(64) void k_frob_final(const ColDesc* __restrict__ descs, int nd,
                                                   const double* __restrict__ partial) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nd) return;
    double s = 0.0;
    for (int q = 0; q < FROB_PARTS; ++q) s += partial[(int64_t)p * FROB_PARTS + q];
    descs[p].sigma[0] = s;
}

// ================================================================ profiling
namespace {
struct ProfState {
    std::mutex mu;
    int level = 0;
    struct Rec { hipEvent_t a, b; int kind; double flops, bytes; };
    // a completed launch's times (ms): duration, start and end since `ref`
    struct Times { float ms, t0, t1; };
    std::vector<Rec> recs;
    bool draining = false;   // one thread files a batch of records unlocked
    std::condition_variable idle;   // signalled when that batch is filed
    // with mu held (through lk): wait until no batch is being filed unlocked
    void wait_idle(std::unique_lock<std::mutex>& lk) {
        idle.wait(lk, [this] { return !draining; });
    }
    std::vector<hipEvent_t> spare;
    ProfTotals tot[PK_COUNT];
    // start / end of every filed launch (ms since `ref`), per kind: their
    // union is the kind's device-busy time (launches on different streams
    // overlap)
    hipEvent_t ref = nullptr;
    std::vector<std::pair<float, float>> spans[PK_COUNT];
    // device counters of the GEMM flops that ran (FP64, split), per device
    std::map<int, unsigned long long*> flop_ctr;
    unsigned long long* counters(int dev) {   // with mu held
        auto it = flop_ctr.find(dev);
        if (it != flop_ctr.end()) return it->second;
        unsigned long long* p = nullptr;
        TNR_HIP(hipMalloc(&p, 2 * sizeof(unsigned long long)));
        TNR_HIP(hipMemset(p, 0, 2 * sizeof(unsigned long long)));
        flop_ctr[dev] = p;
        return p;
    }
    hipEvent_t ev() {   // with mu held
        if (!spare.empty()) { hipEvent_t e = spare.back(); spare.pop_back(); return e; }
        hipEvent_t e;
        TNR_HIP(hipEventCreate(&e));
        return e;
    }
    Times times(const Rec& r) const {   // r.b has completed; no lock needed
        Times x{0, 0, 0};
        TNR_HIP(hipEventElapsedTime(&x.ms, r.a, r.b));
        TNR_HIP(hipEventElapsedTime(&x.t0, ref, r.a));
        TNR_HIP(hipEventElapsedTime(&x.t1, ref, r.b));
        return x;
    }
    void file_times(const Rec& r, const Times& x) {   // with mu held
        tot[r.kind].launches++;
        tot[r.kind].ms += x.ms;
        tot[r.kind].flops += r.flops;
        tot[r.kind].bytes += r.bytes;
        spans[r.kind].emplace_back(x.t0, x.t1);
        spare.push_back(r.a);
        spare.push_back(r.b);
    }
    void file(const Rec& r) { file_times(r, times(r)); }   // with mu held; r.b has completed
    void drain() {      // with mu held; waits for every pending launch
        for (auto& r : recs) {
            TNR_HIP(hipEventSynchronize(r.b));
            file(r);
        }
        recs.clear();
    }
};
ProfState& prof() {
    static ProfState p;
    return p;
}
}  // namespace

void prof_enable(int level) {
    ProfState& p = prof();
    std::unique_lock<std::mutex> lk(p.mu);
    p.wait_idle(lk);
    p.drain();
    p.level = level;
    for (auto& t : p.tot) t = ProfTotals();
    for (auto& v : p.spans) v.clear();
    if (level) {
        // the reference point of the launch spans, and zeroed flop counters
        // (outside any timed region: this synchronises the device)
        TNR_HIP(hipDeviceSynchronize());
        if (!p.ref) TNR_HIP(hipEventCreate(&p.ref));
        TNR_HIP(hipEventRecord(p.ref, 0));
        TNR_HIP(hipEventSynchronize(p.ref));
        for (auto& kv : p.flop_ctr) {
            int dev = 0;
            TNR_HIP(hipGetDevice(&dev));
            TNR_HIP(hipSetDevice(kv.first));
            TNR_HIP(hipMemset(kv.second, 0, 2 * sizeof(unsigned long long)));
            TNR_HIP(hipSetDevice(dev));
        }
    }
}

unsigned long long* prof_flop_counter(bool split32) {
    ProfState& p = prof();
    std::lock_guard<std::mutex> lk(p.mu);
    if (!p.level) return nullptr;
    int dev = 0;
    TNR_HIP(hipGetDevice(&dev));
    return p.counters(dev) + (split32 ? 1 : 0);
}

bool prof_enabled() { return prof_level() > 0; }

int prof_level() {
    ProfState& p = prof();
    std::lock_guard<std::mutex> lk(p.mu);
    return p.level;
}

// Launches may come from several host threads (one stream each): the
// events live in the scope object and the record is filed at its end.
ProfScope::ProfScope(hipStream_t s, int k, double f, double b)
    : st(s), kind(k), active(false), flops(f), bytes(b) {
    ProfState& p = prof();
    std::lock_guard<std::mutex> lk(p.mu);
    if (!p.level) return;
    ea = p.ev();
    eb = p.ev();
    active = true;
}

void ProfScope::bracket() {
    if (active && mode == 0) {
        TNR_HIP(hipEventRecord(ea, st));
        mode = 1;
    }
}

void ProfScope::end() {
    if (active && mode == 1 && !ended) {
        (void)hipEventRecord(eb, st);
        ended = true;
    }
}

ProfScope::~ProfScope() {
    if (!active) return;
    end();
    ProfState& p = prof();
    std::unique_lock<std::mutex> lk(p.mu);
    if (mode == 0) {   // nothing was launched
        p.spare.push_back(ea);
        p.spare.push_back(eb);
        return;
    }
    p.recs.push_back(ProfState::Rec{ea, eb, kind, flops, bytes});
    if (p.recs.size() <= 4096 || p.draining) return;
    // File the completed launches.  The event queries take a few ms for
    // thousands of records, so they run without the lock: the other lanes'
    // threads keep launching meanwhile.
    std::vector<ProfState::Rec> batch;
    batch.swap(p.recs);
    p.draining = true;
    lk.unlock();
    std::vector<std::pair<ProfState::Rec, ProfState::Times>> done;
    std::vector<ProfState::Rec> pending;
    try {
        for (const auto& r : batch) {
            if (hipEventQuery(r.b) == hipSuccess) done.emplace_back(r, p.times(r));
            else pending.push_back(r);
        }
    } catch (...) {
        pending = batch;   // try again later
        done.clear();
    }
    lk.lock();
    for (const auto& d : done) p.file_times(d.first, d.second);
    p.recs.insert(p.recs.end(), pending.begin(), pending.end());
    p.draining = false;
    p.idle.notify_all();
}

ProfTotals prof_read(int kind) {
    ProfState& p = prof();
    std::unique_lock<std::mutex> lk(p.mu);
    p.wait_idle(lk);
    p.drain();
    if (kind < 0 || kind >= PK_COUNT) return ProfTotals();
    ProfTotals t = p.tot[kind];
    // union of the launch spans
    auto v = p.spans[kind];
    std::sort(v.begin(), v.end());
    double busy = 0;
    float a = 0, b = -1;
    for (auto& s : v) {
        if (s.first > b) {
            if (b > a) busy += b - a;
            a = s.first;
            b = s.second;
        } else if (s.second > b) {
            b = s.second;
        }
    }
    if (b > a) busy += b - a;
    t.busy_ms = busy;
    if (kind == PK_GEMM_F64 || kind == PK_GEMM_SPLIT32) {
        // flops of the problems that ran, from the device counters
        TNR_HIP(hipDeviceSynchronize());
        double ran = 0;
        for (auto& kv : p.flop_ctr) {
            unsigned long long c[2] = {0, 0};
            int dev = 0;
            TNR_HIP(hipGetDevice(&dev));
            TNR_HIP(hipSetDevice(kv.first));
            TNR_HIP(hipMemcpy(c, kv.second, sizeof(c), hipMemcpyDeviceToHost));
            TNR_HIP(hipSetDevice(dev));
            ran += (double)c[kind == PK_GEMM_SPLIT32 ? 1 : 0];
        }
        t.flops_planned = t.flops;
        t.flops = ran;
    }
    return t;
}

// =================================================================== launch
static int grid_for(int64_t total, int cap = 4096) {
    int64_t g = (total + 255) / 256;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

void launch_copy(hipStream_t st, const CopyDesc* d_descs, int nd, int64_t chunks,
                 int64_t elems) {
    if (chunks <= 0 || nd <= 0) return;
    ProfScope ps(st, PK_COPY, 0.0, 16.0 * (double)elems);
    const int grid = (int)std::min<int64_t>(chunks, 8192);
    TNR_LAUNCH(ps, k_copy, dim3(grid), dim3(256), 0, st, d_descs, nd, chunks);
    TNR_HIP(hipGetLastError());
}

void launch_transpose(hipStream_t st, const TransDesc* d_descs, int nd, int64_t tiles,
                      int64_t elems) {
    if (tiles <= 0 || nd <= 0) return;
    ProfScope ps(st, PK_COPY, 0.0, 16.0 * (double)elems);
    const int grid = (int)std::min<int64_t>(tiles, 4096);
    TNR_LAUNCH(ps, k_transpose, dim3(grid), dim3(256), 0, st, d_descs, nd, tiles);
    TNR_HIP(hipGetLastError());
}

void launch_gemm(hipStream_t st, const GemmDesc* d_descs, int nd, int64_t tiles, bool split32,
                 double flops, int shape, bool dma) {
    if (tiles <= 0 || nd <= 0) return;
    ProfScope ps(st, split32 ? PK_GEMM_SPLIT32 : PK_GEMM_F64, flops, 0.0);
    unsigned long long* fc = ps.active ? prof_flop_counter(split32) : nullptr;
    if (split32)
        TNR_LAUNCH(ps, k_gemm_split32, dim3((unsigned)tiles), dim3(256), 0, st, d_descs, nd, fc);
    else if (shape == GEMM_128x128 && dma)
        TNR_LAUNCH(ps, (k_gemm_dl<128, 128, 4, 2, 2, 3, 16>), dim3((unsigned)tiles), dim3(512), 0, st,
                           d_descs, nd, fc);
    else if (shape == GEMM_128x64 && dma)
        TNR_LAUNCH(ps, (k_gemm_dl<128, 64, 4, 2, 4, 3, 16>), dim3((unsigned)tiles), dim3(512), 0, st,
                           d_descs, nd, fc);
    else if (shape == GEMM_64x64 && dma)
        TNR_LAUNCH(ps, (k_gemm_dl<64, 64, 2, 2, 3, 3, 16>), dim3((unsigned)tiles), dim3(256), 0, st,
                           d_descs, nd, fc);
    else if (shape == GEMM_128x128)
        TNR_LAUNCH(ps, (k_gemm_f64<128, 128, 4, 2, 16>), dim3((unsigned)tiles), dim3(512), 0,
                           st, d_descs, nd, fc);
    else if (shape == GEMM_128x64)
        TNR_LAUNCH(ps, (k_gemm_f64<128, 64, 4, 2, 16>), dim3((unsigned)tiles), dim3(512), 0,
                           st, d_descs, nd, fc);
    else if (shape == GEMM_64x64_K32)
        TNR_LAUNCH(ps, (k_gemm_f64<64, 64, 2, 2, 32>), dim3((unsigned)tiles), dim3(256), 0,
                           st, d_descs, nd, fc);
    else
        TNR_LAUNCH(ps, (k_gemm_f64<64, 64, 2, 2, 16>), dim3((unsigned)tiles), dim3(256), 0,
                           st, d_descs, nd, fc);
    TNR_HIP(hipGetLastError());
}

void launch_gemm_reduce(hipStream_t st, const GemmDesc* d_descs, int nd, const int64_t* d_starts,
                        int64_t total) {
    if (total <= 0 || nd <= 0) return;
    ProfScope ps(st, PK_GEMM_REDUCE, 0.0, 0.0);
    TNR_LAUNCH(ps, k_gemm_reduce, dim3(grid_for(total, 8192)), dim3(256), 0, st, d_descs, nd,
                       d_starts, total);
    TNR_HIP(hipGetLastError());
}

void launch_trace_normalize(hipStream_t st, const TraceDesc* d_descs, const TraceProb* d_probs,
                            int nprob, int64_t max_size) {
    if (nprob <= 0) return;
    ProfScope ps(st, PK_TRACE, 0.0, 0.0);
    ps.bracket();
    TNR_LAUNCH(ps, k_trace_partial, dim3(TRACE_PARTS, nprob), dim3(256), 0, st, d_descs,
                       d_probs);
    TNR_HIP(hipGetLastError());
    int gx = grid_for(max_size, 1024);
    TNR_LAUNCH(ps, k_normalize, dim3(gx, nprob), dim3(256), 0, st, d_probs);
    TNR_HIP(hipGetLastError());
}

size_t jacobi_lds_bytes(int n) {
    const size_t np = n + (n & 1);
    return 2 * np * (np + 1) * sizeof(double) + (LMAX / 2) * sizeof(double) * 4 +
           LMAX * sizeof(int);
}

void launch_jacobi(hipStream_t st, const JacobiDesc* d_descs, int nd, int nmax, double tol,
                   int max_sweeps) {
    if (nd <= 0) return;
    // once per process (thread-safe static initialisation)
    static const hipError_t attr = hipFuncSetAttribute(
        (const void*)k_jacobi, hipFuncAttributeMaxDynamicSharedMemorySize,
        (int)jacobi_lds_bytes(LMAX));
    TNR_HIP(attr);
    ProfScope ps(st, PK_JACOBI, 0.0, 0.0);
    TNR_LAUNCH(ps, k_jacobi, dim3(nd), dim3(JT), jacobi_lds_bytes(nmax), st, d_descs, nd, tol,
                       max_sweeps);
    TNR_HIP(hipGetLastError());
    if (ps.active && prof_level() >= 2) {
        // profiling only: the work a launch did depends on the sweeps each
        // problem took, read back after the launch (the timing events are
        // already recorded)
        ps.end();
        TNR_HIP(hipStreamSynchronize(st));
        std::vector<JacobiDesc> hd(nd);
        TNR_HIP(hipMemcpy(hd.data(), d_descs, nd * sizeof(JacobiDesc), hipMemcpyDeviceToHost));
        double fl = 0.0, by = 0.0;
        for (const JacobiDesc& d : hd) {
            int sk = 0, sw = 0;
            if (d.skip) TNR_HIP(hipMemcpy(&sk, d.skip, sizeof(int), hipMemcpyDeviceToHost));
            if (sk) continue;
            if (d.sweeps) TNR_HIP(hipMemcpy(&sw, d.sweeps, sizeof(int), hipMemcpyDeviceToHost));
            const double n = d.n;
            // per sweep: (n-1) rounds of n/2 rotations, each 18n flops on
            // G (two rows, two columns) and W (two columns)
            fl += (double)sw * (n - 1.0) * (n / 2.0) * 18.0 * n;
            by += (2.0 * n * n + n) * sizeof(double);
        }
        ps.flops = fl;
        ps.bytes = by;
    }
}


// Problems of at most CQ64 columns run on k_cholqr64, larger ones on
// k_cholqr96 (each kernel skips the other's), so the kernel a problem gets
// depends on its own size only, never on the rest of its launch group.
void launch_cholqr(hipStream_t st, const JacobiDesc* d_descs, int nd, int nmin, int nmax,
                   double null_rel2) {
    if (nd <= 0) return;
    static const hipError_t attr = hipFuncSetAttribute(
        (const void*)k_cholqr96, hipFuncAttributeMaxDynamicSharedMemorySize,
        (int)cholqr96_lds_bytes());
    TNR_HIP(attr);
    static const hipError_t attr64 = hipFuncSetAttribute(
        (const void*)k_cholqr64, hipFuncAttributeMaxDynamicSharedMemorySize,
        (int)cholqr64_lds_bytes());
    TNR_HIP(attr64);
    if (nmin <= CQ64) {
        ProfScope ps(st, PK_JACOBI, 0.0, 0.0);
        TNR_LAUNCH(ps, k_cholqr64, dim3(nd), dim3(CQ64), cholqr64_lds_bytes(), st, d_descs, nd,
                   null_rel2);
        TNR_HIP(hipGetLastError());
    }
    if (nmax > CQ64) {
        ProfScope ps(st, PK_JACOBI, 0.0, 0.0);
        TNR_LAUNCH(ps, k_cholqr96, dim3(nd), dim3(CQ96T), cholqr96_lds_bytes(), st, d_descs, nd,
                   null_rel2);
        TNR_HIP(hipGetLastError());
    }
}

void launch_lincomb(hipStream_t st, const LinDesc* d_descs, int nd, int64_t total) {
    if (total <= 0 || nd <= 0) return;
    ProfScope ps(st, PK_MISC, 0.0, 0.0);
    TNR_LAUNCH(ps, k_lincomb, dim3(grid_for(total, 8192)), dim3(256), 0, st, d_descs, nd,
                       total);
    TNR_HIP(hipGetLastError());
}

void launch_colnorm(hipStream_t st, const ColDesc* d_descs, int nd, int ncol_max, double rtol) {
    if (nd <= 0 || ncol_max <= 0) return;
    const dim3 grid(nd, (ncol_max + CN_COLS - 1) / CN_COLS);
    ProfScope ps(st, PK_COLNORM, 0.0, 0.0);
    ps.bracket();
    TNR_LAUNCH(ps, k_colnorm_part, grid, dim3(64 * CN_COLS), 0, st, d_descs, nd);
    TNR_HIP(hipGetLastError());
    TNR_LAUNCH(ps, k_colnorm_scale, grid, dim3(64 * CN_COLS), 0, st, d_descs, nd, rtol);
    TNR_HIP(hipGetLastError());
}

void launch_canon(hipStream_t st, const CanonDesc* d_descs, int nd, int kmax, double tie) {
    if (nd <= 0) return;
    TNR_REQUIRE(kmax >= 1 && kmax <= LMAX, TNR_ERR_INTERNAL, "gauge block above LMAX");
    static const hipError_t attr = hipFuncSetAttribute(
        (const void*)k_canon, hipFuncAttributeMaxDynamicSharedMemorySize,
        (int)(((size_t)LMAX * LMAX + 2 * (size_t)LMAX) * sizeof(double)));
    TNR_HIP(attr);
    ProfScope ps(st, PK_MISC, 0.0, 0.0);
    const size_t shm = ((size_t)kmax * kmax + 2 * (size_t)kmax) * sizeof(double);
    TNR_LAUNCH(ps, k_canon, dim3(nd), dim3(CNT), shm, st, d_descs, nd, tie);
    TNR_HIP(hipGetLastError());
}

void launch_dot(hipStream_t st, const DotDesc* d_descs, int nd) {
    if (nd <= 0) return;
    TNR_REQUIRE(nd <= 65535, TNR_ERR_INTERNAL, "too many dot products in one launch");
    ProfScope ps(st, PK_MISC, 0.0, 0.0);
    ps.bracket();
    TNR_LAUNCH(ps, k_dot_part, dim3(DOT_PARTS, nd), dim3(256), 0, st, d_descs, nd);
    TNR_HIP(hipGetLastError());
    TNR_LAUNCH(ps, k_dot_final, dim3(nd), dim3(64), 0, st, d_descs, nd);
    TNR_HIP(hipGetLastError());
}

// Problems up to EIG_NMAX run with the matrix in LDS, larger ones (up to
// EIG_NMAX_G, each with its workspace of eig_work_doubles(n)) in global
// memory; each kernel skips the other's, so a problem's kernel depends on
// its own size only.
void launch_eig(hipStream_t st, const EigDesc* d_descs, int nd, int nmin, int nmax) {
    if (nd <= 0) return;
    TNR_REQUIRE(nmin >= 1 && nmax <= EIG_NMAX_G, TNR_ERR_INTERNAL, "eigenproblem above EIG_NMAX_G");
    auto bytes = [](int n) { return (size_t)eig_work_doubles(n) * sizeof(double); };
    static const hipError_t attr = hipFuncSetAttribute(
        (const void*)k_eig_hqr, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes(EIG_NMAX));
    TNR_HIP(attr);
    if (nmin <= EIG_NMAX) {
        ProfScope ps(st, PK_MISC, 0.0, 0.0);
        TNR_LAUNCH(ps, k_eig_hqr, dim3(nd), dim3(EIGT), bytes(std::min(nmax, EIG_NMAX)), st, d_descs,
                   nd);
        TNR_HIP(hipGetLastError());
    }
    if (nmax > EIG_NMAX) {
        ProfScope ps(st, PK_MISC, 0.0, 0.0);
        TNR_LAUNCH(ps, k_eig_hqr_g, dim3(nd), dim3(EIGT), 0, st, d_descs, nd);
        TNR_HIP(hipGetLastError());
    }
}

void launch_symmetrize(hipStream_t st, const ColDesc* d_descs, int nd, int nmax) {
    if (nd <= 0) return;
    int64_t tot = (int64_t)nmax * nmax;
    ProfScope ps(st, PK_MISC, 0.0, 0.0);
    TNR_LAUNCH(ps, k_symmetrize, dim3(grid_for(tot, 1024), nd), dim3(256), 0, st, d_descs, nd);
    TNR_HIP(hipGetLastError());
}

void launch_conv(hipStream_t st, const ConvDesc* d_descs, int nd, double null_rtol, double target,
                 int min_iters, int max_iters, int stall_from) {
    if (nd <= 0) return;
    ProfScope ps(st, PK_MISC, 0.0, 0.0);
    TNR_LAUNCH(ps, k_conv, dim3((nd + 63) / 64), dim3(64), 0, st, d_descs, nd, null_rtol, target,
                       min_iters, max_iters, stall_from);
    TNR_HIP(hipGetLastError());
}

void launch_frob(hipStream_t st, const ColDesc* d_descs, int nd, double* partial) {
    if (nd <= 0) return;
    ProfScope ps(st, PK_MISC, 0.0, 0.0);
    ps.bracket();
    TNR_LAUNCH(ps, k_frob_part, dim3(FROB_PARTS, nd), dim3(256), 0, st, d_descs, partial);
    TNR_LAUNCH(ps, k_frob_final, dim3((nd + 63) / 64), dim3(64), 0, st, d_descs, nd, partial);
    TNR_HIP(hipGetLastError());
}

void launch_randfill(hipStream_t st, const FillDesc* d_descs, int nd, int64_t total) {
    if (total <= 0 || nd <= 0) return;
    ProfScope ps(st, PK_FILL, 0.0, 0.0);
    TNR_LAUNCH(ps, k_randfill, dim3(grid_for(total, 4096)), dim3(256), 0, st, d_descs, nd, total);
    TNR_HIP(hipGetLastError());
}

void launch_gather(hipStream_t st, const GatherDesc* d_descs, int nd, int64_t total) {
    if (total <= 0 || nd <= 0) return;
    ProfScope ps(st, PK_GATHER, 0.0, 16.0 * (double)total);
    TNR_LAUNCH(ps, k_gather, dim3(grid_for(total, 8192)), dim3(256), 0, st, d_descs, nd, total);
    TNR_HIP(hipGetLastError());
}

}  // namespace tnr
