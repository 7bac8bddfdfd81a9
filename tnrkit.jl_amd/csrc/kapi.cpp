// Kernel-level entry points (include/tnr_kernels.h).
#include <algorithm>
#include <cstring>
#include <string>

#include "schemes.h"
#include "solver.h"
#include "../../include/tnr_kernels.h"

using namespace tnr;

namespace {
template <class F>
int kguard(F&& f) {
    try {
        return f();
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return TNR_ERR_INTERNAL;
    }
}

Buf up(Ctx& ctx, const double* h, size_t n) {
    Buf b(&ctx.pool, std::max<size_t>(n, 1) * 8);
    if (n) TNR_HIP(hipMemcpyAsync(b.d(), h, n * 8, hipMemcpyHostToDevice, ctx.stream));
    return b;
}
}  // namespace

extern "C" {

int tnr_kernel_gemm(int M, int N, int K, const double* A, int lda, int ta, const double* B,
                    int ldb, int tb, double* C, int ldc, double alpha, double beta, int mode,
                    int repeat, double* ms_out) {
    return kguard([&] {
        const int kind = mode & 7, split = (mode & 8) ? 0 : (mode & 16) ? 1 : -1;
        const bool dma = (mode & 32) == 0;
        TNR_REQUIRE(M > 0 && N > 0 && K >= 0 && A && B && C && repeat >= 1 && kind <= 5 &&
                        (mode & ~63) == 0 && (mode & 24) != 24,
                    TNR_ERR_ARG, "bad GEMM arguments");
        TNR_REQUIRE(lda >= (ta ? std::max(K, 1) : M) && ldb >= (tb ? N : std::max(K, 1)) && ldc >= M,
                    TNR_ERR_ARG, "leading dimension too small");
        Ctx ctx;
        const size_t na = (size_t)lda * (ta ? M : K), nb = (size_t)ldb * (tb ? K : N),
                     nc = (size_t)ldc * N;
        Buf dA = up(ctx, A, na), dB = up(ctx, B, nb), dC = up(ctx, C, nc);
        Buf dC0 = up(ctx, C, nc);
        std::vector<GemmJob> j{GemmJob{dA.d(), dB.d(), dC.d(), M, N, K, lda, ldb, ldc, ta, tb}};
        j[0].alpha = alpha;
        j[0].beta = beta;
        hipEvent_t e0, e1;
        TNR_HIP(hipEventCreate(&e0));
        TNR_HIP(hipEventCreate(&e1));
        float ms = 0;
        for (int r = 0; r < repeat; ++r) {
            if (r > 0 && beta != 0.0)
                TNR_HIP(hipMemcpyAsync(dC.d(), dC0.d(), nc * 8, hipMemcpyDeviceToDevice, ctx.stream));
            if (r == repeat - 1 || r == 0) TNR_HIP(hipEventRecord(e0, ctx.stream));
            gemm_grouped(ctx, j, kind == 1, kind >= 2 ? kind - 2 : -1, split, dma);
            if (r == repeat - 1 || r == 0) {
                TNR_HIP(hipEventRecord(e1, ctx.stream));
                TNR_HIP(hipEventSynchronize(e1));
                float t = 0;
                TNR_HIP(hipEventElapsedTime(&t, e0, e1));
                if (r == repeat - 1) ms = t;
            }
        }
        ctx.fetch(C, dC.d(), nc * 8);
        if (ms_out) *ms_out = ms;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return TNR_OK;
    });
}

int tnr_autotune_enable(int on) {
    return kguard([&] {
        gemm_autotune_enable(on);
        return TNR_OK;
    });
}

int tnr_solver_config(const char* spec) {
    return kguard([&] {
        set_solver_cfg(spec);
        return TNR_OK;
    });
}

int tnr_kernel_svd(int m, int n, const double* M, int k, double* U, double* s, double* V,
                   int* iters) {
    return kguard([&] {
        TNR_REQUIRE(m > 0 && n > 0 && k > 0 && k <= std::min(m, n) && M && U && s && V,
                    TNR_ERR_ARG, "bad SVD arguments");
        Ctx ctx;
        Buf dM = up(ctx, M, (size_t)m * n);
        std::vector<SvdIn> in{SvdIn{dM.d(), m, n, m, k, 12345}};
        std::vector<SvdOut> out;
        svd_grouped(ctx, in, out);
        std::vector<double> hu((size_t)out[0].ldu * k), hv((size_t)out[0].ldv * k);
        ctx.fetch(hu.data(), out[0].U.d(), hu.size() * 8);
        ctx.fetch(hv.data(), out[0].V.d(), hv.size() * 8);
        for (int j = 0; j < k; ++j) {
            std::memcpy(U + (size_t)j * m, hu.data() + (size_t)j * out[0].ldu, m * 8);
            std::memcpy(V + (size_t)j * n, hv.data() + (size_t)j * out[0].ldv, n * 8);
        }
        std::memcpy(s, out[0].s_host.data(), k * 8);
        if (iters) *iters = out[0].iters;
        TNR_REQUIRE(out[0].status == TNR_OK, out[0].status, "the truncated SVD failed");
        return TNR_OK;
    });
}

int tnr_kernel_pivoted_gauge(int m, int k, const double* U, const double* s, double* Q) {
    return kguard([&] {
        TNR_REQUIRE(m > 0 && k > 0 && k <= std::min(m, LMAX) && U && s && Q, TNR_ERR_ARG,
                    "bad gauge arguments");
        Ctx ctx;
        Buf dU = up(ctx, U, (size_t)m * k);
        Buf dQ(&ctx.pool, (size_t)k * k * 8);
        std::vector<int> cl;
        int knn = 0;
        const int nc = gauge_clusters(s, k, cl, knn);
        if (cl.empty()) cl.push_back(0);
        const int* dcl = static_cast<const int*>(
            ctx.stage->put(ctx.stream, cl.data(), cl.size() * sizeof(int)));
        std::vector<CanonDesc> cd{CanonDesc{dU.d(), dQ.d(), dcl, m, m, k, knn, nc}};
        CanonDesc* dcd = ctx.stage->put(ctx.stream, cd);
        launch_canon(ctx.stream, dcd, 1, k, kGaugeTie);
        ctx.fetch(Q, dQ.d(), (size_t)k * k * 8);
        return TNR_OK;
    });
}

int tnr_kernel_eig(int count, int n, const double* A, double* wr, double* wi) {
    return kguard([&] {
        TNR_REQUIRE(count > 0 && n > 0 && n <= EIG_NMAX_G && A && wr && wi, TNR_ERR_ARG,
                    "bad eigenvalue arguments");
        Ctx ctx;
        Buf dA = up(ctx, A, (size_t)count * n * n);
        Buf dr(&ctx.pool, (size_t)count * n * 8), di(&ctx.pool, (size_t)count * n * 8);
        Buf dinfo(&ctx.pool, (size_t)count * sizeof(int));
        Buf work;
        if (n > EIG_NMAX) work = Buf(&ctx.pool, (size_t)count * eig_work_doubles(n) * 8);
        std::vector<EigDesc> d;
        for (int i = 0; i < count; ++i)
            d.push_back(EigDesc{dA.d() + (size_t)i * n * n, dr.d() + (size_t)i * n,
                                di.d() + (size_t)i * n, dinfo.as<int>() + i, n, n,
                                n > EIG_NMAX ? work.d() + (size_t)i * eig_work_doubles(n)
                                             : nullptr});
        EigDesc* dd = ctx.stage->put(ctx.stream, d);
        launch_eig(ctx.stream, dd, count, n, n);
        std::vector<int> info(count);
        ctx.fetch(wr, dr.d(), (size_t)count * n * 8);
        ctx.fetch(wi, di.d(), (size_t)count * n * 8);
        ctx.fetch(info.data(), dinfo.d(), (size_t)count * sizeof(int));
        for (int x : info)
            TNR_REQUIRE(x == 0, TNR_ERR_NOCONV, "QR iteration did not converge");
        return TNR_OK;
    });
}

int tnr_kernel_eigh(int n, const double* M, int k, double* U, double* lam, int* iters) {
    return kguard([&] {
        TNR_REQUIRE(n > 0 && k > 0 && k <= n && M && U && lam, TNR_ERR_ARG, "bad eigh arguments");
        Ctx ctx;
        Buf dM = up(ctx, M, (size_t)n * n);
        std::vector<EighIn> in{EighIn{dM.d(), n, n, k, 12345}};
        std::vector<EighOut> out;
        eigh_grouped(ctx, in, out);
        std::vector<double> hu((size_t)n * out[0].ncols);
        ctx.fetch(hu.data(), out[0].U.d(), hu.size() * 8);
        std::memcpy(U, hu.data(), (size_t)n * k * 8);
        std::memcpy(lam, out[0].lam_host.data(), k * 8);
        if (iters) *iters = out[0].iters;
        TNR_REQUIRE(out[0].status == TNR_OK, out[0].status, "the truncated eigensolve failed");
        return TNR_OK;
    });
}

int tnr_kernel_jacobi(int count, int n, const double* G, double* W, double* lam, int* sweeps,
                      double* ms_out) {
    return kguard([&] {
        TNR_REQUIRE(count > 0 && n > 0 && n <= LMAX && G && W && lam, TNR_ERR_ARG,
                    "bad Jacobi arguments");
        Ctx ctx;
        const size_t nn = (size_t)n * n;
        Buf dG = up(ctx, G, nn * count);
        Buf dW(&ctx.pool, nn * count * 8), dL(&ctx.pool, (size_t)n * count * 8);
        Buf dS(&ctx.pool, (size_t)count * sizeof(int));
        std::vector<JacobiDesc> jd;
        for (int i = 0; i < count; ++i)
            jd.push_back(JacobiDesc{dG.d() + nn * i, dW.d() + nn * i, dL.d() + (size_t)n * i,
                                    dS.as<int>() + i, n, n, n});
        JacobiDesc* d = ctx.stage->put(ctx.stream, jd);
        hipEvent_t e0, e1;
        TNR_HIP(hipEventCreate(&e0));
        TNR_HIP(hipEventCreate(&e1));
        ctx.sync();
        TNR_HIP(hipEventRecord(e0, ctx.stream));
        launch_jacobi(ctx.stream, d, count, n, solver_cfg().jacobi_tol, solver_cfg().jacobi_sweeps);
        TNR_HIP(hipEventRecord(e1, ctx.stream));
        TNR_HIP(hipEventSynchronize(e1));
        float ms = 0;
        TNR_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (ms_out) *ms_out = ms;
        ctx.fetch(W, dW.d(), nn * count * 8);
        ctx.fetch(lam, dL.d(), (size_t)n * count * 8);
        if (sweeps) ctx.fetch(sweeps, dS.d(), count * sizeof(int));
        return TNR_OK;
    });
}

int tnr_profile_enable(int on) {
    return kguard([&] {
        TNR_REQUIRE(on >= 0 && on <= 2, TNR_ERR_ARG, "profiling level must be 0, 1 or 2");
        prof_enable(on);
        return TNR_OK;
    });
}

int tnr_device_sync(void) {
    return kguard([&] {
        TNR_HIP(hipDeviceSynchronize());
        return TNR_OK;
    });
}

int tnr_profile_read(int kind, int64_t* launches, double* ms, double* flops, double* bytes,
                     double* busy_ms, double* flops_planned) {
    return kguard([&] {
        TNR_REQUIRE(kind >= 0 && kind < PK_COUNT, TNR_ERR_ARG, "unknown kernel class");
        ProfTotals t = prof_read(kind);
        if (launches) *launches = t.launches;
        if (ms) *ms = t.ms;
        if (flops) *flops = t.flops;
        if (bytes) *bytes = t.bytes;
        if (busy_ms) *busy_ms = t.busy_ms;
        if (flops_planned) *flops_planned = t.flops_planned;
        return TNR_OK;
    });
}

}  // extern "C"
