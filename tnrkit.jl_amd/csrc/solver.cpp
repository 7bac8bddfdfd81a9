// Grouped truncated SVD / symmetric eigensolver (see solver.h).
#include "solver.h"

#include <algorithm>
#include <cmath>

namespace tnr {

SolverCfg& solver_cfg() {
    static SolverCfg c;
    return c;
}

namespace {

// One two-pass Gram/Jacobi Rayleigh-Ritz job: rotates the columns of Y
// (rows x l) to the Ritz basis of span(Y), normalises them into U and
// writes the column norms to sigma.  Wtot (optional) receives the total
// l x l rotation.
struct RR {
    double* Y;
    int rows, l, ldy;
    double* Y2;
    double* G;
    double* W;
    double* W2;
    double* lam;
    int* sweeps;
    double* Wtot;
    double* U;
    int ldu;
    double* sigma;
};

void rr_grouped(Ctx& ctx, const std::vector<RR>& jobs) {
    if (jobs.empty()) return;
    const SolverCfg& cfg = solver_cfg();
    int lmax = 0;
    for (auto& j : jobs) lmax = std::max(lmax, j.l);
    std::vector<GemmJob> g;
    std::vector<JacobiDesc> jd;
    auto gram = [&](double* Y, const RR& j) {
        GemmJob q{Y, Y, j.G, j.l, j.l, j.rows, Y == j.Y ? j.ldy : j.rows,
                  Y == j.Y ? j.ldy : j.rows, j.l, 1, 0};
        return q;
    };
    for (int pass = 0; pass < 2; ++pass) {
        g.clear();
        jd.clear();
        for (auto& j : jobs) g.push_back(gram(pass == 0 ? j.Y : j.Y2, j));
        gemm_grouped(ctx, g);
        for (auto& j : jobs) {
            JacobiDesc d;
            d.G = j.G;
            d.W = pass == 0 ? j.W : j.W2;
            d.lam = j.lam;
            d.sweeps = j.sweeps;
            d.n = j.l;
            d.ldg = j.l;
            d.ldw = j.l;
            jd.push_back(d);
        }
        JacobiDesc* djd = ctx.stage->put(ctx.stream, jd);
        launch_jacobi(ctx.stream, djd, (int)jd.size(), lmax, cfg.jacobi_tol, cfg.jacobi_sweeps);
        g.clear();
        for (auto& j : jobs) {
            if (pass == 0)  // Y2 = Y W
                g.push_back(GemmJob{j.Y, j.W, j.Y2, j.rows, j.l, j.l, j.ldy, j.l, j.rows, 0, 0});
            else            // Y = Y2 W2
                g.push_back(GemmJob{j.Y2, j.W2, j.Y, j.rows, j.l, j.l, j.rows, j.l, j.ldy, 0, 0});
            if (pass == 1 && j.Wtot)
                g.push_back(GemmJob{j.W, j.W2, j.Wtot, j.l, j.l, j.l, j.l, j.l, j.l, 0, 0});
        }
        gemm_grouped(ctx, g);
    }
    std::vector<ColDesc> cd;
    for (auto& j : jobs) cd.push_back(ColDesc{j.Y, j.U, j.sigma, j.rows, j.l, j.ldy, j.ldu});
    ColDesc* dcd = ctx.stage->put(ctx.stream, cd);
    launch_colnorm(ctx.stream, dcd, (int)cd.size(), cfg.null_rtol);
}

void randfill(Ctx& ctx, const std::vector<FillDesc>& fills) {
    if (fills.empty()) return;
    std::vector<FillDesc> f(fills);
    int64_t start = 0;
    for (auto& x : f) {
        x.start = start;
        start += (int64_t)x.rows * x.cols;
    }
    FillDesc* d = ctx.stage->put(ctx.stream, f);
    launch_randfill(ctx.stream, d, (int)f.size(), start);
}

int block_size(int kmax) {
    int l = std::min(LMAX, 2 * kmax);
    return std::max(l, std::min(LMAX, kmax + 2));
}

// per-problem convergence bookkeeping of the block iteration
struct Conv {
    int it = 0;
    double prod = 1.0;
    bool done = false;
    void update(const double* sv, int l, int kmax, bool squared) {
        const SolverCfg& cfg = solver_cfg();
        ++it;
        int kk = 0;
        for (int i = 0; i < kmax; ++i)
            if (sv[i] > cfg.null_rtol * sv[0]) kk = i + 1;
        double r = 0.0;
        if (kk > 0 && sv[kk - 1] > 0.0) {
            r = sv[l - 1] / sv[kk - 1];
            if (squared) r *= r;
            r = std::min(r, 1.0);
        }
        if (it >= cfg.min_iters) prod *= r;
        if ((it >= cfg.min_iters && prod < cfg.conv_target) || it >= cfg.max_iters) done = true;
    }
};

}  // namespace

// ---------------------------------------------------------------------- SVD
void svd_grouped(Ctx& ctx, const std::vector<SvdIn>& in, std::vector<SvdOut>& out) {
    out.clear();
    out.resize(in.size());
    std::vector<int> direct, iter;
    for (size_t p = 0; p < in.size(); ++p) {
        const SvdIn& s = in[p];
        TNR_REQUIRE(s.m > 0 && s.n > 0 && s.kmax > 0 && s.kmax <= std::min(s.m, s.n),
                    TNR_ERR_INTERNAL, "bad SVD problem");
        if (std::min(s.m, s.n) <= LMAX) direct.push_back((int)p);
        else {
            TNR_REQUIRE(s.kmax <= LMAX - 8, TNR_ERR_ARG,
                        "chi above 88 is not supported for blocks larger than 96");
            iter.push_back((int)p);
        }
    }
    // ---------------- direct problems
    if (!direct.empty()) {
        struct DW { Buf Y, Y2, G, W, W2, Wt, lam, sig; Buf sw; };
        std::vector<DW> w(direct.size());
        std::vector<GatherDesc> gd;
        std::vector<RR> rr;
        int64_t start = 0;
        for (size_t i = 0; i < direct.size(); ++i) {
            const SvdIn& s = in[direct[i]];
            const bool tr = s.n > s.m;
            const int rows = std::max(s.m, s.n), l = std::min(s.m, s.n);
            DW& x = w[i];
            x.Y = Buf(&ctx.pool, (size_t)rows * l * 8);
            x.Y2 = Buf(&ctx.pool, (size_t)rows * l * 8);
            x.G = Buf(&ctx.pool, (size_t)l * l * 8);
            x.W = Buf(&ctx.pool, (size_t)l * l * 8);
            x.W2 = Buf(&ctx.pool, (size_t)l * l * 8);
            x.Wt = Buf(&ctx.pool, (size_t)l * l * 8);
            x.lam = Buf(&ctx.pool, (size_t)l * 8);
            x.sig = Buf(&ctx.pool, (size_t)l * 8);
            x.sw = Buf(&ctx.pool, 64);
            GatherDesc g{};
            g.in = s.M; g.out = x.Y.d(); g.rows = rows; g.cols = l;
            g.ldi = s.ld; g.ldo = rows; g.trans = tr ? 1 : 0; g.start = start;
            start += (int64_t)rows * l;
            gd.push_back(g);
            rr.push_back(RR{x.Y.d(), rows, l, rows, x.Y2.d(), x.G.d(), x.W.d(), x.W2.d(),
                            x.lam.d(), x.sw.as<int>(), x.Wt.d(), x.Y.d(), rows, x.sig.d()});
        }
        GatherDesc* dgd = ctx.stage->put(ctx.stream, gd);
        launch_gather(ctx.stream, dgd, (int)gd.size(), start);
        rr_grouped(ctx, rr);
        for (size_t i = 0; i < direct.size(); ++i) {
            const SvdIn& s = in[direct[i]];
            SvdOut& o = out[direct[i]];
            DW& x = w[i];
            const bool tr = s.n > s.m;
            const int rows = std::max(s.m, s.n), l = std::min(s.m, s.n);
            if (!tr) {
                o.U = std::move(x.Y); o.ldu = rows;
                o.V = std::move(x.Wt); o.ldv = l;
            } else {
                o.U = std::move(x.Wt); o.ldu = l;
                o.V = std::move(x.Y); o.ldv = rows;
            }
            o.s = std::move(x.sig);
            o.iters = 0;
        }
    }
    // ---------------- iterative problems
    if (!iter.empty()) {
        const size_t P = iter.size();
        struct IW { Buf V, Z, Y, U, Y2, G, W, W2, Wt, lam, sigU, sw; int l; Conv cv; };
        std::vector<IW> w(P);
        Buf sigV(&ctx.pool, P * LMAX * 8);
        std::vector<FillDesc> fills;
        std::vector<RR> rr;
        for (size_t i = 0; i < P; ++i) {
            const SvdIn& s = in[iter[i]];
            IW& x = w[i];
            x.l = block_size(s.kmax);
            const int l = x.l, big = std::max(s.m, s.n);
            x.V = Buf(&ctx.pool, (size_t)s.n * l * 8);
            x.Z = Buf(&ctx.pool, (size_t)s.n * l * 8);
            x.Y = Buf(&ctx.pool, (size_t)s.m * l * 8);
            x.U = Buf(&ctx.pool, (size_t)s.m * l * 8);
            x.Y2 = Buf(&ctx.pool, (size_t)big * l * 8);
            x.G = Buf(&ctx.pool, (size_t)l * l * 8);
            x.W = Buf(&ctx.pool, (size_t)l * l * 8);
            x.W2 = Buf(&ctx.pool, (size_t)l * l * 8);
            x.Wt = Buf(&ctx.pool, (size_t)l * l * 8);
            x.lam = Buf(&ctx.pool, (size_t)l * 8);
            x.sigU = Buf(&ctx.pool, (size_t)l * 8);
            x.sw = Buf(&ctx.pool, 64);
            fills.push_back(FillDesc{x.V.d(), s.n, l, s.n, s.seed, 0});
            rr.push_back(RR{x.V.d(), s.n, l, s.n, x.Y2.d(), x.G.d(), x.W.d(), x.W2.d(),
                            x.lam.d(), x.sw.as<int>(), nullptr, x.V.d(), s.n,
                            sigV.d() + i * LMAX});
        }
        randfill(ctx, fills);
        rr_grouped(ctx, rr);
        std::vector<double> sv(P * LMAX);
        std::vector<int> active(P);
        for (size_t i = 0; i < P; ++i) active[i] = (int)i;
        while (!active.empty()) {
            std::vector<GemmJob> g;
            std::vector<RR> r1, r2;
            for (int i : active) {
                const SvdIn& s = in[iter[i]];
                IW& x = w[i];
                g.push_back(GemmJob{s.M, x.V.d(), x.Y.d(), s.m, x.l, s.n, s.ld, s.n, s.m, 0, 0});
                r1.push_back(RR{x.Y.d(), s.m, x.l, s.m, x.Y2.d(), x.G.d(), x.W.d(), x.W2.d(),
                                x.lam.d(), x.sw.as<int>(), nullptr, x.U.d(), s.m, x.sigU.d()});
            }
            gemm_grouped(ctx, g);
            rr_grouped(ctx, r1);
            g.clear();
            for (int i : active) {
                const SvdIn& s = in[iter[i]];
                IW& x = w[i];
                g.push_back(GemmJob{s.M, x.U.d(), x.Z.d(), s.n, x.l, s.m, s.ld, s.m, s.n, 1, 0});
                r2.push_back(RR{x.Z.d(), s.n, x.l, s.n, x.Y2.d(), x.G.d(), x.W.d(), x.W2.d(),
                                x.lam.d(), x.sw.as<int>(), nullptr, x.V.d(), s.n,
                                sigV.d() + i * LMAX});
            }
            gemm_grouped(ctx, g);
            rr_grouped(ctx, r2);
            ctx.fetch(sv.data(), sigV.d(), sv.size() * 8);
            std::vector<int> still;
            for (int i : active) {
                w[i].cv.update(&sv[i * LMAX], w[i].l, in[iter[i]].kmax, true);
                if (!w[i].cv.done) still.push_back(i);
            }
            active.swap(still);
        }
        // Ritz extraction from Y = M V
        std::vector<GemmJob> g;
        std::vector<RR> r;
        for (size_t i = 0; i < P; ++i) {
            const SvdIn& s = in[iter[i]];
            IW& x = w[i];
            g.push_back(GemmJob{s.M, x.V.d(), x.Y.d(), s.m, x.l, s.n, s.ld, s.n, s.m, 0, 0});
            r.push_back(RR{x.Y.d(), s.m, x.l, s.m, x.Y2.d(), x.G.d(), x.W.d(), x.W2.d(),
                           x.lam.d(), x.sw.as<int>(), x.Wt.d(), x.U.d(), s.m, x.sigU.d()});
        }
        gemm_grouped(ctx, g);
        rr_grouped(ctx, r);
        g.clear();
        for (size_t i = 0; i < P; ++i) {
            const SvdIn& s = in[iter[i]];
            IW& x = w[i];
            g.push_back(GemmJob{x.V.d(), x.Wt.d(), x.Z.d(), s.n, x.l, x.l, s.n, x.l, s.n, 0, 0});
        }
        gemm_grouped(ctx, g);
        for (size_t i = 0; i < P; ++i) {
            SvdOut& o = out[iter[i]];
            IW& x = w[i];
            o.U = std::move(x.U); o.ldu = in[iter[i]].m;
            o.V = std::move(x.Z); o.ldv = in[iter[i]].n;
            o.s = std::move(x.sigU);
            o.iters = x.cv.it;
            o.conv = x.cv.prod;
        }
    }
    // singular values to the host (one transfer per problem, all queued)
    for (size_t p = 0; p < in.size(); ++p) {
        out[p].s_host.resize(in[p].kmax);
        TNR_HIP(hipMemcpyAsync(out[p].s_host.data(), out[p].s.d(), in[p].kmax * 8,
                               hipMemcpyDeviceToHost, ctx.stream));
    }
    ctx.sync();
}

// ---------------------------------------------------------------- eigh
void eigh_grouped(Ctx& ctx, const std::vector<EighIn>& in, std::vector<EighOut>& out) {
    out.clear();
    out.resize(in.size());
    const SolverCfg& cfg = solver_cfg();
    std::vector<int> direct, iter;
    for (size_t p = 0; p < in.size(); ++p) {
        const EighIn& s = in[p];
        TNR_REQUIRE(s.n > 0 && s.kmax > 0 && s.kmax <= s.n, TNR_ERR_INTERNAL, "bad eigh problem");
        if (s.n <= LMAX) direct.push_back((int)p);
        else {
            TNR_REQUIRE(s.kmax <= LMAX - 8, TNR_ERR_ARG,
                        "chi above 88 is not supported for blocks larger than 96");
            iter.push_back((int)p);
        }
    }
    if (!direct.empty()) {
        std::vector<JacobiDesc> jd;
        std::vector<Buf> sw(direct.size());
        int nmax = 0;
        for (size_t i = 0; i < direct.size(); ++i) {
            const EighIn& s = in[direct[i]];
            EighOut& o = out[direct[i]];
            o.U = Buf(&ctx.pool, (size_t)s.n * s.n * 8);
            o.lam = Buf(&ctx.pool, (size_t)s.n * 8);
            o.ncols = s.n;
            o.direct = true;
            sw[i] = Buf(&ctx.pool, 64);
            JacobiDesc d{s.M, o.U.d(), o.lam.d(), sw[i].as<int>(), s.n, s.ld, s.n};
            jd.push_back(d);
            nmax = std::max(nmax, s.n);
        }
        JacobiDesc* djd = ctx.stage->put(ctx.stream, jd);
        launch_jacobi(ctx.stream, djd, (int)jd.size(), nmax, cfg.jacobi_tol, cfg.jacobi_sweeps);
    }
    if (!iter.empty()) {
        const size_t P = iter.size();
        struct IW { Buf V, Y, Y2, G, W, W2, lam, sw; int l; Conv cv; };
        std::vector<IW> w(P);
        Buf sigV(&ctx.pool, P * LMAX * 8);
        std::vector<FillDesc> fills;
        std::vector<RR> rr;
        std::vector<ColDesc> sym;
        int nmax = 0;
        for (size_t i = 0; i < P; ++i) {
            const EighIn& s = in[iter[i]];
            IW& x = w[i];
            x.l = block_size(s.kmax);
            const int l = x.l;
            x.V = Buf(&ctx.pool, (size_t)s.n * l * 8);
            x.Y = Buf(&ctx.pool, (size_t)s.n * l * 8);
            x.Y2 = Buf(&ctx.pool, (size_t)s.n * l * 8);
            x.G = Buf(&ctx.pool, (size_t)l * l * 8);
            x.W = Buf(&ctx.pool, (size_t)l * l * 8);
            x.W2 = Buf(&ctx.pool, (size_t)l * l * 8);
            x.lam = Buf(&ctx.pool, (size_t)l * 8);
            x.sw = Buf(&ctx.pool, 64);
            fills.push_back(FillDesc{x.V.d(), s.n, l, s.n, s.seed, 0});
            rr.push_back(RR{x.V.d(), s.n, l, s.n, x.Y2.d(), x.G.d(), x.W.d(), x.W2.d(),
                            x.lam.d(), x.sw.as<int>(), nullptr, x.V.d(), s.n,
                            sigV.d() + i * LMAX});
            sym.push_back(ColDesc{s.M, nullptr, nullptr, s.n, s.n, s.ld, 0});
            nmax = std::max(nmax, s.n);
        }
        ColDesc* dsym = ctx.stage->put(ctx.stream, sym);
        launch_symmetrize(ctx.stream, dsym, (int)sym.size(), nmax);
        randfill(ctx, fills);
        rr_grouped(ctx, rr);
        std::vector<double> sv(P * LMAX);
        std::vector<int> active(P);
        for (size_t i = 0; i < P; ++i) active[i] = (int)i;
        while (!active.empty()) {
            std::vector<GemmJob> g;
            std::vector<RR> r1;
            for (int i : active) {
                const EighIn& s = in[iter[i]];
                IW& x = w[i];
                g.push_back(GemmJob{s.M, x.V.d(), x.Y.d(), s.n, x.l, s.n, s.ld, s.n, s.n, 0, 0});
                r1.push_back(RR{x.Y.d(), s.n, x.l, s.n, x.Y2.d(), x.G.d(), x.W.d(), x.W2.d(),
                                x.lam.d(), x.sw.as<int>(), nullptr, x.V.d(), s.n,
                                sigV.d() + i * LMAX});
            }
            gemm_grouped(ctx, g);
            rr_grouped(ctx, r1);
            ctx.fetch(sv.data(), sigV.d(), sv.size() * 8);
            std::vector<int> still;
            for (int i : active) {
                w[i].cv.update(&sv[i * LMAX], w[i].l, in[iter[i]].kmax, false);
                if (!w[i].cv.done) still.push_back(i);
            }
            active.swap(still);
        }
        // Ritz pairs: H = V^T (M V), eigendecompose, U = V W
        std::vector<GemmJob> g;
        for (size_t i = 0; i < P; ++i) {
            const EighIn& s = in[iter[i]];
            IW& x = w[i];
            g.push_back(GemmJob{s.M, x.V.d(), x.Y.d(), s.n, x.l, s.n, s.ld, s.n, s.n, 0, 0});
        }
        gemm_grouped(ctx, g);
        g.clear();
        for (size_t i = 0; i < P; ++i) {
            const EighIn& s = in[iter[i]];
            IW& x = w[i];
            g.push_back(GemmJob{x.V.d(), x.Y.d(), x.G.d(), x.l, x.l, s.n, s.n, s.n, x.l, 1, 0});
        }
        gemm_grouped(ctx, g);
        std::vector<JacobiDesc> jd;
        int lmax = 0;
        for (size_t i = 0; i < P; ++i) {
            IW& x = w[i];
            EighOut& o = out[iter[i]];
            o.lam = Buf(&ctx.pool, (size_t)x.l * 8);
            jd.push_back(JacobiDesc{x.G.d(), x.W.d(), o.lam.d(), x.sw.as<int>(), x.l, x.l, x.l});
            lmax = std::max(lmax, x.l);
        }
        JacobiDesc* djd = ctx.stage->put(ctx.stream, jd);
        launch_jacobi(ctx.stream, djd, (int)jd.size(), lmax, cfg.jacobi_tol, cfg.jacobi_sweeps);
        g.clear();
        for (size_t i = 0; i < P; ++i) {
            const EighIn& s = in[iter[i]];
            IW& x = w[i];
            EighOut& o = out[iter[i]];
            o.U = Buf(&ctx.pool, (size_t)s.n * x.l * 8);
            o.ncols = x.l;
            o.direct = false;
            o.iters = x.cv.it;
            o.conv = x.cv.prod;
            g.push_back(GemmJob{x.V.d(), x.W.d(), o.U.d(), s.n, x.l, x.l, s.n, x.l, s.n, 0, 0});
        }
        gemm_grouped(ctx, g);
    }
    for (size_t p = 0; p < in.size(); ++p) {
        out[p].lam_host.resize(out[p].ncols);
        TNR_HIP(hipMemcpyAsync(out[p].lam_host.data(), out[p].lam.d(), out[p].ncols * 8,
                               hipMemcpyDeviceToHost, ctx.stream));
    }
    ctx.sync();
}

}  // namespace tnr
