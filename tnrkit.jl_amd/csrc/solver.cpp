// Grouped truncated SVD / symmetric eigensolver (see solver.h).
//
// The block iteration is prepared once per call: every GEMM / Jacobi /
// column-norm descriptor of one iteration is uploaded once and reused by
// each iteration; a device kernel applies the stopping rule per problem and
// raises a flag that makes every later launch skip that problem, so the
// host only looks at the flags between chunks of iterations.
#include "solver.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <string>

namespace tnr {

// "key=value,..." overrides of the defaults (keys: tol_iter, sweeps_iter,
// sweeps, block_mult, block_extra, conv_target, max_iters, cholqr_y,
// cholqr_z); unknown keys are an argument error.  TNR_SOLVER_CFG in the
// environment is applied first, then tnr_solver_config's string.
static void apply_cfg(SolverCfg& c, const char* e) {
    if (!e) return;
    std::string s(e);
    size_t i = 0;
    while (i < s.size()) {
        size_t j = s.find(',', i);
        if (j == std::string::npos) j = s.size();
        const std::string kv = s.substr(i, j - i);
        const size_t eq = kv.find('=');
        TNR_REQUIRE(eq != std::string::npos || kv.empty(), TNR_ERR_ARG,
                    "solver option '" + kv + "' is not key=value");
        if (eq != std::string::npos) {
            const std::string k = kv.substr(0, eq);
            const double v = std::atof(kv.c_str() + eq + 1);
            if (k == "tol_iter") c.jacobi_tol_iter = v;
            else if (k == "sweeps_iter") c.jacobi_sweeps_iter = (int)v;
            else if (k == "sweeps") c.jacobi_sweeps = (int)v;
            else if (k == "block_mult") c.block_mult = v;
            else if (k == "block_extra") c.block_extra = (int)v;
            else if (k == "conv_target") c.conv_target = v;
            else if (k == "max_iters") c.max_iters = (int)v;
            else if (k == "cholqr_y") c.cholqr_y = v != 0.0;
            else if (k == "cholqr_z") c.cholqr_z = v != 0.0;
            else TNR_REQUIRE(false, TNR_ERR_ARG, "unknown solver option '" + k + "'");
        }
        i = j + 1;
    }
}

static SolverCfg env_cfg() {
    SolverCfg c;
    try {
        apply_cfg(c, std::getenv("TNR_SOLVER_CFG"));
    } catch (const Error&) {
        c = SolverCfg();   // a malformed environment leaves the defaults
    }
    return c;
}

SolverCfg& solver_cfg() {
    static SolverCfg c = env_cfg();
    return c;
}

void set_solver_cfg(const char* spec) {
    SolverCfg c = env_cfg();
    apply_cfg(c, spec);
    TNR_REQUIRE(c.max_iters >= 1 && c.jacobi_sweeps >= 0 && c.jacobi_sweeps_iter >= 0 &&
                    c.block_mult >= 1.0 && c.block_extra >= 0,
                TNR_ERR_ARG, "solver option out of range");
    solver_cfg() = c;
}

namespace {

// One Gram/Jacobi Rayleigh-Ritz job: rotates the columns of Y (rows x l) to
// the Ritz basis of span(Y), normalises them into U and writes the column
// norms to sigma.  Wtot (two-pass jobs only) receives the total rotation.
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
    const int* skip = nullptr;
    double* fail = nullptr;     // the problem's non-convergence flag
};

// `passes` = 1 inside the block iteration (the next iteration corrects
// what one pass leaves), 2 for the final Ritz extraction and direct solves.
struct RRPlan {
    int passes = 1, lmax = 0;
    bool chol = false;          // one pass, pivoted Cholesky QR instead of Jacobi
    GemmPlan gram[2], rot[2];
    std::vector<JacobiDesc> jac[2];
    std::vector<ColDesc> col;
    size_t o_jac[2] = {0, 0}, o_col = 0;
    const JacobiDesc* d_jac[2] = {nullptr, nullptr};
    const ColDesc* d_col = nullptr;
};

RRPlan rr_plan(Ctx& ctx, const std::vector<RR>& jobs, int passes, bool chol = false) {
    TNR_REQUIRE(passes == 1 || passes == 2, TNR_ERR_INTERNAL, "one or two passes");
    TNR_REQUIRE(!chol || passes == 1, TNR_ERR_INTERNAL, "Cholesky QR is a one-pass step");
    RRPlan P;
    P.passes = passes;
    P.chol = chol;
    for (auto& j : jobs) {
        TNR_REQUIRE(passes == 2 || !j.Wtot, TNR_ERR_INTERNAL, "total rotation needs two passes");
        P.lmax = std::max(P.lmax, j.l);
    }
    for (int pass = 0; pass < passes; ++pass) {
        std::vector<GemmJob> gj, rj;
        for (auto& j : jobs) {
            double* Y = pass == 0 ? j.Y : j.Y2;
            const int ld = pass == 0 ? j.ldy : j.rows;
            GemmJob g{Y, Y, j.G, j.l, j.l, j.rows, ld, ld, j.l, 1, 0};   // G = Y^T Y
            g.skip = j.skip;
            gj.push_back(g);
            GemmJob r = pass == 0
                            ? GemmJob{j.Y, j.W, j.Y2, j.rows, j.l, j.l, j.ldy, j.l, j.rows, 0, 0}
                            : GemmJob{j.Y2, j.W2, j.Y, j.rows, j.l, j.l, j.rows, j.l, j.ldy, 0, 0};
            r.skip = j.skip;
            rj.push_back(r);
            if (pass == 1 && j.Wtot) {
                GemmJob w{j.W, j.W2, j.Wtot, j.l, j.l, j.l, j.l, j.l, j.l, 0, 0};
                w.skip = j.skip;
                rj.push_back(w);
            }
            P.jac[pass].push_back(JacobiDesc{j.G, pass == 0 ? j.W : j.W2, j.lam, j.sweeps, j.l, j.l,
                                             j.l, j.skip, j.fail});
        }
        P.gram[pass] = gemm_plan(ctx, gj);
        P.rot[pass] = gemm_plan(ctx, rj);
    }
    // rotated columns: Y2 after one pass, Y after two
    for (auto& j : jobs)
        P.col.push_back(passes == 1 ? ColDesc{j.Y2, j.U, j.sigma, j.rows, j.l, j.rows, j.ldu, j.skip}
                                    : ColDesc{j.Y, j.U, j.sigma, j.rows, j.l, j.ldy, j.ldu, j.skip});
    return P;
}

void rr_stage(RRPlan& P, DescArena& a) {
    for (int pass = 0; pass < P.passes; ++pass) {
        gemm_stage(P.gram[pass], a);
        gemm_stage(P.rot[pass], a);
        P.o_jac[pass] = a.add(P.jac[pass]);
    }
    P.o_col = a.add(P.col);
}

void rr_bind(RRPlan& P, const DescArena& a) {
    for (int pass = 0; pass < P.passes; ++pass) {
        gemm_bind(P.gram[pass], a);
        gemm_bind(P.rot[pass], a);
        P.d_jac[pass] = a.at<JacobiDesc>(P.o_jac[pass]);
    }
    P.d_col = a.at<ColDesc>(P.o_col);
}

void rr_issue(Ctx& ctx, RRPlan& P) {
    const SolverCfg& cfg = solver_cfg();
    for (int pass = 0; pass < P.passes; ++pass) {
        gemm_issue(ctx, P.gram[pass]);
        const bool in_loop = P.passes == 1;
        if (P.chol)
            launch_cholqr(ctx.stream, P.d_jac[pass], (int)P.jac[pass].size(), P.lmax,
                          cfg.null_rtol * cfg.null_rtol);
        else
            launch_jacobi(ctx.stream, P.d_jac[pass], (int)P.jac[pass].size(), P.lmax,
                          in_loop ? cfg.jacobi_tol_iter : cfg.jacobi_tol,
                          in_loop ? cfg.jacobi_sweeps_iter : cfg.jacobi_sweeps);
        gemm_issue(ctx, P.rot[pass]);
    }
    launch_colnorm(ctx.stream, P.d_col, (int)P.col.size(), cfg.null_rtol);
}

void rr_grouped(Ctx& ctx, const std::vector<RR>& jobs, int passes) {
    if (jobs.empty()) return;
    RRPlan P = rr_plan(ctx, jobs, passes);
    DescArena a;
    rr_stage(P, a);
    a.commit(ctx);
    rr_bind(P, a);
    rr_issue(ctx, P);
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

// One double per problem, zeroed on the stream: the Jacobi kernels set a
// problem's entry to 1 when they stop at the sweep cap unconverged.
Buf fail_flags(Ctx& ctx, size_t P) {
    Buf f(&ctx.pool, std::max<size_t>(1, P) * sizeof(double));
    std::vector<double> z(std::max<size_t>(1, P), 0.0);
    void* a = ctx.stage->put(ctx.stream, z.data(), z.size() * sizeof(double));
    TNR_HIP(hipMemcpyAsync(f.d(), a, z.size() * sizeof(double), hipMemcpyDeviceToDevice, ctx.stream));
    return f;
}

// Status of one problem from its read-back values and Jacobi flag; a
// non-finite value is zeroed (a NaN would break the global truncation's
// ordering) and marks the problem TNR_ERR_NONFINITE.
int problem_status(std::vector<double>& v, double fail_flag, int status) {
    bool finite = true;
    for (double& x : v)
        if (!std::isfinite(x)) {
            finite = false;
            x = 0.0;
        }
    if (!finite) return TNR_ERR_NONFINITE;
    if (fail_flag != 0.0) return TNR_ERR_NOCONV;
    return status;
}

// block width of the subspace iteration: max(block_mult * k, k + block_extra),
// at most LMAX
int block_size(int kmax) {
    const SolverCfg& cfg = solver_cfg();
    const int l = std::max((int)std::ceil(cfg.block_mult * kmax), kmax + cfg.block_extra);
    return std::min(LMAX, std::max(l, kmax + 2 <= LMAX ? kmax + 2 : LMAX));
}

// Per-call device state of the stopping rule: it, done (ints), prod (double).
struct ConvState {
    Buf ints, prods;
    int P = 0;
    ConvState(Ctx& ctx, int P_) : P(P_) {
        ints = Buf(&ctx.pool, (size_t)2 * P * sizeof(int));
        prods = Buf(&ctx.pool, (size_t)P * sizeof(double));
        std::vector<int> zi(2 * P, 0);
        std::vector<double> one(P, 1.0);
        void* a = ctx.stage->put(ctx.stream, zi.data(), zi.size() * sizeof(int));
        void* b = ctx.stage->put(ctx.stream, one.data(), one.size() * sizeof(double));
        TNR_HIP(hipMemcpyAsync(ints.d(), a, zi.size() * sizeof(int), hipMemcpyDeviceToDevice,
                               ctx.stream));
        TNR_HIP(hipMemcpyAsync(prods.d(), b, one.size() * sizeof(double),
                               hipMemcpyDeviceToDevice, ctx.stream));
    }
    int* it(int p) const { return ints.as<int>() + p; }
    int* done(int p) const { return ints.as<int>() + P + p; }
    double* prod(int p) const { return prods.d() + p; }
};

// Issue `issue_one` chunk by chunk until the device marks every problem done.
template <class F>
void iterate(Ctx& ctx, const ConvState& cs, F&& issue_one) {
    const SolverCfg& cfg = solver_cfg();
    std::vector<int> done(cs.P);
    int total = 0;
    const int chunk = 3;
    while (total < cfg.max_iters) {
        for (int c = 0; c < chunk; ++c) issue_one();
        total += chunk;
        ctx.fetch(done.data(), cs.done(0), cs.P * sizeof(int));
        if (std::all_of(done.begin(), done.end(), [](int d) { return d != 0; })) break;
    }
}

}  // namespace

// ---------------------------------------------------------------------- SVD
void svd_grouped(Ctx& ctx, const std::vector<SvdIn>& in, std::vector<SvdOut>& out) {
    out.clear();
    out.resize(in.size());
    const SolverCfg& cfg = solver_cfg();
    Buf failv = fail_flags(ctx, in.size());
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
    // ---------------- direct problems: Gram/Jacobi twice on the smaller side
    if (!direct.empty()) {
        struct DW { Buf Y, Y2, G, W, W2, Wt, lam, sig, sw; };
        std::vector<DW> w(direct.size());
        std::vector<GatherDesc> gd;
        std::vector<RR> rr;
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
            g.ldi = s.ld; g.ldo = rows; g.trans = tr ? 1 : 0;
            gd.push_back(g);
            rr.push_back(RR{x.Y.d(), rows, l, rows, x.Y2.d(), x.G.d(), x.W.d(), x.W2.d(),
                            x.lam.d(), x.sw.as<int>(), x.Wt.d(), x.Y.d(), rows, x.sig.d()});
            rr.back().fail = failv.d() + direct[i];
        }
        gather_many(ctx, gd);
        rr_grouped(ctx, rr, 2);
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
    // ---------------- iterative problems: block subspace iteration
    if (!iter.empty()) {
        const int P = (int)iter.size();
        struct IW { Buf V, Z, Y, U, Y2, G, W, W2, Wt, lam, sigU, sw; int l; };
        std::vector<IW> w(P);
        Buf sigV(&ctx.pool, (size_t)P * LMAX * 8);
        ConvState cs(ctx, P);
        std::vector<FillDesc> fills;
        std::vector<RR> rY, rZ, rV, rF;
        std::vector<GemmJob> jMV, jMtU, jF, jV;
        std::vector<ConvDesc> conv;
        for (int i = 0; i < P; ++i) {
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
            const int* skip = cs.done(i);
            double* sv = sigV.d() + (size_t)i * LMAX;
            double* fl = failv.d() + iter[i];
            fills.push_back(FillDesc{x.V.d(), s.n, l, s.n, s.seed, 0});
            GemmJob mv{s.M, x.V.d(), x.Y.d(), s.m, l, s.n, s.ld, s.n, s.m, 0, 0};   // Y = M V
            mv.skip = skip;
            jMV.push_back(mv);
            rY.push_back(RR{x.Y.d(), s.m, l, s.m, x.Y2.d(), x.G.d(), x.W.d(), x.W2.d(), x.lam.d(),
                            x.sw.as<int>(), nullptr, x.U.d(), s.m, x.sigU.d(), skip, fl});
            GemmJob mtu{s.M, x.U.d(), x.Z.d(), s.n, l, s.m, s.ld, s.m, s.n, 1, 0};  // Z = M^T U
            mtu.skip = skip;
            jMtU.push_back(mtu);
            rZ.push_back(RR{x.Z.d(), s.n, l, s.n, x.Y2.d(), x.G.d(), x.W.d(), x.W2.d(), x.lam.d(),
                            x.sw.as<int>(), nullptr, x.V.d(), s.n, sv, skip, fl});
            // Ritz values of the Z side, or with Cholesky QR the pivots R_jj
            conv.push_back(ConvDesc{cfg.cholqr_z ? x.lam.d() : sv, l, s.kmax, 1, cs.it(i),
                                    cs.prod(i), cs.done(i)});
            rV.push_back(RR{x.Z.d(), s.n, l, s.n, x.Y2.d(), x.G.d(), x.W.d(), x.W2.d(), x.lam.d(),
                            x.sw.as<int>(), nullptr, x.V.d(), s.n, sv, nullptr, fl});
            // Ritz extraction from Y = M V (two passes), V <- V Wtot
            jF.push_back(GemmJob{s.M, x.V.d(), x.Y.d(), s.m, l, s.n, s.ld, s.n, s.m, 0, 0});
            rF.push_back(RR{x.Y.d(), s.m, l, s.m, x.Y2.d(), x.G.d(), x.W.d(), x.W2.d(), x.lam.d(),
                            x.sw.as<int>(), x.Wt.d(), x.U.d(), s.m, x.sigU.d(), nullptr, fl});
            jV.push_back(GemmJob{x.V.d(), x.Wt.d(), x.Z.d(), s.n, l, l, s.n, l, s.n, 0, 0});
        }
        randfill(ctx, fills);
        // the random start block is not orthonormalised: the first
        // Rayleigh-Ritz step on Y = M V builds an orthonormal basis of its span
        GemmPlan pMV = gemm_plan(ctx, jMV), pMtU = gemm_plan(ctx, jMtU);
        // Y side: an orthonormal basis of span(M V) is all the next product
        // needs (pivoted Cholesky QR); Z side: Rayleigh-Ritz, whose values
        // drive the stopping rule
        RRPlan pY = rr_plan(ctx, rY, 1, cfg.cholqr_y), pZ = rr_plan(ctx, rZ, 1, cfg.cholqr_z);
        DescArena a;
        gemm_stage(pMV, a);
        gemm_stage(pMtU, a);
        rr_stage(pY, a);
        rr_stage(pZ, a);
        const size_t o_conv = a.add(conv);
        a.commit(ctx);
        gemm_bind(pMV, a);
        gemm_bind(pMtU, a);
        rr_bind(pY, a);
        rr_bind(pZ, a);
        const ConvDesc* d_conv = a.at<ConvDesc>(o_conv);
        iterate(ctx, cs, [&] {
            gemm_issue(ctx, pMV);
            rr_issue(ctx, pY);
            gemm_issue(ctx, pMtU);
            rr_issue(ctx, pZ);
            launch_conv(ctx.stream, d_conv, P, cfg.null_rtol, cfg.conv_target, cfg.min_iters,
                        cfg.max_iters);
        });
        // the loop orthonormalises only to its looser Jacobi threshold: V is
        // rebuilt from the last Z = M^T U at full accuracy before the final
        // extraction relies on V^T V = I
        rr_grouped(ctx, rV, 2);
        gemm_grouped(ctx, jF);
        rr_grouped(ctx, rF, 2);
        gemm_grouped(ctx, jV);
        std::vector<int> its(P);
        std::vector<double> prods(P);
        ctx.fetch(its.data(), cs.it(0), P * sizeof(int));
        ctx.fetch(prods.data(), cs.prod(0), P * sizeof(double));
        for (int i = 0; i < P; ++i) {
            SvdOut& o = out[iter[i]];
            IW& x = w[i];
            o.U = std::move(x.U); o.ldu = in[iter[i]].m;
            o.V = std::move(x.Z); o.ldv = in[iter[i]].n;
            o.s = std::move(x.sigU);
            o.iters = its[i];
            o.conv = prods[i];
            if (!(prods[i] < cfg.conv_target)) o.status = TNR_ERR_NOCONV;   // stopped at max_iters
        }
    }
    // singular values and the non-convergence flags to the host: packed on
    // the device, one copy
    std::vector<const double*> src(in.size() + 1);
    std::vector<int> cnt(in.size() + 1);
    std::vector<std::vector<double>*> dst(in.size() + 1);
    for (size_t p = 0; p < in.size(); ++p) {
        src[p] = out[p].s.d();
        cnt[p] = in[p].kmax;
        dst[p] = &out[p].s_host;
    }
    std::vector<double> failh;
    src[in.size()] = failv.d();
    cnt[in.size()] = (int)in.size();
    dst[in.size()] = &failh;
    readback_vectors(ctx, src, cnt, dst);
    for (size_t p = 0; p < in.size(); ++p)
        out[p].status = problem_status(out[p].s_host, failh[p], out[p].status);
}

// ---------------------------------------------------------------- eigh
void eigh_grouped(Ctx& ctx, const std::vector<EighIn>& in, std::vector<EighOut>& out) {
    out.clear();
    out.resize(in.size());
    const SolverCfg& cfg = solver_cfg();
    Buf failv = fail_flags(ctx, in.size());
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
            jd.push_back(JacobiDesc{s.M, o.U.d(), o.lam.d(), sw[i].as<int>(), s.n, s.ld, s.n, nullptr,
                                    failv.d() + direct[i]});
            nmax = std::max(nmax, s.n);
        }
        JacobiDesc* djd = ctx.stage->put(ctx.stream, jd);
        launch_jacobi(ctx.stream, djd, (int)jd.size(), nmax, cfg.jacobi_tol, cfg.jacobi_sweeps);
    }
    if (!iter.empty()) {
        const int P = (int)iter.size();
        struct IW { Buf V, Y, Y2, G, W, W2, lam, sw; int l; };
        std::vector<IW> w(P);
        Buf sigV(&ctx.pool, (size_t)P * LMAX * 8);
        ConvState cs(ctx, P);
        std::vector<FillDesc> fills;
        std::vector<RR> rY, rV;
        std::vector<ColDesc> sym;
        std::vector<GemmJob> jMV, jF, jH;
        std::vector<ConvDesc> conv;
        int nmax = 0;
        for (int i = 0; i < P; ++i) {
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
            const int* skip = cs.done(i);
            double* sv = sigV.d() + (size_t)i * LMAX;
            fills.push_back(FillDesc{x.V.d(), s.n, l, s.n, s.seed, 0});
            sym.push_back(ColDesc{s.M, nullptr, nullptr, s.n, s.n, s.ld, 0});
            GemmJob mv{s.M, x.V.d(), x.Y.d(), s.n, l, s.n, s.ld, s.n, s.n, 0, 0};   // Y = M V
            mv.skip = skip;
            jMV.push_back(mv);
            rY.push_back(RR{x.Y.d(), s.n, l, s.n, x.Y2.d(), x.G.d(), x.W.d(), x.W2.d(), x.lam.d(),
                            x.sw.as<int>(), nullptr, x.V.d(), s.n, sv, skip, failv.d() + iter[i]});
            conv.push_back(ConvDesc{sv, l, s.kmax, 0, cs.it(i), cs.prod(i), cs.done(i)});
            rV.push_back(RR{x.Y.d(), s.n, l, s.n, x.Y2.d(), x.G.d(), x.W.d(), x.W2.d(), x.lam.d(),
                            x.sw.as<int>(), nullptr, x.V.d(), s.n, sv, nullptr, failv.d() + iter[i]});
            jF.push_back(GemmJob{s.M, x.V.d(), x.Y.d(), s.n, l, s.n, s.ld, s.n, s.n, 0, 0});
            jH.push_back(GemmJob{x.V.d(), x.Y.d(), x.G.d(), l, l, s.n, s.n, s.n, l, 1, 0});
            nmax = std::max(nmax, s.n);
        }
        ColDesc* dsym = ctx.stage->put(ctx.stream, sym);
        launch_symmetrize(ctx.stream, dsym, (int)sym.size(), nmax);
        randfill(ctx, fills);
        // (random start block used as is, see svd_grouped)
        GemmPlan pMV = gemm_plan(ctx, jMV);
        RRPlan pY = rr_plan(ctx, rY, 1);
        DescArena a;
        gemm_stage(pMV, a);
        rr_stage(pY, a);
        const size_t o_conv = a.add(conv);
        a.commit(ctx);
        gemm_bind(pMV, a);
        rr_bind(pY, a);
        const ConvDesc* d_conv = a.at<ConvDesc>(o_conv);
        iterate(ctx, cs, [&] {
            gemm_issue(ctx, pMV);
            rr_issue(ctx, pY);
            launch_conv(ctx.stream, d_conv, P, cfg.null_rtol, cfg.conv_target, cfg.min_iters,
                        cfg.max_iters);
        });
        // V at full orthonormality from the last Y = M V (the loop uses a
        // looser Jacobi threshold)
        rr_grouped(ctx, rV, 2);
        // Ritz pairs: H = V^T (M V), eigendecompose, U = V W
        gemm_grouped(ctx, jF);
        gemm_grouped(ctx, jH);
        std::vector<JacobiDesc> jd;
        int lmax = 0;
        for (int i = 0; i < P; ++i) {
            IW& x = w[i];
            EighOut& o = out[iter[i]];
            o.lam = Buf(&ctx.pool, (size_t)x.l * 8);
            jd.push_back(JacobiDesc{x.G.d(), x.W.d(), o.lam.d(), x.sw.as<int>(), x.l, x.l, x.l, nullptr,
                                    failv.d() + iter[i]});
            lmax = std::max(lmax, x.l);
        }
        JacobiDesc* djd = ctx.stage->put(ctx.stream, jd);
        launch_jacobi(ctx.stream, djd, (int)jd.size(), lmax, cfg.jacobi_tol, cfg.jacobi_sweeps);
        std::vector<GemmJob> jU;
        for (int i = 0; i < P; ++i) {
            const EighIn& s = in[iter[i]];
            IW& x = w[i];
            EighOut& o = out[iter[i]];
            o.U = Buf(&ctx.pool, (size_t)s.n * x.l * 8);
            o.ncols = x.l;
            o.direct = false;
            jU.push_back(GemmJob{x.V.d(), x.W.d(), o.U.d(), s.n, x.l, x.l, s.n, x.l, s.n, 0, 0});
        }
        gemm_grouped(ctx, jU);
        std::vector<int> its(P);
        std::vector<double> prods(P);
        ctx.fetch(its.data(), cs.it(0), P * sizeof(int));
        ctx.fetch(prods.data(), cs.prod(0), P * sizeof(double));
        for (int i = 0; i < P; ++i) {
            out[iter[i]].iters = its[i];
            out[iter[i]].conv = prods[i];
            if (!(prods[i] < cfg.conv_target)) out[iter[i]].status = TNR_ERR_NOCONV;   // at max_iters
        }
    }
    std::vector<const double*> lsrc(in.size() + 1);
    std::vector<int> lcnt(in.size() + 1);
    std::vector<std::vector<double>*> ldst(in.size() + 1);
    for (size_t p = 0; p < in.size(); ++p) {
        lsrc[p] = out[p].lam.d();
        lcnt[p] = out[p].ncols;
        ldst[p] = &out[p].lam_host;
    }
    std::vector<double> failh;
    lsrc[in.size()] = failv.d();
    lcnt[in.size()] = (int)in.size();
    ldst[in.size()] = &failh;
    readback_vectors(ctx, lsrc, lcnt, ldst);
    for (size_t p = 0; p < in.size(); ++p)
        out[p].status = problem_status(out[p].lam_host, failh[p], out[p].status);
}

}  // namespace tnr
