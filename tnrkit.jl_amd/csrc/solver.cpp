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
#include <numeric>
#include <string>

namespace tnr {

// "key=value,..." overrides of the defaults (keys: tol_iter, tol_first, sweeps_iter,
// sweeps, block_mult, block_extra, conv_target, max_iters, stall_from,
// cholqr_y, cholqr_z, cholqr_eigh, rv_passes, rf_passes); unknown keys are an
// argument error.  TNR_SOLVER_CFG in the
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
            else if (k == "tol_first") c.jacobi_tol_first = v;
            else if (k == "sweeps_iter") c.jacobi_sweeps_iter = (int)v;
            else if (k == "sweeps") c.jacobi_sweeps = (int)v;
            else if (k == "block_mult") c.block_mult = v;
            else if (k == "block_extra") c.block_extra = (int)v;
            else if (k == "conv_target") c.conv_target = v;
            else if (k == "max_iters") c.max_iters = (int)v;
            else if (k == "stall_from") c.stall_from = (int)v;
            else if (k == "cholqr_y") c.cholqr_y = v != 0.0;
            else if (k == "cholqr_z") c.cholqr_z = v != 0.0;
            else if (k == "cholqr_eigh") c.cholqr_eigh = v != 0.0;
            else if (k == "rv_passes") c.rv_passes = (int)v;
            else if (k == "rf_passes") c.rf_passes = (int)v;
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
    TNR_REQUIRE(c.max_iters >= 1 && c.stall_from >= 0 && c.jacobi_sweeps >= 0 && c.jacobi_sweeps_iter >= 0 &&
                    c.block_mult >= 1.0 && c.block_extra >= 0 && c.rv_passes >= 1 &&
                    c.rv_passes <= 2 && c.rf_passes >= 1 && c.rf_passes <= 2,
                TNR_ERR_ARG, "solver option out of range");
    solver_cfg() = c;
}

namespace {

// One Gram/Jacobi Rayleigh-Ritz job: rotates the columns of Y (rows x l) to
// the Ritz basis of span(Y), normalises them into U and writes the column
// norms to sigma.  Wtot (optional) receives the total rotation.
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
// what one pass leaves), 2 for the final Ritz extraction and direct solves;
// `tight` = one pass at the final threshold, for columns that are already
// orthogonal to the loop's threshold (the second pass of a two-pass step).
struct RRPlan {
    int passes = 1, lmax = 0, lmin = 0;
    bool chol = false;          // one pass, pivoted Cholesky QR instead of Jacobi
    bool tight = false;
    GemmPlan gram[2], rot[2];
    std::vector<JacobiDesc> jac[2];
    std::vector<ColDesc> col;
    size_t o_jac[2] = {0, 0}, o_col = 0;
    const JacobiDesc* d_jac[2] = {nullptr, nullptr};
    const ColDesc* d_col = nullptr;
};

RRPlan rr_plan(Ctx& ctx, const std::vector<RR>& jobs, int passes, bool chol = false,
               bool tight = false) {
    TNR_REQUIRE(passes == 1 || passes == 2, TNR_ERR_INTERNAL, "one or two passes");
    TNR_REQUIRE(!chol || passes == 1, TNR_ERR_INTERNAL, "Cholesky QR is a one-pass step");
    TNR_REQUIRE(!tight || (passes == 1 && !chol), TNR_ERR_INTERNAL, "tight is a one-pass Jacobi step");
    RRPlan P;
    P.passes = passes;
    P.chol = chol;
    P.tight = tight;
    P.lmin = jobs.empty() ? 0 : jobs[0].l;
    for (auto& j : jobs) {
        TNR_REQUIRE(!chol || !j.Wtot, TNR_ERR_INTERNAL, "no total rotation with Cholesky QR");
        P.lmax = std::max(P.lmax, j.l);
        P.lmin = std::min(P.lmin, j.l);
    }
    for (int pass = 0; pass < passes; ++pass) {
        std::vector<GemmJob> gj, rj;
        for (auto& j : jobs) {
            double* Y = pass == 0 ? j.Y : j.Y2;
            const int ld = pass == 0 ? j.ldy : j.rows;
            // one pass: its rotation is the total one
            double* W1 = passes == 1 && j.Wtot ? j.Wtot : j.W;
            GemmJob g{Y, Y, j.G, j.l, j.l, j.rows, ld, ld, j.l, 1, 0};   // G = Y^T Y
            g.skip = j.skip;
            gj.push_back(g);
            GemmJob r = pass == 0
                            ? GemmJob{j.Y, W1, j.Y2, j.rows, j.l, j.l, j.ldy, j.l, j.rows, 0, 0}
                            : GemmJob{j.Y2, j.W2, j.Y, j.rows, j.l, j.l, j.rows, j.l, j.ldy, 0, 0};
            r.skip = j.skip;
            rj.push_back(r);
            if (pass == 1 && j.Wtot) {
                GemmJob w{j.W, j.W2, j.Wtot, j.l, j.l, j.l, j.l, j.l, j.l, 0, 0};
                w.skip = j.skip;
                rj.push_back(w);
            }
            P.jac[pass].push_back(JacobiDesc{j.G, pass == 0 ? W1 : j.W2, j.lam, j.sweeps, j.l, j.l,
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

// `first`: the block iteration's first step, whose Rayleigh-Ritz step only
// has to align a random basis roughly (jacobi_tol_first)
void rr_issue(Ctx& ctx, RRPlan& P, bool first = false) {
    const SolverCfg& cfg = solver_cfg();
    for (int pass = 0; pass < P.passes; ++pass) {
        gemm_issue(ctx, P.gram[pass]);
        const bool in_loop = P.passes == 1 && !P.tight;
        if (P.chol)
            launch_cholqr(ctx.stream, P.d_jac[pass], (int)P.jac[pass].size(), P.lmin, P.lmax,
                          cfg.null_rtol * cfg.null_rtol);
        else
            launch_jacobi(ctx.stream, P.d_jac[pass], (int)P.jac[pass].size(), P.lmax,
                          !in_loop ? cfg.jacobi_tol
                                   : (first ? cfg.jacobi_tol_first : cfg.jacobi_tol_iter),
                          in_loop ? cfg.jacobi_sweeps_iter : cfg.jacobi_sweeps);
        gemm_issue(ctx, P.rot[pass]);
    }
    launch_colnorm(ctx.stream, P.d_col, (int)P.col.size(), P.lmax, cfg.null_rtol);
}

void rr_grouped(Ctx& ctx, const std::vector<RR>& jobs, int passes, bool tight = false) {
    if (jobs.empty()) return;
    RRPlan P = rr_plan(ctx, jobs, passes, false, tight);
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
        for (int c = 0; c < chunk; ++c) issue_one(total + c);
        total += chunk;
        ctx.fetch(done.data(), cs.done(0), cs.P * sizeof(int));
        if (std::all_of(done.begin(), done.end(), [](int d) { return d != 0; })) break;
    }
}

// ------------------------------------------------------------ wide blocks
// Blocks whose Rayleigh-Ritz step is wider than LMAX (chi above LMAX - 8 on
// a block larger than LMAX) run the same block subspace iteration with a
// block one-sided Jacobi as the Rayleigh-Ritz step: the l columns of Y are
// cut into blocks of at most LMAX / 2; in each round of a tournament over
// the blocks every disjoint block pair [Y_I Y_J] is gathered, its Gram
// matrix (MFMA GEMM) diagonalised by k_jacobi and the pair rotated by a GEMM
// (and the accumulated rotation W with it); sweeps repeat until no pair
// rotates.  Columns are then normalised and put in order of decreasing
// norm.  The host steers the loop (one readback per sweep and per
// iteration): wide blocks only occur at large chi, where the arithmetic
// dwarfs the synchronisation.
constexpr int WIDE_B = LMAX / 2;
// Largest side factorised whole when a block iteration stalls.
constexpr int kWholeMax = 4 * LMAX;

struct WideRR {
    double* Y;           // rows x l, ld ldy: rotated in place
    int rows, l, ldy;
    double* U;           // rows x l (ld rows): normalised columns, by decreasing norm
    double* sigma;       // l norms, decreasing (0 for null columns)
    double* Wtot;        // optional l x l (ld l): Y_in Wtot = U diag(sigma)
};

int block_count(int l) {
    const int nb = (l + WIDE_B - 1) / WIDE_B;
    return nb + (nb & 1);   // even, for the tournament (the last block may be empty)
}

// columns [b0, b1) of block b of an l-wide job
void block_cols(int l, int b, int& b0, int& b1) {
    b0 = std::min(l, b * WIDE_B);
    b1 = std::min(l, (b + 1) * WIDE_B);
}

// A job is converged after a sweep in which no pair needed a rotation, i.e.
// every pair's Gram matrix had relative off-diagonals below the threshold.
// The Gram matrices are formed anew from the rotated columns each round, so
// their roundoff (~1e-16 sqrt(rows) relative) sets a floor: the pair solves
// use at least WIDE_TOL.
constexpr double WIDE_TOL = 1e-13;

// One-sided block Jacobi on every job; returns TNR_OK or TNR_ERR_NOCONV per job.
std::vector<int> wide_rr(Ctx& ctx, const std::vector<WideRR>& jobs, double tol, int max_sweeps) {
    const SolverCfg& cfg = solver_cfg();
    tol = std::max(tol, WIDE_TOL);
    const size_t J = jobs.size();
    std::vector<int> status(J, TNR_OK);
    // accumulated rotations start at the identity
    for (size_t j = 0; j < J; ++j) {
        if (!jobs[j].Wtot) continue;
        const int l = jobs[j].l;
        std::vector<double> id((size_t)l * l, 0.0);
        for (int i = 0; i < l; ++i) id[(size_t)i * l + i] = 1.0;
        void* t = ctx.stage->put(ctx.stream, id.data(), id.size() * sizeof(double));
        TNR_HIP(hipMemcpyAsync(jobs[j].Wtot, t, id.size() * sizeof(double), hipMemcpyDeviceToDevice,
                               ctx.stream));
    }
    std::vector<char> live(J, 1);
    for (int sw = 0;; ++sw) {
        int nlive = 0;
        for (size_t j = 0; j < J; ++j) nlive += live[j];
        if (!nlive) break;
        if (sw == max_sweeps) {
            for (size_t j = 0; j < J; ++j)
                if (live[j]) status[j] = TNR_ERR_NOCONV;
            break;
        }
        int maxnb = 0;
        for (size_t j = 0; j < J; ++j)
            if (live[j]) maxnb = std::max(maxnb, block_count(jobs[j].l));
        // sweep counts of every pair solve of this sweep, and its job
        std::vector<int> owner;
        Buf counts(&ctx.pool, std::max<size_t>(1, J * maxnb * maxnb) * sizeof(int));
        int npairs = 0;
        for (int rnd = 0; rnd + 1 < maxnb; ++rnd) {
            struct Pair { size_t j; int c0[2], c1[2]; int c; Buf Z, Z2, G, Q, lam, Wz, Wz2; };
            std::vector<Pair> P;
            for (size_t j = 0; j < J; ++j) {
                if (!live[j]) continue;
                const int nb = block_count(jobs[j].l);
                if (rnd + 1 >= nb) continue;
                const int m = nb - 1;
                for (int slot = 0; slot < nb / 2; ++slot) {
                    // circle method: slot 0 fixed at position 0, the rest rotate
                    int a = slot == 0 ? 0 : 1 + (slot - 1 + rnd) % m;
                    int b = 1 + (nb - 2 - slot + rnd) % m;
                    Pair q;
                    q.j = j;
                    block_cols(jobs[j].l, a, q.c0[0], q.c1[0]);
                    block_cols(jobs[j].l, b, q.c0[1], q.c1[1]);
                    const int wa = q.c1[0] - q.c0[0], wb = q.c1[1] - q.c0[1];
                    if (wa <= 0 || wb <= 0) continue;
                    q.c = wa + wb;
                    P.push_back(std::move(q));
                }
            }
            if (P.empty()) continue;
            std::vector<GatherDesc> gin, gout;
            std::vector<GemmJob> gram, rot;
            std::vector<JacobiDesc> jd;
            int cmax = 0;
            for (auto& q : P) {
                const WideRR& R = jobs[q.j];
                q.Z = Buf(&ctx.pool, (size_t)R.rows * q.c * 8);
                q.Z2 = Buf(&ctx.pool, (size_t)R.rows * q.c * 8);
                q.G = Buf(&ctx.pool, (size_t)q.c * q.c * 8);
                q.Q = Buf(&ctx.pool, (size_t)q.c * q.c * 8);
                q.lam = Buf(&ctx.pool, (size_t)q.c * 8);
                int o = 0;
                for (int h = 0; h < 2; ++h) {
                    const int w = q.c1[h] - q.c0[h];
                    GatherDesc g{};
                    g.in = R.Y + (int64_t)q.c0[h] * R.ldy; g.out = q.Z.d() + (int64_t)o * R.rows;
                    g.rows = R.rows; g.cols = w; g.ldi = R.ldy; g.ldo = R.rows;
                    gin.push_back(g);
                    GatherDesc b{};
                    b.in = q.Z2.d() + (int64_t)o * R.rows; b.out = R.Y + (int64_t)q.c0[h] * R.ldy;
                    b.rows = R.rows; b.cols = w; b.ldi = R.rows; b.ldo = R.ldy;
                    gout.push_back(b);
                    o += w;
                }
                gram.push_back(GemmJob{q.Z.d(), q.Z.d(), q.G.d(), q.c, q.c, R.rows, R.rows, R.rows,
                                       q.c, 1, 0});
                // eigenvectors left in place: a pair's rotation stays close to
                // the identity, so columns do not migrate between blocks (with
                // the eigenpairs sorted, the sweeps hardly converge)
                JacobiDesc d{q.G.d(), q.Q.d(), q.lam.d(), counts.as<int>() + npairs, q.c, q.c, q.c};
                d.natural = 1;
                jd.push_back(d);
                owner.push_back((int)q.j);
                ++npairs;
                rot.push_back(GemmJob{q.Z.d(), q.Q.d(), q.Z2.d(), R.rows, q.c, q.c, R.rows, q.c,
                                      R.rows, 0, 0});
                if (R.Wtot) {
                    q.Wz = Buf(&ctx.pool, (size_t)R.l * q.c * 8);
                    q.Wz2 = Buf(&ctx.pool, (size_t)R.l * q.c * 8);
                    int o2 = 0;
                    for (int h = 0; h < 2; ++h) {
                        const int w = q.c1[h] - q.c0[h];
                        GatherDesc g{};
                        g.in = R.Wtot + (int64_t)q.c0[h] * R.l; g.out = q.Wz.d() + (int64_t)o2 * R.l;
                        g.rows = R.l; g.cols = w; g.ldi = R.l; g.ldo = R.l;
                        gin.push_back(g);
                        GatherDesc b{};
                        b.in = q.Wz2.d() + (int64_t)o2 * R.l; b.out = R.Wtot + (int64_t)q.c0[h] * R.l;
                        b.rows = R.l; b.cols = w; b.ldi = R.l; b.ldo = R.l;
                        gout.push_back(b);
                        o2 += w;
                    }
                    rot.push_back(GemmJob{q.Wz.d(), q.Q.d(), q.Wz2.d(), R.l, q.c, q.c, R.l, q.c, R.l,
                                          0, 0});
                }
                cmax = std::max(cmax, q.c);
            }
            gather_many(ctx, gin);
            gemm_grouped(ctx, gram);
            JacobiDesc* djd = ctx.stage->put(ctx.stream, jd);
            launch_jacobi(ctx.stream, djd, (int)jd.size(), cmax, tol, cfg.jacobi_sweeps);
            gemm_grouped(ctx, rot);
            gather_many(ctx, gout);
        }
        // a job whose pairs needed no rotation in this sweep is done
        std::vector<int> h(std::max(1, npairs));
        if (npairs) ctx.fetch(h.data(), counts.d(), (size_t)npairs * sizeof(int));
        std::vector<char> rotated(J, 0);
        for (int i = 0; i < npairs; ++i)
            if (h[i] > 0) rotated[owner[i]] = 1;
        for (size_t j = 0; j < J; ++j)
            if (live[j] && !rotated[j]) live[j] = 0;
    }
    // norms, then columns in order of decreasing norm
    std::vector<Buf> Yn(J), sg(J);
    std::vector<ColDesc> cd;
    for (size_t j = 0; j < J; ++j) {
        const WideRR& R = jobs[j];
        Yn[j] = Buf(&ctx.pool, (size_t)R.rows * R.l * 8);
        sg[j] = Buf(&ctx.pool, (size_t)R.l * 8);
        cd.push_back(ColDesc{R.Y, Yn[j].d(), sg[j].d(), R.rows, R.l, R.ldy, R.rows});
    }
    int lcol = 0;
    for (const ColDesc& c : cd) lcol = std::max(lcol, (int)c.ncol);
    ColDesc* dcd = ctx.stage->put(ctx.stream, cd);
    launch_colnorm(ctx.stream, dcd, (int)cd.size(), lcol, cfg.null_rtol);
    std::vector<const double*> src;
    std::vector<int> cnt;
    std::vector<std::vector<double>> hs(J);
    std::vector<std::vector<double>*> dst;
    for (size_t j = 0; j < J; ++j) {
        src.push_back(sg[j].d());
        cnt.push_back(jobs[j].l);
        dst.push_back(&hs[j]);
    }
    readback_vectors(ctx, src, cnt, dst);
    std::vector<GatherDesc> perm;
    std::vector<double> sorted_all;
    std::vector<size_t> soff(J);
    std::vector<Buf> Wsrc(J);
    for (size_t j = 0; j < J; ++j) {
        const WideRR& R = jobs[j];
        std::vector<int> ord(R.l);
        std::iota(ord.begin(), ord.end(), 0);
        std::stable_sort(ord.begin(), ord.end(),
                         [&](int a, int b) { return hs[j][a] > hs[j][b]; });
        soff[j] = sorted_all.size();
        if (R.Wtot) {   // permute a copy of the accumulated rotation
            Wsrc[j] = Buf(&ctx.pool, (size_t)R.l * R.l * 8);
            GatherDesc g{};
            g.in = R.Wtot; g.out = Wsrc[j].d(); g.rows = R.l; g.cols = R.l; g.ldi = R.l; g.ldo = R.l;
            perm.push_back(g);
        }
        for (int c = 0; c < R.l; ++c) sorted_all.push_back(hs[j][ord[c]]);
        hs[j].assign(ord.begin(), ord.end());   // keep the order
    }
    gather_many(ctx, perm);
    perm.clear();
    for (size_t j = 0; j < J; ++j) {
        const WideRR& R = jobs[j];
        for (int c = 0; c < R.l; ++c) {
            const int from = (int)hs[j][c];
            GatherDesc g{};
            g.in = Yn[j].d() + (int64_t)from * R.rows; g.out = R.U + (int64_t)c * R.rows;
            g.rows = R.rows; g.cols = 1; g.ldi = R.rows; g.ldo = R.rows;
            perm.push_back(g);
            if (R.Wtot) {
                GatherDesc w{};
                w.in = Wsrc[j].d() + (int64_t)from * R.l; w.out = R.Wtot + (int64_t)c * R.l;
                w.rows = R.l; w.cols = 1; w.ldi = R.l; w.ldo = R.l;
                perm.push_back(w);
            }
        }
        void* t = ctx.stage->put(ctx.stream, sorted_all.data() + soff[j], (size_t)R.l * 8);
        TNR_HIP(hipMemcpyAsync(R.sigma, t, (size_t)R.l * 8, hipMemcpyDeviceToDevice, ctx.stream));
    }
    gather_many(ctx, perm);
    return status;
}

// block width of a wide problem
int wide_block_size(int kmax, int cap) {
    const SolverCfg& cfg = solver_cfg();
    const int l = std::max((int)std::ceil(cfg.block_mult * kmax), kmax + cfg.block_extra);
    return std::min(l, cap);
}

// Host-side stopping rule of the wide iteration (the device rule of k_conv):
// prod *= min(1, sigma_l / sigma_kk)^(squared ? 2 : 1) from min_iters on; a
// block as wide as the matrix (`whole`) spans its whole range and is exact.
bool wide_converged(const std::vector<double>& sv, int kmax, bool squared, bool whole, int it,
                    double& prod) {
    const SolverCfg& cfg = solver_cfg();
    if (whole) {
        prod = 0.0;
        return it >= cfg.min_iters;
    }
    int kk = 0;
    for (int i = 0; i < kmax && i < (int)sv.size(); ++i)
        if (sv[i] > cfg.null_rtol * sv[0]) kk = i + 1;
    double r = 0.0;
    if (kk > 0 && sv[kk - 1] > 0.0) {
        r = std::min(1.0, sv.back() / sv[kk - 1]);
        if (squared) r *= r;
    }
    if (it >= cfg.min_iters) prod *= r;
    return it >= cfg.min_iters && prod < cfg.conv_target;
}

// `whole`: the block spans the whole smaller side (an exact factorisation;
// the retry of a block iteration that stalled)
void svd_wide(Ctx& ctx, const std::vector<SvdIn>& in, const std::vector<int>& idx,
              std::vector<SvdOut>& out, bool whole = false) {
    const SolverCfg& cfg = solver_cfg();
    struct WS { Buf V, Y, Z, U, sY, sZ, Wt; int l; double prod = 1.0; bool done = false; };
    std::vector<WS> w(idx.size());
    std::vector<FillDesc> fills;
    for (size_t i = 0; i < idx.size(); ++i) {
        const SvdIn& s = in[idx[i]];
        WS& x = w[i];
        x.l = whole ? std::min(s.m, s.n) : wide_block_size(s.kmax, std::min(s.m, s.n));
        x.V = Buf(&ctx.pool, (size_t)s.n * x.l * 8);
        x.Z = Buf(&ctx.pool, (size_t)s.n * x.l * 8);
        x.Y = Buf(&ctx.pool, (size_t)s.m * x.l * 8);
        x.U = Buf(&ctx.pool, (size_t)s.m * x.l * 8);
        x.sY = Buf(&ctx.pool, (size_t)x.l * 8);
        x.sZ = Buf(&ctx.pool, (size_t)x.l * 8);
        fills.push_back(FillDesc{x.V.d(), s.n, x.l, s.n, s.seed, 0});
    }
    randfill(ctx, fills);
    for (int it = 1; it <= cfg.max_iters; ++it) {
        std::vector<size_t> act;
        for (size_t i = 0; i < idx.size(); ++i)
            if (!w[i].done) act.push_back(i);
        if (act.empty()) break;
        std::vector<GemmJob> mv, mtu;
        std::vector<WideRR> ry, rz;
        for (size_t i : act) {
            const SvdIn& s = in[idx[i]];
            WS& x = w[i];
            mv.push_back(GemmJob{s.M, x.V.d(), x.Y.d(), s.m, x.l, s.n, s.ld, s.n, s.m, 0, 0});
            ry.push_back(WideRR{x.Y.d(), s.m, x.l, s.m, x.U.d(), x.sY.d(), nullptr});
            mtu.push_back(GemmJob{s.M, x.U.d(), x.Z.d(), s.n, x.l, s.m, s.ld, s.m, s.n, 1, 0});
            rz.push_back(WideRR{x.Z.d(), s.n, x.l, s.n, x.V.d(), x.sZ.d(), nullptr});
        }
        gemm_grouped(ctx, mv);
        auto sty = wide_rr(ctx, ry, cfg.jacobi_tol_iter, cfg.jacobi_sweeps_iter);
        gemm_grouped(ctx, mtu);
        auto stz = wide_rr(ctx, rz, cfg.jacobi_tol_iter, cfg.jacobi_sweeps_iter);
        std::vector<const double*> src;
        std::vector<int> cnt;
        std::vector<std::vector<double>> hv(act.size());
        std::vector<std::vector<double>*> dst;
        for (size_t a = 0; a < act.size(); ++a) {
            src.push_back(w[act[a]].sZ.d());
            cnt.push_back(w[act[a]].l);
            dst.push_back(&hv[a]);
        }
        readback_vectors(ctx, src, cnt, dst);
        for (size_t a = 0; a < act.size(); ++a) {
            WS& x = w[act[a]];
            SvdOut& o = out[idx[act[a]]];
            if (sty[a] != TNR_OK || stz[a] != TNR_OK) o.status = TNR_ERR_NOCONV;
            o.iters = it;
            const SvdIn& sa = in[idx[act[a]]];
            x.done = wide_converged(hv[a], sa.kmax, true, x.l >= std::min(sa.m, sa.n), it, x.prod);
            if (!x.done && it == cfg.max_iters) o.status = TNR_ERR_NOCONV;
            o.conv = x.prod;
        }
    }
    // the loop orthogonalises to its looser threshold: V is rebuilt from the
    // last Z = M^T U at full accuracy before the extraction relies on
    // V^T V = I; then the Ritz triplets from Y = M V: U, sigma, V <- V Wtot
    {
        std::vector<WideRR> rv;
        for (size_t i = 0; i < idx.size(); ++i) {
            const SvdIn& s = in[idx[i]];
            WS& x = w[i];
            rv.push_back(WideRR{x.Z.d(), s.n, x.l, s.n, x.V.d(), x.sZ.d(), nullptr});
        }
        auto stv = wide_rr(ctx, rv, cfg.jacobi_tol, cfg.jacobi_sweeps);
        for (size_t i = 0; i < idx.size(); ++i)
            if (stv[i] != TNR_OK) out[idx[i]].status = TNR_ERR_NOCONV;
    }
    std::vector<GemmJob> mv, vw;
    std::vector<WideRR> rf;
    for (size_t i = 0; i < idx.size(); ++i) {
        const SvdIn& s = in[idx[i]];
        WS& x = w[i];
        x.Wt = Buf(&ctx.pool, (size_t)x.l * x.l * 8);
        mv.push_back(GemmJob{s.M, x.V.d(), x.Y.d(), s.m, x.l, s.n, s.ld, s.n, s.m, 0, 0});
        rf.push_back(WideRR{x.Y.d(), s.m, x.l, s.m, x.U.d(), x.sY.d(), x.Wt.d()});
        vw.push_back(GemmJob{x.V.d(), x.Wt.d(), x.Z.d(), s.n, x.l, x.l, s.n, x.l, s.n, 0, 0});
    }
    gemm_grouped(ctx, mv);
    auto stf = wide_rr(ctx, rf, cfg.jacobi_tol, cfg.jacobi_sweeps);
    gemm_grouped(ctx, vw);
    for (size_t i = 0; i < idx.size(); ++i) {
        SvdOut& o = out[idx[i]];
        WS& x = w[i];
        if (stf[i] != TNR_OK) o.status = TNR_ERR_NOCONV;
        o.U = std::move(x.U); o.ldu = in[idx[i]].m;
        o.V = std::move(x.Z); o.ldv = in[idx[i]].n;
        o.s = std::move(x.sY);
    }
}

// `whole`: the block spans the whole matrix (all n eigenpairs, exact; the
// retry of a block iteration that stalled)
void eigh_wide(Ctx& ctx, const std::vector<EighIn>& in, const std::vector<int>& idx,
               std::vector<EighOut>& out, bool whole = false) {
    const SolverCfg& cfg = solver_cfg();
    struct WS { Buf V, Y, sY, H, Uh, sH, Wt; int l; double prod = 1.0; bool done = false; };
    std::vector<WS> w(idx.size());
    std::vector<FillDesc> fills;
    std::vector<ColDesc> sym;
    int nmax = 0;
    for (size_t i = 0; i < idx.size(); ++i) {
        const EighIn& s = in[idx[i]];
        WS& x = w[i];
        x.l = whole ? s.n : wide_block_size(s.kmax, s.n);
        x.V = Buf(&ctx.pool, (size_t)s.n * x.l * 8);
        x.Y = Buf(&ctx.pool, (size_t)s.n * x.l * 8);
        x.sY = Buf(&ctx.pool, (size_t)x.l * 8);
        fills.push_back(FillDesc{x.V.d(), s.n, x.l, s.n, s.seed, 0});
        sym.push_back(ColDesc{s.M, nullptr, nullptr, s.n, s.n, s.ld, 0});
        nmax = std::max(nmax, s.n);
    }
    ColDesc* dsym = ctx.stage->put(ctx.stream, sym);
    launch_symmetrize(ctx.stream, dsym, (int)sym.size(), nmax);
    randfill(ctx, fills);
    for (int it = 1; it <= cfg.max_iters; ++it) {
        std::vector<size_t> act;
        for (size_t i = 0; i < idx.size(); ++i)
            if (!w[i].done) act.push_back(i);
        if (act.empty()) break;
        std::vector<GemmJob> mv;
        std::vector<WideRR> ry;
        for (size_t i : act) {
            const EighIn& s = in[idx[i]];
            WS& x = w[i];
            mv.push_back(GemmJob{s.M, x.V.d(), x.Y.d(), s.n, x.l, s.n, s.ld, s.n, s.n, 0, 0});
            ry.push_back(WideRR{x.Y.d(), s.n, x.l, s.n, x.V.d(), x.sY.d(), nullptr});
        }
        gemm_grouped(ctx, mv);
        auto sty = wide_rr(ctx, ry, cfg.jacobi_tol_iter, cfg.jacobi_sweeps_iter);
        std::vector<const double*> src;
        std::vector<int> cnt;
        std::vector<std::vector<double>> hv(act.size());
        std::vector<std::vector<double>*> dst;
        for (size_t a = 0; a < act.size(); ++a) {
            src.push_back(w[act[a]].sY.d());
            cnt.push_back(w[act[a]].l);
            dst.push_back(&hv[a]);
        }
        readback_vectors(ctx, src, cnt, dst);
        for (size_t a = 0; a < act.size(); ++a) {
            WS& x = w[act[a]];
            EighOut& o = out[idx[act[a]]];
            if (sty[a] != TNR_OK) o.status = TNR_ERR_NOCONV;
            o.iters = it;
            x.done = wide_converged(hv[a], in[idx[act[a]]].kmax, false, x.l >= in[idx[act[a]]].n, it,
                                    x.prod);
            if (!x.done && it == cfg.max_iters) o.status = TNR_ERR_NOCONV;
            o.conv = x.prod;
        }
    }
    // V at full orthonormality from the last Y = M V (the loop uses a looser
    // threshold)
    {
        std::vector<WideRR> rv;
        for (size_t i = 0; i < idx.size(); ++i) {
            const EighIn& s = in[idx[i]];
            WS& x = w[i];
            rv.push_back(WideRR{x.Y.d(), s.n, x.l, s.n, x.V.d(), x.sY.d(), nullptr});
        }
        auto stv = wide_rr(ctx, rv, cfg.jacobi_tol, cfg.jacobi_sweeps);
        for (size_t i = 0; i < idx.size(); ++i)
            if (stv[i] != TNR_OK) out[idx[i]].status = TNR_ERR_NOCONV;
    }
    // Ritz pairs: H = V^T M V; one-sided Jacobi H W = Uh diag(sigma) gives the
    // eigenvectors W (sigma = |lambda|) and lambda_i = sigma_i (w_i . uh_i)
    std::vector<GemmJob> mv, hh;
    std::vector<WideRR> rh;
    for (size_t i = 0; i < idx.size(); ++i) {
        const EighIn& s = in[idx[i]];
        WS& x = w[i];
        x.H = Buf(&ctx.pool, (size_t)x.l * x.l * 8);
        x.Uh = Buf(&ctx.pool, (size_t)x.l * x.l * 8);
        x.sH = Buf(&ctx.pool, (size_t)x.l * 8);
        x.Wt = Buf(&ctx.pool, (size_t)x.l * x.l * 8);
        mv.push_back(GemmJob{s.M, x.V.d(), x.Y.d(), s.n, x.l, s.n, s.ld, s.n, s.n, 0, 0});
        hh.push_back(GemmJob{x.V.d(), x.Y.d(), x.H.d(), x.l, x.l, s.n, s.n, s.n, x.l, 1, 0});
        rh.push_back(WideRR{x.H.d(), x.l, x.l, x.l, x.Uh.d(), x.sH.d(), x.Wt.d()});
    }
    gemm_grouped(ctx, mv);
    gemm_grouped(ctx, hh);
    auto sth = wide_rr(ctx, rh, cfg.jacobi_tol, cfg.jacobi_sweeps);
    std::vector<GemmJob> vu;
    for (size_t i = 0; i < idx.size(); ++i) {
        const EighIn& s = in[idx[i]];
        WS& x = w[i];
        EighOut& o = out[idx[i]];
        if (sth[i] != TNR_OK) o.status = TNR_ERR_NOCONV;
        // signs of the eigenvalues on the host (l x l, a large-chi path)
        const size_t ll = (size_t)x.l * x.l;
        std::vector<double> hW(ll), hU(ll), hs(x.l);
        ctx.fetch(hW.data(), x.Wt.d(), ll * 8);
        ctx.fetch(hU.data(), x.Uh.d(), ll * 8);
        ctx.fetch(hs.data(), x.sH.d(), (size_t)x.l * 8);
        for (int c = 0; c < x.l; ++c) {
            double d = 0.0;
            for (int r = 0; r < x.l; ++r) d += hW[(size_t)c * x.l + r] * hU[(size_t)c * x.l + r];
            if (d < 0.0) hs[c] = -hs[c];
        }
        o.lam = Buf(&ctx.pool, (size_t)x.l * 8);
        void* t = ctx.stage->put(ctx.stream, hs.data(), (size_t)x.l * 8);
        TNR_HIP(hipMemcpyAsync(o.lam.d(), t, (size_t)x.l * 8, hipMemcpyDeviceToDevice, ctx.stream));
        o.U = Buf(&ctx.pool, (size_t)s.n * x.l * 8);
        o.ncols = x.l;
        o.direct = whole;     // all n eigenvalues: the truncation error is exact
        vu.push_back(GemmJob{x.V.d(), x.Wt.d(), o.U.d(), s.n, x.l, x.l, s.n, x.l, s.n, 0, 0});
    }
    gemm_grouped(ctx, vu);
}

}  // namespace

// ---------------------------------------------------------------------- SVD
void svd_grouped(Ctx& ctx, const std::vector<SvdIn>& in, std::vector<SvdOut>& out) {
    out.clear();
    out.resize(in.size());
    const SolverCfg& cfg = solver_cfg();
    Buf failv = fail_flags(ctx, in.size());
    std::vector<char> retried(in.size(), 0);   // refactorised whole after a stall
    std::vector<int> direct, iter, wide;
    for (size_t p = 0; p < in.size(); ++p) {
        const SvdIn& s = in[p];
        TNR_REQUIRE(s.m > 0 && s.n > 0 && s.kmax > 0 && s.kmax <= std::min(s.m, s.n),
                    TNR_ERR_INTERNAL, "bad SVD problem");
        if (std::min(s.m, s.n) <= LMAX) direct.push_back((int)p);
        else if (s.kmax <= LMAX - 8) iter.push_back((int)p);
        else wide.push_back((int)p);
    }
    if (!wide.empty()) svd_wide(ctx, in, wide, out);
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
                                    cs.prod(i), cs.done(i), std::min(s.m, s.n) <= kWholeMax});
            // the loop's last Rayleigh-Ritz step left Z W in Y2; one pass at the
            // final threshold (into Z, free until V Wtot) rebuilds V.  With
            // Cholesky QR on the Z side Y2 is not a Ritz basis: two passes on Z.
            if (cfg.cholqr_z)
                rV.push_back(RR{x.Z.d(), s.n, l, s.n, x.Y2.d(), x.G.d(), x.W.d(), x.W2.d(),
                                x.lam.d(), x.sw.as<int>(), nullptr, x.V.d(), s.n, sv, nullptr, fl});
            else
                rV.push_back(RR{x.Y2.d(), s.n, l, s.n, x.Z.d(), x.G.d(), x.W.d(), x.W2.d(),
                                x.lam.d(), x.sw.as<int>(), nullptr, x.V.d(), s.n, sv, nullptr, fl});
            // Ritz extraction from Y = M V over the whole basis (the top
            // triplets recombine all l columns: extracting them from the
            // leading columns only moves them by the subspace's residual
            // angle, ~1e-7), V <- V Wtot
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
        iterate(ctx, cs, [&](int it) {
            gemm_issue(ctx, pMV);
            rr_issue(ctx, pY, it == 0);
            gemm_issue(ctx, pMtU);
            rr_issue(ctx, pZ, it == 0);
            launch_conv(ctx.stream, d_conv, P, cfg.null_rtol, cfg.conv_target, cfg.min_iters,
                        cfg.max_iters, cfg.stall_from);
        });
        // the loop orthonormalises only to its looser Jacobi threshold: V is
        // rebuilt at full accuracy before the final extraction relies on
        // V^T V = I.  Its last rotated Z has columns orthogonal to that
        // threshold, so one pass at the final threshold does what the second
        // pass of a two-pass step does.
        if (cfg.cholqr_z) rr_grouped(ctx, rV, cfg.rv_passes, cfg.rv_passes == 1);
        else rr_grouped(ctx, rV, 1, true);
        gemm_grouped(ctx, jF);
        rr_grouped(ctx, rF, cfg.rf_passes, cfg.rf_passes == 1);
        gemm_grouped(ctx, jV);
        std::vector<int> its(P);
        std::vector<double> prods(P);
        ctx.fetch(its.data(), cs.it(0), P * sizeof(int));
        ctx.fetch(prods.data(), cs.prod(0), P * sizeof(double));
        std::vector<int> stalled;
        for (int i = 0; i < P; ++i) {
            SvdOut& o = out[iter[i]];
            IW& x = w[i];
            o.U = std::move(x.U); o.ldu = in[iter[i]].m;
            o.V = std::move(x.Z); o.ldv = in[iter[i]].n;
            o.s = std::move(x.sigU);
            o.iters = its[i];
            o.conv = prods[i];
            if (!(prods[i] < cfg.conv_target)) {   // stopped at max_iters
                const SvdIn& si = in[iter[i]];
                if (std::min(si.m, si.n) <= kWholeMax) stalled.push_back(iter[i]);
                else o.status = TNR_ERR_NOCONV;
            }
        }
        // a slowly decaying spectrum (sigma_l / sigma_k near 1) can stall the
        // iteration: such problems, when small enough, are factorised whole
        if (!stalled.empty()) {
            for (int p : stalled) {
                out[p] = SvdOut();
                retried[p] = 1;
            }
            svd_wide(ctx, in, stalled, out, true);
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
        out[p].status = problem_status(out[p].s_host, retried[p] ? 0.0 : failh[p], out[p].status);
}

// ---------------------------------------------------------------- eigh
void eigh_grouped(Ctx& ctx, const std::vector<EighIn>& in, std::vector<EighOut>& out) {
    out.clear();
    out.resize(in.size());
    const SolverCfg& cfg = solver_cfg();
    Buf failv = fail_flags(ctx, in.size());
    std::vector<char> retried(in.size(), 0);   // refactorised whole after a stall
    std::vector<int> direct, iter, wide;
    for (size_t p = 0; p < in.size(); ++p) {
        const EighIn& s = in[p];
        TNR_REQUIRE(s.n > 0 && s.kmax > 0 && s.kmax <= s.n, TNR_ERR_INTERNAL, "bad eigh problem");
        if (s.n <= LMAX) direct.push_back((int)p);
        else if (s.kmax <= LMAX - 8) iter.push_back((int)p);
        else wide.push_back((int)p);
    }
    if (!wide.empty()) eigh_wide(ctx, in, wide, out);
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
            // Ritz values, or with Cholesky QR the pivots R_jj
            conv.push_back(ConvDesc{cfg.cholqr_eigh ? x.lam.d() : sv, l, s.kmax, 0, cs.it(i),
                                    cs.prod(i), cs.done(i), s.n <= kWholeMax});
            // one tight pass on the loop's last rotated Y (in Y2), output into
            // Y; after Cholesky QR steps Y2 is no Ritz basis: two passes on Y
            if (cfg.cholqr_eigh)
                rV.push_back(RR{x.Y.d(), s.n, l, s.n, x.Y2.d(), x.G.d(), x.W.d(), x.W2.d(),
                                x.lam.d(), x.sw.as<int>(), nullptr, x.V.d(), s.n, sv, nullptr,
                                failv.d() + iter[i]});
            else
                rV.push_back(RR{x.Y2.d(), s.n, l, s.n, x.Y.d(), x.G.d(), x.W.d(), x.W2.d(),
                                x.lam.d(), x.sw.as<int>(), nullptr, x.V.d(), s.n, sv, nullptr,
                                failv.d() + iter[i]});
            jF.push_back(GemmJob{s.M, x.V.d(), x.Y.d(), s.n, l, s.n, s.ld, s.n, s.n, 0, 0});
            jH.push_back(GemmJob{x.V.d(), x.Y.d(), x.G.d(), l, l, s.n, s.n, s.n, l, 1, 0});
            nmax = std::max(nmax, s.n);
        }
        ColDesc* dsym = ctx.stage->put(ctx.stream, sym);
        launch_symmetrize(ctx.stream, dsym, (int)sym.size(), nmax);
        randfill(ctx, fills);
        // (random start block used as is, see svd_grouped)
        GemmPlan pMV = gemm_plan(ctx, jMV);
        RRPlan pY = rr_plan(ctx, rY, 1, cfg.cholqr_eigh);
        DescArena a;
        gemm_stage(pMV, a);
        rr_stage(pY, a);
        const size_t o_conv = a.add(conv);
        a.commit(ctx);
        gemm_bind(pMV, a);
        rr_bind(pY, a);
        const ConvDesc* d_conv = a.at<ConvDesc>(o_conv);
        iterate(ctx, cs, [&](int it) {
            gemm_issue(ctx, pMV);
            rr_issue(ctx, pY, it == 0);
            launch_conv(ctx.stream, d_conv, P, cfg.null_rtol, cfg.conv_target, cfg.min_iters,
                        cfg.max_iters, cfg.stall_from);
        });
        // V at full orthonormality from the loop's last rotated Y = M V W
        // (the loop uses a looser Jacobi threshold; one pass at the final
        // threshold, as in svd_grouped), or from the last Y in two passes
        if (cfg.cholqr_eigh) rr_grouped(ctx, rV, 2);
        else rr_grouped(ctx, rV, 1, true);
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
        std::vector<int> stalled;
        for (int i = 0; i < P; ++i) {
            out[iter[i]].iters = its[i];
            out[iter[i]].conv = prods[i];
            if (!(prods[i] < cfg.conv_target)) {   // at max_iters
                if (in[iter[i]].n <= kWholeMax) stalled.push_back(iter[i]);
                else out[iter[i]].status = TNR_ERR_NOCONV;
            }
        }
        // a stalled iteration (|lambda_l / lambda_k| near 1) of a small enough
        // matrix: the whole eigendecomposition instead
        if (!stalled.empty()) {
            for (int p : stalled) {
                out[p] = EighOut();
                retried[p] = 1;
            }
            eigh_wide(ctx, in, stalled, out, true);
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
        out[p].status = problem_status(out[p].lam_host, retried[p] ? 0.0 : failh[p], out[p].status);
}

}  // namespace tnr
