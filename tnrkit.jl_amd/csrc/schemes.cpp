// TRG / BTRG / HOTRG steps (see schemes.h).  Networks follow
// src/schemes/trg.jl:35-41, src/schemes/btrg.jl:60-94 and
// src/schemes/hotrg.jl:42-160; finalize follows src/utility/finalize.jl:1-14.
#include "schemes.h"

#include <algorithm>
#include <cmath>
#include <functional>
#include <numeric>

namespace tnr {

namespace {

uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// seed of a solver problem: a function of its shape only, so a run gets the
// same start vectors alone and inside a batch
uint64_t problem_seed(int m, int n, int c, int tag) {
    return mix64(((uint64_t)m << 40) ^ ((uint64_t)n << 16) ^ ((uint64_t)c << 8) ^ (uint64_t)tag);
}

// Bond vectors of many schemes in one device store, one upload.
std::vector<BondVec> make_bonds(Ctx& ctx, const std::vector<int>& Ns,
                                const std::vector<const int*>& dims,
                                const std::vector<const std::vector<double>*>& vals) {
    std::vector<BondVec> out(dims.size());
    std::vector<double> all;
    for (size_t b = 0; b < dims.size(); ++b) {
        BondVec& v = out[b];
        const int N = Ns[b];
        v.N = N;
        int o = 0;
        for (int c = 0; c < N; ++c) {
            v.dim[c] = dims[b][c];
            v.off[c] = o;
            o += dims[b][c];
        }
        TNR_REQUIRE((size_t)o == vals[b]->size(), TNR_ERR_INTERNAL, "bond vector size mismatch");
        v.host = *vals[b];
        v.base = all.size();
        all.insert(all.end(), vals[b]->begin(), vals[b]->end());
    }
    auto store = std::make_shared<Buf>(&ctx.pool, std::max<size_t>(1, all.size()) * sizeof(double));
    if (!all.empty()) {
        void* tmp = ctx.stage->put(ctx.stream, all.data(), all.size() * sizeof(double));
        TNR_HIP(hipMemcpyAsync(store->d(), tmp, all.size() * sizeof(double),
                               hipMemcpyDeviceToDevice, ctx.stream));
    }
    for (auto& v : out) v.store = store;
    return out;
}

BondVec make_bond(Ctx& ctx, int N, const int* dim, const std::vector<double>& vals) {
    return std::move(make_bonds(ctx, {N}, {dim}, {&vals})[0]);
}

// s^a where s >= tol, else s unchanged (src/schemes/btrg.jl:50-58)
double pseudopow(double s, double a) {
    const double tol = std::pow(2.220446049250313e-16, 0.75);
    return s < tol ? s : std::pow(s, a);
}

struct Factor {
    BT U, V;                 // U-side (codomain legs | new), V-side (new | domain legs)
    std::vector<double> s;   // kept singular values, by sector c then descending
    int dim[MAXN] = {0};     // kept count per sector c
    double kept_min = 0, dropped_max = 0;
    int iters = 0;
    double conv = 0;
    int status = 0;          // worst block status of the solver (TNR_OK if none failed)
};

// weights (wu on U's new leg, wv on V's new leg) of tensor t from its kept
// singular values
using Weigh = std::function<void(size_t t, const std::vector<double>&, std::vector<double>&,
                                  std::vector<double>&)>;
using SvdWeigh = Weigh;

// svd_trunc(T; trunc = truncrank(chi)) per coupled-charge block with the
// chi largest values kept across blocks (stable, descending), then
// U * diag(wu) and diag(wv) * V^T as block tensors.
// Pivoted-row gauge (canonical = true; the impurity TRG needs it because
// its impurity norm contracts factors of different SVDs): per block, kept
// values within GAUGE_CLUSTER * s_max of their neighbour form a cluster,
// values at or below GAUGE_NULL * s_max are null; inside each cluster the
// basis is rebuilt by k_canon (the normalised projection of the pivot
// coordinate vector onto what is left of the span, pivot = first row whose
// residual norm is within GAUGE_TIE of the largest), then U Q and V Q are
// formed with one grouped GEMM.  Result: a function of each cluster's span,
// positive at every pivot row, never mixing coordinates a diagonal symmetry
// separates (the hidden Z2 of a Trivial Ising tensor).
constexpr double GAUGE_CLUSTER = 1e-9;
constexpr double GAUGE_NULL = 1e-14;


std::vector<Factor> svd_trunc_many(Ctx& ctx, const std::vector<const BT*>& Ts,
                                   const std::vector<int>& chis, const std::vector<int>& tags,
                                   const Weigh& weigh, bool canonical = false) {
    std::vector<SvdIn> in;
    std::vector<std::pair<int, int>> owner;
    for (size_t t = 0; t < Ts.size(); ++t) {
        const Layout& L = Ts[t]->L;
        TNR_REQUIRE(chis[t] >= 1, TNR_ERR_ARG, "chi must be >= 1");
        for (int c = 0; c < L.N; ++c) {
            const int R = (int)L.R[c], C = (int)L.C[c];
            if (R == 0 || C == 0) continue;
            SvdIn p;
            p.M = Ts[t]->d() + L.off[c];
            p.m = R; p.n = C; p.ld = R;
            p.kmax = std::min(chis[t], std::min(R, C));
            p.seed = problem_seed(R, C, c, tags[t]);
            in.push_back(p);
            owner.emplace_back((int)t, c);
        }
    }
    std::vector<SvdOut> out;
    svd_grouped(ctx, in, out);
    std::vector<Factor> F(Ts.size());
    std::vector<GatherDesc> gd;
    std::vector<int64_t> gw;     // per descriptor: offset of its weights in wall
    std::vector<double> wall;    // all weight vectors, uploaded once
    // pivoted-row gauge: descriptors, cluster bounds (offsets into allcl),
    // rotation GEMMs and the buffers they write
    std::vector<CanonDesc> cd;
    std::vector<size_t> cl_off;
    std::vector<int> allcl;
    std::vector<GemmJob> rot;
    std::vector<Buf> keep_alive;
    int canon_kmax = 0;
    size_t p0 = 0;
    for (size_t t = 0; t < Ts.size(); ++t) {
        const Layout& L = Ts[t]->L;
        const int N = L.N;
        size_t p1 = p0;
        while (p1 < owner.size() && owner[p1].first == (int)t) ++p1;
        // global cut
        struct Cand { double v; int blk, idx; };
        std::vector<Cand> flat;
        for (size_t p = p0; p < p1; ++p)
            for (int i = 0; i < in[p].kmax; ++i) flat.push_back(Cand{out[p].s_host[i], (int)p, i});
        std::stable_sort(flat.begin(), flat.end(),
                         [](const Cand& a, const Cand& b) { return a.v > b.v; });
        const size_t keep = std::min<size_t>(chis[t], flat.size());
        TNR_REQUIRE(keep > 0, TNR_ERR_ARG, "tensor has no non-empty block");
        Factor& f = F[t];
        std::vector<int> kc(in.size(), 0);
        for (size_t i = 0; i < keep; ++i) kc[flat[i].blk]++;
        f.kept_min = flat[keep - 1].v;
        f.dropped_max = keep < flat.size() ? flat[keep].v : 0.0;
        for (size_t p = p0; p < p1; ++p) {
            f.dim[owner[p].second] = kc[p];
            f.iters = std::max(f.iters, out[p].iters);
            f.conv = std::max(f.conv, out[p].conv);
            if (out[p].status != TNR_OK) f.status = out[p].status;
        }
        int off[MAXN], o = 0;
        for (int c = 0; c < N; ++c) { off[c] = o; o += f.dim[c]; }
        f.s.assign(o, 0.0);
        for (size_t p = p0; p < p1; ++p) {
            const int c = owner[p].second;
            for (int i = 0; i < kc[p]; ++i) f.s[off[c] + i] = out[p].s_host[i];
        }
        std::vector<double> wu, wv;
        weigh(t, f.s, wu, wv);
        // device addresses are patched in after the single upload below
        const size_t owu = wall.size();
        wall.insert(wall.end(), wu.begin(), wu.end());
        const size_t owv = wall.size();
        wall.insert(wall.end(), wv.begin(), wv.end());
        // layouts of the factors
        Leg nl_u, nl_v;
        for (int c = 0; c < N; ++c) {
            nl_u.d[(N - c) % N] = f.dim[c];   // domain leg with label c
            nl_v.d[c] = f.dim[c];             // codomain leg with label c
        }
        std::vector<Leg> lu, lv;
        for (int i = 0; i < L.ncod; ++i) lu.push_back(L.legs[i]);
        lu.push_back(nl_u);
        lv.push_back(nl_v);
        for (int i = L.ncod; i < L.nleg; ++i) lv.push_back(L.legs[i]);
        f.U = BT(ctx, Layout(N, L.ncod, lu));
        f.V = BT(ctx, Layout(N, 1, lv));
        for (size_t p = p0; p < p1; ++p) {
            const int c = owner[p].second;
            if (!kc[p]) continue;
            TNR_REQUIRE(f.U.L.R[c] == in[p].m && f.U.L.C[c] == kc[p] && f.V.L.R[c] == kc[p] &&
                            f.V.L.C[c] == in[p].n,
                        TNR_ERR_INTERNAL, "factor layout mismatch");
            const double* usrc = out[p].U.d();
            const double* vsrc = out[p].V.d();
            int ldus = out[p].ldu, ldvs = out[p].ldv;
            if (canonical) {
                TNR_REQUIRE(kc[p] <= LMAX, TNR_ERR_ARG,
                            "the pivoted-row SVD gauge of ImpurityTRG keeps at most 96 values "
                            "per symmetry block");
                const int k = kc[p], m = in[p].m, n = in[p].n;
                int knn = 0;
                cl_off.push_back(allcl.size());
                const int nc = gauge_clusters(out[p].s_host.data(), k, allcl, knn);
                Buf Q(&ctx.pool, (size_t)k * k * 8), Ur(&ctx.pool, (size_t)m * k * 8),
                    Vr(&ctx.pool, (size_t)n * k * 8);
                cd.push_back(CanonDesc{usrc, Q.d(), nullptr, m, ldus, k, knn, nc});
                rot.push_back(GemmJob{usrc, Q.d(), Ur.d(), m, k, k, ldus, k, m, 0, 0});
                rot.push_back(GemmJob{vsrc, Q.d(), Vr.d(), n, k, k, ldvs, k, n, 0, 0});
                canon_kmax = std::max(canon_kmax, k);
                usrc = Ur.d(); ldus = m;
                vsrc = Vr.d(); ldvs = n;
                keep_alive.push_back(std::move(Q));
                keep_alive.push_back(std::move(Ur));
                keep_alive.push_back(std::move(Vr));
            }
            // U * diag(wu): column weights; diag(wv) * V^T: row weights
            GatherDesc gu{};
            gu.in = usrc; gu.out = f.U.d() + f.U.L.off[c];
            gu.rows = in[p].m; gu.cols = kc[p]; gu.ldi = ldus; gu.ldo = in[p].m;
            gd.push_back(gu);
            gw.push_back((int64_t)(owu + off[c]));
            GatherDesc gv{};
            gv.in = vsrc; gv.out = f.V.d() + f.V.L.off[c];
            gv.rows = kc[p]; gv.cols = in[p].n; gv.ldi = ldvs; gv.ldo = kc[p]; gv.trans = 1;
            gd.push_back(gv);
            gw.push_back((int64_t)(owv + off[c]));
        }
        p0 = p1;
    }
    if (gd.empty()) return F;
    if (!cd.empty()) {
        if (allcl.empty()) allcl.push_back(0);
        const int* dcl = static_cast<const int*>(
            ctx.stage->put(ctx.stream, allcl.data(), std::max<size_t>(1, allcl.size()) * sizeof(int)));
        for (size_t i = 0; i < cd.size(); ++i) cd[i].cl = dcl + cl_off[i];
        CanonDesc* dcd = ctx.stage->put(ctx.stream, cd);
        launch_canon(ctx.stream, dcd, (int)cd.size(), canon_kmax, kGaugeTie);
        gemm_grouped(ctx, rot);
    }
    const double* dw = static_cast<const double*>(
        ctx.stage->put(ctx.stream, wall.data(), std::max<size_t>(1, wall.size()) * sizeof(double)));
    for (size_t i = 0; i < gd.size(); ++i) {
        if (gd[i].trans) gd[i].rw = dw + gw[i];   // V side
        else gd[i].cw = dw + gw[i];               // U side
    }
    gather_many(ctx, gd);
    return F;
}

struct EFactor {
    BT U;                    // (codomain legs | new)
    double eps = 0;          // 2-norm of the discarded eigenvalues
    double kept_min = 0, dropped_max = 0;
    int iters = 0;
    double conv = 0;
    int status = 0;          // worst block status of the solver (TNR_OK if none failed)
};

// eigh_trunc!(project_hermitian!(MM); trunc = truncrank(chi)): the chi
// largest |lambda| across blocks; MM is overwritten.
std::vector<EFactor> eigh_trunc_many(Ctx& ctx, const std::vector<BT*>& Ms,
                                     const std::vector<int>& chis, int tag) {
    std::vector<EighIn> in;
    std::vector<std::pair<int, int>> owner;
    for (size_t t = 0; t < Ms.size(); ++t) {
        const Layout& L = Ms[t]->L;
        for (int c = 0; c < L.N; ++c) {
            const int R = (int)L.R[c];
            if (R == 0) continue;
            TNR_REQUIRE(L.C[c] == R, TNR_ERR_INTERNAL, "environment block is not square");
            EighIn p;
            p.M = Ms[t]->d() + L.off[c];
            p.n = R; p.ld = R;
            p.kmax = std::min(chis[t], R);
            p.seed = problem_seed(R, R, c, tag);
            in.push_back(p);
            owner.emplace_back((int)t, c);
        }
    }
    std::vector<EighOut> out;
    eigh_grouped(ctx, in, out);
    std::vector<EFactor> F(Ms.size());
    std::vector<int> kc(in.size(), 0);
    std::vector<double> eps2(Ms.size(), 0.0);
    // cut and the directly known part of the truncation error
    size_t p0 = 0;
    std::vector<std::pair<size_t, size_t>> ranges;
    for (size_t t = 0; t < Ms.size(); ++t) {
        size_t p1 = p0;
        while (p1 < owner.size() && owner[p1].first == (int)t) ++p1;
        ranges.emplace_back(p0, p1);
        struct Cand { double a; int blk, idx; };
        std::vector<Cand> flat;
        for (size_t p = p0; p < p1; ++p)
            for (int i = 0; i < in[p].kmax; ++i)
                flat.push_back(Cand{std::fabs(out[p].lam_host[i]), (int)p, i});
        std::stable_sort(flat.begin(), flat.end(),
                         [](const Cand& a, const Cand& b) { return a.a > b.a; });
        const size_t keep = std::min<size_t>(chis[t], flat.size());
        TNR_REQUIRE(keep > 0, TNR_ERR_ARG, "environment has no non-empty block");
        for (size_t i = 0; i < keep; ++i) kc[flat[i].blk]++;
        F[t].kept_min = flat[keep - 1].a;
        F[t].dropped_max = keep < flat.size() ? flat[keep].a : 0.0;
        for (size_t p = p0; p < p1; ++p) {
            F[t].iters = std::max(F[t].iters, out[p].iters);
            F[t].conv = std::max(F[t].conv, out[p].conv);
            if (out[p].status != TNR_OK) F[t].status = out[p].status;
            if (out[p].direct)
                for (int i = kc[p]; i < in[p].n; ++i) eps2[t] += out[p].lam_host[i] * out[p].lam_host[i];
        }
        p0 = p1;
    }
    // iterative blocks: |MM - U_k L_k U_k^T|_F^2 on the device
    {
        std::vector<GatherDesc> gd;
        std::vector<GemmJob> gj;
        std::vector<ColDesc> fd;
        std::vector<Buf> tmp;
        std::vector<size_t> which;
        Buf fro(&ctx.pool, std::max<size_t>(1, in.size()) * sizeof(double));
        for (size_t p = 0; p < in.size(); ++p) {
            if (out[p].direct) continue;
            const int n = in[p].n, k = kc[p];
            if (k > 0) {
                tmp.emplace_back(&ctx.pool, (size_t)n * k * 8);
                GatherDesc g{};
                g.in = out[p].U.d(); g.out = tmp.back().d(); g.cw = out[p].lam.d();
                g.rows = n; g.cols = k; g.ldi = n; g.ldo = n;
                gd.push_back(g);
                GemmJob j{tmp.back().d(), out[p].U.d(), in[p].M, n, n, k, n, n, in[p].ld, 0, 1};
                j.alpha = -1.0;
                j.beta = 1.0;
                gj.push_back(j);
            }
            fd.push_back(ColDesc{in[p].M, nullptr, fro.d() + which.size(), n, n, in[p].ld, 0});
            which.push_back(p);
        }
        if (!which.empty()) {
            gather_many(ctx, gd);
            gemm_grouped(ctx, gj);
            ColDesc* dfd = ctx.stage->put(ctx.stream, fd);
            Buf part(&ctx.pool, fd.size() * FROB_PARTS * sizeof(double));
            launch_frob(ctx.stream, dfd, (int)fd.size(), part.d());
            std::vector<double> h(which.size());
            ctx.fetch(h.data(), fro.d(), h.size() * 8);
            for (size_t i = 0; i < which.size(); ++i) eps2[owner[which[i]].first] += h[i];
        }
    }
    std::vector<GatherDesc> gd;
    for (size_t t = 0; t < Ms.size(); ++t) {
        const Layout& L = Ms[t]->L;
        const int N = L.N;
        F[t].eps = std::sqrt(eps2[t]);
        Leg nl;
        for (size_t p = ranges[t].first; p < ranges[t].second; ++p)
            nl.d[(N - owner[p].second) % N] = kc[p];
        F[t].U = BT(ctx, Layout(N, 2, {L.legs[0], L.legs[1], nl}));
        for (size_t p = ranges[t].first; p < ranges[t].second; ++p) {
            const int c = owner[p].second;
            if (!kc[p]) continue;
            TNR_REQUIRE(F[t].U.L.R[c] == in[p].n && F[t].U.L.C[c] == kc[p], TNR_ERR_INTERNAL,
                        "projector layout mismatch");
            GatherDesc g{};
            g.in = out[p].U.d(); g.out = F[t].U.d() + F[t].U.L.off[c];
            g.rows = in[p].n; g.cols = kc[p]; g.ldi = in[p].n; g.ldo = in[p].n;
            gd.push_back(g);
        }
    }
    gather_many(ctx, gd);
    return F;
}

// a factorization whose solver failed (unconverged, non-finite values)
// marks the scheme failed: the reference's LAPACK-backed factorizations
// raise there
void log_factor(Scheme& s, double kept, double dropped, int iters, double conv, int status) {
    s.tlog.push_back({kept, dropped, (double)iters, conv});
    if (status != TNR_OK && s.status == TNR_OK) s.status = status;
}

// ------------------------------------------------------------------- TRG
// T[-1 -2; -3 -4] := D[-2; 1 2] * B[-1; 4 1] * C[4 3; -3] * A[3 2; -4]
// (src/schemes/trg.jl:39, impuritytrg.jl:84-88) for every (A, B, C, D)
struct TrgNet { const BT *A, *B, *C, *D; };

std::vector<BT> trg_networks(Ctx& ctx, const std::vector<TrgNet>& nets) {
    std::vector<ContractJob> cj;
    for (const TrgNet& t : nets) {
        cj.push_back(ContractJob{{t.B, {-1, -4, 1}}, {t.D, {-2, 1, -3}}, {-1, -2, -4, -3}, 2});
        cj.push_back(ContractJob{{t.C, {-1, 1, -3}}, {t.A, {1, -2, -4}}, {-1, -2, -3, -4}, 2});
    }
    auto XY = contract_many(ctx, cj);
    cj.clear();
    for (size_t i = 0; i < nets.size(); ++i)
        cj.push_back(ContractJob{{&XY[2 * i], {-1, -2, 1, 2}}, {&XY[2 * i + 1], {1, 2, -3, -4}},
                                 {-1, -2, -3, -4}, 2});
    return contract_many(ctx, cj);
}

SvdWeigh sqrt_weights() {
    return [](size_t, const std::vector<double>& s, std::vector<double>& wu,
              std::vector<double>& wv) {
        wu.resize(s.size());
        for (size_t i = 0; i < s.size(); ++i) wu[i] = std::sqrt(s[i]);
        wv = wu;
    };
}

void step_trg(Ctx& ctx, const std::vector<Scheme*>& G, const std::vector<int>& chis) {
    const size_t n = G.size();
    auto sqrtw = [](size_t, const std::vector<double>& s, std::vector<double>& wu,
                    std::vector<double>& wv) {
        wu.resize(s.size());
        for (size_t i = 0; i < s.size(); ++i) wu[i] = std::sqrt(s[i]);
        wv = wu;
    };
    // Tp = transpose(T, ((2, 4), (1, 3)))
    std::vector<PermJob> pj;
    for (auto* s : G) pj.push_back(PermJob{&s->T, {1, 3, 0, 2}, 2, false, {}});
    auto Tp = permute_many(ctx, pj);
    // A, B = SVD12(T); C, D = SVD12(Tp) (src/utility/projectors.jl:211-217),
    // both factorizations of every scheme in one grouped solve
    std::vector<const BT*> Ts;
    std::vector<int> ch, tags;
    for (size_t i = 0; i < n; ++i) { Ts.push_back(&G[i]->T); ch.push_back(chis[i]); tags.push_back(11); }
    for (size_t i = 0; i < n; ++i) { Ts.push_back(&Tp[i]); ch.push_back(chis[i]); tags.push_back(12); }
    auto F = svd_trunc_many(ctx, Ts, ch, tags, sqrtw);
    std::vector<Factor> F1(std::make_move_iterator(F.begin()), std::make_move_iterator(F.begin() + n));
    std::vector<Factor> F2(std::make_move_iterator(F.begin() + n), std::make_move_iterator(F.end()));
    std::vector<TrgNet> nets;
    for (size_t i = 0; i < n; ++i) nets.push_back(TrgNet{&F1[i].U, &F1[i].V, &F2[i].U, &F2[i].V});
    auto Tn = trg_networks(ctx, nets);
    for (size_t i = 0; i < n; ++i) {
        G[i]->tlog.clear();
        log_factor(*G[i], F1[i].kept_min, F1[i].dropped_max, F1[i].iters, F1[i].conv,
                   F1[i].status);
        log_factor(*G[i], F2[i].kept_min, F2[i].dropped_max, F2[i].iters, F2[i].conv,
                   F2[i].status);
        G[i]->T = std::move(Tn[i]);
    }
}

// ------------------------------------------------------------------ BTRG
void step_btrg(Ctx& ctx, const std::vector<Scheme*>& G, const std::vector<int>& chis) {
    const size_t n = G.size();
    std::vector<const BT*> Ts;
    for (auto* s : G) Ts.push_back(&s->T);
    // U, S, V = svd_trunc(T): A = U S_a, B = S_a V, S1' = S_b
    // U, S, V = svd_trunc(permute(T, ((3, 1), (4, 2)))): C, D, S2'
    // (both factorizations of every scheme in one grouped solve)
    std::vector<PermJob> pj;
    for (auto* s : G) pj.push_back(PermJob{&s->T, {2, 0, 3, 1}, 2, false, {}});
    auto P = permute_many(ctx, pj);
    std::vector<const BT*> all(Ts);
    std::vector<int> ch(chis), tags(n, 21);
    for (size_t i = 0; i < n; ++i) { all.push_back(&P[i]); ch.push_back(chis[i]); tags.push_back(22); }
    std::vector<std::vector<double>> sb(2 * n);
    auto weigh = [&G, &sb, n](size_t t, const std::vector<double>& s, std::vector<double>& wu,
                              std::vector<double>& wv) {
        const double k = G[t % n]->k;
        wu.resize(s.size());
        sb[t].resize(s.size());
        for (size_t i = 0; i < s.size(); ++i) {
            wu[i] = pseudopow(s[i], (1.0 - k) / 2.0);   // S_a
            sb[t][i] = pseudopow(s[i], k);               // S_b
        }
        wv = wu;
    };
    auto F = svd_trunc_many(ctx, all, ch, tags, weigh);
    std::vector<Factor> F1(std::make_move_iterator(F.begin()), std::make_move_iterator(F.begin() + n));
    std::vector<Factor> F2(std::make_move_iterator(F.begin() + n), std::make_move_iterator(F.end()));
    // T[-1 -2; -3 -4] := D[-1; 4 7] S1[1; 7] B[-2; 1 3] S2[3; 2]
    //                    C[8 2; -4] S1[8; 5] A[6 5; -3] S2[4; 6]
    // with diagonal S1 (vertical) and S2 (horizontal) absorbed as leg weights:
    //   X[p,q,f,c] = sum_g D[p,f,g] s2[f] s1[g] B[q,g,c]
    //   Y[f,c,r,s] = sum_h C[h,c,s] s1[h] s2[c] A[f,h,r]
    //   T'[p,q,r,s] = sum_{f,c} X[p,q,f,c] Y[f,c,r,s]
    std::vector<LegWeights> w1s(n), w1d(n), w2s(n), w2d(n);
    for (size_t i = 0; i < n; ++i) {
        w1s[i] = G[i]->S1.same_side(); w1d[i] = G[i]->S1.dual_side();
        w2s[i] = G[i]->S2.same_side(); w2d[i] = G[i]->S2.dual_side();
    }
    std::vector<ContractJob> cj;
    for (size_t i = 0; i < n; ++i) {
        const BT& A = F1[i].U; const BT& B = F1[i].V;
        const BT& C = F2[i].U; const BT& D = F2[i].V;
        Operand d{&D, {-1, -2, 1}, false, {WeightSpec{1, w2d[i]}, WeightSpec{2, w1s[i]}}};
        Operand b{&B, {-3, 1, -4}};
        cj.push_back(ContractJob{d, b, {-1, -3, -2, -4}, 2});
        Operand c{&C, {1, -2, -3}, false, {WeightSpec{0, w1d[i]}, WeightSpec{1, w2s[i]}}};
        Operand a{&A, {-1, 1, -4}};
        cj.push_back(ContractJob{c, a, {-1, -2, -4, -3}, 2});
    }
    auto XY = contract_many(ctx, cj);
    cj.clear();
    for (size_t i = 0; i < n; ++i)
        cj.push_back(ContractJob{{&XY[2 * i], {-1, -2, 1, 2}}, {&XY[2 * i + 1], {1, 2, -3, -4}},
                                 {-1, -2, -3, -4}, 2});
    auto Tn = contract_many(ctx, cj);
    for (size_t i = 0; i < n; ++i) {
        Scheme& s = *G[i];
        s.tlog.clear();
        log_factor(s, F1[i].kept_min, F1[i].dropped_max, F1[i].iters, F1[i].conv,
                   F1[i].status);
        log_factor(s, F2[i].kept_min, F2[i].dropped_max, F2[i].iters, F2[i].conv,
                   F2[i].status);
        s.T = std::move(Tn[i]);
    }
    std::vector<int> Ns;
    std::vector<const int*> dims;
    std::vector<const std::vector<double>*> vals;
    for (size_t i = 0; i < n; ++i) {
        Ns.push_back(G[i]->N); dims.push_back(F1[i].dim); vals.push_back(&sb[i]);
        Ns.push_back(G[i]->N); dims.push_back(F2[i].dim); vals.push_back(&sb[n + i]);
    }
    auto bonds = make_bonds(ctx, Ns, dims, vals);
    for (size_t i = 0; i < n; ++i) {
        G[i]->S1 = std::move(bonds[2 * i]);
        G[i]->S2 = std::move(bonds[2 * i + 1]);
    }
}

// ----------------------------------------------------------------- HOTRG
// environments of src/schemes/hotrg.jl:82-116 (x) and :118-152 (y); pick
// U unless eps > eps' (:112-115, :148-151)
std::vector<BT> hotrg_proj(Ctx& ctx, const std::vector<Scheme*>& G, const std::vector<int>& chis,
                           bool xdir, int tag) {
    const size_t n = G.size();
    std::vector<ContractJob> cj;
    for (size_t i = 0; i < n; ++i) {
        const BT* T = &G[i]->T;
        if (xdir) {
            // MM  = A2[-1 5; 1 2] A1[-2 3; 5 4] conj(A2[-3 6; 1 2]) conj(A1[-4 3; 6 4])
            cj.push_back(ContractJob{{T, {-1, -2, 1, 2}}, {T, {-3, -4, 1, 2}, true}, {-1, -2, -3, -4}, 2});
            cj.push_back(ContractJob{{T, {-1, 3, -2, 4}}, {T, {-3, 3, -4, 4}, true}, {-1, -2, -3, -4}, 2});
            // MM' = conj(A2[2 5; 1 -1]) conj(A1[4 3; 5 -2]) A2[2 6; 1 -3] A1[4 3; 6 -4]
            cj.push_back(ContractJob{{T, {2, -1, 1, -2}, true}, {T, {2, -3, 1, -4}}, {-1, -2, -3, -4}, 2});
            cj.push_back(ContractJob{{T, {4, 3, -1, -2}, true}, {T, {4, 3, -3, -4}}, {-1, -2, -3, -4}, 2});
        } else {
            // MM  = A1[1 -1; 2 5] A2[5 -2; 4 3] conj(A1[1 -3; 2 6]) conj(A2[6 -4; 4 3])
            cj.push_back(ContractJob{{T, {1, -1, 2, -2}}, {T, {1, -3, 2, -4}, true}, {-1, -2, -3, -4}, 2});
            cj.push_back(ContractJob{{T, {-1, -2, 4, 3}}, {T, {-3, -4, 4, 3}, true}, {-1, -2, -3, -4}, 2});
            // MM' = conj(A1[1 2; -1 5]) conj(A2[5 4; -2 3]) A1[1 2; -3 6] A2[6 4; -4 3]
            cj.push_back(ContractJob{{T, {1, 2, -1, -2}, true}, {T, {1, 2, -3, -4}}, {-1, -2, -3, -4}, 2});
            cj.push_back(ContractJob{{T, {-1, 4, -2, 3}, true}, {T, {-3, 4, -4, 3}}, {-1, -2, -3, -4}, 2});
        }
    }
    auto E = contract_many(ctx, cj);
    cj.clear();
    for (size_t i = 0; i < n; ++i) {
        if (xdir) {
            cj.push_back(ContractJob{{&E[4 * i], {-1, 5, -3, 6}}, {&E[4 * i + 1], {-2, 5, -4, 6}}, {-1, -2, -3, -4}, 2});
            cj.push_back(ContractJob{{&E[4 * i + 2], {5, -1, 6, -3}}, {&E[4 * i + 3], {5, -2, 6, -4}}, {-1, -2, -3, -4}, 2});
        } else {
            cj.push_back(ContractJob{{&E[4 * i], {-1, 5, -3, 6}}, {&E[4 * i + 1], {5, -2, 6, -4}}, {-1, -2, -3, -4}, 2});
            cj.push_back(ContractJob{{&E[4 * i + 2], {-1, 5, -3, 6}}, {&E[4 * i + 3], {5, -2, 6, -4}}, {-1, -2, -3, -4}, 2});
        }
    }
    auto MM = contract_many(ctx, cj);
    E.clear();
    std::vector<BT*> Ms;
    std::vector<int> ch;
    for (size_t i = 0; i < n; ++i) {
        Ms.push_back(&MM[2 * i]);
        Ms.push_back(&MM[2 * i + 1]);
        ch.push_back(chis[i]);
        ch.push_back(chis[i]);
    }
    auto F = eigh_trunc_many(ctx, Ms, ch, tag);
    std::vector<BT> U;
    for (size_t i = 0; i < n; ++i) {
        EFactor& a = F[2 * i];
        EFactor& b = F[2 * i + 1];
        log_factor(*G[i], a.kept_min, a.dropped_max, a.iters, a.conv,
                   a.status);
        log_factor(*G[i], b.kept_min, b.dropped_max, b.iters, b.conv,
                   b.status);
        U.push_back(a.eps > b.eps ? std::move(b.U) : std::move(a.U));
    }
    return U;
}

// _step_hotrg_y(A1, A2, U) (hotrg.jl:42-60) or _step_hotrg_x(A1, A2, U)
// (hotrg.jl:62-80) for every (A1, A2, U), as three grouped contractions
struct HotrgApply { const BT *A1, *A2, *U; };

std::vector<BT> hotrg_apply(Ctx& ctx, const std::vector<HotrgApply>& jobs, bool ydir) {
    std::vector<ContractJob> cj;
    for (const HotrgApply& j : jobs) {
        if (ydir)   // P[m1, m3, x3 | x2, x5] = conj(U[1 2; -1]) A2[1 5; -3 3]
            cj.push_back(ContractJob{{j.U, {1, -4, -1}, true}, {j.A2, {1, -5, -2, -3}},
                                     {-1, -2, -3, -4, -5}, 3});
        else        // R[m1, x3, m2 | x2, x5] = A1[-1 1; 3 5] conj(U[1 2; -2])
            cj.push_back(ContractJob{{j.A1, {-1, 1, -2, -5}}, {j.U, {1, -4, -3}, true},
                                     {-1, -2, -3, -4, -5}, 3});
    }
    auto P = contract_many(ctx, cj);
    cj.clear();
    for (size_t i = 0; i < jobs.size(); ++i) {
        if (ydir)   // Q[m1, m2, m3 | x3, x4] = P A1[2 -2; 5 4]
            cj.push_back(ContractJob{{&P[i], {-1, -3, -4, 2, 5}}, {jobs[i].A1, {2, -2, 5, -5}},
                                     {-1, -2, -3, -4, -5}, 3});
        else        // S[m1, m2, m4 | x3, x4] = R A2[5 2; 4 -4]
            cj.push_back(ContractJob{{&P[i], {-1, -4, -2, 2, 5}}, {jobs[i].A2, {5, 2, -5, -3}},
                                     {-1, -2, -3, -4, -5}, 3});
    }
    auto Q = contract_many(ctx, cj);
    P.clear();
    cj.clear();
    for (size_t i = 0; i < jobs.size(); ++i) {
        if (ydir)
            cj.push_back(ContractJob{{&Q[i], {-1, -2, -3, 3, 4}}, {jobs[i].U, {3, 4, -4}},
                                     {-1, -2, -3, -4}, 2});
        else
            cj.push_back(ContractJob{{&Q[i], {-1, -2, -4, 3, 4}}, {jobs[i].U, {3, 4, -3}},
                                     {-1, -2, -3, -4}, 2});
    }
    return contract_many(ctx, cj);
}

// ------------------------------------------------------------ ImpurityTRG
// src/schemes/impuritytrg.jl:63-91: six SVD12 factorizations per scheme
// (pure T and its transpose, T_imp1, transpose(T_imp2), T_imp3,
// transpose(T_imp4)) in one grouped solve, then five TRG networks
void step_impurity_trg(Ctx& ctx, const std::vector<Scheme*>& G, const std::vector<int>& chis) {
    const size_t n = G.size();
    std::vector<PermJob> pj;
    for (auto* s : G) {
        pj.push_back(PermJob{&s->T, {1, 3, 0, 2}, 2, false, {}});
        pj.push_back(PermJob{&s->imp[1], {1, 3, 0, 2}, 2, false, {}});
        pj.push_back(PermJob{&s->imp[3], {1, 3, 0, 2}, 2, false, {}});
    }
    auto P = permute_many(ctx, pj);
    std::vector<const BT*> Ts;
    std::vector<int> ch, tags;
    for (size_t i = 0; i < n; ++i) {
        const BT* src[6] = {&G[i]->T, &P[3 * i], &G[i]->imp[0], &P[3 * i + 1], &G[i]->imp[2],
                            &P[3 * i + 2]};
        for (int f = 0; f < 6; ++f) {
            Ts.push_back(src[f]);
            ch.push_back(chis[i]);
            tags.push_back(40 + f);
        }
    }
    auto F = svd_trunc_many(ctx, Ts, ch, tags, sqrt_weights(), /*canonical=*/true);
    std::vector<TrgNet> nets;
    for (size_t i = 0; i < n; ++i) {
        const Factor* f = &F[6 * i];
        const BT *Ap = &f[0].U, *Bp = &f[0].V, *Cp = &f[1].U, *Dp = &f[1].V;
        const BT *A1 = &f[2].U, *B1 = &f[2].V, *C2 = &f[3].U, *D2 = &f[3].V;
        const BT *A3 = &f[4].U, *B3 = &f[4].V, *C4 = &f[5].U, *D4 = &f[5].V;
        nets.push_back(TrgNet{Ap, Bp, Cp, Dp});   // pure
        nets.push_back(TrgNet{Ap, B1, C2, Dp});   // T_imp1
        nets.push_back(TrgNet{Ap, B3, Cp, D2});   // T_imp2
        nets.push_back(TrgNet{A3, Bp, Cp, D4});   // T_imp3
        nets.push_back(TrgNet{A1, Bp, C4, Dp});   // T_imp4
    }
    auto out = trg_networks(ctx, nets);
    for (size_t i = 0; i < n; ++i) {
        G[i]->tlog.clear();
        for (int f = 0; f < 6; ++f) {
            const Factor& x = F[6 * i + f];
            log_factor(*G[i], x.kept_min, x.dropped_max, x.iters, x.conv,
                   x.status);
        }
        G[i]->T = std::move(out[5 * i]);
        for (int k = 0; k < 4; ++k) G[i]->imp[k] = std::move(out[5 * i + 1 + k]);
    }
}

// ------------------------------------------------------------------ ATRG
// src/schemes/atrg.jl:48-76, one half-step for every scheme of the group:
//   A, S, B = svd_trunc(permute(T, ((1, 3), (2, 4)))); D = B; B = S B; C = A S
//   M[-1 -2; -3 -4] := B[-3; 1 -4] C[-1 1; -2]
//   X, S, Y = svd_trunc(permute(M, ((1, 3), (2, 4)))); X = X sqrt(S); Y = sqrt(S) Y
//   Q[-1 -2; -3 -4] := A[3 -3; 2] D[1; -2 4] X[4 2; -4] Y[-1 1; 3]
//                    = AX[3, -3, 4, -4] DY[-2, 4, -1, 3]
//   H, S, G = svd_trunc(Q); H = H sqrt(S); G = sqrt(S) G
//   T[-1 -2; -3 -4] := G[-1; -3 1] H[1 -2; -4]
void atrg_half(Ctx& ctx, const std::vector<Scheme*>& G, const std::vector<int>& chis, int tag) {
    const size_t n = G.size();
    std::vector<PermJob> pj;
    for (auto* s : G) pj.push_back(PermJob{&s->T, {0, 2, 1, 3}, 2, false, {}});
    auto P = permute_many(ctx, pj);
    std::vector<const BT*> Ts;
    std::vector<int> tags;
    for (size_t i = 0; i < n; ++i) { Ts.push_back(&P[i]); tags.push_back(tag); }
    auto unit = [](size_t, const std::vector<double>& s, std::vector<double>& wu,
                   std::vector<double>& wv) {
        wu.assign(s.size(), 1.0);
        wv = wu;
    };
    auto F1 = svd_trunc_many(ctx, Ts, chis, tags, unit);
    // S on the new legs: V's is a codomain leg (q = c), U's a domain leg (q = -c)
    std::vector<int> Ns;
    std::vector<const int*> dims;
    std::vector<const std::vector<double>*> vals;
    for (size_t i = 0; i < n; ++i) {
        Ns.push_back(G[i]->N);
        dims.push_back(F1[i].dim);
        vals.push_back(&F1[i].s);
    }
    auto S = make_bonds(ctx, Ns, dims, vals);
    std::vector<ContractJob> cj;
    for (size_t i = 0; i < n; ++i) {
        Operand b{&F1[i].V, {-3, 1, -4}, false, {WeightSpec{0, S[i].same_side()}}};
        Operand c{&F1[i].U, {-1, 1, -2}, false, {WeightSpec{2, S[i].dual_side()}}};
        cj.push_back(ContractJob{b, c, {-1, -2, -3, -4}, 2});
    }
    auto M = contract_many(ctx, cj);
    pj.clear();
    for (size_t i = 0; i < n; ++i) pj.push_back(PermJob{&M[i], {0, 2, 1, 3}, 2, false, {}});
    auto PM = permute_many(ctx, pj);
    Ts.clear();
    tags.clear();
    for (size_t i = 0; i < n; ++i) { Ts.push_back(&PM[i]); tags.push_back(tag + 1); }
    auto F2 = svd_trunc_many(ctx, Ts, chis, tags, sqrt_weights());
    cj.clear();
    for (size_t i = 0; i < n; ++i) {
        cj.push_back(ContractJob{{&F1[i].U, {-1, -2, 1}}, {&F2[i].U, {-3, 1, -4}},
                                 {-1, -2, -3, -4}, 2});
        cj.push_back(ContractJob{{&F1[i].V, {1, -1, -2}}, {&F2[i].V, {-3, 1, -4}},
                                 {-1, -2, -3, -4}, 2});
    }
    auto AD = contract_many(ctx, cj);
    cj.clear();
    for (size_t i = 0; i < n; ++i)
        cj.push_back(ContractJob{{&AD[2 * i], {3, -3, 4, -4}}, {&AD[2 * i + 1], {-2, 4, -1, 3}},
                                 {-1, -2, -3, -4}, 2});
    auto Q = contract_many(ctx, cj);
    AD.clear();
    Ts.clear();
    tags.clear();
    for (size_t i = 0; i < n; ++i) { Ts.push_back(&Q[i]); tags.push_back(tag + 2); }
    auto F3 = svd_trunc_many(ctx, Ts, chis, tags, sqrt_weights());
    cj.clear();
    for (size_t i = 0; i < n; ++i)
        cj.push_back(ContractJob{{&F3[i].V, {-1, -3, 1}}, {&F3[i].U, {1, -2, -4}},
                                 {-1, -2, -3, -4}, 2});
    auto Tn = contract_many(ctx, cj);
    for (size_t i = 0; i < n; ++i) {
        for (const Factor* f : {&F1[i], &F2[i], &F3[i]})
            log_factor(*G[i], f->kept_min, f->dropped_max, f->iters, f->conv,
                   f->status);
        G[i]->T = std::move(Tn[i]);
    }
}

// src/schemes/atrg.jl:35-43: half-step, T = permute(T, ((2, 4), (1, 3))),
// half-step, T = permute(T, ((3, 1), (4, 2)))
void step_atrg(Ctx& ctx, const std::vector<Scheme*>& G, const std::vector<int>& chis) {
    for (auto* s : G) s->tlog.clear();
    auto rotate = [&](std::vector<int> perm) {
        std::vector<PermJob> pj;
        for (auto* s : G) pj.push_back(PermJob{&s->T, perm, 2, false, {}});
        auto R = permute_many(ctx, pj);
        for (size_t i = 0; i < G.size(); ++i) G[i]->T = std::move(R[i]);
    };
    atrg_half(ctx, G, chis, 71);
    rotate({1, 3, 0, 2});
    atrg_half(ctx, G, chis, 74);
    rotate({2, 0, 3, 1});
}

// ------------------------------------------------------------- HOTRG_3D
// Tensors T[z z'; Y X y x] (2|4); bosonic sectors, so the fermionic twists of
// src/schemes/hotrg3d.jl:36-46 are identities.
// Projector of the x pair (hotrg3d.jl:86-99) of each tensor A (= A1 = A2;
// for the y pair A is permuted by ((1, 2), (4, 3, 6, 5)) first, :101-107):
//   E2[x2 z z' x2'] = A2[z z2; Y2 X2 y2 x2] conj(A2[z' z2; Y2 X2 y2 x2'])
//   E1[z x1 z' x1'] = A1[z1 z; Y1 X1 y1 x1] conj(A1[z1 z'; Y1 X1 y1 x1'])
//   MM[x1 x2; x1' x2'] = E2[x2 z z' x2'] E1[z x1 z' x1']        (:48-65)
// and MM' from the conjugated pair with x and X exchanged (:67-84); U of MM
// unless eps > eps'.
std::vector<BT> hotrg3d_proj(Ctx& ctx, const std::vector<const BT*>& As, const std::vector<Scheme*>& G,
                             const std::vector<int>& chis, int tag) {
    const size_t n = As.size();
    std::vector<ContractJob> cj;
    for (size_t i = 0; i < n; ++i) {
        const BT* A = As[i];
        cj.push_back(ContractJob{{A, {-2, 1, 2, 3, 4, -1}}, {A, {-3, 1, 2, 3, 4, -4}, true},
                                 {-1, -2, -3, -4}, 2});
        cj.push_back(ContractJob{{A, {1, -1, 2, 3, 4, -2}}, {A, {1, -3, 2, 3, 4, -4}, true},
                                 {-1, -2, -3, -4}, 2});
        cj.push_back(ContractJob{{A, {-2, 1, 2, -1, 3, 4}, true}, {A, {-3, 1, 2, -4, 3, 4}},
                                 {-1, -2, -3, -4}, 2});
        cj.push_back(ContractJob{{A, {1, -1, 2, -2, 3, 4}, true}, {A, {1, -3, 2, -4, 3, 4}},
                                 {-1, -2, -3, -4}, 2});
    }
    auto E = contract_many(ctx, cj);
    cj.clear();
    for (size_t i = 0; i < n; ++i)
        for (int k = 0; k < 2; ++k)
            cj.push_back(ContractJob{{&E[4 * i + 2 * k], {-2, 5, 6, -4}},
                                     {&E[4 * i + 2 * k + 1], {5, -1, 6, -3}}, {-1, -2, -3, -4}, 2});
    auto MM = contract_many(ctx, cj);
    E.clear();
    std::vector<BT*> Ms;
    std::vector<int> ch;
    for (size_t i = 0; i < n; ++i) {
        Ms.push_back(&MM[2 * i]);
        Ms.push_back(&MM[2 * i + 1]);
        ch.push_back(chis[i]);
        ch.push_back(chis[i]);
    }
    auto F = eigh_trunc_many(ctx, Ms, ch, tag);
    std::vector<BT> U;
    for (size_t i = 0; i < n; ++i) {
        EFactor& a = F[2 * i];
        EFactor& b = F[2 * i + 1];
        log_factor(*G[i], a.kept_min, a.dropped_max, a.iters, a.conv,
                   a.status);
        log_factor(*G[i], b.kept_min, b.dropped_max, b.iters, b.conv,
                   b.status);
        U.push_back(a.eps > b.eps ? std::move(b.U) : std::move(a.U));
    }
    return U;
}

// one z compression (hotrg3d.jl:121-126) for every scheme:
//   T[-1 -2; -3 -4 -5 -6] := conj(Ux[x1 x2; -6]) Ux[x1' x2'; -4] conj(Uy[y1 y2; -5])
//        Uy[y1' y2'; -3] A1[-1 z; y1' x1' y1 x1] A2[z -2; y2' x2' y2 x2]
// (hotrg3d.jl:109-119) pairwise, at most eight legs at a time
void hotrg3d_substep(Ctx& ctx, const std::vector<Scheme*>& G, const std::vector<int>& chis) {
    const size_t n = G.size();
    std::vector<const BT*> As;
    std::vector<PermJob> pj;
    for (auto* s : G) {
        As.push_back(&s->T);
        pj.push_back(PermJob{&s->T, {0, 1, 3, 2, 5, 4}, 2, false, {}});
    }
    auto Ux = hotrg3d_proj(ctx, As, G, chis, 81);
    auto Ay = permute_many(ctx, pj);
    std::vector<const BT*> Ays;
    for (auto& a : Ay) Ays.push_back(&a);
    auto Uy = hotrg3d_proj(ctx, Ays, G, chis, 82);
    Ay.clear();
    std::vector<ContractJob> cj;
    for (size_t i = 0; i < n; ++i)   // B1[a z y1' y1 x1 x2' n4]
        cj.push_back(ContractJob{{&G[i]->T, {-1, -2, -3, 1, -4, -5}}, {&Ux[i], {1, -6, -7}},
                                 {-1, -2, -3, -4, -5, -6, -7}, 2});
    auto B1 = contract_many(ctx, cj);
    cj.clear();
    for (size_t i = 0; i < n; ++i)   // B2[a z y1' y1 x2' n4 x2 n6]
        cj.push_back(ContractJob{{&B1[i], {-1, -2, -3, -4, 1, -5, -6}}, {&Ux[i], {1, -7, -8}, true},
                                 {-1, -2, -3, -4, -5, -6, -7, -8}, 2});
    auto B2 = contract_many(ctx, cj);
    B1.clear();
    cj.clear();
    for (size_t i = 0; i < n; ++i)   // B3[a y1' y1 n4 n6 b y2' y2]
        cj.push_back(ContractJob{{&B2[i], {-1, 1, -2, -3, 2, -4, 3, -5}},
                                 {&G[i]->T, {1, -6, -7, 2, -8, 3}},
                                 {-1, -2, -3, -4, -5, -6, -7, -8}, 2});
    auto B3 = contract_many(ctx, cj);
    B2.clear();
    cj.clear();
    for (size_t i = 0; i < n; ++i)   // B4[a y1 n4 n6 b y2 n3]
        cj.push_back(ContractJob{{&B3[i], {-1, 1, -2, -3, -4, -5, 2, -6}}, {&Uy[i], {1, 2, -7}},
                                 {-1, -2, -3, -4, -5, -6, -7}, 2});
    auto B4 = contract_many(ctx, cj);
    B3.clear();
    cj.clear();
    for (size_t i = 0; i < n; ++i)   // T[a b; n3 n4 n5 n6]
        cj.push_back(ContractJob{{&B4[i], {-1, 1, -4, -6, -2, 2, -3}}, {&Uy[i], {1, 2, -5}, true},
                                 {-1, -2, -3, -4, -5, -6}, 2});
    auto Tn = contract_many(ctx, cj);
    for (size_t i = 0; i < n; ++i) G[i]->T = std::move(Tn[i]);
}

// hotrg3d.jl:128-136: three z compressions, T permuted by ((6, 4), (2, 3, 1, 5))
// after each
void step_hotrg3d(Ctx& ctx, const std::vector<Scheme*>& G, const std::vector<int>& chis) {
    for (auto* s : G) s->tlog.clear();
    for (int k = 0; k < 3; ++k) {
        hotrg3d_substep(ctx, G, chis);
        std::vector<PermJob> pj;
        for (auto* s : G) pj.push_back(PermJob{&s->T, {5, 3, 1, 2, 0, 4}, 2, false, {}});
        auto R = permute_many(ctx, pj);
        for (size_t i = 0; i < G.size(); ++i) G[i]->T = std::move(R[i]);
    }
}

// ---------------------------------------------------------- ImpurityHOTRG
// src/schemes/impurityhotrg.jl:48-84: projector from (T, T), then nine
// placements of the tensors through it, averaged:
//   T' = X(T, T), T11' = (X(T11, T) + X(T, T11)) / 2, likewise T12',
//   T2' = (X(T2, T) + X(T, T2) + X(T11, T12) + X(T12, T11)) / 4
// (scaling by 1/2, 1/4 is exact, so summing the scaled terms left to right
// gives the reference's bits); first along x (y projector), then along y.
void impurity_hotrg_half(Ctx& ctx, const std::vector<Scheme*>& G, const std::vector<int>& chis,
                         bool ydir) {
    const size_t n = G.size();
    auto U = hotrg_proj(ctx, G, chis, /*xdir=*/ydir, ydir ? 51 : 52);
    std::vector<HotrgApply> jobs;
    for (size_t i = 0; i < n; ++i) {
        const BT *T = &G[i]->T, *T11 = &G[i]->imp[0], *T12 = &G[i]->imp[1], *T2 = &G[i]->imp[2];
        const BT* u = &U[i];
        const BT* pairs[9][2] = {{T, T},    {T11, T}, {T, T11}, {T12, T},  {T, T12},
                                 {T2, T},   {T, T2},  {T11, T12}, {T12, T11}};
        for (auto& p : pairs) jobs.push_back(HotrgApply{p[0], p[1], u});
    }
    auto R = hotrg_apply(ctx, jobs, ydir);
    std::vector<LinJob> lj;
    for (size_t i = 0; i < n; ++i) {
        BT* r = &R[9 * i];
        lj.push_back(LinJob{&r[1], {{0.5, &r[1]}, {0.5, &r[2]}}});
        lj.push_back(LinJob{&r[3], {{0.5, &r[3]}, {0.5, &r[4]}}});
        lj.push_back(LinJob{&r[5], {{0.25, &r[5]}, {0.25, &r[6]}, {0.25, &r[7]}, {0.25, &r[8]}}});
    }
    lincomb_many(ctx, lj);
    for (size_t i = 0; i < n; ++i) {
        G[i]->T = std::move(R[9 * i]);
        G[i]->imp[0] = std::move(R[9 * i + 1]);
        G[i]->imp[1] = std::move(R[9 * i + 3]);
        G[i]->imp[2] = std::move(R[9 * i + 5]);
    }
}

// src/schemes/hotrg.jl:154-160: x projector, y step, y projector, x step
void step_hotrg(Ctx& ctx, const std::vector<Scheme*>& G, const std::vector<int>& chis) {
    const size_t n = G.size();
    for (auto* s : G) s->tlog.clear();
    auto Ux = hotrg_proj(ctx, G, chis, true, 31);
    std::vector<HotrgApply> jobs;
    for (size_t i = 0; i < n; ++i) jobs.push_back(HotrgApply{&G[i]->T, &G[i]->T, &Ux[i]});
    auto Ty = hotrg_apply(ctx, jobs, /*ydir=*/true);
    for (size_t i = 0; i < n; ++i) G[i]->T = std::move(Ty[i]);
    auto Uy = hotrg_proj(ctx, G, chis, false, 32);
    jobs.clear();
    for (size_t i = 0; i < n; ++i) jobs.push_back(HotrgApply{&G[i]->T, &G[i]->T, &Uy[i]});
    auto Tx = hotrg_apply(ctx, jobs, /*ydir=*/false);
    for (size_t i = 0; i < n; ++i) G[i]->T = std::move(Tx[i]);
}

void step_impurity_hotrg(Ctx& ctx, const std::vector<Scheme*>& G, const std::vector<int>& chis) {
    for (auto* s : G) s->tlog.clear();
    impurity_hotrg_half(ctx, G, chis, /*ydir=*/false);   // _get_hotrg_yproj + _step_hotrg_x
    impurity_hotrg_half(ctx, G, chis, /*ydir=*/true);    // _get_hotrg_xproj + _step_hotrg_y
}

}  // namespace

int gauge_clusters(const double* sv, int k, std::vector<int>& cl, int& knn) {
    knn = 0;
    while (knn < k && sv[knn] > GAUGE_NULL * sv[0]) ++knn;
    if (knn == 0) return 0;
    const size_t b = cl.size();
    cl.push_back(0);
    for (int i = 1; i < knn; ++i)
        if (sv[i - 1] - sv[i] > GAUGE_CLUSTER * sv[0]) cl.push_back(i);
    const int nc = (int)(cl.size() - b);
    cl.push_back(knn);
    return nc;
}

// ------------------------------------------------------------------ public
void scheme_set_tensor(Ctx& ctx, Scheme& s, BT&& T) {
    s.status = TNR_OK;   // a new tensor clears a failed step
    if (s.kind == TNR_HOTRG_3D) {
        const int N = T.L.N;
        TNR_REQUIRE(T.L.nleg == 6 && T.L.ncod == 2, TNR_ERR_ARG, "3D tensors have 2+4 legs");
        TNR_REQUIRE(T.L.legs[1] == T.L.legs[0].dual(N) && T.L.legs[4] == T.L.legs[2].dual(N) &&
                        T.L.legs[5] == T.L.legs[3].dual(N),
                    TNR_ERR_ARG, "legs 1/2, 3/5 and 4/6 must carry the same spaces");
        s.N = N;
        s.T = std::move(T);
        return;
    }
    TNR_REQUIRE(T.L.nleg == 4 && T.L.ncod == 2, TNR_ERR_ARG, "scheme tensors have 2+2 legs");
    TNR_REQUIRE(T.L.legs[0] == T.L.legs[3].dual(T.L.N) && T.L.legs[1] == T.L.legs[2].dual(T.L.N),
                TNR_ERR_ARG, "legs 1/4 and 2/3 must carry the same spaces");
    s.N = T.L.N;
    s.T = std::move(T);
    if (s.kind == TNR_BTRG) {
        // S1 = id(space(T, 2)), S2 = id(space(T, 1)) (src/schemes/btrg.jl:41-43)
        std::vector<double> o1(s.T.L.legs[1].total(s.N), 1.0), o2(s.T.L.legs[0].total(s.N), 1.0);
        s.S1 = make_bond(ctx, s.N, s.T.L.legs[1].d, o1);
        s.S2 = make_bond(ctx, s.N, s.T.L.legs[0].d, o2);
    }
}

bool single_tensor_kind(int kind) {
    return kind == TNR_TRG || kind == TNR_BTRG || kind == TNR_HOTRG || kind == TNR_ATRG ||
           kind == TNR_HOTRG_3D;
}

int scheme_tensors(int kind) {
    switch (kind) {
        case TNR_TRG: case TNR_BTRG: case TNR_HOTRG: case TNR_ATRG: case TNR_HOTRG_3D: return 1;
        case TNR_IMPURITY_TRG: return 5;
        case TNR_IMPURITY_HOTRG: return 4;
        default: return 0;
    }
}

int scheme_components(int kind) {
    switch (kind) {
        case TNR_TRG: case TNR_BTRG: case TNR_HOTRG: case TNR_ATRG: case TNR_HOTRG_3D: return 1;
        case TNR_IMPURITY_TRG: return 2;
        case TNR_IMPURITY_HOTRG: return 4;
        default: return 0;
    }
}

void scheme_init_multi(Ctx& ctx, Scheme& s, int kind, std::vector<BT>&& Ts) {
    TNR_REQUIRE(kind == TNR_IMPURITY_TRG || kind == TNR_IMPURITY_HOTRG, TNR_ERR_ARG,
                "unknown impurity scheme kind");
    TNR_REQUIRE((int)Ts.size() == scheme_tensors(kind), TNR_ERR_ARG,
                "wrong number of tensors for this scheme");
    // impuritytrg.jl:52-56, impurityhotrg.jl:40-44: every tensor on T's spaces
    for (size_t i = 1; i < Ts.size(); ++i)
        TNR_REQUIRE(Ts[i].L.same(Ts[0].L), TNR_ERR_ARG,
                    "the impurity tensors must carry the spaces of T");
    s.kind = kind;
    s.k = 0.0;
    s.steps = 0;
    s.imp.clear();
    for (size_t i = 1; i < Ts.size(); ++i) s.imp.push_back(std::move(Ts[i]));
    scheme_set_tensor(ctx, s, std::move(Ts[0]));
}

void scheme_init(Ctx& ctx, Scheme& s, int kind, BT&& T, double k) {
    TNR_REQUIRE(single_tensor_kind(kind), TNR_ERR_ARG, "unknown scheme kind");
    s.kind = kind;
    s.k = k;
    s.steps = 0;
    scheme_set_tensor(ctx, s, std::move(T));
}

namespace {
// Contraction precision of a context for the duration of a scope.
struct PrecisionScope {
    Ctx& ctx;
    int saved;
    PrecisionScope(Ctx& c, int p) : ctx(c), saved(c.precision) { c.precision = p; }
    ~PrecisionScope() { ctx.precision = saved; }
};

void step_kind(Ctx& ctx, int kind, const std::vector<Scheme*>& g, const std::vector<int>& c) {
    if (g.empty()) return;
    if (kind == TNR_TRG) step_trg(ctx, g, c);
    else if (kind == TNR_BTRG) step_btrg(ctx, g, c);
    else if (kind == TNR_HOTRG) step_hotrg(ctx, g, c);
    else if (kind == TNR_ATRG) step_atrg(ctx, g, c);
    else if (kind == TNR_HOTRG_3D) step_hotrg3d(ctx, g, c);
    else if (kind == TNR_IMPURITY_TRG) step_impurity_trg(ctx, g, c);
    else step_impurity_hotrg(ctx, g, c);
}
}  // namespace

void step_many(Ctx& ctx, const std::vector<Scheme*>& group, const std::vector<int>& chis) {
    for (int kind : {TNR_TRG, TNR_BTRG, TNR_HOTRG, TNR_ATRG, TNR_HOTRG_3D, TNR_IMPURITY_TRG,
                     TNR_IMPURITY_HOTRG}) {
        // split precision: a scheme's first kSplitFromStep steps stay FP64
        std::vector<Scheme*> g[2];
        std::vector<int> c[2];
        for (size_t i = 0; i < group.size(); ++i)
            if (group[i]->kind == kind) {
                const int late =
                    ctx.precision == TNR_PREC_SPLIT32 && group[i]->steps >= kSplitFromStep;
                g[late].push_back(group[i]);
                c[late].push_back(chis[i]);
            }
        if (!g[0].empty()) {
            PrecisionScope fp64(ctx, TNR_PREC_FP64);
            step_kind(ctx, kind, g[0], c[0]);
        }
        step_kind(ctx, kind, g[1], c[1]);
        for (auto& gg : g)
            for (auto* s : gg) s->steps++;
    }
}

void finalize_many(Ctx& ctx, const std::vector<Scheme*>& group, const std::vector<double*>& slots) {
    std::vector<TraceJob> tj(group.size());
    std::vector<LegWeights> w1(group.size()), w2(group.size());
    for (size_t i = 0; i < group.size(); ++i) {
        Scheme& s = *group[i];
        tj[i].t = &s.T;
        tj[i].norm_slot = slots[i];
        if (s.kind == TNR_BTRG) {
            // n = |T[1 2; 4 3] S1[4; 2] S2[3; 1]| (src/utility/finalize.jl:10-14)
            w1[i] = s.S2.same_side();
            w2[i] = s.S1.same_side();
            tj[i].w_x1 = &w1[i];
            tj[i].w_x2 = &w2[i];
        }
    }
    trace_normalize_many(ctx, tj);
}

void finalize_host(Ctx& ctx, Scheme& s, double* out) {
    Buf slots(&ctx.pool, 8 * sizeof(double));
    double* sl = slots.d();
    if (single_tensor_kind(s.kind)) {
        finalize_many(ctx, {&s}, {sl});
        ctx.fetch(out, sl, sizeof(double));
        return;
    }
    // npure = |tr T[1 2; 2 1]|, T /= npure; then every impurity tensor's own
    // trace (ImpurityHOTRG reports them) and division by npure
    trace_normalize_many(ctx, {TraceJob{&s.T, nullptr, nullptr, sl, nullptr}});
    std::vector<TraceJob> tj;
    for (size_t i = 0; i < s.imp.size(); ++i)
        tj.push_back(TraceJob{&s.imp[i], nullptr, nullptr, sl + 1 + i, sl});
    trace_normalize_many(ctx, tj);
    if (s.kind == TNR_IMPURITY_HOTRG) {   // finalize.jl:90-101
        ctx.fetch(out, sl, 4 * sizeof(double));
        return;
    }
    // ImpurityTRG (finalize.jl:74-88): nimp = |T1[5 4; 6 1] T2[1 2; 7 5]
    // T3[3 7; 2 8] T4[8 6; 4 3]|, as P12[4,6,2,7] and P34[7,2,6,4], then
    // M[4, 4'] = P12 . P34 over (6, 2, 7) and the trace of M
    std::vector<ContractJob> cj{
        ContractJob{{&s.imp[0], {5, -1, -2, 1}}, {&s.imp[1], {1, -3, -4, 5}}, {-1, -2, -3, -4}, 2},
        ContractJob{{&s.imp[2], {3, -1, -2, 8}}, {&s.imp[3], {8, -3, -4, 3}}, {-1, -2, -3, -4}, 2}};
    auto PP = contract_many(ctx, cj);
    cj = {ContractJob{{&PP[0], {-1, 1, 2, 3}}, {&PP[1], {3, 2, 1, -2}}, {-1, -2}, 1}};
    auto M = contract_many(ctx, cj);
    const Layout& L = M[0].L;
    std::vector<double> h((size_t)std::max<int64_t>(L.size, 1));
    ctx.fetch(h.data(), M[0].d(), (size_t)L.size * sizeof(double));
    double tr = 0.0;
    for (int c = 0; c < L.N; ++c) {
        TNR_REQUIRE(L.R[c] == L.C[c], TNR_ERR_INTERNAL, "plaquette trace of a non-square block");
        for (int64_t i = 0; i < L.R[c]; ++i) tr += h[(size_t)(L.off[c] + i + i * L.R[c])];
    }
    ctx.fetch(out, sl, sizeof(double));
    out[1] = std::fabs(tr);
}

double free_energy(const double* norms, int n, double beta, double scalefactor,
                   double initial_size) {
    // src/utility/free_energy.jl:13-20
    double lnz = 0.0;
    const double x = 1.0 - std::log(initial_size) / std::log(scalefactor);
    for (int i = 1; i <= n; ++i) lnz += std::log(norms[i - 1]) * std::pow(scalefactor, x - i);
    return -lnz / beta;
}

size_t scheme_hbm_estimate(int kind, int N, int chi) {
    // steady-state live set: a handful of chi^4 tensors (T, permuted T,
    // contraction intermediates) spread over N sectors, the solver blocks,
    // and for HOTRG the chi^5 intermediates
    const double c4 = std::pow((double)chi, 4) / std::max(1, N) * 8.0;
    const double c5 = std::pow((double)chi, 5) / std::max(1, N) * 8.0;
    double b = 8.0 * c4 + 4.0 * (double)chi * chi * LMAX * 8.0 * 6.0;
    // impurity methods carry their impurity tensors through every step
    if (kind == TNR_IMPURITY_TRG || kind == TNR_IMPURITY_HOTRG)
        b += 2.0 * c4 * (scheme_tensors(kind) - 1);
    if (kind == TNR_HOTRG || kind == TNR_IMPURITY_HOTRG) b += 3.0 * c5;
    if (kind == TNR_HOTRG_3D) b += 4.0 * std::pow((double)chi, 8) / std::max(1, N) * 8.0;
    return (size_t)b + (size_t(8) << 20);
}

}  // namespace tnr
