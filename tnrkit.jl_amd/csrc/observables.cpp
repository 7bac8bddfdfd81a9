// Observables on a scheme's tensor; see observables.h.
#include "observables.h"

#include <algorithm>
#include <cmath>

namespace tnr {

namespace {

using cplx = std::complex<double>;

// Identity on a matrix layout whose domain legs are the duals of its
// codomain legs, in order (then every coupled-charge block is square and
// rows and columns enumerate the same sector tuples): uploaded from the host.
BT identity_of(Ctx& ctx, const Layout& L) {
    TNR_REQUIRE(L.nleg == 2 * L.ncod, TNR_ERR_INTERNAL, "identity needs (k|k) legs");
    for (int i = 0; i < L.ncod; ++i)
        TNR_REQUIRE(L.legs[L.ncod + i] == L.legs[i].dual(L.N), TNR_ERR_ARG,
                    "the transfer matrix maps a space to a different one");
    BT I(ctx, L);
    std::vector<double> h((size_t)std::max<int64_t>(L.size, 1), 0.0);
    for (int c = 0; c < L.N; ++c) {
        TNR_REQUIRE(L.R[c] == L.C[c], TNR_ERR_INTERNAL, "identity block not square");
        for (int64_t i = 0; i < L.R[c]; ++i) h[(size_t)(L.off[c] + i + i * L.R[c])] = 1.0;
    }
    if (L.size > 0)
        TNR_HIP(hipMemcpyAsync(I.d(), h.data(), (size_t)L.size * sizeof(double),
                               hipMemcpyHostToDevice, ctx.stream));
    ctx.sync();
    return I;
}

// Identity on a leg pair (a, dual a): contracting T{.., 1, .., 2, ..} with
// it over labels {2, 1} traces the two legs of T it meets.
BT identity_leg(Ctx& ctx, int N, const Leg& a) {
    return identity_of(ctx, Layout(N, 1, {a, a.dual(N)}));
}

// sum_i x_i y_i for each pair (tensors of one layout), one device launch and
// one readback
std::vector<double> dots(Ctx& ctx, const std::vector<std::pair<const BT*, const BT*>>& pairs) {
    Buf out(&ctx.pool, std::max<size_t>(1, pairs.size()) * sizeof(double));
    Buf part(&ctx.pool, std::max<size_t>(1, pairs.size()) * DOT_PARTS * sizeof(double));
    std::vector<DotDesc> d;
    for (size_t i = 0; i < pairs.size(); ++i) {
        TNR_REQUIRE(pairs[i].first->L.same(pairs[i].second->L), TNR_ERR_INTERNAL,
                    "dot of tensors with different layouts");
        d.push_back(DotDesc{pairs[i].first->d(), pairs[i].second->d(), out.d() + i,
                            part.d() + i * DOT_PARTS, pairs[i].first->L.size});
    }
    DotDesc* dd = ctx.stage->put(ctx.stream, d);
    launch_dot(ctx.stream, dd, (int)d.size());
    std::vector<double> h(pairs.size());
    ctx.fetch(h.data(), out.d(), pairs.size() * sizeof(double));
    return h;
}

// tr X of a (k|k) tensor
double trace_of(Ctx& ctx, const BT& X) {
    BT I = identity_of(ctx, X.L);
    return dots(ctx, {{&X, &I}})[0];
}

// BTRG's T_unit = T[1 2; -3 -4] S1[-2; 2] S2[-1; 1] (cft.jl:44-45): bond
// weights on legs 0 (S2) and 1 (S1); no weights for the other schemes
std::vector<WeightSpec> unit_weights(const Scheme& s) {
    if (s.kind != TNR_BTRG) return {};
    return {WeightSpec{0, s.S2.same_side()}, WeightSpec{1, s.S1.same_side()}};
}

BT unit_tensor(Ctx& ctx, const Scheme& s) {
    return permute(ctx, s.T, {0, 1, 2, 3}, 2, false, unit_weights(s));
}

// Transfer matrix of `unitcell` tensors in a row, periodic (cft.jl:7-16,
// 280-295): M[(b_1..b_u), (t_1..t_u)] = sum T[a_1 b_1 t_1 a_2] ... T[a_u b_u t_u a_1]
BT transfer_matrix(Ctx& ctx, const Scheme& s, int unitcell) {
    TNR_REQUIRE(s.T.L.nleg == 4, TNR_ERR_ARG, "observables are defined for 2D (4-leg) tensors");
    TNR_REQUIRE(unitcell == 1 || unitcell == 2, TNR_ERR_ARG,
                "unitcell must be 1 or 2 (tensors of at most six legs)");
    const BT& T = s.T;
    const int N = T.L.N;
    TNR_REQUIRE(T.L.legs[3] == T.L.legs[0].dual(N) && T.L.legs[2] == T.L.legs[1].dual(N),
                TNR_ERR_ARG, "the transfer matrix needs equal left/right and top/bottom spaces");
    BT Id = identity_leg(ctx, N, T.L.legs[3].dual(N));
    const std::vector<WeightSpec> ws = unit_weights(s);
    if (unitcell == 1) {
        return std::move(contract_many(ctx, {ContractJob{{&T, {1, -1, -2, 2}, false, ws},
                                                        {&Id, {2, 1}}, {-1, -2}, 1}})[0]);
    }
    BT X = std::move(contract_many(ctx, {ContractJob{{&T, {-1, -2, -3, 1}, false, ws},
                                                    {&T, {1, -4, -5, -6}, false, ws},
                                                    {-1, -2, -3, -4, -5, -6}, 3}})[0]);
    return std::move(contract_many(ctx, {ContractJob{{&X, {1, -1, -2, -3, -4, 2}}, {&Id, {2, 1}},
                                                    {-1, -3, -2, -4}, 2}})[0]);
}

// sum A[a,b,c,d] B[d,c,b,a] in O(size): B' = B with its legs reversed and
// its charges negated (conj) has A's layout when A's legs 3, 2 carry the
// duals of legs 0, 1, and B'[a,b,c,d] = B[d,c,b,a] sits where A[a,b,c,d]
// does; then one dot product
double trace_pair(Ctx& ctx, const BT& A, const BT& B) {
    BT Br = permute(ctx, B, {3, 2, 1, 0}, 2, /*conj=*/true);
    TNR_REQUIRE(Br.L.same(A.L), TNR_ERR_ARG,
                "the torus needs equal left/right and top/bottom spaces");
    return dots(ctx, {{&A, &Br}})[0];
}

bool by_modulus(const cplx& a, const cplx& b) {
    const double ma = std::abs(a), mb = std::abs(b);
    if (ma != mb) return ma > mb;
    if (a.real() != b.real()) return a.real() > b.real();
    return a.imag() > b.imag();
}

}  // namespace

std::vector<cplx> transfer_spectrum(Ctx& ctx, const Scheme& s, int unitcell) {
    BT M = transfer_matrix(ctx, s, unitcell);
    const Layout& L = M.L;
    std::vector<EigDesc> d;
    std::vector<int> at;
    int64_t total = 0, wtotal = 0;
    int nmin = EIG_NMAX_G, nmax = 0;
    std::vector<int64_t> wat;
    for (int c = 0; c < L.N; ++c) {
        if (L.R[c] == 0) continue;
        TNR_REQUIRE(L.R[c] == L.C[c], TNR_ERR_INTERNAL, "transfer block not square");
        TNR_REQUIRE(L.R[c] <= EIG_NMAX_G, TNR_ERR_ARG,
                    "transfer-matrix block of dimension " + std::to_string(L.R[c]) +
                        " is above the eigensolver's limit of " + std::to_string(EIG_NMAX_G));
        at.push_back((int)total);
        wat.push_back(L.R[c] > EIG_NMAX ? wtotal : -1);
        if (L.R[c] > EIG_NMAX) wtotal += eig_work_doubles((int)L.R[c]);
        d.push_back(EigDesc{M.d() + L.off[c], nullptr, nullptr, nullptr, (int32_t)L.R[c],
                            (int32_t)L.R[c]});
        total += L.R[c];
        nmin = std::min(nmin, (int)L.R[c]);
        nmax = std::max(nmax, (int)L.R[c]);
    }
    std::vector<cplx> out;
    if (d.empty()) return out;
    Buf wr(&ctx.pool, total * sizeof(double)), wi(&ctx.pool, total * sizeof(double));
    Buf info(&ctx.pool, d.size() * sizeof(int));
    Buf work;
    if (wtotal) work = Buf(&ctx.pool, (size_t)wtotal * sizeof(double));
    for (size_t i = 0; i < d.size(); ++i) {
        d[i].wr = wr.d() + at[i];
        d[i].wi = wi.d() + at[i];
        d[i].info = info.as<int>() + i;
        d[i].work = wat[i] >= 0 ? work.d() + wat[i] : nullptr;
    }
    EigDesc* dd = ctx.stage->put(ctx.stream, d);
    launch_eig(ctx.stream, dd, (int)d.size(), nmin, nmax);
    std::vector<double> hr(total), hi(total);
    std::vector<int> hinfo(d.size());
    ctx.fetch(hr.data(), wr.d(), total * sizeof(double));
    ctx.fetch(hi.data(), wi.d(), total * sizeof(double));
    ctx.fetch(hinfo.data(), info.d(), d.size() * sizeof(int));
    for (int x : hinfo)
        TNR_REQUIRE(x == 0, TNR_ERR_NOCONV, "the transfer-matrix QR iteration did not converge");
    out.resize(total);
    for (int64_t i = 0; i < total; ++i) out[i] = cplx(hr[i], hi[i]);
    std::stable_sort(out.begin(), out.end(), by_modulus);
    return out;
}

std::vector<cplx> cft_data(Ctx& ctx, const Scheme& s, int unitcell, double v, bool is_real) {
    std::vector<cplx> ev = transfer_spectrum(ctx, s, unitcell);
    // cft.jl:28-36: keep real part > 0 and modulus > 1e-12, then
    // unitcell / (2 pi v) log(d_1 / d_i)
    std::vector<cplx> data;
    for (const cplx& z : ev)
        if (z.real() > 0.0 && std::abs(z) > 1e-12) data.push_back(is_real ? cplx(z.real(), 0.0) : z);
    std::vector<cplx> out;
    if (data.empty()) return out;
    const double pref = unitcell * (1.0 / (2.0 * M_PI * v));
    for (const cplx& z : data) {
        const cplx r = data[0] / z;
        out.push_back(is_real ? cplx(pref * std::log(r.real()), 0.0) : pref * std::log(r));
    }
    return out;
}

double ground_state_degeneracy(Ctx& ctx, const Scheme& s, int unitcell) {
    std::vector<cplx> ev = transfer_spectrum(ctx, s, unitcell);
    // cft.jl:297-308: D / tr(D), S = -sum |d| log |d| over nonzero |d|, exp(S)
    cplx tr(0.0, 0.0);
    for (const cplx& z : ev) tr += z;
    double S = 0.0;
    for (const cplx& z : ev) {
        const double a = std::abs(z / tr);
        if (a > 0.0) S -= a * std::log(a);
    }
    return std::exp(S);
}

void gu_wen_ratio(Ctx& ctx, const Scheme& s, double* x1, double* x2) {
    TNR_REQUIRE(s.T.L.nleg == 4, TNR_ERR_ARG, "observables are defined for 2D (4-leg) tensors");
    // cft.jl:374-383 on T (BTRG: T_unit, cft.jl:386-387)
    BT Tu = unit_tensor(ctx, s);
    const int N = Tu.L.N;
    Scheme su;
    su.kind = TNR_TRG;
    su.T = std::move(Tu);
    BT M = transfer_matrix(ctx, su, 1);
    const double one = std::fabs(trace_of(ctx, M));
    // T[1 2; 2 3] T[3 4; 4 1] = tr(N N), N[a, b] = sum_i T[a, i, i, b]
    BT Iv = identity_leg(ctx, N, su.T.L.legs[2].dual(N));
    BT Nm = std::move(contract_many(ctx, {ContractJob{{&su.T, {-1, 1, 2, -2}}, {&Iv, {2, 1}},
                                                     {-1, -2}, 1}})[0]);
    BT NN = std::move(contract_many(ctx, {ContractJob{{&Nm, {-1, 1}}, {&Nm, {1, -2}}, {-1, -2},
                                                     1}})[0]);
    const double two1 = std::fabs(trace_of(ctx, NN));
    // T[1 2; 3 4] T[4 3; 2 1]
    const double two2 = std::fabs(trace_pair(ctx, su.T, su.T));
    *x1 = one * one / two1;
    *x2 = one * one / two2;
}

double central_charge(Ctx& ctx, const Scheme& s, double n) {
    TNR_REQUIRE(s.T.L.nleg == 4, TNR_ERR_ARG, "observables are defined for 2D (4-leg) tensors");
    // M[-1; -2] = T[1 -1; -2 1] (cft.jl:257), BTRG T[1 -1; 3 2] S1[3; -2]
    // S2[2; 1] (cft.jl:263-266); the largest singular value of the first
    // non-empty block (S.data[1]) of M / n
    const BT& T = s.T;
    const int N = T.L.N;
    TNR_REQUIRE(T.L.legs[3] == T.L.legs[0].dual(N), TNR_ERR_ARG,
                "central_charge needs equal left and right spaces");
    std::vector<WeightSpec> ws;
    if (s.kind == TNR_BTRG) ws = {WeightSpec{0, s.S2.same_side()}, WeightSpec{2, s.S1.dual_side()}};
    BT Id = identity_leg(ctx, N, T.L.legs[3].dual(N));
    BT M = std::move(contract_many(ctx, {ContractJob{{&T, {1, -1, -2, 2}, false, ws},
                                                    {&Id, {2, 1}}, {-1, -2}, 1}})[0]);
    int c0 = -1;
    for (int c = 0; c < M.L.N && c0 < 0; ++c)
        if (M.L.R[c] > 0 && M.L.C[c] > 0) c0 = c;
    TNR_REQUIRE(c0 >= 0, TNR_ERR_ARG, "empty transfer matrix");
    std::vector<SvdIn> in{SvdIn{M.d() + M.L.off[c0], (int)M.L.R[c0], (int)M.L.C[c0],
                                (int)M.L.R[c0], 1, 0x63656e7472616cULL}};
    std::vector<SvdOut> out;
    svd_grouped(ctx, in, out);
    return std::log(out[0].s_host[0] / n) * 6.0 / M_PI;
}

double finalize_two_by_two(Ctx& ctx, Scheme& s) {
    TNR_REQUIRE(s.T.L.nleg == 4, TNR_ERR_ARG, "observables are defined for 2D (4-leg) tensors");
    // n = |T[7 1; 5 4] T[4 2; 6 7] T[3 6; 2 8] T[8 5; 1 3]| (finalize.jl:18-21),
    // BTRG with S2 and S1 on the eight inner bonds (finalize.jl:29-38): as
    // P[1,5,2,6] = Ta Tb, Q[6,2,5,1] = Tc Td and sum P[a,b,c,d] Q[d,c,b,a]
    const BT& T = s.T;
    std::vector<WeightSpec> wa, wb, wc;
    if (s.kind == TNR_BTRG) {
        wa = {WeightSpec{1, s.S1.same_side()}};
        wb = {WeightSpec{0, s.S2.same_side()}, WeightSpec{1, s.S1.same_side()}};
        wc = {WeightSpec{0, s.S2.same_side()}};
    }
    auto PQ = contract_many(
        ctx, {ContractJob{{&T, {7, -1, -2, 4}, false, wa}, {&T, {4, -3, -4, 7}, false, wb},
                          {-1, -2, -3, -4}, 2},
              ContractJob{{&T, {3, -1, -2, 8}, false, wc}, {&T, {8, -3, -4, 3}},
                          {-1, -2, -3, -4}, 2}});
    const double n = std::fabs(trace_pair(ctx, PQ[0], PQ[1]));
    const double r = std::pow(n, 0.25);
    lincomb_many(ctx, {LinJob{&s.T, {{1.0 / r, &s.T}}}});
    return r;
}

}  // namespace tnr
