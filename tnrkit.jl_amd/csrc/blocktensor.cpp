// Block-sparse tensor layouts, permutes, contractions and grouped GEMM issue.
#include "blocktensor.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <numeric>
#include <sstream>

namespace tnr {

static int64_t ipow(int b, int e) {
    int64_t r = 1;
    for (int i = 0; i < e; ++i) r *= b;
    return r;
}

// ------------------------------------------------------------------ Layout
Layout::Layout(int N_, int ncod_, const std::vector<Leg>& legs_) {
    TNR_REQUIRE(N_ >= 1 && N_ <= MAXN, TNR_ERR_ARG, "Z_N order out of range");
    TNR_REQUIRE((int)legs_.size() <= MAXLEG, TNR_ERR_ARG, "too many legs");
    TNR_REQUIRE(ncod_ >= 0 && ncod_ <= (int)legs_.size(), TNR_ERR_ARG, "bad codomain size");
    N = N_;
    ncod = ncod_;
    nleg = (int)legs_.size();
    for (int i = 0; i < nleg; ++i) legs[i] = legs_[i];
    build();
}

void Layout::build() {
    R.fill(0);
    C.fill(0);
    off.fill(0);
    const int nd = nleg - ncod;
    roff.assign(ipow(N, ncod), 0);
    coff.assign(ipow(N, nd), 0);
    for (int64_t t = 0; t < (int64_t)roff.size(); ++t) {
        int64_t x = t, c = 0, sz = 1;
        for (int k = 0; k < ncod; ++k) {
            int l = (int)(x % N);
            x /= N;
            c += l;
            sz *= legs[k].d[l];
        }
        c %= N;
        roff[t] = R[c];
        R[c] += sz;
    }
    for (int64_t t = 0; t < (int64_t)coff.size(); ++t) {
        int64_t x = t, c = 0, sz = 1;
        for (int k = 0; k < nd; ++k) {
            int l = (int)(x % N);
            x /= N;
            c += l;
            sz *= legs[ncod + k].d[(N - l) % N];
        }
        c %= N;
        coff[t] = C[c];
        C[c] += sz;
    }
    int64_t o = 0;
    for (int c = 0; c < N; ++c) {
        off[c] = o;
        o += R[c] * C[c];
    }
    size = o;
}

bool Layout::same(const Layout& o) const {
    if (N != o.N || nleg != o.nleg || ncod != o.ncod) return false;
    for (int i = 0; i < nleg; ++i)
        if (legs[i] != o.legs[i]) return false;
    return true;
}

std::string Layout::str() const {
    std::ostringstream s;
    s << "Z" << N << " (";
    for (int i = 0; i < nleg; ++i) {
        if (i == ncod) s << "| ";
        s << "[";
        for (int q = 0; q < N; ++q) s << legs[i].d[q] << (q + 1 < N ? "," : "");
        s << "] ";
    }
    s << ") size " << size;
    return s.str();
}

BT::BT(Ctx& ctx, const Layout& L_) : L(L_), buf(&ctx.pool, std::max<int64_t>(L_.size, 1) * sizeof(double)) {}

// ---------------------------------------------------------------- BondVec
LegWeights BondVec::same_side() const {
    LegWeights w;
    w.w = ptr();
    for (int q = 0; q < N; ++q) w.off[q] = off[q];
    return w;
}

LegWeights BondVec::dual_side() const {
    LegWeights w;
    w.w = ptr();
    for (int q = 0; q < N; ++q) w.off[q] = off[(N - q) % N];
    return w;
}

Leg BondVec::leg_same() const {
    Leg l;
    for (int c = 0; c < N; ++c) l.d[c] = dim[c];
    return l;
}

Leg BondVec::leg_dual() const {
    Leg l;
    for (int c = 0; c < N; ++c) l.d[(N - c) % N] = dim[c];
    return l;
}

// --------------------------------------------------- sub-block enumeration
// Calls f(q[], sizes[]) for every outgoing-charge assignment with sum 0 and
// all leg sectors non-empty.
template <class F>
static void for_each_sector(const Layout& L, F&& f) {
    const int n = L.nleg, N = L.N;
    int q[MAXLEG] = {0};
    const int64_t tot = ipow(N, n > 0 ? n - 1 : 0);
    for (int64_t t = 0; t < tot; ++t) {
        int64_t x = t, s = 0;
        for (int k = 0; k + 1 < n; ++k) {
            q[k] = (int)(x % N);
            x /= N;
            s += q[k];
        }
        if (n > 0) q[n - 1] = (int)((N - s % N) % N);
        bool ok = true;
        for (int k = 0; k < n; ++k)
            if (L.legs[k].d[q[k]] == 0) { ok = false; break; }
        if (ok) f(q);
    }
}

// element offset of a sub-block and the stride of every leg inside it
static int64_t subblock(const Layout& L, const int* q, int64_t* stride) {
    const int N = L.N;
    int64_t tc = 0, td = 0, mul = 1, c = 0;
    int64_t st = 1;
    for (int k = 0; k < L.ncod; ++k) {
        int l = q[k];
        tc += l * mul;
        mul *= N;
        c += l;
        stride[k] = st;
        st *= L.legs[k].d[q[k]];
    }
    c %= N;
    mul = 1;
    st = L.R[c];
    for (int k = L.ncod; k < L.nleg; ++k) {
        int l = (N - q[k]) % N;
        td += l * mul;
        mul *= N;
        stride[k] = st;
        st *= L.legs[k].d[q[k]];
    }
    return L.off[c] + L.roff[tc] + L.R[c] * L.coff[td];
}

// ------------------------------------------------------------------ permute
static Layout permuted_layout(const Layout& S, const std::vector<int>& perm, int ncod, bool conj) {
    const int n = S.nleg;
    TNR_REQUIRE((int)perm.size() == n, TNR_ERR_INTERNAL, "permutation size mismatch");
    std::vector<Leg> dl(n);
    std::vector<int> seen(n, 0);
    for (int j = 0; j < n; ++j) {
        TNR_REQUIRE(perm[j] >= 0 && perm[j] < n && !seen[perm[j]], TNR_ERR_INTERNAL,
                    "invalid permutation");
        seen[perm[j]] = 1;
        dl[j] = conj ? S.legs[perm[j]].dual(S.N) : S.legs[perm[j]];
    }
    return Layout(S.N, ncod, dl);
}

// append the sub-block copy descriptors of one permute job
static void permute_descs(const BT& src, BT& dst, const std::vector<int>& perm, bool conj,
                          const std::vector<WeightSpec>& ws, std::vector<CopyDesc>& descs,
                          int64_t& start) {
    const Layout& S = src.L;
    const int n = S.nleg, N = S.N;
    TNR_REQUIRE(ws.size() <= 2, TNR_ERR_INTERNAL, "at most two weighted legs");
    std::vector<int> inv(n);
    for (int j = 0; j < n; ++j) inv[perm[j]] = j;
    for_each_sector(S, [&](const int* q) {
        int64_t ss[MAXLEG], ds[MAXLEG];
        int qd[MAXLEG];
        int64_t sbase = subblock(S, q, ss);
        for (int j = 0; j < n; ++j) qd[j] = conj ? (N - q[perm[j]]) % N : q[perm[j]];
        int64_t dbase = subblock(dst.L, qd, ds);
        CopyDesc d;
        std::memset(&d, 0, sizeof(d));
        d.src = src.d() + sbase;
        d.dst = dst.d() + dbase;
        d.nd = n;
        int64_t cnt = 1;
        for (int j = 0; j < n; ++j) {
            d.dim[j] = S.legs[perm[j]].d[q[perm[j]]];
            d.sstr[j] = ss[perm[j]];
            d.dstr[j] = ds[j];
            cnt *= d.dim[j];
        }
        // every access of the job must stay inside both tensors
        int64_t smax = sbase, dmax = dbase;
        for (int j = 0; j < n; ++j) {
            smax += (int64_t)(d.dim[j] - 1) * d.sstr[j];
            dmax += (int64_t)(d.dim[j] - 1) * d.dstr[j];
        }
        TNR_REQUIRE(sbase >= 0 && dbase >= 0 && smax < S.size && dmax < dst.L.size,
                    TNR_ERR_INTERNAL, "permute descriptor out of bounds");
        d.wleg[0] = d.wleg[1] = -1;
        for (size_t w = 0; w < ws.size(); ++w) {
            d.wleg[w] = inv[ws[w].leg];
            d.w[w] = ws[w].w.w + ws[w].w.off[q[ws[w].leg]];
        }
        d.start = start;
        start += cnt;
        descs.push_back(d);
    });
}

// Sub-blocks whose source- and destination-contiguous dims differ (and are
// both long enough) go to the tiled transpose; the rest to the element copy.
static void issue_copies(Ctx& ctx, const std::vector<CopyDesc>& descs, int64_t total) {
    if (total <= 0 || descs.empty()) return;
    std::vector<CopyDesc> elem;
    std::vector<TransDesc> tr;
    int64_t estart = 0, eelems = 0, tiles = 0, telems = 0;
    for (const CopyDesc& d : descs) {
        int s0 = -1, d0 = -1;
        for (int j = 0; j < d.nd; ++j) {
            if (d.dim[j] <= 1) continue;
            if (d.sstr[j] == 1) s0 = j;
            if (d.dstr[j] == 1) d0 = j;
        }
        int64_t cnt = 1;
        for (int j = 0; j < d.nd; ++j) cnt *= d.dim[j];
        const bool weighted = d.wleg[0] >= 0 || d.wleg[1] >= 0;
        if (!weighted && s0 >= 0 && d0 >= 0 && s0 != d0 && d.dim[s0] >= 8 && d.dim[d0] >= 8 &&
            d.nd - 2 <= TR_BATCH) {
            TransDesc t;
            std::memset(&t, 0, sizeof(t));
            t.src = d.src;
            t.dst = d.dst;
            t.n0 = d.dim[s0];
            t.ds0 = d.dstr[s0];
            t.n1 = d.dim[d0];
            t.ss1 = d.sstr[d0];
            int k = 0;
            int64_t nb = 1;
            for (int j = 0; j < d.nd; ++j) {
                if (j == s0 || j == d0) continue;
                t.bdim[k] = d.dim[j];
                t.bs[k] = d.sstr[j];
                t.bd[k] = d.dstr[j];
                nb *= d.dim[j];
                ++k;
            }
            for (; k < TR_BATCH; ++k) t.bdim[k] = 1;
            t.nb = (int32_t)nb;
            auto pow2 = [](int v) { int p = 1; while (p < v && p < 64) p <<= 1; return p; };
            t.tx = pow2(t.n0);
            t.ty = pow2(t.n1);
            t.sub = 4096 / (t.tx * t.ty);
            t.t0 = (t.n0 + t.tx - 1) / t.tx;
            t.t1 = (t.n1 + t.ty - 1) / t.ty;
            t.tile_start = tiles;
            tiles += (int64_t)t.t0 * t.t1 * ((nb + t.sub - 1) / t.sub);
            telems += cnt;
            tr.push_back(t);
        } else {
            TNR_REQUIRE(cnt < ((int64_t)1 << 31), TNR_ERR_INTERNAL, "copy job too large");
            CopyDesc e = d;
            e.start = estart;              // in 4096-element chunks
            estart += (cnt + 4095) / 4096;
            eelems += cnt;
            elem.push_back(e);
        }
    }
    (void)total;
    if (!tr.empty()) {
        TransDesc* dt = ctx.stage->put(ctx.stream, tr);
        launch_transpose(ctx.stream, dt, (int)tr.size(), tiles, telems);
    }
    if (!elem.empty()) {
        CopyDesc* dd = ctx.stage->put(ctx.stream, elem);
        launch_copy(ctx.stream, dd, (int)elem.size(), estart, eelems);
    }
}

std::vector<BT> permute_many(Ctx& ctx, const std::vector<PermJob>& jobs) {
    std::vector<BT> out;
    out.reserve(jobs.size());
    std::vector<CopyDesc> descs;
    int64_t start = 0;
    for (const PermJob& j : jobs) {
        out.emplace_back(ctx, permuted_layout(j.src->L, j.perm, j.ncod, j.conj));
        permute_descs(*j.src, out.back(), j.perm, j.conj, j.ws, descs, start);
    }
    issue_copies(ctx, descs, start);
    return out;
}

BT permute(Ctx& ctx, const BT& src, const std::vector<int>& perm, int ncod, bool conj,
           const std::vector<WeightSpec>& ws) {
    std::vector<PermJob> j(1);
    j[0].src = &src;
    j[0].perm = perm;
    j[0].ncod = ncod;
    j[0].conj = conj;
    j[0].ws = ws;
    return std::move(permute_many(ctx, j)[0]);
}

// -------------------------------------------------------------- grouped GEMM
// The staging ring recycles a chunk only after synchronising the stream, so
// the uploaded descriptors stay valid for every launch queued before the
// next wrap-around.
void DescArena::commit(Ctx& ctx) {
    if (host_.empty()) return;
    base_ = static_cast<char*>(ctx.stage->put(ctx.stream, host_.data(), host_.size()));
}

// K chunking: a function of K alone (every chunk is summed from zero and
// the chunk sums are added in order), so whether a launch splits K over
// workgroups changes no bits and can depend on the whole launch.
static int kchunk_of(int K) { return K >= 512 ? 256 : std::max(32, (K + 31) / 32 * 32); }
static int nchunks_of(int K) { return K >= 512 ? (K + 255) / 256 : 1; }

// Tile shape of one grouped launch: the largest tile that still gives the
// chip >= 4 workgroups per CU and pads the problems by at most 15 % more
// than 64 x 64 tiles do.  Larger tiles halve (128 x 64) or quarter
// (128 x 128) the operand bytes per flop the workgroups pull through L2.
static int pick_shape(const std::vector<GemmJob>& jobs, bool split32) {
    if (split32) return GEMM_64x64;
    const int bm[3] = {64, 128, 128}, bn[3] = {64, 64, 128};
    double pad[3] = {0, 0, 0};
    int64_t tiles[3] = {0, 0, 0};
    for (const GemmJob& j : jobs) {
        if (j.M <= 0 || j.N <= 0) continue;
        const int ks = nchunks_of(j.K);
        for (int s = 0; s < 3; ++s) {
            const int64_t tm = (j.M + bm[s] - 1) / bm[s], tn = (j.N + bn[s] - 1) / bn[s];
            tiles[s] += tm * tn * ks;
            pad[s] += (double)(tm * bm[s]) * (double)(tn * bn[s]) * std::max(j.K, 1);
        }
    }
    for (int s = 2; s >= 1; --s)
        if (tiles[s] >= 1024 && pad[s] <= 1.15 * pad[0]) return s;
    return GEMM_64x64;
}


// ------------------------------------------------------ tile-shape autotuner
// Every output element is summed in the same order under each tile shape,
// so the choice changes speed only, never results.  Launches are keyed by a
// coarse signature (device, operand transposes, log2 of the largest M, N,
// K, of the problem count and of the flops).  The first occurrences of a
// key are timed once per shape (HIP events around the launch, read back
// later without waiting); once every shape has TUNE_SAMPLES timings the
// fastest is used for that key.  Until then the heuristic above decides.
// The tuner is off by default (TNR_AUTOTUNE=1 or gemm_autotune_enable
// turns it on): the heuristic alone then picks every shape.
namespace {
constexpr int TUNE_SAMPLES = 3;
using TuneKey = std::array<int, 9>;
struct TuneTrial {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int shape = 0;
};
struct TuneEntry {
    bool cand[GEMM_NSHAPES] = {};      // shapes this key chooses among
    float ms[GEMM_NSHAPES] = {};
    int samples[GEMM_NSHAPES] = {};
    int requested[GEMM_NSHAPES] = {};
    std::vector<TuneTrial> pending;
    int best = -1;
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneEntry> g_tune;
int g_tune_on = -1;   // -1: read TNR_AUTOTUNE on first use

int ilog2(double v) { return v < 1.0 ? 0 : (int)std::floor(std::log2(v)); }

bool tune_enabled() {
    if (g_tune_on < 0) {
        // off unless asked for: timed on alternating runs of the 128-run
        // BTRG batch, the static rule was 3 % faster (the trials run slower
        // shapes, and timings under the other lanes' load pick noisily)
        const char* e = std::getenv("TNR_AUTOTUNE");
        g_tune_on = (e && e[0] == '1') ? 1 : 0;
    }
    return g_tune_on == 1;
}

void tune_poll(TuneEntry& t) {
    for (size_t i = 0; i < t.pending.size();) {
        TuneTrial& tr = t.pending[i];
        if (hipEventQuery(tr.e1) != hipSuccess) { ++i; continue; }
        float ms = 0;
        if (hipEventElapsedTime(&ms, tr.e0, tr.e1) == hipSuccess) {
            const int s = tr.shape;
            t.ms[s] = t.samples[s] ? std::min(t.ms[s], ms) : ms;
            ++t.samples[s];
        } else {
            --t.requested[tr.shape];   // measure again
        }
        (void)hipEventDestroy(tr.e0);
        (void)hipEventDestroy(tr.e1);
        t.pending.erase(t.pending.begin() + (long)i);
    }
    bool all = true;
    for (int s = 0; s < GEMM_NSHAPES; ++s) all = all && (!t.cand[s] || t.samples[s] >= TUNE_SAMPLES);
    if (all) {
        int best = -1;
        for (int s = 0; s < GEMM_NSHAPES; ++s)
            if (t.cand[s] && (best < 0 || t.ms[s] < t.ms[best])) best = s;
        t.best = best;
    }
}
}  // namespace

void gemm_autotune_enable(int on) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_on = on ? 1 : 0;
    for (auto& kv : g_tune)
        for (auto& tr : kv.second.pending) {
            (void)hipEventSynchronize(tr.e1);
            (void)hipEventDestroy(tr.e0);
            (void)hipEventDestroy(tr.e1);
        }
    g_tune.clear();
}

// Shape for a launch and whether this launch is a timing trial.
// The 32-deep register-staged 64 x 64 tile competes only where the LDS-DMA
// kernels cannot take the launch: measured alone it is 1.3-1.5x slower than
// the LDS-DMA tiles on every product of the block iteration and the
// contractions, and under the timing noise of concurrent lanes the tuner
// picked it anyway.
static int tuned_shape(const Ctx& ctx, const std::vector<GemmJob>& jobs, int heuristic,
                       bool dma_ok, TuneKey* key_out, bool* trial) {
    *trial = false;
    if (!tune_enabled()) return heuristic;
    int mM = 0, mN = 0, mK = 0, cnt = 0, ta = 0, tb = 0;
    double fl = 0;
    for (const GemmJob& j : jobs) {
        if (j.M <= 0 || j.N <= 0) continue;
        if (!cnt) { ta = j.ta; tb = j.tb; }
        mM = std::max(mM, j.M); mN = std::max(mN, j.N); mK = std::max(mK, j.K);
        fl += 2.0 * j.M * j.N * j.K;
        ++cnt;
    }
    if (!cnt) return heuristic;
    const TuneKey key{ctx.device, ta, tb, ilog2(mM), ilog2(mN), ilog2(mK), ilog2(cnt),
                      ilog2(fl), dma_ok ? 1 : 0};
    std::lock_guard<std::mutex> lk(g_tune_mu);
    TuneEntry& t = g_tune[key];
    if (t.best >= 0) return t.best;
    for (int s = 0; s < GEMM_NSHAPES; ++s) t.cand[s] = !(dma_ok && s == GEMM_64x64_K32);
    tune_poll(t);
    if (t.best >= 0) return t.best;
    for (int s = GEMM_NSHAPES - 1; s >= 0; --s)
        if (t.cand[s] && t.requested[s] < TUNE_SAMPLES) {
            ++t.requested[s];
            *key_out = key;
            *trial = true;
            return s;
        }
    return heuristic;   // trials in flight
}

static void tune_record_begin(Ctx& ctx, GemmPlan& p, TuneTrial& tr) {
    TNR_HIP(hipEventCreate(&tr.e0));
    TNR_HIP(hipEventCreate(&tr.e1));
    tr.shape = p.shape;
    TNR_HIP(hipEventRecord(tr.e0, ctx.stream));
}

static void tune_record_end(Ctx& ctx, const GemmPlan& p, TuneTrial& tr) {
    TNR_HIP(hipEventRecord(tr.e1, ctx.stream));
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(p.tune_key);
    if (it == g_tune.end()) {   // the tuner was reset meanwhile
        (void)hipEventSynchronize(tr.e1);
        (void)hipEventDestroy(tr.e0);
        (void)hipEventDestroy(tr.e1);
        return;
    }
    it->second.pending.push_back(tr);
}

// TNR_GEMM_DMA=0 keeps the 16-deep tile shapes on register-staged loads
// (the LDS-DMA kernels give the same bits)
static bool gemm_dma_enabled() {
    static const bool on = [] {
        const char* e = std::getenv("TNR_GEMM_DMA");
        return !(e && e[0] == '0');
    }();
    return on;
}

GemmPlan gemm_plan(Ctx& ctx, const std::vector<GemmJob>& jobs, bool split32, int force_shape,
                   int force_split) {
    GemmPlan P;
    P.split32 = split32;
    bool dma_ok = gemm_dma_enabled();
    for (const GemmJob& j : jobs)
        if (j.M > 0 && j.N > 0) dma_ok = dma_ok && gemm_dma_ok(j.ta, j.lda, j.tb, j.ldb);
    if (force_shape >= 0 && !split32) P.shape = force_shape;
    else if (split32) P.shape = GEMM_64x64;
    else P.shape = tuned_shape(ctx, jobs, pick_shape(jobs, false), dma_ok, &P.tune_key, &P.trial);
    const int BMs = gemm_shape_bm(P.shape), BNs = gemm_shape_bn(P.shape);
    // split K over workgroups only when the launch is too small to fill
    // the chip without it (results are the same either way)
    int64_t whole = 0;
    for (const GemmJob& j : jobs)
        if (j.M > 0 && j.N > 0)
            whole += (int64_t)((j.M + BMs - 1) / BMs) * ((j.N + BNs - 1) / BNs);
    const bool split = force_split >= 0 ? force_split != 0 : whole < 512;
    size_t work_elems = 0;
    P.descs.reserve(jobs.size());
    for (const GemmJob& j : jobs) {
        if (j.M <= 0 || j.N <= 0) continue;
        TNR_REQUIRE(j.K >= 0 && j.lda >= 1 && j.ldb >= 1 && j.ldc >= j.M, TNR_ERR_INTERNAL,
                    "bad GEMM job");
        GemmDesc g;
        std::memset(&g, 0, sizeof(g));
        g.A = j.A; g.B = j.B; g.C = j.C;
        g.M = j.M; g.N = j.N; g.K = j.K;
        g.lda = j.lda; g.ldb = j.ldb; g.ldc = j.ldc;
        g.ta = j.ta; g.tb = j.tb;
        g.alpha = j.alpha; g.beta = j.beta;
        g.skip = j.skip;
        g.tiles_m = (j.M + BMs - 1) / BMs;
        g.tiles_n = (j.N + BNs - 1) / BNs;
        g.kchunk = kchunk_of(j.K);
        const int ks = split ? nchunks_of(j.K) : 1;
        const int64_t mn = (int64_t)g.tiles_m * g.tiles_n;
        g.ksplit = ks;
        g.tile_start = P.tiles;
        P.tiles += mn * ks;
        if (ks > 1) work_elems += (size_t)j.M * j.N * ks;
        P.flops += 2.0 * j.M * j.N * j.K;
        P.dma = P.dma && gemm_dma_ok(g);
        P.descs.push_back(g);
    }
    if (work_elems) {
        P.work = Buf(&ctx.pool, work_elems * sizeof(double));
        size_t o = 0;
        for (auto& g : P.descs)
            if (g.ksplit > 1) {
                g.work = P.work.d() + o;
                o += (size_t)g.M * g.N * g.ksplit;
                P.red.push_back(g);
                P.red_starts.push_back(P.red_total);
                P.red_total += (int64_t)g.M * g.N;
            }
    }
    return P;
}

void gemm_stage(GemmPlan& p, DescArena& arena) {
    p.off_descs = arena.add(p.descs);
    p.off_red = arena.add(p.red);
    p.off_starts = arena.add(p.red_starts);
}

void gemm_bind(GemmPlan& p, const DescArena& arena) {
    p.d_descs = arena.at<GemmDesc>(p.off_descs);
    p.d_red = arena.at<GemmDesc>(p.off_red);
    p.d_starts = arena.at<int64_t>(p.off_starts);
}

void gemm_issue(Ctx& ctx, GemmPlan& p) {
    if (p.descs.empty()) return;
    TuneTrial tr;
    const bool trial = p.trial;
    p.trial = false;   // one timing per plan
    if (trial) tune_record_begin(ctx, p, tr);
    launch_gemm(ctx.stream, p.d_descs, (int)p.descs.size(), p.tiles, p.split32, p.flops,
                p.shape, p.dma && gemm_dma_enabled());
    if (trial) tune_record_end(ctx, p, tr);
    if (!p.red.empty())
        launch_gemm_reduce(ctx.stream, p.d_red, (int)p.red.size(), p.d_starts, p.red_total);
}

void gemm_grouped(Ctx& ctx, const std::vector<GemmJob>& jobs, bool split32, int force_shape,
                  int force_split, bool allow_dma) {
    GemmPlan p = gemm_plan(ctx, jobs, split32, force_shape, force_split);
    p.dma = p.dma && allow_dma;
    if (p.descs.empty()) return;
    DescArena arena;
    gemm_stage(p, arena);
    arena.commit(ctx);
    gemm_bind(p, arena);
    gemm_issue(ctx, p);
    // the arena and the split-K workspace return to the pool here; their
    // next users are queued after these launches on the same stream
}

// --------------------------------------------------------------- contract
static bool is_identity(const std::vector<int>& p) {
    for (size_t i = 0; i < p.size(); ++i)
        if (p[i] != (int)i) return false;
    return true;
}

namespace {
struct Plan {
    std::vector<int> ca, cb, oa, ob;   // contracted / open legs
    int transA = 0, transB = 0;
    int permA = -1, permB = -1;        // index into the operand-permute batch
    std::vector<int> pc;               // output permutation
    Layout Cl;
};
}  // namespace

std::vector<BT> contract_many(Ctx& ctx, const std::vector<ContractJob>& jobs) {
    const size_t J = jobs.size();
    std::vector<Plan> plans(J);
    std::vector<PermJob> pj;
    for (size_t x = 0; x < J; ++x) {
        const Operand& a = jobs[x].a;
        const Operand& b = jobs[x].b;
        Plan& P = plans[x];
        const Layout& LA = a.t->L;
        const Layout& LB = b.t->L;
        const int N = LA.N;
        TNR_REQUIRE(LB.N == N, TNR_ERR_INTERNAL, "mixed symmetry groups");
        TNR_REQUIRE((int)a.labels.size() == LA.nleg && (int)b.labels.size() == LB.nleg,
                    TNR_ERR_INTERNAL, "label count mismatch");
        for (int i = 0; i < LA.nleg; ++i) {
            int l = a.labels[i];
            if (l > 0) {
                auto it = std::find(b.labels.begin(), b.labels.end(), l);
                TNR_REQUIRE(it != b.labels.end(), TNR_ERR_INTERNAL, "unmatched contracted label");
                P.ca.push_back(i);
                P.cb.push_back((int)(it - b.labels.begin()));
            }
        }
        for (int l : jobs[x].out) {
            auto ia = std::find(a.labels.begin(), a.labels.end(), l);
            if (ia != a.labels.end()) { P.oa.push_back((int)(ia - a.labels.begin())); continue; }
            auto ib = std::find(b.labels.begin(), b.labels.end(), l);
            TNR_REQUIRE(ib != b.labels.end(), TNR_ERR_INTERNAL, "unknown output label");
            P.ob.push_back((int)(ib - b.labels.begin()));
        }
        TNR_REQUIRE(P.oa.size() + P.ca.size() == (size_t)LA.nleg &&
                        P.ob.size() + P.cb.size() == (size_t)LB.nleg,
                    TNR_ERR_INTERNAL, "every leg must be open or contracted once");
        const bool conj_noop = N <= 2;   // -q == q
        const bool cja = a.conj && !conj_noop, cjb = b.conj && !conj_noop;
        std::vector<int> pa(P.oa), pb(P.cb);
        pa.insert(pa.end(), P.ca.begin(), P.ca.end());
        pb.insert(pb.end(), P.ob.begin(), P.ob.end());
        // A as (open | contracted): as stored, as a transposed view (only when
        // tuple orders coincide, N <= 2), or permuted
        std::vector<int> ta(P.ca), tb(P.ob);
        ta.insert(ta.end(), P.oa.begin(), P.oa.end());
        tb.insert(tb.end(), P.cb.begin(), P.cb.end());
        if (!cja && a.ws.empty() && is_identity(pa) && LA.ncod == (int)P.oa.size()) {
        } else if (conj_noop && a.ws.empty() && is_identity(ta) && LA.ncod == (int)P.ca.size()) {
            P.transA = 1;
        } else {
            P.permA = (int)pj.size();
            pj.push_back(PermJob{a.t, pa, (int)P.oa.size(), cja, a.ws});
        }
        if (!cjb && b.ws.empty() && is_identity(pb) && LB.ncod == (int)P.cb.size()) {
        } else if (conj_noop && b.ws.empty() && is_identity(tb) && LB.ncod == (int)P.ob.size()) {
            P.transB = 1;
        } else {
            P.permB = (int)pj.size();
            pj.push_back(PermJob{b.t, pb, (int)P.cb.size(), cjb, b.ws});
        }
        auto leg_of = [&](const Operand& o, int i) {
            return (o.conj && !conj_noop) ? o.t->L.legs[i].dual(N) : o.t->L.legs[i];
        };
        for (size_t i = 0; i < P.ca.size(); ++i)
            TNR_REQUIRE(leg_of(a, P.ca[i]) == leg_of(b, P.cb[i]).dual(N), TNR_ERR_INTERNAL,
                        "contracted legs do not carry dual spaces");
        std::vector<Leg> cl;
        for (int i : P.oa) cl.push_back(leg_of(a, i));
        for (int i : P.ob) cl.push_back(leg_of(b, i));
        P.Cl = Layout(N, (int)P.oa.size(), cl);
        for (int l : jobs[x].out) {
            int pos = -1;
            for (size_t i = 0; i < P.oa.size(); ++i)
                if (a.labels[P.oa[i]] == l) pos = (int)i;
            for (size_t i = 0; i < P.ob.size(); ++i)
                if (b.labels[P.ob[i]] == l) pos = (int)(P.oa.size() + i);
            P.pc.push_back(pos);
        }
    }
    std::vector<BT> ops = permute_many(ctx, pj);
    std::vector<BT> Cs;
    Cs.reserve(J);
    std::vector<GemmJob> gj;
    for (size_t x = 0; x < J; ++x) {
        Plan& P = plans[x];
        const BT* A = P.permA >= 0 ? &ops[P.permA] : jobs[x].a.t;
        const BT* B = P.permB >= 0 ? &ops[P.permB] : jobs[x].b.t;
        Cs.emplace_back(ctx, P.Cl);
        BT& C = Cs.back();
        for (int c = 0; c < C.L.N; ++c) {
            const int64_t M = C.L.R[c], Nn = C.L.C[c];
            if (M == 0 || Nn == 0) continue;
            const int64_t K = P.transA ? A->L.R[c] : A->L.C[c];
            const int64_t KB = P.transB ? B->L.C[c] : B->L.R[c];
            TNR_REQUIRE(K == KB, TNR_ERR_INTERNAL, "contracted dimension mismatch");
            GemmJob j;
            j.A = A->d() + A->L.off[c];
            j.B = B->d() + B->L.off[c];
            j.C = C.d() + C.L.off[c];
            j.M = (int)M; j.N = (int)Nn; j.K = (int)K;
            j.lda = (int)std::max<int64_t>(A->L.R[c], 1);
            j.ldb = (int)std::max<int64_t>(B->L.R[c], 1);
            j.ldc = (int)M;
            j.ta = P.transA; j.tb = P.transB;
            gj.push_back(j);
        }
    }
    gemm_grouped(ctx, gj, ctx.precision == TNR_PREC_SPLIT32);
    // output arrangement
    std::vector<PermJob> fj;
    std::vector<int> fidx(J, -1);
    for (size_t x = 0; x < J; ++x) {
        const Plan& P = plans[x];
        if (is_identity(P.pc) && jobs[x].out_ncod == (int)P.oa.size()) continue;
        fidx[x] = (int)fj.size();
        fj.push_back(PermJob{&Cs[x], P.pc, jobs[x].out_ncod, false, {}});
    }
    std::vector<BT> fin = permute_many(ctx, fj);
    std::vector<BT> out;
    out.reserve(J);
    for (size_t x = 0; x < J; ++x)
        out.push_back(fidx[x] >= 0 ? std::move(fin[fidx[x]]) : std::move(Cs[x]));
    return out;
}

// -------------------------------------------------------- trace/normalise
void trace_normalize_many(Ctx& ctx, const std::vector<TraceJob>& jobs) {
    std::vector<TraceDesc> descs;
    std::vector<TraceProb> probs;
    Buf partial(&ctx.pool, std::max<size_t>(1, jobs.size()) * TRACE_PARTS * sizeof(double));
    int64_t max_size = 0;
    for (size_t x = 0; x < jobs.size(); ++x) {
        BT& t = *jobs[x].t;
        const Layout& L = t.L;
        const int N = L.N;
        TraceProb P;
        P.first = (int)descs.size();
        int64_t start = 0;
        if (L.nleg == 6) {
            // n = sum T[a a; b c b c] (src/utility/finalize.jl:62-66), no weights
            TNR_REQUIRE(L.ncod == 2 && L.legs[1] == L.legs[0].dual(N) &&
                            L.legs[4] == L.legs[2].dual(N) && L.legs[5] == L.legs[3].dual(N),
                        TNR_ERR_ARG, "legs 1/2, 3/5 and 4/6 must carry the same spaces");
            TNR_REQUIRE(!jobs[x].w_x1 && !jobs[x].w_x2, TNR_ERR_INTERNAL, "weighted 3D trace");
            for (int q0 = 0; q0 < N; ++q0)
                for (int q2 = 0; q2 < N; ++q2)
                    for (int q3 = 0; q3 < N; ++q3) {
                        int q[6] = {q0, (N - q0) % N, q2, q3, (N - q2) % N, (N - q3) % N};
                        if (!L.legs[0].d[q0] || !L.legs[2].d[q2] || !L.legs[3].d[q3]) continue;
                        int64_t st[6];
                        int64_t base = subblock(L, q, st);
                        TraceDesc d;
                        d.T = t.d() + base;
                        d.s1 = st[0] + st[1];
                        d.s2 = st[2] + st[4];
                        d.s3 = st[3] + st[5];
                        d.d1 = L.legs[0].d[q0];
                        d.d2 = L.legs[2].d[q2];
                        d.d3 = L.legs[3].d[q3];
                        TNR_REQUIRE(base + (int64_t)(d.d1 - 1) * d.s1 + (int64_t)(d.d2 - 1) * d.s2 +
                                            (int64_t)(d.d3 - 1) * d.s3 < L.size,
                                    TNR_ERR_INTERNAL, "trace descriptor out of bounds");
                        d.w1 = d.w2 = nullptr;
                        d.start = start;
                        start += (int64_t)d.d1 * d.d2 * d.d3;
                        descs.push_back(d);
                    }
        } else {
        TNR_REQUIRE(L.nleg == 4 && L.ncod == 2, TNR_ERR_INTERNAL, "finalize needs a (2,2) tensor");
        TNR_REQUIRE(L.legs[0] == L.legs[3].dual(N) && L.legs[1] == L.legs[2].dual(N),
                    TNR_ERR_ARG, "legs 1/4 and 2/3 must carry the same spaces");
        for (int q1 = 0; q1 < N; ++q1)
            for (int q2 = 0; q2 < N; ++q2) {
                int q[4] = {q1, q2, (N - q2) % N, (N - q1) % N};
                if (!L.legs[0].d[q1] || !L.legs[1].d[q2]) continue;
                int64_t st[4];
                int64_t base = subblock(L, q, st);
                TNR_REQUIRE(base + (int64_t)(L.legs[0].d[q1] - 1) * (st[0] + st[3]) +
                                    (int64_t)(L.legs[1].d[q2] - 1) * (st[1] + st[2]) < L.size,
                            TNR_ERR_INTERNAL, "trace descriptor out of bounds");
                TraceDesc d;
                d.T = t.d() + base;
                d.s1 = st[0] + st[3];
                d.s2 = st[1] + st[2];
                d.d1 = L.legs[0].d[q1];
                d.d2 = L.legs[1].d[q2];
                d.w1 = jobs[x].w_x1 ? jobs[x].w_x1->w + jobs[x].w_x1->off[q1] : nullptr;
                d.w2 = jobs[x].w_x2 ? jobs[x].w_x2->w + jobs[x].w_x2->off[q2] : nullptr;
                d.s3 = 0;
                d.d3 = 1;
                d.start = start;
                start += (int64_t)d.d1 * d.d2;
                descs.push_back(d);
            }
        }
        TNR_REQUIRE(start > 0, TNR_ERR_ARG, "empty tensor");
        P.count = (int)descs.size() - P.first;
        P.total = start;
        P.T = t.d();
        P.size = L.size;
        P.partial = partial.d() + x * TRACE_PARTS;
        P.norm_slot = jobs[x].norm_slot;
        P.divide_by = jobs[x].divide_by;
        probs.push_back(P);
        max_size = std::max(max_size, L.size);
    }
    if (probs.empty()) return;
    TraceDesc* dd = ctx.stage->put(ctx.stream, descs);
    TraceProb* dp = ctx.stage->put(ctx.stream, probs);
    launch_trace_normalize(ctx.stream, dd, dp, (int)probs.size(), max_size);
}

// ------------------------------------------------------ linear combinations
void lincomb_many(Ctx& ctx, const std::vector<LinJob>& jobs) {
    std::vector<LinDesc> d;
    int64_t start = 0;
    for (const LinJob& j : jobs) {
        TNR_REQUIRE(j.out && !j.terms.empty() && j.terms.size() <= 4, TNR_ERR_INTERNAL,
                    "bad linear combination");
        LinDesc x{};
        x.out = j.out->d();
        x.n = j.out->L.size;
        x.nin = (int32_t)j.terms.size();
        for (size_t i = 0; i < j.terms.size(); ++i) {
            TNR_REQUIRE(j.terms[i].second->L.same(j.out->L), TNR_ERR_ARG,
                        "combined tensors must share their spaces");
            x.in[i] = j.terms[i].second->d();
            x.coef[i] = j.terms[i].first;
        }
        x.start = start;
        start += x.n;
        if (x.n > 0) d.push_back(x);
    }
    if (d.empty()) return;
    LinDesc* dd = ctx.stage->put(ctx.stream, d);
    launch_lincomb(ctx.stream, dd, (int)d.size(), start);
}

// ------------------------------------------------------------------ gather
void gather_many(Ctx& ctx, std::vector<GatherDesc> descs) {
    int64_t start = 0;
    std::vector<GatherDesc> d;
    for (auto& g : descs) {
        if (g.rows <= 0 || g.cols <= 0) continue;
        g.start = start;
        start += (int64_t)g.rows * g.cols;
        d.push_back(g);
    }
    if (d.empty()) return;
    GatherDesc* dd = ctx.stage->put(ctx.stream, d);
    launch_gather(ctx.stream, dd, (int)d.size(), start);
}

void readback_vectors(Ctx& ctx, const std::vector<const double*>& src,
                      const std::vector<int>& cnt, const std::vector<std::vector<double>*>& dst) {
    size_t total = 0;
    for (size_t p = 0; p < src.size(); ++p) {
        dst[p]->resize(std::max(cnt[p], 0));
        total += (size_t)std::max(cnt[p], 0);
    }
    if (!total) return;
    Buf packed(&ctx.pool, total * sizeof(double));
    std::vector<GatherDesc> gd;
    size_t o = 0;
    for (size_t p = 0; p < src.size(); ++p) {
        if (cnt[p] <= 0) continue;
        GatherDesc g{};
        g.in = src[p];
        g.out = packed.d() + o;
        g.rows = cnt[p];
        g.cols = 1;
        g.ldi = cnt[p];
        g.ldo = cnt[p];
        gd.push_back(g);
        o += (size_t)cnt[p];
    }
    gather_many(ctx, gd);
    std::vector<double> h(total);
    ctx.fetch(h.data(), packed.d(), total * sizeof(double));
    o = 0;
    for (size_t p = 0; p < src.size(); ++p) {
        if (cnt[p] <= 0) continue;
        std::copy(h.begin() + o, h.begin() + o + cnt[p], dst[p]->begin());
        o += (size_t)cnt[p];
    }
}

// ---------------------------------------------------------- ABI transfer
Layout abi_layout(int N, const int32_t* leg_dims, int nleg, int ncod) {
    TNR_REQUIRE(N >= 1 && N <= MAXN, TNR_ERR_ARG, "zn must be in 1..8");
    TNR_REQUIRE(nleg >= 2 && nleg <= MAXLEG && ncod >= 1 && ncod < nleg, TNR_ERR_ARG,
                "bad leg counts");
    std::vector<Leg> legs(nleg);
    for (int i = 0; i < nleg; ++i)
        for (int a = 0; a < N; ++a) {
            int d = leg_dims[i * N + a];
            TNR_REQUIRE(d >= 0, TNR_ERR_ARG, "negative sector dimension");
            // codomain legs: q = a; domain legs: q = -a
            int q = (i < ncod) ? a : (N - a) % N;
            legs[i].d[q] = d;
        }
    return Layout(N, ncod, legs);
}

void abi_dims(const Layout& L, int32_t* leg_dims) {
    const int N = L.N;
    for (int i = 0; i < L.nleg; ++i)
        for (int a = 0; a < N; ++a) {
            int q = (i < L.ncod) ? a : (N - a) % N;
            leg_dims[i * N + a] = L.legs[i].d[q];
        }
}

BT upload(Ctx& ctx, int N, const int32_t* leg_dims, const double* data, size_t n, int nleg,
          int ncod) {
    TNR_REQUIRE(leg_dims && data, TNR_ERR_ARG, "null pointer");
    Layout L = abi_layout(N, leg_dims, nleg, ncod);
    TNR_REQUIRE((int64_t)n == L.size, TNR_ERR_ARG,
                "ndata = " + std::to_string(n) + " does not match the sector dimensions (" +
                    std::to_string(L.size) + ")");
    TNR_REQUIRE(L.size > 0, TNR_ERR_ARG, "empty tensor");
    // NaN / Inf entries would only surface as non-finite norms later (and,
    // in a grouped launch, inside other runs' solves): refused here
    for (size_t i = 0; i < n; ++i)
        TNR_REQUIRE(std::isfinite(data[i]), TNR_ERR_NONFINITE,
                    "tensor entry " + std::to_string(i) + " is not finite");
    BT t(ctx, L);
    TNR_HIP(hipMemcpyAsync(t.d(), data, n * sizeof(double), hipMemcpyHostToDevice, ctx.stream));
    ctx.sync();
    return t;
}

void download(Ctx& ctx, const BT& t, double* data) {
    ctx.fetch(data, t.d(), t.L.size * sizeof(double));
}

}  // namespace tnr
