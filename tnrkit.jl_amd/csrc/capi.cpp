// extern "C" entry points of libtnr_hip.so (include/tnr.h).  No exception
// crosses this boundary: every call maps failures to a status code and keeps
// the message for tnr_last_error().
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "observables.h"
#include "schemes.h"

using namespace tnr;

namespace {
thread_local std::string g_last_error;

}  // namespace

void tnr::set_last_error(const std::string& msg) { g_last_error = msg; }

namespace {
int fail(int code, const char* msg) {
    g_last_error = msg;
    return code;
}

template <class F>
int guarded(F&& f) {
    try {
        return f();
    } catch (const Error& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return TNR_ERR_NOMEM;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return TNR_ERR_INTERNAL;
    } catch (...) {
        g_last_error = "unknown failure";
        return TNR_ERR_INTERNAL;
    }
}

int steps_of(int maxiter) { return maxiter < 1 ? 1 : maxiter; }   // run! steps at least once

void check_norms(const double* v, int n) {
    for (int i = 0; i < n; ++i)
        if (!std::isfinite(v[i]) || v[i] == 0.0)
            throw Error(TNR_ERR_NONFINITE, "norm " + std::to_string(i) + " is " + std::to_string(v[i]));
}

// A scheme whose step failed (sticky until its tensor is replaced): the
// reference's step! raises there, e.g. from a LAPACK-backed svd_trunc.
void check_status(const Scheme& s) {
    if (s.status == TNR_ERR_NOCONV)
        throw Error(s.status, "a truncated SVD or eigensolve did not converge; the tensor is "
                              "undefined until tnr_set_tensor replaces it");
    if (s.status != TNR_OK)
        throw Error(s.status, "a truncated SVD or eigensolve met non-finite values; the tensor "
                              "is undefined until tnr_set_tensor replaces it");
}

// one step! of a single scheme, failing like the reference's step!
void step_one(Ctx& ctx, Scheme& s, int chi) {
    check_status(s);
    step_many(ctx, {&s}, {chi});
    check_status(s);
}

double now_s(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}
}  // namespace

// Contexts of destroyed handles are kept (a few per process) and handed to
// the next handle on the same device.  A context per handle from scratch
// means a newly pinned staging ring and an empty pool, so every buffer of
// the handle's first steps would be allocated (and later freed) with the
// device synchronised each time.  A kept context holds no stream: idle
// streams would still take the device's few hardware queues.  The list of
// spares is never destroyed: at process exit the HIP runtime may already be
// gone.
namespace {
constexpr size_t kSpareContexts = 4;
constexpr size_t kSparePoolBytes = size_t(8) << 30;   // cache kept with a spare context
std::mutex g_ctx_mu;
std::vector<std::unique_ptr<Ctx>>& spare_contexts() {
    static auto* v = new std::vector<std::unique_ptr<Ctx>>();
    return *v;
}

std::unique_ptr<Ctx> take_context(int precision) {
    ensure_device();
    int dev = 0;
    TNR_HIP(hipGetDevice(&dev));
    std::unique_ptr<Ctx> c;
    {
        std::lock_guard<std::mutex> lk(g_ctx_mu);
        auto& v = spare_contexts();
        for (size_t i = 0; i < v.size(); ++i)
            if (v[i]->device == dev) {
                c = std::move(v[i]);
                v.erase(v.begin() + (long)i);
                break;
            }
    }
    if (!c) c.reset(new Ctx());
    c->ensure_stream();
    c->precision = precision;
    return c;
}

// never throws: a context whose stream reports an error is destroyed
void give_context(std::unique_ptr<Ctx> c) noexcept {
    if (!c) return;
    try {
        c->drop_stream();
        if (c->pool.cached() > kSparePoolBytes) c->pool.trim();
        std::lock_guard<std::mutex> lk(g_ctx_mu);
        auto& v = spare_contexts();
        if (v.size() < kSpareContexts) v.push_back(std::move(c));
    } catch (...) {
    }
}
}  // namespace

struct tnr_scheme {
    std::unique_ptr<Ctx> ctx;
    Scheme s;
    Buf norms;
    int cap = 0;
    ~tnr_scheme() {
        try {
            if (ctx) ctx->sync();
        } catch (...) {
        }
        s = Scheme();
        norms.release();
        give_context(std::move(ctx));
    }
    double* slots(int n) {
        if (n > cap) {
            norms = Buf(&ctx->pool, (size_t)n * sizeof(double));
            cap = n;
        }
        return norms.d();
    }
};

extern "C" {

int tnr_create(tnr_scheme** out, int kind, int zn, const int32_t* leg_dims, const double* data,
               size_t ndata, double k, int precision) {
    return guarded([&] {
        if (!out) return fail(TNR_ERR_ARG, "out is null");
        *out = nullptr;
        if (!single_tensor_kind(kind) || kind == TNR_HOTRG_3D)
            return fail(TNR_ERR_ARG, "kind must be TNR_TRG, TNR_BTRG, TNR_HOTRG or TNR_ATRG "
                                     "(TNR_HOTRG_3D: tnr_create_3d)");
        if (zn < 1 || zn > MAXN) return fail(TNR_ERR_ARG, "zn must be in 1..8");
        if (precision != TNR_PREC_FP64 && precision != TNR_PREC_SPLIT32)
            return fail(TNR_ERR_ARG, "unknown precision");
        if (!std::isfinite(k)) return fail(TNR_ERR_ARG, "k must be finite");
        std::unique_ptr<tnr_scheme> h(new tnr_scheme());
        h->ctx = take_context(precision);
        BT T = upload(*h->ctx, zn, leg_dims, data, ndata);
        scheme_init(*h->ctx, h->s, kind, std::move(T), k);
        *out = h.release();
        return TNR_OK;
    });
}

int tnr_create_3d(tnr_scheme** out, int zn, const int32_t* leg_dims, const double* data,
                  size_t ndata, int precision) {
    return guarded([&] {
        if (!out) return fail(TNR_ERR_ARG, "out is null");
        *out = nullptr;
        if (zn < 1 || zn > MAXN) return fail(TNR_ERR_ARG, "zn must be in 1..8");
        if (precision != TNR_PREC_FP64 && precision != TNR_PREC_SPLIT32)
            return fail(TNR_ERR_ARG, "unknown precision");
        std::unique_ptr<tnr_scheme> h(new tnr_scheme());
        h->ctx = take_context(precision);
        BT T = upload(*h->ctx, zn, leg_dims, data, ndata, 6, 2);
        scheme_init(*h->ctx, h->s, TNR_HOTRG_3D, std::move(T), 0.0);
        *out = h.release();
        return TNR_OK;
    });
}

int tnr_get_nlegs(const tnr_scheme* s, int* nleg, int* ncod) {
    return guarded([&] {
        if (!s) return fail(TNR_ERR_ARG, "null handle");
        if (nleg) *nleg = s->s.T.L.nleg;
        if (ncod) *ncod = s->s.T.L.ncod;
        return TNR_OK;
    });
}

int tnr_create_multi(tnr_scheme** out, int kind, int zn, const int32_t* leg_dims,
                     const double* const* data, const size_t* ndata, int ntensors,
                     int precision) {
    return guarded([&] {
        if (!out) return fail(TNR_ERR_ARG, "out is null");
        *out = nullptr;
        if (kind != TNR_IMPURITY_TRG && kind != TNR_IMPURITY_HOTRG)
            return fail(TNR_ERR_ARG, "kind must be TNR_IMPURITY_TRG or TNR_IMPURITY_HOTRG");
        if (zn < 1 || zn > MAXN) return fail(TNR_ERR_ARG, "zn must be in 1..8");
        if (precision != TNR_PREC_FP64 && precision != TNR_PREC_SPLIT32)
            return fail(TNR_ERR_ARG, "unknown precision");
        if (ntensors != scheme_tensors(kind))
            return fail(TNR_ERR_ARG, "ImpurityTRG takes 5 tensors, ImpurityHOTRG 4");
        if (!data || !ndata) return fail(TNR_ERR_ARG, "null tensor list");
        std::unique_ptr<tnr_scheme> h(new tnr_scheme());
        h->ctx = take_context(precision);
        std::vector<BT> Ts;
        for (int i = 0; i < ntensors; ++i)
            Ts.push_back(upload(*h->ctx, zn, leg_dims, data[i], ndata[i]));
        scheme_init_multi(*h->ctx, h->s, kind, std::move(Ts));
        *out = h.release();
        return TNR_OK;
    });
}

int tnr_ncomponents(int kind) {
    const int c = scheme_components(kind);
    if (!c) {
        g_last_error = "unknown scheme kind";
        return TNR_ERR_ARG;
    }
    return c;
}

int tnr_finalize_multi(tnr_scheme* s, double* out) {
    return guarded([&] {
        if (!s || !out) return fail(TNR_ERR_ARG, "null pointer");
        check_status(s->s);
        finalize_host(*s->ctx, s->s, out);
        check_norms(out, 1);
        return TNR_OK;
    });
}

int tnr_get_tensor_at(tnr_scheme* s, int index, double* data, size_t ndata) {
    return guarded([&] {
        if (!s || !data) return fail(TNR_ERR_ARG, "null pointer");
        if (index < 0 || index > (int)s->s.imp.size()) return fail(TNR_ERR_ARG, "no such tensor");
        const BT& t = index == 0 ? s->s.T : s->s.imp[index - 1];
        if ((int64_t)ndata != t.L.size) return fail(TNR_ERR_ARG, "ndata does not match");
        download(*s->ctx, t, data);
        return TNR_OK;
    });
}

// ---- observables on the scheme's tensor
namespace {
int write_complex(const std::vector<std::complex<double>>& v, double* re, double* im, size_t cap,
                  size_t* count) {
    if (count) *count = v.size();
    if (v.size() > cap)
        return fail(TNR_ERR_ARG, ("output holds " + std::to_string(cap) + " values, " +
                                  std::to_string(v.size()) + " needed").c_str());
    for (size_t i = 0; i < v.size(); ++i) {
        re[i] = v[i].real();
        if (im) im[i] = v[i].imag();
    }
    return TNR_OK;
}
}  // namespace

int tnr_transfer_spectrum(tnr_scheme* s, int unitcell, double* re, double* im, size_t cap,
                          size_t* count) {
    return guarded([&] {
        if (!s || !re || !im || !count) return fail(TNR_ERR_ARG, "null pointer");
        return write_complex(transfer_spectrum(*s->ctx, s->s, unitcell), re, im, cap, count);
    });
}

int tnr_cft_data(tnr_scheme* s, int unitcell, double v, int is_real, double* re, double* im,
                 size_t cap, size_t* count) {
    return guarded([&] {
        if (!s || !re || !im || !count) return fail(TNR_ERR_ARG, "null pointer");
        if (!(v > 0.0)) return fail(TNR_ERR_ARG, "v must be positive");
        return write_complex(cft_data(*s->ctx, s->s, unitcell, v, is_real != 0), re, im, cap,
                             count);
    });
}

int tnr_ground_state_degeneracy(tnr_scheme* s, int unitcell, double* gsd) {
    return guarded([&] {
        if (!s || !gsd) return fail(TNR_ERR_ARG, "null pointer");
        *gsd = ground_state_degeneracy(*s->ctx, s->s, unitcell);
        return TNR_OK;
    });
}

int tnr_gu_wen_ratio(tnr_scheme* s, double* x1, double* x2) {
    return guarded([&] {
        if (!s || !x1 || !x2) return fail(TNR_ERR_ARG, "null pointer");
        gu_wen_ratio(*s->ctx, s->s, x1, x2);
        return TNR_OK;
    });
}

int tnr_central_charge(tnr_scheme* s, double n, double* c) {
    return guarded([&] {
        if (!s || !c) return fail(TNR_ERR_ARG, "null pointer");
        if (!(n > 0.0)) return fail(TNR_ERR_ARG, "n must be positive");
        *c = central_charge(*s->ctx, s->s, n);
        return TNR_OK;
    });
}

int tnr_finalize_two_by_two(tnr_scheme* s, double* out) {
    return guarded([&] {
        if (!s || !out) return fail(TNR_ERR_ARG, "null pointer");
        if (!single_tensor_kind(s->s.kind) || s->s.kind == TNR_HOTRG_3D)
            return fail(TNR_ERR_STATE,
                        "finalize_two_by_two! is defined for TRG, BTRG, HOTRG and ATRG");
        *out = finalize_two_by_two(*s->ctx, s->s);
        check_norms(out, 1);
        return TNR_OK;
    });
}

int tnr_step(tnr_scheme* s, int chi) {
    return guarded([&] {
        if (!s) return fail(TNR_ERR_ARG, "null handle");
        if (chi < 1) return fail(TNR_ERR_ARG, "chi must be >= 1");
        step_one(*s->ctx, s->s, chi);
        s->ctx->sync();
        return TNR_OK;
    });
}

int tnr_finalize(tnr_scheme* s, double* norm) {
    return guarded([&] {
        if (!s || !norm) return fail(TNR_ERR_ARG, "null pointer");
        if (scheme_components(s->s.kind) != 1)
            return fail(TNR_ERR_STATE, "impurity schemes return several values: tnr_finalize_multi");
        check_status(s->s);
        double* slot = s->slots(1);
        finalize_many(*s->ctx, {&s->s}, {slot});
        s->ctx->fetch(norm, slot, sizeof(double));
        check_norms(norm, 1);
        return TNR_OK;
    });
}

int tnr_run(tnr_scheme* s, int chi, int maxiter, int finalize_beginning, double* norms,
            int* n_out) {
    return guarded([&] {
        if (!s || !norms) return fail(TNR_ERR_ARG, "null pointer");
        if (chi < 1) return fail(TNR_ERR_ARG, "chi must be >= 1");
        const int steps = steps_of(maxiter);
        const int n = steps + (finalize_beginning ? 1 : 0);
        Ctx& ctx = *s->ctx;
        const int nc = scheme_components(s->s.kind);
        if (nc > 1) {   // impurity schemes: each finalize! fetches its values
            int w = 0;
            if (finalize_beginning) finalize_host(ctx, s->s, norms + nc * w++);
            check_status(s->s);
            for (int i = 0; i < steps; ++i) {
                step_one(ctx, s->s, chi);
                finalize_host(ctx, s->s, norms + nc * w++);
            }
            if (n_out) *n_out = n;
            for (int i = 0; i < n; ++i) check_norms(norms + nc * i, 1);
            return TNR_OK;
        }
        double* slot = s->slots(n);
        int w = 0;
        check_status(s->s);
        if (finalize_beginning) finalize_many(ctx, {&s->s}, {slot + w++});
        for (int i = 0; i < steps; ++i) {
            step_one(ctx, s->s, chi);
            finalize_many(ctx, {&s->s}, {slot + w++});
        }
        ctx.fetch(norms, slot, n * sizeof(double));
        if (n_out) *n_out = n;
        check_norms(norms, n);
        return TNR_OK;
    });
}

int tnr_free_energy(const double* norms, int n, double beta, double scalefactor,
                    double initial_size, double* f_out) {
    return guarded([&] {
        if (!norms || !f_out) return fail(TNR_ERR_ARG, "null pointer");
        if (n < 1 || !(scalefactor > 0) || scalefactor == 1.0 || !(initial_size > 0) || beta == 0)
            return fail(TNR_ERR_ARG, "invalid free-energy arguments");
        *f_out = free_energy(norms, n, beta, scalefactor, initial_size);
        return TNR_OK;
    });
}

int tnr_get_dims(const tnr_scheme* s, int32_t* leg_dims, size_t* ndata) {
    return guarded([&] {
        if (!s) return fail(TNR_ERR_ARG, "null handle");
        if (leg_dims) abi_dims(s->s.T.L, leg_dims);
        if (ndata) *ndata = (size_t)s->s.T.L.size;
        return TNR_OK;
    });
}

int tnr_get_tensor(tnr_scheme* s, double* data, size_t ndata) {
    return guarded([&] {
        if (!s || !data) return fail(TNR_ERR_ARG, "null pointer");
        if ((int64_t)ndata != s->s.T.L.size) return fail(TNR_ERR_ARG, "ndata does not match");
        download(*s->ctx, s->s.T, data);
        return TNR_OK;
    });
}

int tnr_set_tensor(tnr_scheme* s, const int32_t* leg_dims, const double* data, size_t ndata) {
    return guarded([&] {
        if (!s) return fail(TNR_ERR_ARG, "null handle");
        BT T = upload(*s->ctx, s->s.N, leg_dims, data, ndata, s->s.T.L.nleg, s->s.T.L.ncod);
        scheme_set_tensor(*s->ctx, s->s, std::move(T));
        return TNR_OK;
    });
}

int tnr_get_bond_weights(tnr_scheme* s, int which, double* w, size_t n) {
    return guarded([&] {
        if (!s || !w) return fail(TNR_ERR_ARG, "null pointer");
        if (s->s.kind != TNR_BTRG) return fail(TNR_ERR_STATE, "only BTRG has bond weights");
        if (which != 1 && which != 2) return fail(TNR_ERR_ARG, "which must be 1 or 2");
        const BondVec& b = which == 1 ? s->s.S1 : s->s.S2;
        if (n != b.host.size()) return fail(TNR_ERR_ARG, "size does not match");
        std::memcpy(w, b.host.data(), n * sizeof(double));
        return TNR_OK;
    });
}

int tnr_get_trunc_log(tnr_scheme* s, double* out, int max_entries, int* n_entries) {
    return guarded([&] {
        if (!s || !n_entries) return fail(TNR_ERR_ARG, "null pointer");
        const int n = (int)s->s.tlog.size();
        *n_entries = n;
        if (out)
            for (int i = 0; i < std::min(n, max_entries); ++i)
                std::memcpy(out + 4 * i, s->s.tlog[i].data(), 4 * sizeof(double));
        return TNR_OK;
    });
}

int tnr_destroy(tnr_scheme* s) {
    return guarded([&] {
        delete s;
        return TNR_OK;
    });
}

const char* tnr_strerror(int status) {
    switch (status) {
        case TNR_OK: return "success";
        case TNR_ERR_ARG: return "invalid argument";
        case TNR_ERR_HIP: return "HIP runtime error";
        case TNR_ERR_NOMEM: return "out of device memory";
        case TNR_ERR_NONFINITE: return "non-finite or zero norm";
        case TNR_ERR_NOCONV: return "solver did not converge";
        case TNR_ERR_STATE: return "invalid state";
        case TNR_ERR_NODEVICE: return "no usable gfx950 device";
        case TNR_ERR_INTERNAL: return "internal error";
        case TNR_ERR_BUDGET: return "request exceeds the HBM budget";
        case TNR_ERR_BUSY: return "result not ready";
        default: return "unknown status";
    }
}

const char* tnr_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"

// ======================================================== serving / batches
namespace {

struct Req {
    int64_t id = 0;
    int kind = 0, zn = 1, chi = 1, maxiter = 1, fb = 1;
    int nleg = 4, ncomp = 1;           // legs of T, values per finalize!
    double k = -0.5;
    std::vector<int32_t> dims;
    std::vector<double> data;
    std::vector<std::vector<double>> imp;   // impurity tensors after T
    size_t est = 0;
    double t_submit = 0, t_done = 0;
    std::vector<double> norms;
    int status = TNR_ERR_BUSY;
};

struct Slot {
    Req* r = nullptr;
    Scheme s;
    Buf norms;                  // single-tensor kinds: one norm per finalize!, on the device
    std::vector<double> hnorms; // impurity kinds: the finalize! rows, on the host
    int nsteps = 0, done = 0, written = 0;
    bool host_finalize() const { return r->ncomp > 1; }
};

}  // namespace

// A server spreads its resident schemes over a few lanes.  Each lane has its
// own stream, memory pool and staging ring and is driven by its own host
// thread during a pump, so one lane's host work (planning, the truncation
// cut between the SVD and the contractions) and its kernels overlap the
// other lanes' kernels.  Every computation is local to one scheme, so the
// lane a scheme lands in changes nothing in its results.
// The batch (tnr_run_batch) the current thread works for, if any: set on the
// calling thread and on its lane threads.
struct tnr_server;
namespace {
thread_local const tnr_server* t_batch_running = nullptr;
}  // namespace

struct Lane {
    Ctx ctx;
    std::vector<std::unique_ptr<Slot>> active;
    double load = 0;   // estimated work of the resident schemes
};

static double work_estimate(const Req& r) {
    const double c = (double)r.chi;
    return (double)steps_of(r.maxiter) * c * c * c * c * c * c / (double)r.zn;
}

static int default_lanes() {
    // four lanes: on alternating runs of the 128-run BTRG chi=32 batch, 3 / 4
    // lanes gave 282-284 / 292-295 free energies per second and serving p99
    // 0.32 / 0.28 s.  With the caller's stream that is five streams on the
    // device's four hardware queues, so two of them share a queue.
    const char* e = std::getenv("TNR_LANES");
    int n = e ? std::atoi(e) : 4;
    return std::max(1, std::min(n, 4));
}

struct tnr_server {
    std::vector<std::unique_ptr<Lane>> lanes;
    size_t budget = 0, in_use = 0;
    int max_slots = 1;
    int64_t next_id = 1;
    std::map<int64_t, std::unique_ptr<Req>> reqs;
    std::deque<Req*> queue;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    explicit tnr_server(int nlanes, int precision) {
        for (int i = 0; i < nlanes; ++i) {
            lanes.emplace_back(new Lane());
            lanes.back()->ctx.precision = precision;
        }
    }

    int resident() const {
        int n = 0;
        for (auto& l : lanes) n += (int)l->active.size();
        return n;
    }

    size_t pool_in_use() const {
        size_t n = 0;
        for (auto& l : lanes) n += l->ctx.pool.in_use();
        return n;
    }

    size_t pool_cached() const {
        size_t n = 0;
        for (auto& l : lanes) n += l->ctx.pool.cached();
        return n;
    }

    void admit() {
        std::vector<std::vector<Slot*>> fresh(lanes.size());
        int res = resident();
        while (!queue.empty() && res < max_slots) {
            Req* r = queue.front();
            if (res > 0 && in_use + r->est > budget) break;   // wait for memory
            queue.pop_front();
            size_t li = 0;
            for (size_t i = 1; i < lanes.size(); ++i)
                if (lanes[i]->load < lanes[li]->load) li = i;
            Lane& L = *lanes[li];
            std::unique_ptr<Slot> sl(new Slot());
            sl->r = r;
            sl->nsteps = steps_of(r->maxiter);
            const int n = sl->nsteps + (r->fb ? 1 : 0);
            try {
                if (r->ncomp > 1) {
                    std::vector<BT> Ts;
                    Ts.push_back(upload(L.ctx, r->zn, r->dims.data(), r->data.data(),
                                        r->data.size()));
                    for (auto& t : r->imp)
                        Ts.push_back(upload(L.ctx, r->zn, r->dims.data(), t.data(), t.size()));
                    scheme_init_multi(L.ctx, sl->s, r->kind, std::move(Ts));
                    sl->hnorms.reserve((size_t)n * r->ncomp);
                } else {
                    BT T = upload(L.ctx, r->zn, r->dims.data(), r->data.data(), r->data.size(),
                                  r->nleg, 2);
                    scheme_init(L.ctx, sl->s, r->kind, std::move(T), r->k);
                    sl->norms = Buf(&L.ctx.pool, (size_t)n * sizeof(double));
                }
            } catch (const Error& e) {
                r->status = e.code;
                r->t_done = now_s(t0);
                continue;
            }
            in_use += r->est;
            L.load += work_estimate(*r);
            fresh[li].push_back(sl.get());
            L.active.push_back(std::move(sl));
            ++res;
        }
        // finalize at the beginning, grouped over the newly admitted of a lane
        for (size_t li = 0; li < lanes.size(); ++li) {
            std::vector<Slot*> fb;
            for (Slot* sl : fresh[li])
                if (sl->r->fb) fb.push_back(sl);
            finalize_slots(lanes[li]->ctx, fb);
        }
    }

    // finalize! of the given slots: single-tensor kinds grouped on the
    // device, impurity kinds one by one with their rows on the host
    static void finalize_slots(Ctx& ctx, const std::vector<Slot*>& sls) {
        std::vector<Scheme*> g;
        std::vector<double*> slots;
        for (Slot* sl : sls) {
            if (sl->host_finalize()) {
                const size_t o = sl->hnorms.size();
                sl->hnorms.resize(o + sl->r->ncomp);
                finalize_host(ctx, sl->s, sl->hnorms.data() + o);
                sl->written++;
            } else {
                g.push_back(&sl->s);
                slots.push_back(sl->norms.d() + sl->written++);
            }
        }
        if (!g.empty()) finalize_many(ctx, g, slots);
    }

    // one coarse-graining step of every resident scheme of a lane
    static void step_lane(Lane& L) {
        if (L.active.empty()) return;
        std::vector<Scheme*> g;
        std::vector<int> chis;
        for (auto& sl : L.active) {
            g.push_back(&sl->s);
            chis.push_back(sl->r->chi);
        }
        step_many(L.ctx, g, chis);
        std::vector<Slot*> fin;
        for (auto& sl : L.active) {
            sl->done++;
            // a failed step ends the run; the others in the group go on
            if (sl->s.status != TNR_OK) sl->done = sl->nsteps;
            fin.push_back(sl.get());
        }
        finalize_slots(L.ctx, fin);
    }

    void step_all() {
        std::vector<Lane*> busy;
        for (auto& l : lanes)
            if (!l->active.empty()) busy.push_back(l.get());
        if (busy.size() <= 1) {
            if (!busy.empty()) step_lane(*busy[0]);
            return;
        }
        std::vector<std::thread> th;
        std::vector<std::exception_ptr> err(busy.size());
        const tnr_server* running = t_batch_running;
        for (size_t i = 1; i < busy.size(); ++i)
            th.emplace_back([&, i, running] {
                t_batch_running = running;
                try {
                    TNR_HIP(hipSetDevice(busy[i]->ctx.device));
                    step_lane(*busy[i]);
                } catch (...) {
                    err[i] = std::current_exception();
                }
            });
        try {
            step_lane(*busy[0]);
        } catch (...) {
            err[0] = std::current_exception();
        }
        for (auto& t : th) t.join();
        for (auto& e : err)
            if (e) std::rethrow_exception(e);
    }

    void retire() {
        for (auto& lp : lanes) {
            Lane& L = *lp;
            std::vector<std::unique_ptr<Slot>> keep;
            for (auto& sl : L.active) {
                if (sl->done < sl->nsteps) {
                    keep.push_back(std::move(sl));
                    continue;
                }
                Req* r = sl->r;
                if (sl->host_finalize()) {
                    r->norms = sl->hnorms;
                } else {
                    r->norms.resize(sl->written);
                    L.ctx.fetch(r->norms.data(), sl->norms.d(), sl->written * sizeof(double));
                }
                r->status = sl->s.status;
                // the first value of each row is the norm (impurity rows also
                // carry impurity traces, which may vanish)
                if (r->status == TNR_OK)
                    for (size_t i = 0; i < r->norms.size(); i += r->ncomp) {
                        const double v = r->norms[i];
                        if (!std::isfinite(v) || v == 0.0) r->status = TNR_ERR_NONFINITE;
                    }
                r->t_done = now_s(t0);
                in_use -= r->est;
                L.load -= work_estimate(*r);
            }
            L.active.swap(keep);
            if (L.active.empty()) L.load = 0;
        }
    }

    int pump() {
        admit();
        step_all();
        retire();
        return resident() + (int)queue.size();
    }

    void sync_all() {
        for (auto& l : lanes) l->ctx.sync();
    }

    void clear() {
        for (auto& l : lanes) l->active.clear();
    }

    // back to an empty server; streams, pools and staging rings are kept
    void reset(int slots) {
        clear();
        for (auto& l : lanes) l->load = 0;
        reqs.clear();
        queue.clear();
        in_use = 0;
        max_slots = slots;
        t0 = std::chrono::steady_clock::now();
    }
};

// tnr_run_batch keeps one server per (device, precision, lane count) for the
// life of the process.  A fresh server per call would allocate every working
// buffer anew and free it again at the end (both synchronise the device) and
// pin a new staging ring: about half a second of idle device per call at the
// BTRG chi = 32 batch of 128 runs.
// The map is never destroyed: at process exit the HIP runtime may already be
// gone, and the driver reclaims the device memory anyway.
namespace {
std::mutex g_batch_mu;
std::map<std::array<int, 3>, std::unique_ptr<tnr_server>>& batch_servers() {
    static auto* m = new std::map<std::array<int, 3>, std::unique_ptr<tnr_server>>();
    return *m;
}

// With g_batch_mu held by this thread or by the batch it works for; lane
// threads of one batch may get here together, hence the second lock.
std::mutex g_trim_mu;
void trim_idle_batch_servers(const tnr_server* busy) {
    std::lock_guard<std::mutex> lk(g_trim_mu);
    for (auto& kv : batch_servers())
        if (kv.second && kv.second.get() != busy)
            for (auto& l : kv.second->lanes) l->ctx.pool.trim();
}

// Pool fallback on a failed allocation: idle spare contexts and batch engines
// give their cached buffers back.  A batch running on this thread keeps its
// own (its pools are in use); one running on another thread is waited for.
void release_idle_memory() {
    {
        std::lock_guard<std::mutex> lk(g_ctx_mu);
        for (auto& c : spare_contexts()) c->pool.trim();
    }
    if (t_batch_running) {
        trim_idle_batch_servers(t_batch_running);   // its batch holds g_batch_mu
    } else {
        std::lock_guard<std::mutex> lk(g_batch_mu);
        trim_idle_batch_servers(nullptr);
    }
}

struct ReleaseHook {
    ReleaseHook() { g_release_idle_memory = &release_idle_memory; }
} g_release_hook;
}  // namespace

extern "C" {

int tnr_server_create(tnr_server** out, size_t hbm_budget_bytes, int max_slots, int precision) {
    return guarded([&] {
        if (!out) return fail(TNR_ERR_ARG, "out is null");
        *out = nullptr;
        if (max_slots < 1) return fail(TNR_ERR_ARG, "max_slots must be >= 1");
        if (precision != TNR_PREC_FP64 && precision != TNR_PREC_SPLIT32)
            return fail(TNR_ERR_ARG, "unknown precision");
        std::unique_ptr<tnr_server> s(new tnr_server(default_lanes(), precision));
        size_t fr = 0, tot = 0;
        TNR_HIP(hipMemGetInfo(&fr, &tot));
        s->budget = hbm_budget_bytes ? hbm_budget_bytes : (size_t)(0.8 * (double)fr);
        s->max_slots = max_slots;
        *out = s.release();
        return TNR_OK;
    });
}

int tnr_server_submit(tnr_server* s, const tnr_run_desc* req, int64_t* id) {
    return guarded([&] {
        if (!s || !req || !id) return fail(TNR_ERR_ARG, "null pointer");
        const int ntens = scheme_tensors(req->kind);
        if (!ntens) return fail(TNR_ERR_ARG, "unknown scheme kind");
        if (req->zn < 1 || req->zn > MAXN) return fail(TNR_ERR_ARG, "zn must be in 1..8");
        if (req->chi < 1) return fail(TNR_ERR_ARG, "chi must be >= 1");
        if (!req->leg_dims || !req->data) return fail(TNR_ERR_ARG, "null tensor");
        if (ntens > 1 && !req->impurity) return fail(TNR_ERR_ARG, "null impurity tensor list");
        const int nleg = req->kind == TNR_HOTRG_3D ? 6 : 4;
        Layout L = abi_layout(req->zn, req->leg_dims, nleg, 2);
        if ((int64_t)req->ndata != L.size || L.size == 0)
            return fail(TNR_ERR_ARG, "ndata does not match the sector dimensions");
        std::vector<const double*> tens{req->data};
        for (int i = 1; i < ntens; ++i) {
            if (!req->impurity[i - 1]) return fail(TNR_ERR_ARG, "null impurity tensor");
            tens.push_back(req->impurity[i - 1]);
        }
        for (const double* t : tens)
            for (size_t i = 0; i < req->ndata; ++i)
                if (!std::isfinite(t[i]))
                    return fail(TNR_ERR_NONFINITE, "the tensor has non-finite entries");
        std::unique_ptr<Req> r(new Req());
        r->id = s->next_id++;
        r->kind = req->kind; r->zn = req->zn; r->chi = req->chi;
        r->maxiter = req->maxiter; r->fb = req->finalize_beginning ? 1 : 0; r->k = req->k;
        r->nleg = nleg;
        r->ncomp = scheme_components(req->kind);
        r->dims.assign(req->leg_dims, req->leg_dims + nleg * req->zn);
        r->data.assign(req->data, req->data + req->ndata);
        for (int i = 1; i < ntens; ++i) r->imp.emplace_back(tens[i], tens[i] + req->ndata);
        r->est = scheme_hbm_estimate(r->kind, r->zn, r->chi);
        if (r->est > s->budget) return fail(TNR_ERR_BUDGET, "request exceeds the HBM budget");
        r->t_submit = now_s(s->t0);
        *id = r->id;
        s->queue.push_back(r.get());
        s->reqs[r->id] = std::move(r);
        return TNR_OK;
    });
}

int tnr_server_pump(tnr_server* s, int* active) {
    return guarded([&] {
        if (!s) return fail(TNR_ERR_ARG, "null handle");
        int a = s->pump();
        if (active) *active = a;
        return TNR_OK;
    });
}

int tnr_server_poll(tnr_server* s, int64_t id, double* norms, int* n_out, double* t_submit,
                    double* t_done) {
    return guarded([&] {
        if (!s) return fail(TNR_ERR_ARG, "null handle");
        auto it = s->reqs.find(id);
        if (it == s->reqs.end()) return fail(TNR_ERR_ARG, "unknown request id");
        Req& r = *it->second;
        if (r.status == TNR_ERR_BUSY) return TNR_ERR_BUSY;
        if (t_submit) *t_submit = r.t_submit;
        if (t_done) *t_done = r.t_done;
        if (n_out) *n_out = (int)(r.norms.size() / r.ncomp);
        if (norms && !r.norms.empty()) std::memcpy(norms, r.norms.data(), r.norms.size() * sizeof(double));
        return r.status;
    });
}

int tnr_server_stats(tnr_server* s, size_t* hbm_in_use, int* resident, int* queued) {
    return guarded([&] {
        if (!s) return fail(TNR_ERR_ARG, "null handle");
        if (hbm_in_use) *hbm_in_use = s->pool_in_use();
        if (resident) *resident = s->resident();
        if (queued) *queued = (int)s->queue.size();
        return TNR_OK;
    });
}

int tnr_server_destroy(tnr_server* s) {
    return guarded([&] {
        if (s) {
            s->sync_all();
            s->clear();
        }
        delete s;
        return TNR_OK;
    });
}

int tnr_run_batch(int nruns, const tnr_run_desc* runs, int precision, double* norms) {
    std::lock_guard<std::mutex> lk(g_batch_mu);
    std::array<int, 3> key{-1, precision, 0};
    const int status = guarded([&] {
        if (nruns < 0 || (nruns > 0 && (!runs || !norms))) return fail(TNR_ERR_ARG, "null pointer");
        if (precision != TNR_PREC_FP64 && precision != TNR_PREC_SPLIT32)
            return fail(TNR_ERR_ARG, "unknown precision");
        ensure_device();
        TNR_HIP(hipGetDevice(&key[0]));
        key[2] = default_lanes();
        std::unique_ptr<tnr_server>& held = batch_servers()[key];
        if (!held) {
            tnr_server* s = nullptr;
            int st = tnr_server_create(&s, 0, 1, precision);
            if (st != TNR_OK) return st;
            held.reset(s);
        }
        tnr_server* s = held.get();
        t_batch_running = s;
        struct Clear {
            ~Clear() { t_batch_running = nullptr; }
        } clear_running;
        s->reset(nruns > 0 ? nruns : 1);
        {
            // the budget follows what is free now: memory that handles, spare
            // contexts or other users took or gave back since the last batch,
            // plus the engine's own cached buffers, which it reuses
            size_t fr = 0, tot = 0;
            TNR_HIP(hipMemGetInfo(&fr, &tot));
            s->budget = (size_t)(0.8 * (double)(fr + s->pool_cached()));
        }
        std::vector<int64_t> ids(nruns);
        for (int i = 0; i < nruns; ++i) {
            int st = tnr_server_submit(s, &runs[i], &ids[i]);
            if (st != TNR_OK) return st;
        }
        int active = 1;
        while (active > 0) active = s->pump();
        size_t o = 0;
        for (int i = 0; i < nruns; ++i) {
            Req& r = *s->reqs[ids[i]];
            if (r.status != TNR_OK) {
                g_last_error = "run " + std::to_string(i) + " failed: " + tnr_strerror(r.status);
                return r.status;
            }
            std::memcpy(norms + o, r.norms.data(), r.norms.size() * sizeof(double));
            o += r.norms.size();
        }
        s->reset(1);
        return TNR_OK;
    });
    if (status != TNR_OK && key[0] >= 0) {
        // a failed batch may leave work or state behind: start the next one afresh
        auto it = batch_servers().find(key);
        if (it != batch_servers().end()) {
            try {
                if (it->second) it->second->sync_all();
            } catch (...) {
            }
            batch_servers().erase(it);
        }
    }
    return status;
}

}  // extern "C"
