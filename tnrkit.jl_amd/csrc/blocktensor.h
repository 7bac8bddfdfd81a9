// Block-sparse tensors with abelian Z_N symmetry on the device.
//
// A tensor has `nleg` legs split into a codomain (first `ncod` legs) and a
// domain.  Each leg has a dimension per *outgoing* charge q in Z_N (a
// codomain leg carrying sector a has q = a, a domain leg carrying sector a
// has q = -a).  Data is stored as one column-major matrix per coupled charge
// c: rows enumerate codomain sector tuples (labels l = q) with sum = c in
// colexicographic order (first leg fastest), columns enumerate domain sector
// tuples (labels l = -q) with sum = c likewise; inside a tuple the element
// order is column-major over the tuple's legs.  N = 1 (Trivial) is the
// plain column-major dense array.  This is TensorKit's block layout for
// abelian sectors with trivial F/R symbols (ZNIrrep, Z2Irrep), for which
// permute/transpose only move data.
#pragma once

#include <array>
#include <memory>
#include <cstring>
#include <utility>
#include <vector>

#include "common.h"
#include "kernels.h"

namespace tnr {

struct Leg {
    int d[MAXN] = {0};              // dimension per outgoing charge q
    int total(int N) const {
        int t = 0;
        for (int q = 0; q < N; ++q) t += d[q];
        return t;
    }
    int offset(int q) const {       // start of charge-q indices in leg order
        int t = 0;
        for (int i = 0; i < q; ++i) t += d[i];
        return t;
    }
    Leg dual(int N) const {         // conj/dual: q -> -q
        Leg o;
        for (int q = 0; q < N; ++q) o.d[(N - q) % N] = d[q];
        return o;
    }
    bool operator==(const Leg& o) const {
        for (int q = 0; q < MAXN; ++q)
            if (d[q] != o.d[q]) return false;
        return true;
    }
    bool operator!=(const Leg& o) const { return !(*this == o); }
};

struct Layout {
    int N = 1, nleg = 0, ncod = 0;
    std::array<Leg, MAXLEG> legs{};
    // derived
    std::array<int64_t, MAXN> R{}, C{}, off{};
    int64_t size = 0;
    std::vector<int64_t> roff, coff;   // tuple offsets (index = colex tuple id of labels)

    Layout() = default;
    Layout(int N_, int ncod_, const std::vector<Leg>& legs_);
    void build();
    int label(int leg, int q) const { return leg < ncod ? q : (N - q) % N; }
    int ndom() const { return nleg - ncod; }
    bool same(const Layout& o) const;
    std::string str() const;
};

// A device tensor: layout + data in a pooled buffer.
struct BT {
    Layout L;
    Buf buf;
    double* d() const { return buf.d(); }
    BT() = default;
    BT(Ctx& ctx, const Layout& L_);
};

// Diagonal weights on a leg: w[off[q] + i] multiplies index i of the sector
// with outgoing charge q.
struct LegWeights {
    const double* w = nullptr;
    int off[MAXN] = {0};
};

// A bond vector (singular values, BTRG bond matrix diagonal) stored per
// sector label c of the side where the bond leg is outgoing with q = c.
struct BondVec {
    std::shared_ptr<Buf> store;      // device values (possibly shared by many bonds)
    size_t base = 0;                 // first element of this bond in the store
    std::vector<double> host;        // same values on the host
    const double* ptr() const { return store->d() + base; }
    int N = 1;
    int dim[MAXN] = {0};
    int off[MAXN] = {0};
    // weights for a leg that carries this bond with outgoing q = +c (same
    // side) or q = -c (dual side)
    LegWeights same_side() const;
    LegWeights dual_side() const;
    Leg leg_same() const;            // Leg with d[q=c] = dim[c]
    Leg leg_dual() const;            // Leg with d[q=-c] = dim[c]
};

// ---------------------------------------------------------------- operations
// dst legs j = src legs perm[j]; dst codomain = first ncod legs; conj negates
// all charges.  Optional weights multiply along given *src* legs.  All jobs
// of one call are issued as a single grouped launch.
struct WeightSpec {
    int leg = -1;                    // src leg
    LegWeights w;
};
struct PermJob {
    const BT* src = nullptr;
    std::vector<int> perm;
    int ncod = 0;
    bool conj = false;
    std::vector<WeightSpec> ws;
};
std::vector<BT> permute_many(Ctx& ctx, const std::vector<PermJob>& jobs);
BT permute(Ctx& ctx, const BT& src, const std::vector<int>& perm, int ncod,
           bool conj = false, const std::vector<WeightSpec>& ws = {});

// Tensor contraction by labels (negative = open; positive labels appear once
// in each operand and are summed).  The result legs follow out_labels and
// its codomain is the first `out_ncod` of them.  A batch of contractions is
// issued as one grouped permute launch, one grouped GEMM launch and one
// grouped output-permute launch.
struct Operand {
    const BT* t = nullptr;
    std::vector<int> labels;
    bool conj = false;
    std::vector<WeightSpec> ws;      // optional diagonal weights on legs
};
struct ContractJob {
    Operand a, b;
    std::vector<int> out;
    int out_ncod = 0;
};
std::vector<BT> contract_many(Ctx& ctx, const std::vector<ContractJob>& jobs);

// Grouped GEMM issue: C_i = alpha op(A_i) op(B_i) + beta C_i.
struct GemmJob {
    const double* A;
    const double* B;
    double* C;
    int M, N, K, lda, ldb, ldc, ta, tb;
    double alpha = 1.0, beta = 0.0;
    const int* skip = nullptr;       // device flag: problem finished
};
// allow_dma = false keeps the 16-deep tile shapes on register-staged loads
void gemm_grouped(Ctx& ctx, const std::vector<GemmJob>& jobs, bool split32 = false,
                  int force_shape = -1, int force_split = -1, bool allow_dma = true);

// Descriptor arrays gathered on the host and uploaded to the device in one
// copy (through the staging ring); every launch queued before the ring's
// next wrap-around may read them, any number of times.
class DescArena {
  public:
    template <class T> size_t add(const std::vector<T>& v) {
        size_t off = (host_.size() + 255) & ~size_t(255);
        host_.resize(off + v.size() * sizeof(T));
        if (!v.empty()) std::memcpy(host_.data() + off, v.data(), v.size() * sizeof(T));
        return off;
    }
    void commit(Ctx& ctx);
    template <class T> T* at(size_t off) const { return reinterpret_cast<T*>(base_ + off); }

  private:
    std::vector<char> host_;
    char* base_ = nullptr;
};

// A grouped GEMM prepared once and issued many times.
struct GemmPlan {
    std::vector<GemmDesc> descs, red;
    std::vector<int64_t> red_starts;
    int64_t tiles = 0, red_total = 0;
    double flops = 0;
    bool split32 = false;
    int shape = 0;                     // GemmShape of the launch
    bool dma = true;                   // every descriptor fits the LDS-DMA kernels
    bool trial = false;                // next issue is timed for the autotuner
    std::array<int, 9> tune_key{};
    Buf work;
    size_t off_descs = 0, off_red = 0, off_starts = 0;
    const GemmDesc* d_descs = nullptr;
    const GemmDesc* d_red = nullptr;
    const int64_t* d_starts = nullptr;
};
// force_shape >= 0 fixes the FP64 tile shape (GemmShape) instead of choosing
// it; force_split 0 / 1 turns splitting K over workgroups off / on instead
GemmPlan gemm_plan(Ctx& ctx, const std::vector<GemmJob>& jobs, bool split32 = false,
                   int force_shape = -1, int force_split = -1);
void gemm_stage(GemmPlan& p, DescArena& arena);          // before arena.commit
void gemm_bind(GemmPlan& p, const DescArena& arena);     // after arena.commit
void gemm_issue(Ctx& ctx, GemmPlan& p);
// on = 0 freezes tile shapes to the static rule; any call clears what the
// autotuner has learned
void gemm_autotune_enable(int on);

// n = |sum T[x1, x2, x2, x1] w1[x1] w2[x2]| (legs 1<->4, 2<->3), written to
// the device slot; then T /= n, all on the device; grouped over tensors.
struct TraceJob {
    BT* t = nullptr;
    const LegWeights* w_x1 = nullptr;
    const LegWeights* w_x2 = nullptr;
    double* norm_slot = nullptr;     // device
    const double* divide_by = nullptr;   // device: divide by this instead of the own norm
};

// out = sum_i c_i * t_i over tensors of one layout (out may be one of them)
struct LinJob {
    BT* out = nullptr;
    std::vector<std::pair<double, const BT*>> terms;   // at most four
};
void lincomb_many(Ctx& ctx, const std::vector<LinJob>& jobs);
void trace_normalize_many(Ctx& ctx, const std::vector<TraceJob>& jobs);

// Grouped strided matrix copies with optional row/column weights.
void gather_many(Ctx& ctx, std::vector<GatherDesc> descs);
// cnt[p] doubles at device src[p] into *dst[p]: packed on the device, one
// copy to the host, one synchronisation
void readback_vectors(Ctx& ctx, const std::vector<const double*>& src,
                      const std::vector<int>& cnt, const std::vector<std::vector<double>*>& dst);

// Upload / download in the ABI format (sector dims per space label, first
// ncod legs in the codomain; 2D tensors have 4 legs, 3D tensors 6).
Layout abi_layout(int N, const int32_t* leg_dims, int nleg = 4, int ncod = 2);
BT upload(Ctx& ctx, int N, const int32_t* leg_dims, const double* data, size_t n, int nleg = 4,
          int ncod = 2);
void download(Ctx& ctx, const BT& t, double* data);
void abi_dims(const Layout& L, int32_t* leg_dims);     // nleg * N values

}  // namespace tnr
