// TRG / BTRG / HOTRG coarse-graining steps and finalize on the device.
//
// Every operation is written for a group of independent schemes: all their
// symmetry blocks go into the same grouped launches.  A single run is a
// group of one, so a run's arithmetic is the same alone and in a batch.
#pragma once

#include <array>
#include <vector>

#include "blocktensor.h"
#include "solver.h"

namespace tnr {

struct Scheme {
    int kind = TNR_BTRG;
    int N = 1;
    double k = -0.5;           // BTRG bond-weight exponent
    BT T;                      // current tensor, (12|34) layout
    BondVec S1, S2;            // BTRG bond weights (vertical, horizontal)
    std::vector<BT> imp;       // ImpurityTRG: T_imp1..4; ImpurityHOTRG: T_imp_order1_1, _1_2, _2
    int steps = 0;
    // TNR_OK, or the status of a failed step (TNR_ERR_NOCONV: a truncated
    // SVD / eigensolve stopped unconverged; TNR_ERR_NONFINITE: it met NaN or
    // infinite values); sticky until the tensor is replaced
    int status = TNR_OK;
    // per factorization of the last step: smallest kept value, largest
    // dropped value, solver iterations, convergence bound
    std::vector<std::array<double, 4>> tlog;
};

void scheme_init(Ctx& ctx, Scheme& s, int kind, BT&& T, double k);
// impurity schemes: Ts = {T, impurity tensors...}, all on T's spaces
void scheme_init_multi(Ctx& ctx, Scheme& s, int kind, std::vector<BT>&& Ts);
// TRG, BTRG, HOTRG, ATRG: one tensor, one norm per finalize!
bool single_tensor_kind(int kind);
int scheme_tensors(int kind);      // tensors a scheme holds (0: unknown kind)
int scheme_components(int kind);   // values one finalize! returns
// finalize! of one scheme with its values on the host (scheme_components)
void finalize_host(Ctx& ctx, Scheme& s, double* out);
void scheme_set_tensor(Ctx& ctx, Scheme& s, BT&& T);

// Split precision (TNR_PREC_SPLIT32) starts at this step of a run: the
// free energy weighs step i's log-norm by scalefactor^-i, so the first steps
// dominate it; run in FP64 they keep |df/f| within 1e-9 of the FP64 path
// (the FP32 hi/lo products are good to ~1e-8 per step).
constexpr int kSplitFromStep = 8;

// step!(scheme, truncrank(chi)) for every scheme of the group
void step_many(Ctx& ctx, const std::vector<Scheme*>& group, const std::vector<int>& chis);

// finalize!(scheme): the norm of scheme i goes to the device slot slots[i]
void finalize_many(Ctx& ctx, const std::vector<Scheme*>& group, const std::vector<double*>& slots);

// free_energy(data, beta; scalefactor, initial_size)
double free_energy(const double* norms, int n, double beta, double scalefactor,
                   double initial_size);

// Pivoted-row gauge of a truncated SVD (used by the impurity TRG): for k
// kept singular values sv (descending) appends the cluster bounds
// 0 = b_0 < ... < b_nc = knn to cl and returns nc; knn = values above the
// null threshold.  kGaugeTie is k_canon's relative pivot tie.
int gauge_clusters(const double* sv, int k, std::vector<int>& cl, int& knn);
constexpr double kGaugeTie = 1e-10;

// HBM estimate (bytes) of a run at bond dimension chi, for admission
size_t scheme_hbm_estimate(int kind, int N, int chi);

}  // namespace tnr
