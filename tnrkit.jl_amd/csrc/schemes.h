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
    int steps = 0;
    // per factorization of the last step: smallest kept value, largest
    // dropped value, solver iterations, convergence bound
    std::vector<std::array<double, 4>> tlog;
};

void scheme_init(Ctx& ctx, Scheme& s, int kind, BT&& T, double k);
void scheme_set_tensor(Ctx& ctx, Scheme& s, BT&& T);

// step!(scheme, truncrank(chi)) for every scheme of the group
void step_many(Ctx& ctx, const std::vector<Scheme*>& group, const std::vector<int>& chis);

// finalize!(scheme): the norm of scheme i goes to the device slot slots[i]
void finalize_many(Ctx& ctx, const std::vector<Scheme*>& group, const std::vector<double*>& slots);

// free_energy(data, beta; scalefactor, initial_size)
double free_energy(const double* norms, int n, double beta, double scalefactor,
                   double initial_size);

// HBM estimate (bytes) of a run at bond dimension chi, for admission
size_t scheme_hbm_estimate(int kind, int N, int chi);

}  // namespace tnr
