// Observables on the (fixed-point) tensor of a scheme: the reference's
// src/utility/cft.jl (cft_data, central_charge, ground_state_degeneracy,
// gu_wen_ratio) and the 2x2 finalizer of src/utility/finalize.jl:17-41.
// Contractions, traces, the SVD and the eigenvalues run on the device; the
// scalar epilogues (sorting, filtering, logarithms) on the host.
#pragma once

#include <complex>
#include <vector>

#include "schemes.h"

namespace tnr {

// All eigenvalues of the unitcell-site transfer matrix (cft.jl:7-16), every
// coupled-charge block in sector order; BTRG absorbs S1 and S2 first
// (cft.jl:44-45).  Sorted by decreasing modulus, ties by decreasing real
// part, then decreasing imaginary part.
std::vector<std::complex<double>> transfer_spectrum(Ctx& ctx, const Scheme& s, int unitcell);

// cft_data(scheme; v, unitcell, is_real) (cft.jl:5-71)
std::vector<std::complex<double>> cft_data(Ctx& ctx, const Scheme& s, int unitcell, double v,
                                           bool is_real);
// ground_state_degeneracy(scheme; unitcell) (cft.jl:278-339)
double ground_state_degeneracy(Ctx& ctx, const Scheme& s, int unitcell);
// gu_wen_ratio(scheme) -> X1, X2 (cft.jl:374-396)
void gu_wen_ratio(Ctx& ctx, const Scheme& s, double* x1, double* x2);
// central_charge(scheme, n) (cft.jl:256-268)
double central_charge(Ctx& ctx, const Scheme& s, double n);
// finalize_two_by_two!(scheme) (finalize.jl:17-41): T /= n^(1/4), returns n^(1/4)
double finalize_two_by_two(Ctx& ctx, Scheme& s);

}  // namespace tnr
