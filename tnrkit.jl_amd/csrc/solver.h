// Grouped truncated SVD and symmetric eigensolver on the device (FP64).
//
// Small problems (min(m, n) <= LMAX) are solved directly: Gram matrix on
// MFMA, parallel-order Jacobi in LDS, a second Gram/Jacobi pass on the
// rotated columns (which are then nearly orthogonal, so the second pass is
// relatively accurate), column norms.  Larger problems run block subspace
// iteration (Y = M V, Z = M^T U) with the same two-pass Gram/Jacobi
// Rayleigh-Ritz step after every product, until the top-k subspace has
// converged to working precision; the Ritz triplets are then extracted from
// Y = M V.  All problems of a call advance together in grouped launches.
#pragma once

#include <vector>

#include "blocktensor.h"

namespace tnr {

struct SvdIn {
    const double* M;     // m x n, column-major, leading dim ld
    int m, n, ld;
    int kmax;            // triplets wanted, <= min(m, n)
    uint64_t seed;
};

struct SvdOut {
    Buf U, V, s;         // U: m x kmax (ld m), V: n x kmax (ld n), s: kmax
    int ldu = 0, ldv = 0;
    std::vector<double> s_host;
    int iters = 0;       // subspace iterations (0 = direct)
    double conv = 0.0;   // accumulated contraction bound of the top-k subspace
    int status = 0;      // TNR_OK, TNR_ERR_NOCONV or TNR_ERR_NONFINITE (see below)
};

void svd_grouped(Ctx& ctx, const std::vector<SvdIn>& in, std::vector<SvdOut>& out);

struct EighIn {
    double* M;           // n x n symmetric (symmetrised in place), leading dim ld
    int n, ld;
    int kmax;
    uint64_t seed;
};

struct EighOut {
    Buf U, lam;          // U: n x ncols (ld n), lam: ncols, by |lambda| descending
    int ncols = 0;       // n (direct) or the block size (iterative)
    std::vector<double> lam_host;
    bool direct = false; // all n eigenvalues available
    int iters = 0;
    double conv = 0.0;
    int status = 0;      // TNR_OK, TNR_ERR_NOCONV or TNR_ERR_NONFINITE (see below)
};

void eigh_grouped(Ctx& ctx, const std::vector<EighIn>& in, std::vector<EighOut>& out);

// Per-problem status: TNR_ERR_NOCONV for a problem whose block iteration
// reaches max_iters with the accumulated contraction bound still above
// conv_target, or any of whose Jacobi solves reaches its sweep cap with a
// rotation still due (the reference's LAPACK-backed svd_trunc and
// eigh_trunc! raise there); TNR_ERR_NONFINITE for a problem with a NaN or
// infinite singular value / eigenvalue, whose host values are then zeroed so
// the global truncation stays well defined.  Other problems of the group are
// unaffected.
//
// Solver knobs (the autotuner may override them).
struct SolverCfg {
    double jacobi_tol = 1e-15;   // relative off-diagonal threshold (final Ritz extraction, direct)
    int jacobi_sweeps = 40;
    // Rayleigh-Ritz inside the block iteration.  Measured on BTRG chi = 32
    // beta scans: looser thresholds or sweep caps here make the iteration
    // and the final extraction take longer than they save, so the loop
    // keeps a tight threshold; V is still rebuilt at full accuracy (1e-15)
    // once the loop ends.
    double jacobi_tol_iter = 1e-10;
    // The first iteration's Rayleigh-Ritz step only aligns the random start
    // basis; later steps refine it.  Measured on single BTRG chi = 32 runs
    // (beta 0.44): 1e-4 takes 3 % less time than 1e-10, 1e-6 .. 1e-2 2 %.
    double jacobi_tol_first = 1e-4;
    int jacobi_sweeps_iter = 40;
    double null_rtol = 1e-14;    // columns below this x sigma_max are null
    double conv_target = 1e-14;  // accumulated contraction of the top-k subspace angle
    double block_mult = 2.0;     // block width max(block_mult k, k + block_extra) <= LMAX
    // Block iteration orthonormalisation by pivoted Cholesky QR instead of
    // Rayleigh-Ritz: the Y side (cholqr_y), the Z side (cholqr_z; the
    // stopping rule then reads the pivots R_jj) and the eigensolver's loop
    // (cholqr_eigh).  Without Rayleigh-Ritz in the loop the iteration is a
    // simultaneous iteration: its basis vectors converge to the singular
    // vectors one by one, so the Gram matrices of the final extraction are
    // nearly diagonal and cheap for Jacobi.  Measured on single BTRG
    // chi = 32 runs: 17 % less time with the Z side on Cholesky QR.  The
    // eigensolver keeps Rayleigh-Ritz: with Cholesky QR its stopping rule on
    // the pivots ended HOTRG chi = 48 solves early (1e-7 norm errors).
    bool cholqr_y = true;
    bool cholqr_z = true;
    bool cholqr_eigh = false;
    // Rayleigh-Ritz passes that rebuild V from the last Z (with Cholesky QR
    // on the Z side) and that extract the Ritz triplets from Y = M V.  V is
    // a Ritz basis of the converged subspace, so the columns of M V are
    // already orthogonal to roundoff relative to their own norms: one pass
    // has the relative accuracy a second pass gives (measured: the parity
    // and LAPACK tests unchanged, single BTRG chi = 32 runs 4 % faster).  The
    // last Z is not orthogonal enough for one pass (1 fails parity).
    int rv_passes = 2;
    int rf_passes = 1;
    int block_extra = 2;
    int min_iters = 2;
    int max_iters = 80;
    // from this iteration on, a block iteration of a matrix small enough to
    // be refactorised whole stops as soon as even twice its current
    // convergence rate could not reach conv_target by max_iters (it would
    // stall and be factorised whole anyway); 0 iterates to max_iters
    int stall_from = 8;
};
SolverCfg& solver_cfg();
// Defaults, then TNR_SOLVER_CFG, then `spec` ("key=value,...", see
// solver.cpp); not while other calls are running.
void set_solver_cfg(const char* spec);

}  // namespace tnr
