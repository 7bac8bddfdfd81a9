/*
 * tnr.h - C ABI of libtnr_hip.so, the MI355X (gfx950) engine for the
 * TRG / BTRG / HOTRG coarse-graining loop of TNRKit.jl v0.5.1.
 *
 * Every entry point returns TNR_OK (0) or a negative status; no C++
 * exception crosses this boundary.  Host pointers are borrowed for the
 * duration of the call only.  A handle is not thread-safe; distinct handles
 * may be driven from different threads (each owns its own HIP stream).
 *
 * Tensor data format.  A scheme tensor T lives in V1 (x) V2 <- V3 (x) V4 with
 * legs (left, bottom, top, right) (docs/src/index.md:107-119).  With
 * Z_N symmetry (zn = N >= 2; zn = 1 is Trivial) every space is a direct sum
 * of sectors a = 0..N-1 and
 *     leg_dims[i*zn + a] = dimension of sector a in V_{i+1},  i = 0..3
 * (V3, V4 as written in the domain).  `data` is the concatenation, for
 * coupled charge c = 0..N-1, of the column-major block matrix of T at c.
 * Block rows run over codomain sector pairs (a1, a2) with a1 + a2 = c mod N
 * in colexicographic order (a1 fastest), each pair contributing
 * d1(a1)*d2(a2) rows (index of leg 1 fastest); columns likewise over
 * (a3, a4) with a3 + a4 = c.  For zn = 1 this is exactly the column-major
 * dense array T[i1, i2, i3, i4] (TensorKit's `t.data` for a Trivial map).
 */
#ifndef TNR_H
#define TNR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define TNR_OK              0
#define TNR_ERR_ARG        -1  /* invalid argument (dims, chi, sizes, kind) */
#define TNR_ERR_HIP        -2  /* HIP runtime error */
#define TNR_ERR_NOMEM      -3  /* device allocation failed */
#define TNR_ERR_NONFINITE  -4  /* a norm came out NaN/Inf/zero */
#define TNR_ERR_NOCONV     -5  /* an eigen/singular solver did not converge */
#define TNR_ERR_STATE      -6  /* call not valid in the handle's state */
#define TNR_ERR_NODEVICE   -7  /* no usable gfx950 device */
#define TNR_ERR_INTERNAL   -8  /* internal consistency check failed */
#define TNR_ERR_BUDGET     -9  /* request exceeds the admission budget */
#define TNR_ERR_BUSY      -10  /* result not ready yet (tnr_server_poll) */

/* ---- scheme kinds and precision -------------------------------------- */
#define TNR_TRG    0   /* src/schemes/trg.jl:26-41   */
#define TNR_BTRG   1   /* src/schemes/btrg.jl:26-94  */
#define TNR_HOTRG  2   /* src/schemes/hotrg.jl:27-160 */
#define TNR_IMPURITY_TRG    3   /* src/schemes/impuritytrg.jl:40-103 */
#define TNR_IMPURITY_HOTRG  4   /* src/schemes/impurityhotrg.jl:27-95 */
#define TNR_ATRG  5   /* src/schemes/atrg.jl:26-76 */
#define TNR_HOTRG_3D  6   /* src/schemes/hotrg3d.jl:27-141 (6-leg tensors) */

#define TNR_PREC_FP64     0  /* all contractions FP64 MFMA */
#define TNR_PREC_SPLIT32  1  /* contractions on FP32 MFMA, hi/lo split
                                products, from a run's 9th step on (the
                                first 8, which dominate the free energy,
                                stay FP64); truncation SVD/eigh, log-norms
                                and finalize stay FP64 */

typedef struct tnr_scheme tnr_scheme;

/* Replaces the constructors TRG(T) (src/schemes/trg.jl:30-32),
 * BTRG(T, k) (src/schemes/btrg.jl:39-48; k = -0.5 is the reference default)
 * and HOTRG(T) (src/schemes/hotrg.jl:31-33).  Copies T to HBM.  `k` is read
 * for BTRG only.  Legs 1/4 and 2/3 must carry the same spaces (the finalize
 * trace connects them).  A tensor with NaN or infinite entries is refused
 * with TNR_ERR_NONFINITE (here and in every call that takes a tensor). */
int tnr_create(tnr_scheme** out, int kind, int zn, const int32_t* leg_dims,
               const double* data, size_t ndata, double k, int precision);

/* step!(scheme, truncrank(chi)): src/schemes/trg.jl:35-41,
 * src/schemes/btrg.jl:60-94, src/schemes/hotrg.jl:154-160.
 * Where the reference's LAPACK-backed svd_trunc / eigh_trunc! would raise
 * (src/schemes/btrg.jl:61,72, src/schemes/hotrg.jl:103,110), the step fails:
 * TNR_ERR_NOCONV when a truncated SVD or eigensolve stops at its iteration
 * or sweep cap unconverged (tnr_solver_config in tnr_kernels.h sets the
 * caps), TNR_ERR_NONFINITE when it meets NaN / infinite values.  The handle's
 * tensor is then undefined, and tnr_step, tnr_finalize and tnr_run return
 * the same status until tnr_set_tensor replaces the tensor. */
int tnr_step(tnr_scheme* s, int chi);

/* finalize!(scheme): src/utility/finalize.jl:4-8 (TRG, HOTRG) and :10-14
 * (BTRG).  Writes the norm n and divides T by n on the device. */
int tnr_finalize(tnr_scheme* s, double* norm);

/* run!(scheme, truncrank(chi), maxiter(maxiter); finalize_beginning):
 * src/schemes/tnrscheme.jl:33-59 with src/utility/stopping.jl:3-5,18.
 * `norms` must hold max(maxiter, 1) + finalize_beginning rows of
 * tnr_ncomponents(kind) doubles (one per finalize!); *n_out receives the
 * number of rows written. */
int tnr_run(tnr_scheme* s, int chi, int maxiter, int finalize_beginning,
            double* norms, int* n_out);

/* free_energy(data, beta; scalefactor, initial_size):
 * src/utility/free_energy.jl:13-20. */
int tnr_free_energy(const double* norms, int n, double beta,
                    double scalefactor, double initial_size, double* f_out);

/* HOTRG_3D(T) (src/schemes/hotrg3d.jl:27-34) for a 3D tensor
 * T in V1 (x) V2 <- V3 (x) V4 (x) V5 (x) V6 (legs z, z', Y, X, y, x):
 * leg_dims holds 6*zn sector dimensions, data the blocks per coupled charge
 * c with rows over codomain pairs (a1, a2) and columns over domain 4-tuples
 * (a3, a4, a5, a6), a3 + a4 + a5 + a6 = c, both colexicographic (first leg
 * fastest).  tnr_step, tnr_finalize (n = |T[1 1; 2 3 2 3]|,
 * src/utility/finalize.jl:62-66) and tnr_run work as for the 2D kinds;
 * free energies use scalefactor 8. */
int tnr_create_3d(tnr_scheme** out, int zn, const int32_t* leg_dims,
                  const double* data, size_t ndata, int precision);
/* Number of legs (4, or 6 for HOTRG_3D) and codomain legs of T. */
int tnr_get_nlegs(const tnr_scheme* s, int* nleg, int* ncod);
/* Current sector dimensions (nleg*zn int32) and element count of T. */
int tnr_get_dims(const tnr_scheme* s, int32_t* leg_dims, size_t* ndata);
/* Copy T out (format above) / replace T (checkpoint, resume, debugging). */
int tnr_get_tensor(tnr_scheme* s, double* data, size_t ndata);

/* ---- observables on the scheme's tensor --------------------------------
 * Each works on the scheme's current T (BTRG: T with S1, S2 absorbed as
 * the reference does); the transfer matrix of `unitcell` (1 or 2) tensors
 * in a row needs equal left/right and top/bottom spaces, and every
 * coupled-charge block of it at most 512 wide (TNR_ERR_ARG above). */
/* All eigenvalues of the transfer matrix (eig_full in cft.jl:16, 50 and
 * 296, 326): re/im hold `cap` values; *count receives the number written
 * (or needed).  Order: decreasing modulus, ties by decreasing real part,
 * then decreasing imaginary part. */
int tnr_transfer_spectrum(tnr_scheme* s, int unitcell, double* re, double* im,
                          size_t cap, size_t* count);
/* cft_data(scheme; v, unitcell, is_real): src/utility/cft.jl:5-71 (the
 * spectrum above, real part > 0 and modulus > 1e-12 kept, then
 * unitcell / (2 pi v) log(d_1 / d_i)); im is 0 when is_real. */
int tnr_cft_data(tnr_scheme* s, int unitcell, double v, int is_real,
                 double* re, double* im, size_t cap, size_t* count);
/* ground_state_degeneracy(scheme; unitcell): cft.jl:278-339. */
int tnr_ground_state_degeneracy(tnr_scheme* s, int unitcell, double* gsd);
/* gu_wen_ratio(scheme) -> X1, X2: cft.jl:374-396. */
int tnr_gu_wen_ratio(tnr_scheme* s, double* x1, double* x2);
/* central_charge(scheme, n): cft.jl:256-268. */
int tnr_central_charge(tnr_scheme* s, double n, double* c);
/* finalize_two_by_two!(scheme): src/utility/finalize.jl:17-41 (TRG, HOTRG
 * and BTRG): *out = n^(1/4) of the 2x2 torus, T /= n^(1/4) on the device. */
int tnr_finalize_two_by_two(tnr_scheme* s, double* out);

/* ---- impurity schemes ------------------------------------------------- */
/* ImpurityTRG(T, T_imp1, T_imp2, T_imp3, T_imp4)
 * (src/schemes/impuritytrg.jl:40-58, ntensors = 5) or
 * ImpurityHOTRG(T, T_imp_order1_1, T_imp_order1_2, T_imp_order2)
 * (src/schemes/impurityhotrg.jl:27-46, ntensors = 4): data[i] / ndata[i] is
 * tensor i in the format of tnr_create; all share leg_dims (the reference
 * asserts equal spaces).  tnr_step and tnr_run work as for the other kinds.
 * ImpurityTRG's impurity norm contracts factors of different SVDs, so it
 * depends on the singular-vector basis, which the reference leaves to
 * LAPACK; here every ImpurityTRG SVD is put in the pivoted-row gauge of
 * tnr_kernel_pivoted_gauge (include/tnr_kernels.h). */
int tnr_create_multi(tnr_scheme** out, int kind, int zn, const int32_t* leg_dims,
                     const double* const* data, const size_t* ndata, int ntensors,
                     int precision);
/* Values per finalize!: 1 (TRG, BTRG, HOTRG), 2 (ImpurityTRG: npure, nimp,
 * src/utility/finalize.jl:74-88), 4 (ImpurityHOTRG: n, n_11, n_12, n_2,
 * finalize.jl:90-101); negative status for an unknown kind. */
int tnr_ncomponents(int kind);
/* finalize! of any scheme: writes tnr_ncomponents(kind) values. */
int tnr_finalize_multi(tnr_scheme* s, double* out);
/* Tensor `index` of the scheme (0 = T, 1.. = the impurity tensors in
 * constructor order), in the format of tnr_create. */
int tnr_get_tensor_at(tnr_scheme* s, int index, double* data, size_t ndata);
int tnr_set_tensor(tnr_scheme* s, const int32_t* leg_dims,
                   const double* data, size_t ndata);
/* BTRG bond weights: which = 1 -> S1 (vertical), 2 -> S2 (horizontal);
 * diagonal entries ordered by sector, then index (src/schemes/btrg.jl:33-36). */
int tnr_get_bond_weights(tnr_scheme* s, int which, double* w, size_t n);

/* Per-factorization diagnostics of the last step: smallest kept and largest
 * dropped singular value (or |eigenvalue|) of each truncation, and the
 * solver iteration count, 4 doubles per factorization. */
int tnr_get_trunc_log(tnr_scheme* s, double* out, int max_entries, int* n_entries);

int tnr_destroy(tnr_scheme* s);
const char* tnr_strerror(int status);
/* Thread-local message of the last failing call. */
const char* tnr_last_error(void);

/* ---- batched runs ------------------------------------------------------
 * Many independent run! calls (beta scans, mixed chi, mixed models) driven
 * together: every coarse-graining step of every run in the batch is issued
 * as grouped launches (one grouped GEMM / permute / solver launch per phase
 * covering all runs and all symmetry blocks).  Results equal the one-run
 * path within the documented tolerance.  Every kind runs here: TRG, BTRG,
 * HOTRG, ATRG, HOTRG_3D (6-leg tensors, as tnr_create_3d) and the impurity
 * methods (several tensors on one set of spaces, as tnr_create_multi). */
typedef struct {
    int kind;                 /* any TNR_* scheme kind           */
    int zn;                   /* 1 = Trivial, N = Z_N            */
    const int32_t* leg_dims;  /* 4*zn (6*zn for TNR_HOTRG_3D)    */
    const double* data;       /* initial tensor T                */
    size_t ndata;
    double k;                 /* BTRG exponent                   */
    int chi;                  /* truncrank(chi)                  */
    int maxiter;              /* maxiter(n)                      */
    int finalize_beginning;
    /* impurity kinds: the tensors after T (4 for TNR_IMPURITY_TRG, 3 for
     * TNR_IMPURITY_HOTRG, in constructor order), each of ndata doubles on
     * T's spaces; ignored (may be NULL) for the other kinds */
    const double* const* impurity;
} tnr_run_desc;

/* norms: concatenation over runs of (maxiter + finalize_beginning) rows of
 * tnr_ncomponents(kind) doubles.
 * If a run fails (see tnr_server_poll) the call returns its status and
 * tnr_last_error names the run.
 * The batch engine behind this call (its streams, device memory pools and
 * pinned staging rings) stays alive for the life of the process and is
 * reused by the next call on the same device with the same precision and
 * lane count; concurrent calls run one after another. */
int tnr_run_batch(int nruns, const tnr_run_desc* runs, int precision,
                  double* norms);

/* ---- serving loop --------------------------------------------------------
 * Continuous batching: requests are admitted as HBM and slots free up
 * (estimate ~ chi^4 x sectors per scheme, checked against hbm_budget_bytes),
 * each admitted scheme stays resident on the device across its iterations,
 * and all resident schemes advance together one grouped step at a time. */
typedef struct tnr_server tnr_server;

int tnr_server_create(tnr_server** out, size_t hbm_budget_bytes, int max_slots,
                      int precision);
/* Queue one run; *id receives a ticket.  Fails with TNR_ERR_BUDGET if the
 * request alone exceeds the budget, TNR_ERR_NONFINITE for a tensor with NaN
 * or infinite entries. */
int tnr_server_submit(tnr_server* s, const tnr_run_desc* req, int64_t* id);
/* Advance every resident scheme by one coarse-graining step, admitting
 * queued requests first.  *active receives the number still in flight. */
int tnr_server_pump(tnr_server* s, int* active);
/* Fetch a finished result: TNR_ERR_BUSY if not finished.  norms must hold
 * maxiter + finalize_beginning rows of tnr_ncomponents(kind) doubles and
 * *n_out receives the rows written; t_admit/t_done are seconds since
 * server creation.  A run whose step failed (TNR_ERR_NOCONV,
 * TNR_ERR_NONFINITE, see tnr_step) ends there and returns that status with
 * the norms written so far; the other resident runs go on unaffected. */
int tnr_server_poll(tnr_server* s, int64_t id, double* norms, int* n_out,
                    double* t_submit, double* t_done);
int tnr_server_stats(tnr_server* s, size_t* hbm_in_use, int* resident,
                     int* queued);
int tnr_server_destroy(tnr_server* s);

#ifdef __cplusplus
}
#endif
#endif /* TNR_H */
