"""Parity at the benchmark configurations' sizes: the plain FP64 HIP path
against the FP64 oracle's stored norm vectors (tests/golden/oracle_configs.json,
written by tests/golden/make_golden.py configs).

* BTRG on the Trivial Ising tensor, chi 32, 30 steps, at three betas of the
  seeded beta scan the headline benchmark runs (src/schemes/btrg.jl:60-94).
* TRG and HOTRG on the Trivial Ising tensor at beta_c, chi 32 and 48, 30
  steps (src/schemes/trg.jl:35-41, src/schemes/hotrg.jl:82-160): HOTRG's
  2304 x 2304 environments run the block-iterated eigensolver, and its
  projector choice (eps > eps', hotrg.jl:112) uses the device's eps.
* BTRG on the Z2 Ising tensor at beta_c and the Z3 Potts tensor at
  ln(1 + sqrt 3), chi 64, 30 steps: block-sparse sectors of uneven sizes.
* The first 16 runs of the seeded 512-run scan (beta ~ U[0.35, 0.55],
  chi ~ U{16, 24, 32, 48}) batched, in FP64 and in split precision.

Where a truncation cuts through a (near-)degenerate multiplet, the kept
subspace is not a function of the tensor, and no two implementations need
agree past that step.  Each fixture carries the oracle's truncation log; the
norms are compared at relative 1e-10 up to the last step whose cuts are all
determined (see determined_steps), and the free energy of the whole run is
compared at the loose bound the whole-run agreement supports."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-10
# A cut is determined when nothing is dropped, when the kept part at the cut
# is numerical noise (the largest value at or below NULL_REL of the top one:
# which noise directions are kept changes the tensor at roundoff only), or
# when the relative gap between the smallest kept and the largest dropped
# value is at least GAP_REL (a roundoff perturbation of the matrix then moves
# the kept subspace by ~1e-16 / GAP_REL, and the critical RG flow grows such
# perturbations by ~sqrt(2) per step: two LAPACK drivers differ by ~1e-13
# after 30 steps).
NULL_REL = 1e-13
GAP_REL = 1e-3
# HOTRG keeps the projector of the environment with the smaller truncation
# error (hotrg.jl:112).  The choice is determined when the two errors differ
# by more than EPS_REL of the larger one (the device's error of an iterated
# block is a Frobenius residual, good to ~1e-8 of itself).  Errors within
# EPS_TIE of the environment's largest eigenvalue are equal up to roundoff:
# the two environments of a reflection-symmetric tensor, which have the same
# spectrum by the symmetry, and either projector then gives the same norms
# (the symmetry maps one run onto the other up to a gauge on the legs).
EPS_REL = 1e-6
EPS_TIE = 1e-12


def _fixtures():
    p = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_configs.json")
    if not os.path.exists(p):
        return {}
    with open(p) as f:
        return json.load(f)


def cut_determined(n, top, kept, dropped, err):
    if dropped == 0.0 or top == 0.0:
        return True
    if kept <= NULL_REL * top:
        return True
    return (kept - dropped) / kept >= GAP_REL


# The oracle's own roundoff sensitivity ("sens" in the fixtures: the same run
# with every factorized matrix perturbed by 1e-15 of its largest entry): a
# step counts as determined only while the norms move by at most SENS_MAX
# under that perturbation.  At beta_c the RG flow is unstable (HOTRG doubles
# a perturbation per step), and past the determined steps the norms are
# compared within AMP times the oracle's sensitivity, up to the step where
# that sensitivity passes CHAOS: from there on two runs that differ in the
# last bit follow different trajectories, and no norm is a function of the
# input tensor (the free energy, dominated by the early steps, still is).
SENS_MAX = 1e-12
AMP = 1e3
CHAOS = 1e-6


def determined_steps(case, use_sens=True):
    """Number of leading steps whose truncations (and, for HOTRG, projector
    choices) are all determined by the tensor and (use_sens) whose norms
    the oracle's roundoff sensitivity leaves within SENS_MAX."""
    cuts = case["cuts"]
    per = 4 if case["scheme"] == "HOTRG" else 2
    steps = len(cuts) // per
    sens = case.get("sens") if use_sens else None
    if sens is not None and sens[0] > SENS_MAX:
        return 0
    for s in range(steps):
        if sens is not None and sens[s + 1] > SENS_MAX:
            return s
        c = cuts[s * per:(s + 1) * per]
        if not all(cut_determined(*x) for x in c):
            return s
        if case["scheme"] == "HOTRG":
            for a, b in ((c[0], c[1]), (c[2], c[3])):
                e1, e2 = a[4], b[4]
                m = max(e1, e2)
                if m > 0 and EPS_TIE * max(a[1], b[1]) < abs(e1 - e2) <= EPS_REL * m:
                    return s
    return steps


def _model(tk, name, beta):
    if name == "ising_trivial":
        return tk.classical_ising(beta, symmetry="Trivial")
    if name == "ising_trivial_betac":
        return tk.classical_ising(None, symmetry="Trivial")
    if name == "ising_z2":
        return tk.classical_ising(None, symmetry="Z2")
    if name == "potts_z3":
        return tk.classical_potts(3, None)
    raise ValueError(name)


def _beta(tk, case):
    if case["beta"] is not None:
        return case["beta"]
    return float(tk.potts_betac(3)) if case["model"] == "potts_z3" else float(tk.ising_betac)


def _compare(tk, case, d, label):
    ref = np.array(case["norms"])
    d = np.asarray(d, dtype=float)
    assert d.shape == ref.shape, label
    k = determined_steps(case)
    err = np.abs(d / ref - 1)
    head = float(np.max(err[:k + 1]))
    sf = 4.0 if case["scheme"] == "HOTRG" else 2.0
    f = tk.free_energy(d, _beta(tk, case), scalefactor=sf)
    df = abs(f / case["f"] - 1)
    print(f"{label}: determined steps {k}/{len(ref) - 1}, max rel norm err {head:.2e} "
          f"(all steps {float(np.max(err)):.2e}), |df/f| {df:.2e}")
    assert head < TOL, label + f" (determined steps {k}): step, rel err: " + "; ".join(
        f"{i} {err[i]:.2e}" for i in range(len(err)))
    sens = case.get("sens")
    if sens is not None:
        # past the determined steps: within AMP times the oracle's own
        # roundoff sensitivity
        # only while every cut keeps its gap: past a split multiplet one
        # perturbed oracle run does not bound how far a correct run may go
        sens = np.asarray(sens)
        bound = np.maximum(TOL, AMP * sens)
        over = np.nonzero(sens > CHAOS)[0]
        stop = min(int(over[0]) if over.size else len(err),
                   determined_steps(case, use_sens=False) + 1)
        bad = np.nonzero(err[:stop] > bound[:stop])[0]
        assert bad.size == 0, label + ": step, rel err, oracle sensitivity: " + "; ".join(
            f"{i} {err[i]:.2e} {sens[i]:.2e}" for i in range(len(err)))
    return k, df


SINGLE = sorted(n for n in _fixtures() if not n.startswith("c4_"))


@pytest.mark.parametrize("name", SINGLE)
def test_config_size_parity(tk, name):
    """One run on the plain FP64 path (tnr_run) against the oracle."""
    case = _fixtures()[name]
    sc = getattr(tk, case["scheme"])(_model(tk, case["model"], case["beta"]))
    d = tk.run(sc, tk.truncrank(case["chi"]), tk.maxiter(case["maxiter"]), verbosity=0)
    sc.close()
    k, df = _compare(tk, case, d, name)
    if k == case["maxiter"]:
        assert df < 1e-11, (name, df)
    else:
        # past a degenerate cut the two runs keep different (equally valid)
        # subspaces; the free energies still agree to the truncation level
        assert df < 1e-7, (name, df)


C4 = sorted(n for n in _fixtures() if n.startswith("c4_"))
# Split precision (TNR_PREC_SPLIT32): FP32-MFMA hi/lo products accumulated
# in FP32 per k-tile from a run's 9th step on; its free energies against the
# oracle's, per run, at the 1e-9 the engine promises.
SPLIT_DF = 1e-9


@pytest.mark.skipif(not C4, reason="no 512-run scan fixtures")
def test_c4_scan_subset_batched(tk):
    """The first 16 runs of the seeded 512-run scan (chi 16-48) in one batch:
    FP64 against the oracle per determined step; split precision within
    SPLIT_DF of the oracle's free energy; the batch equals each run alone."""
    g = _fixtures()
    cases = [g[n] for n in C4]
    reqs = [tk.Request("BTRG", _model(tk, c["model"], c["beta"]), c["chi"], c["maxiter"])
            for c in cases]
    fp64 = tk.run_batch(reqs)
    split = tk.run_batch(reqs, tk.SPLIT32)
    worst_split = 0.0
    for name, c, a, b in zip(C4, cases, fp64, split):
        _compare(tk, c, a, name)
        fs = tk.free_energy(b, c["beta"])
        worst_split = max(worst_split, abs(fs / c["f"] - 1))
    print(f"c4 subset: split-precision max |df/f| vs oracle {worst_split:.2e}")
    assert worst_split < SPLIT_DF
    # one chi=48 run alone on the plain path equals its batch result bitwise
    i = next(i for i, c in enumerate(cases) if c["chi"] == 48)
    sc = tk.BTRG(reqs[i].T)
    alone = np.array(tk.run(sc, tk.truncrank(48), tk.maxiter(cases[i]["maxiter"]), verbosity=0))
    sc.close()
    assert np.array_equal(alone, fp64[i])


def test_serving_mix_with_queueing_and_budget(tk):
    """Continuous batching on the serving mix (Z2 Ising, Z3 Potts, Z6 clock;
    chi 8, 32, 64; maxiter 20-24) with requests arriving while others run.
    A 1 GB budget holds one chi=64 run with a few small ones, so later chi=64
    requests queue until memory frees up; a chi=128 Trivial request alone
    exceeds the budget and is refused (TNR_ERR_BUDGET).  Every result equals
    the same run alone on the plain path, bit for bit."""
    models = [lambda: tk.classical_ising(0.45, symmetry="Z2"),
              lambda: tk.classical_potts(3, 1.0),
              lambda: tk.classical_clock(6, 1.1)]
    reqs = []
    for i, chi in enumerate((8, 32, 64, 64, 8, 32, 64, 32, 8)):
        reqs.append(tk.Request("BTRG", models[i % 3](), chi, 20 + (i % 5)))
    srv = tk.Server(hbm_budget_bytes=1 << 30, max_slots=16)
    too_big = tk.Request("BTRG", tk.classical_ising(0.44, symmetry="Trivial"), 128, 2)
    with pytest.raises(tk.TNRError) as e:
        srv.submit(too_big)
    assert e.value.status == -9
    ids = [srv.submit(r) for r in reqs[:4]]
    srv.pump()
    queued = srv.stats()["queued"]
    srv.pump()
    ids += [srv.submit(r) for r in reqs[4:]]          # arrive while others run
    max_queued = queued
    while srv.pump() > 0:
        max_queued = max(max_queued, srv.stats()["queued"])
    assert max_queued > 0
    for r, i in zip(reqs, ids):
        res = srv.poll(i)
        assert res is not None
        sc = tk.BTRG(r.T)
        alone = np.array(tk.run(sc, tk.truncrank(r.chi), tk.maxiter(r.maxiter), verbosity=0))
        sc.close()
        assert np.array_equal(res[0], alone), (r.chi, r.maxiter)
    srv.close()
