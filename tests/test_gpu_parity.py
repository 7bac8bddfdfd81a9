"""Scheme-level parity on the GPU through the C ABI: per-step norms and free
energies against the FP64 oracle (relative 1e-10 on the plain FP64 path),
the published BTRG free energy, exact small tori for random tensors, and
batch / serving results equal to the one-run results."""
import numpy as np
import pytest

from oracle import models as om
from oracle import schemes as osch
from oracle import torus
from oracle.ztensor import ZT
from conftest import rel, to_sym

pytestmark = pytest.mark.gpu
TOL = 1e-10


def gpu_run(tk, cls, T, chi, n, **kw):
    sc = cls(T, **kw)
    out = np.array(tk.run(sc, tk.truncrank(chi), tk.maxiter(n), verbosity=0))
    sc.close()
    return out


def _golden_norms():
    import json
    import os
    p = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_norms.json")
    with open(p) as f:
        return json.load(f)


def _tk_model(tk, name, beta):
    if name == "ising_z2":
        return tk.classical_ising(beta, symmetry="Z2")
    if name == "ising_trivial":
        return tk.classical_ising(beta, symmetry="Trivial")
    if name == "potts_z3":
        return tk.classical_potts(3, beta)
    if name == "clock_z6":
        return tk.classical_clock(6, beta)
    raise ValueError(name)


@pytest.mark.parametrize("name", sorted(_golden_norms()))
def test_against_golden_oracle_norms(tk, name):
    """Stored oracle norm vectors (tests/golden/oracle_norms.json): every
    scheme, Z2 / Z3 / Z6 / Trivial sectors, relative 1e-10."""
    case = _golden_norms()[name]
    d = gpu_run(tk, getattr(tk, case["scheme"]), _tk_model(tk, case["model"], case["beta"]),
                case["chi"], case["maxiter"])
    assert rel(d, np.array(case["norms"])) < TOL


def test_readme_btrg_z2_chi16(tk):
    """README.md:78-87: f = -2.1096504926141826902647832 (26 norms)."""
    d = gpu_run(tk, tk.BTRG, tk.classical_ising(symmetry="Z2"), 16, 25)
    assert len(d) == 26
    ref = np.array(osch.run(osch.BTRG(om.classical_ising_z2(om.ising_betac_mp)), 16, 25))
    assert rel(d, ref) < TOL
    f = tk.free_energy(d, tk.ising_betac)
    assert abs(f / -2.1096504926141826902647832 - 1) < 1e-12


@pytest.mark.parametrize("name,chi,n", [("TRG", 16, 20), ("BTRG", 24, 20), ("HOTRG", 8, 12)])
@pytest.mark.parametrize("sym", ["Trivial", "Z2"])
def test_ising_schemes_vs_oracle(tk, name, chi, n, sym):
    T = tk.classical_ising(symmetry=sym)
    zt = (om.classical_ising_z2 if sym == "Z2" else om.classical_ising_trivial)(om.ising_betac_mp)
    d = gpu_run(tk, getattr(tk, name), T, chi, n)
    ref = np.array(osch.run(getattr(osch, name)(zt), chi, n))
    assert rel(d, ref) < TOL


@pytest.mark.parametrize("seed", [1, 2])
def test_exact_first_steps_random_tensor(tk, seed):
    rng = np.random.default_rng(seed)
    T = rng.standard_normal((2, 2, 2, 2)) + 1.0
    z1, zt2, z22, z44 = (torus.z_1site(T), torus.z_tilted2(T), torus.z_torus(T, 2),
                         torus.z_torus(T, 4))
    for cls in (tk.TRG, tk.BTRG):
        n = gpu_run(tk, cls, tk.trivial(T), 16, 2)
        assert abs(n[0] / abs(z1) - 1) < 1e-13
        assert abs(n[0] ** 2 * n[1] / abs(zt2) - 1) < 1e-12
        assert abs(n[0] ** 4 * n[1] ** 2 * n[2] / abs(z22) - 1) < 1e-12
    n = gpu_run(tk, tk.HOTRG, tk.trivial(T), 16, 2)
    assert abs(n[0] ** 4 * n[1] / abs(z22) - 1) < 1e-12
    assert abs(n[0] ** 16 * n[1] ** 4 * n[2] / abs(z44) - 1) < 1e-11


@pytest.mark.parametrize("name,chi,n", [("BTRG", 16, 10), ("TRG", 16, 10), ("HOTRG", 9, 6)])
def test_potts_z3_vs_oracle(tk, name, chi, n):
    b = om.potts_betac(3)
    zt = om.classical_potts_zn(3, b)
    d = gpu_run(tk, getattr(tk, name), to_sym(tk, zt), chi, n)
    ref = np.array(osch.run(getattr(osch, name)(zt), chi, n))
    assert rel(d, ref) < TOL


def test_clock_z6_vs_oracle(tk):
    zt = om.classical_clock_zn(6, 1.1)
    d = gpu_run(tk, tk.BTRG, to_sym(tk, zt), 12, 8)
    ref = np.array(osch.run(osch.BTRG(zt), 12, 8))
    assert rel(d, ref) < TOL


def test_batch_equals_single_runs(tk):
    reqs = [tk.Request("BTRG", tk.classical_ising(0.40, symmetry="Trivial"), 16, 12),
            tk.Request("TRG", tk.classical_ising(0.45, symmetry="Z2"), 12, 10),
            tk.Request("BTRG", tk.classical_potts(3, 1.0), 16, 8),
            tk.Request("HOTRG", tk.classical_ising(0.50, symmetry="Trivial"), 8, 6)]
    batch = tk.run_batch(reqs)
    for r, b in zip(reqs, batch):
        cls = {0: tk.TRG, 1: tk.BTRG, 2: tk.HOTRG}[r.kind]
        single = gpu_run(tk, cls, r.T, r.chi, r.maxiter)
        assert np.array_equal(single, b), (r.kind, rel(single, b))


def test_server_continuous_batching(tk):
    srv = tk.Server(max_slots=3)
    reqs = [tk.Request("BTRG", tk.classical_ising(0.35 + 0.02 * i, symmetry="Z2"), 8, 5 + i)
            for i in range(6)]
    ids = [srv.submit(r) for r in reqs[:2]]
    srv.pump()
    ids += [srv.submit(r) for r in reqs[2:]]          # arrive while others run
    while srv.pump() > 0:
        pass
    for r, i in zip(reqs, ids):
        res = srv.poll(i)
        assert res is not None
        alone = gpu_run(tk, tk.BTRG, r.T, r.chi, r.maxiter)
        assert np.array_equal(res[0], alone)
    srv.close()


def test_batch_mixed_symmetries_equals_single_runs(tk):
    """Z2 Ising, Z3 Potts and Z6 clock BTRG runs stepped in the same grouped
    launches give the same norms as each run alone."""
    reqs = [tk.Request("BTRG", tk.classical_ising(0.44, symmetry="Z2"), 16, 8),
            tk.Request("BTRG", tk.classical_potts(3, 1.0), 16, 8),
            tk.Request("BTRG", tk.classical_clock(6, 1.1), 12, 8),
            tk.Request("BTRG", tk.classical_ising(0.40, symmetry="Trivial"), 16, 8)]
    batch = tk.run_batch(reqs)
    for r, b in zip(reqs, batch):
        single = gpu_run(tk, tk.BTRG, r.T, r.chi, r.maxiter)
        assert rel(single, b) < 1e-12


def test_lanes_do_not_change_results(tk, monkeypatch):
    """The server spreads runs over 1-4 lanes (stream + host thread each);
    every run's norms are bitwise the same for any lane count."""
    reqs = [tk.Request("BTRG", tk.classical_ising(0.36 + 0.03 * i, symmetry="Trivial"), 16, 6)
            for i in range(5)]
    reqs.append(tk.Request("HOTRG", tk.classical_ising(0.44, symmetry="Z2"), 8, 4))
    out = {}
    for lanes in ("1", "2", "3", "4"):
        monkeypatch.setenv("TNR_LANES", lanes)
        out[lanes] = tk.run_batch(reqs)
    for a, b, c, d in zip(out["1"], out["2"], out["3"], out["4"]):
        assert np.array_equal(a, b) and np.array_equal(a, c) and np.array_equal(a, d)


def test_device_run_equals_host_loop(tk):
    """run!(.., maxiter(n)) as one device call (tnr_run) and the generic host
    loop of step!/finalize! (taken for any non-default Finalizer) give the
    same norms, bit for bit, for every scheme."""
    for cls, chi in ((tk.TRG, 12), (tk.BTRG, 16), (tk.HOTRG, 8)):
        T = tk.classical_ising(0.42, symmetry="Z2")
        a = gpu_run(tk, cls, T, chi, 7)
        sc = cls(T)
        b = np.array(tk.run(sc, tk.truncrank(chi), tk.maxiter(7), tk.Finalizer(tk.finalize),
                            verbosity=0))
        sc.close()
        assert np.array_equal(a, b)


def test_convcrit_run_stops_like_the_reference(tk):
    """trivial_convcrit on the device path: the loop stops once the last norm
    falls to delta (stopping.jl:19, 42); norms match a maxiter run of the
    same length."""
    T = tk.classical_ising(0.30, symmetry="Trivial")          # high temperature: norms -> 1
    sc = tk.BTRG(T)
    d = np.array(tk.run(sc, tk.truncrank(8), tk.maxiter(40) & tk.trivial_convcrit(1.0 + 1e-9),
                        verbosity=0))
    sc.close()
    ref = gpu_run(tk, tk.BTRG, T, 8, len(d) - 1)
    assert np.array_equal(d, ref)


def _close_rows(got, ref, cols=None, rtol=1e-10):
    """Row-wise: every compared component within rtol of the oracle, or
    within rtol * 1e-4 of the row's first value when it is smaller."""
    got, ref = np.asarray(got, dtype=float), np.asarray(ref, dtype=float)
    assert got.shape == ref.shape
    if cols is not None:
        got, ref = got[:, cols], ref[:, cols]
    tol = rtol * (np.abs(ref) + 1e-4 * np.abs(ref[:, :1]))
    return bool(np.all(np.abs(got - ref) <= tol)), float(np.max(np.abs(got - ref) / tol))


# The impurity TRG norm contracts factors of different SVDs (B1 with Ap,
# A1 with Bp), so it depends on each SVD's singular-vector basis: the
# device fixes it with the pivoted-row gauge and is compared with the
# oracle in the same gauge.  At h = 0, <m> = 0 by the Z2 symmetry; the gauge
# keeps it at roundoff (~1e-16 npure) for the first steps, compared with the
# floor of 1e-14 npure; a later step can lift the odd-sector roundoff (to
# 3e-7 npure at beta 0.3, step 6), which the cases stop short of.
@pytest.mark.parametrize("beta,h,chi,n,cols", [
    (0.44, 0.1, 8, 8, None), (0.40, 0.1, 12, 6, None), (0.6, 0.05, 8, 6, None),
    (0.3, 0.0, 8, 5, None), (0.44, 0.0, 8, 8, None)])
def test_impurity_trg_vs_oracle(tk, beta, h, chi, n, cols):
    """src/schemes/impuritytrg.jl:63-91 with finalize.jl:74-88: (npure, nimp)
    per step against the oracle, magnetisation impurity at site 1."""
    T, Ti = tk.classical_ising(beta, "Trivial", h), tk.classical_ising_impurity(beta, h)
    sc = tk.ImpurityTRG(T, Ti, T, T, T)
    got = tk.run(sc, tk.truncrank(chi), tk.maxiter(n), verbosity=0)
    sc.close()
    To, Tio = om.classical_ising_trivial(beta, h), om.classical_ising_impurity(beta, h)
    ref = osch.run(osch.ImpurityTRG(To, Tio, To, To, To, gauge="pivoted"), chi, n)
    ok, worst = _close_rows(got, ref, cols)
    assert ok, worst


@pytest.mark.parametrize("beta,h,chi,n,cols", [
    (0.44, 0.0, 8, 5, [0, 3]), (0.35, 0.05, 6, 5, None), (0.5, 0.1, 8, 4, None)])
def test_impurity_hotrg_vs_oracle(tk, beta, h, chi, n, cols):
    """src/schemes/impurityhotrg.jl:48-84 with finalize.jl:90-101:
    (n, n_11, n_12, n_2) per step against the oracle (n_2 = <sigma sigma>
    is Z2-even and compared at h = 0 too)."""
    T, Ti = tk.classical_ising(beta, "Trivial", h), tk.classical_ising_impurity(beta, h)
    sc = tk.ImpurityHOTRG(T, Ti, Ti, T)
    got = tk.run(sc, tk.truncrank(chi), tk.maxiter(n), verbosity=0)
    sc.close()
    To, Tio = om.classical_ising_trivial(beta, h), om.classical_ising_impurity(beta, h)
    ref = osch.run(osch.ImpurityHOTRG(To, Tio, Tio, To), chi, n)
    ok, worst = _close_rows(got, ref, cols)
    assert ok, worst


def test_impurity_reference_gates(tk):
    """The reference's own impurity tests on the device path:
    test/schemes.jl:443-450 (ImpurityTRG chi 24, Onsager 2e-6),
    :400-410 (ImpurityHOTRG chi 16, Onsager 6e-7), :454-475 (<m> -> 0 at
    beta 0.1, -> 1 at beta 2 with h 1e-6) and :412-438 (<m^2> -> 0 at
    beta 0.2, -> 1 at beta 1)."""
    T, Ti = tk.classical_ising(None, "Trivial"), tk.classical_ising_impurity(None)
    d = tk.run(tk.ImpurityTRG(T, Ti, T, T, T), tk.truncrank(24), tk.maxiter(25), verbosity=0)
    f = tk.free_energy([x[0] for x in d], tk.ising_betac)
    assert abs(f / tk.f_onsager - 1) < 2e-6
    d = tk.run(tk.ImpurityHOTRG(T, Ti, Ti, T), tk.truncrank(16), tk.maxiter(25), verbosity=0)
    f = tk.free_energy([x[0] for x in d], tk.ising_betac, scalefactor=4.0)
    assert abs(f / tk.f_onsager - 1) < 6e-7
    T, Ti = tk.classical_ising(0.1, "Trivial"), tk.classical_ising_impurity(0.1)
    d = tk.run(tk.ImpurityTRG(T, Ti, T, T, T), tk.truncrank(16), tk.maxiter(25), verbosity=0)
    assert abs(d[-1][1] / d[-1][0]) < 1e-4
    T, Ti = tk.classical_ising(2.0, "Trivial", 1e-6), tk.classical_ising_impurity(2.0, 1e-6)
    d = tk.run(tk.ImpurityTRG(T, Ti, T, T, T), tk.truncrank(16), tk.maxiter(25), verbosity=0)
    assert abs(d[-1][1] / d[-1][0] - 1) < 1e-4
    T, Ti = tk.classical_ising(0.2, "Trivial"), tk.classical_ising_impurity(0.2)
    d = tk.run(tk.ImpurityHOTRG(T, Ti, Ti, T), tk.truncrank(8), tk.maxiter(25), verbosity=0)
    assert abs(d[-1][3] / d[-1][0]) < 1e-14
    T, Ti = tk.classical_ising(1.0, "Trivial"), tk.classical_ising_impurity(1.0)
    d = tk.run(tk.ImpurityHOTRG(T, Ti, Ti, T), tk.truncrank(8), tk.maxiter(25), verbosity=0)
    assert abs(d[-1][3] / d[-1][0] - 1) < 1e-2


def test_impurity_host_loop_equals_device_run(tk):
    T, Ti = tk.classical_ising(0.42, "Trivial"), tk.classical_ising_impurity(0.42)
    a = tk.run(tk.ImpurityTRG(T, Ti, T, T, T), tk.truncrank(8), tk.maxiter(4), verbosity=0)
    b = tk.run(tk.ImpurityTRG(T, Ti, T, T, T), tk.truncrank(8), tk.maxiter(4),
               tk.Finalizer(tk.finalize), verbosity=0)
    assert a == b



def test_against_golden_extensions(tk):
    """tests/golden/oracle_extensions.json on the device: ATRG and HOTRG_3D
    norms (rtol 1e-10), ImpurityTRG / ImpurityHOTRG rows, observables after
    the stored TRG / BTRG runs (the tolerances of test_gpu_observables)."""
    import json
    import os
    p = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_extensions.json")
    with open(p) as f:
        g = json.load(f)
    c = g["atrg_z2_ising_betac_chi12_8"]
    sc = tk.ATRG(tk.classical_ising(symmetry="Z2"))
    d = tk.run(sc, tk.truncrank(c["chi"]), tk.maxiter(c["maxiter"]), verbosity=0)
    sc.close()
    assert np.max(np.abs(np.array(d) / np.array(c["norms"]) - 1)) < 1e-10
    c = g["hotrg3d_z2_ising3d_betac_chi6_3"]
    sc = tk.HOTRG_3D(tk.classical_ising_3D())
    d = tk.run(sc, tk.truncrank(c["chi"]), tk.maxiter(c["maxiter"]), verbosity=0)
    sc.close()
    assert np.max(np.abs(np.array(d) / np.array(c["norms"]) - 1)) < 1e-10
    c = g["impurity_trg_b0.44_h0.1_chi8_6"]
    T = tk.classical_ising(c["beta"], "Trivial", c["h"])
    Ti = tk.classical_ising_impurity(c["beta"], c["h"])
    sc = tk.ImpurityTRG(T, Ti, T, T, T)
    d = tk.run(sc, tk.truncrank(c["chi"]), tk.maxiter(c["maxiter"]), verbosity=0)
    sc.close()
    ok, worst = _close_rows(d, c["rows"])
    assert ok, worst
    c = g["impurity_hotrg_b0.5_h0.1_chi8_4"]
    T = tk.classical_ising(c["beta"], "Trivial", c["h"])
    Ti = tk.classical_ising_impurity(c["beta"], c["h"])
    sc = tk.ImpurityHOTRG(T, Ti, Ti, T)
    d = tk.run(sc, tk.truncrank(c["chi"]), tk.maxiter(c["maxiter"]), verbosity=0)
    sc.close()
    ok, worst = _close_rows(d, c["rows"])
    assert ok, worst
    for kind in ("TRG", "BTRG"):
        c = g[f"observables_{kind.lower()}_z2_ising_betac_chi16_6"]
        sc = getattr(tk, kind)(tk.classical_ising(symmetry="Z2"))
        d = tk.run(sc, tk.truncrank(c["chi"]), tk.maxiter(c["maxiter"]), verbosity=0)
        assert np.max(np.abs(tk.cft_data(sc)[:4] - np.array(c["cft_data_first6"][:4]))) < 1e-7
        assert abs(tk.ground_state_degeneracy(sc) / c["ground_state_degeneracy"] - 1) < 1e-8
        assert np.allclose(tk.gu_wen_ratio(sc), c["gu_wen_ratio"], rtol=1e-8)
        assert abs(tk.central_charge(sc, d[-1]) - c["central_charge_n_last"]) < 1e-8
        sc.close()


def test_reused_engine_and_contexts_repeat_bitwise(tk):
    """The batch engine and the contexts of destroyed handles are reused with
    whatever their pools cached from earlier work: a batch, then a batch of
    other sizes and symmetries, then the first batch again gives the first
    norms bit for bit; the same holds for single handles created after
    handles of other sizes were destroyed.  The last check is a batch that
    fails (a solver capped to non-convergence) followed by the first batch
    again."""
    a = [tk.Request("BTRG", tk.classical_ising(0.41 + 0.02 * i, symmetry="Trivial"), 24, 10)
         for i in range(6)]
    b = [tk.Request("HOTRG", tk.classical_ising(0.44, symmetry="Z2"), 12, 6),
         tk.Request("BTRG", tk.classical_potts(3, 1.0), 32, 7),
         tk.Request("TRG", tk.classical_clock(6, 1.1), 20, 5)]
    first = tk.run_batch(a)
    tk.run_batch(b)
    again = tk.run_batch(a)
    for x, y in zip(first, again):
        assert np.array_equal(x, y)

    def single(T, chi, n):
        return gpu_run(tk, tk.BTRG, T, chi, n)
    T1 = tk.classical_ising(0.43, symmetry="Z2")
    s1 = single(T1, 16, 8)
    single(tk.classical_ising(0.5, symmetry="Trivial"), 32, 6)
    single(tk.classical_potts(3, 0.9), 24, 5)
    assert np.array_equal(single(T1, 16, 8), s1)

    # a batch in which a run fails mid-way (its block iteration capped at
    # two iterations: TNR_ERR_NOCONV), then the first batch again
    fail = tk.Request("BTRG", tk.classical_ising(0.44, symmetry="Trivial"), 32, 8)
    tk.solver_config("max_iters=2")
    try:
        with pytest.raises(tk.TNRError):
            tk.run_batch([fail] + a[:2])
    finally:
        tk.solver_config("")
    after = tk.run_batch(a)
    for x, y in zip(first, after):
        assert np.array_equal(x, y)


def test_batch_runs_every_scheme_kind(tk):
    """ImpurityTRG, ImpurityHOTRG and HOTRG_3D runs go through the batch
    engine (tnr_run_batch) and the server beside TRG / BTRG / HOTRG / ATRG
    runs, each giving the same rows as the run alone (impuritytrg.jl:103,
    hotrg3d.jl:128-136: the reference runs them through the same run!)."""
    T, Ti = tk.classical_ising(0.44, "Trivial", 0.1), tk.classical_ising_impurity(0.44, 0.1)
    T3 = tk.classical_ising_3D()
    reqs = [tk.Request("ImpurityTRG", [T, Ti, T, T, T], 8, 5),
            tk.Request("BTRG", tk.classical_ising(0.42, symmetry="Z2"), 16, 6),
            tk.Request("ImpurityHOTRG", [T, Ti, Ti, T], 8, 4),
            tk.Request("HOTRG_3D", T3, 6, 3),
            tk.Request("ATRG", tk.classical_ising(0.44, symmetry="Z2"), 12, 4),
            tk.Request("HOTRG", tk.classical_ising(0.44, symmetry="Trivial"), 8, 4)]
    batch = tk.run_batch(reqs)
    alone = []
    for cls, args, chi, n in ((tk.ImpurityTRG, (T, Ti, T, T, T), 8, 5),
                              (tk.BTRG, (reqs[1].T,), 16, 6),
                              (tk.ImpurityHOTRG, (T, Ti, Ti, T), 8, 4),
                              (tk.HOTRG_3D, (T3,), 6, 3),
                              (tk.ATRG, (reqs[4].T,), 12, 4),
                              (tk.HOTRG, (reqs[5].T,), 8, 4)):
        sc = cls(*args)
        alone.append(np.array(tk.run(sc, tk.truncrank(chi), tk.maxiter(n), verbosity=0)))
        sc.close()
    for r, b, a in zip(reqs, batch, alone):
        assert b.shape == a.shape and np.array_equal(b, a), r.kind
    srv = tk.Server(max_slots=8)
    ids = [srv.submit(r) for r in reqs]
    while srv.pump() > 0:
        pass
    for i, a in zip(ids, alone):
        assert np.array_equal(srv.poll(i)[0], a)
    srv.close()
