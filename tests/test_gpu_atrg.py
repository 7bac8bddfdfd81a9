"""ATRG on the device (src/schemes/atrg.jl:35-76) through the C ABI: per-step
norms against the oracle (rtol 1e-10) on Z2, Trivial and Z3 tensors, the
exact 2x2 torus of an untruncated step, the reference's own ATRG tests, and
batched runs equal to single runs."""
import numpy as np
import pytest

from oracle import models as om
from oracle import schemes as osch
from oracle import torus
from oracle.ztensor import ZT

pytestmark = pytest.mark.gpu


def _oracle_T(model, beta):
    if model == "ising_z2":
        return om.classical_ising_z2(beta)
    if model == "ising_trivial":
        return om.classical_ising_trivial(beta)
    return om.classical_potts_zn(3, beta)


def _tk_T(tk, model, beta):
    b = None if beta is om.ising_betac_mp else float(beta)
    if model == "ising_z2":
        return tk.classical_ising(b, symmetry="Z2")
    if model == "ising_trivial":
        return tk.classical_ising(b, symmetry="Trivial")
    return tk.classical_potts(3, b)


@pytest.mark.parametrize("model,beta,chi,n", [
    ("ising_z2", om.ising_betac_mp, 16, 6), ("ising_trivial", 0.44, 8, 6),
    ("potts_z3", 1.0, 12, 5), ("ising_z2", om.ising_betac_mp, 24, 4)])
def test_atrg_vs_oracle(tk, model, beta, chi, n):
    ref = np.array(osch.run(osch.ATRG(_oracle_T(model, beta)), chi, n))
    sc = tk.ATRG(_tk_T(tk, model, beta))
    got = np.array(tk.run(sc, tk.truncrank(chi), tk.maxiter(n), verbosity=0))
    sc.close()
    assert np.max(np.abs(got / ref - 1)) < 1e-10


def test_atrg_untruncated_step_is_exact(tk):
    rng = np.random.default_rng(7)
    T = rng.standard_normal((2, 2, 2, 2)) + 1.0
    sc = tk.ATRG(tk.trivial(T))
    n = tk.run(sc, tk.truncrank(16), tk.maxiter(1), verbosity=0)
    sc.close()
    assert abs(n[0] ** 4 * n[1] / abs(torus.z_torus(T, 2)) - 1) < 1e-12


def test_atrg_reference_gates(tk):
    """test/schemes.jl:131-165 on the device."""
    sc = tk.ATRG(tk.classical_ising(symmetry="Z2"))
    d = tk.run(sc, tk.truncrank(24), tk.maxiter(25), verbosity=0)
    sc.close()
    f = tk.free_energy(d, tk.ising_betac, scalefactor=4.0)
    assert abs(f / tk.f_onsager - 1) < 3e-6
    sc = tk.ATRG(tk.classical_ising(symmetry="Z2"))
    tk.run(sc, tk.truncrank(24), tk.maxiter(3), verbosity=0)
    c = tk.cft_data(sc)[1:]
    sc.close()
    assert abs(c[0] / tk.ising_cft_exact[0] - 1) < 1e-2 and abs(c[1] - 1) < 1e-2
    for db, want in ((-0.01, 1.0), (0.01, 2.0)):
        sc = tk.ATRG(tk.classical_ising(tk.ising_betac + db, symmetry="Z2"))
        tk.run(sc, tk.truncrank(16), tk.maxiter(20), verbosity=0)
        x1, x2 = tk.gu_wen_ratio(sc)
        gsd = tk.ground_state_degeneracy(sc)
        sc.close()
        for v in (gsd, x1, x2):
            assert abs(v / want - 1) < 1e-2


def test_atrg_batch_equals_single_runs(tk):
    reqs = [tk.Request("ATRG", tk.classical_ising(0.40 + 0.02 * i, symmetry="Trivial"), 12, 5)
            for i in range(3)]
    reqs.append(tk.Request("BTRG", tk.classical_ising(0.44, symmetry="Z2"), 16, 5))
    batch = tk.run_batch(reqs)
    for i in range(3):
        sc = tk.ATRG(reqs[i].T)
        one = np.array(tk.run(sc, tk.truncrank(12), tk.maxiter(5), verbosity=0))
        sc.close()
        assert np.array_equal(one, np.asarray(batch[i]))
