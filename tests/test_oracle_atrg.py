"""The oracle's ATRG (oracle/schemes.py: ATRG, src/schemes/atrg.jl:35-76)
against every pin there is (CPU only):

* exactness: with chi above every bond dimension one step loses nothing,
  so n0^4 n1 = Z(2x2 torus) of a random, symmetry-free tensor (brute-force
  sum, oracle/torus.py), which fixes every index of the three
  factorizations and four contractions;
* the reference's tests: Onsager at chi 24 (rtol 3e-6, scalefactor 4,
  test/schemes.jl:131-136), cft_data after 3 steps (rtol 1e-2, :138-145),
  ground_state_degeneracy / gu_wen_ratio -> 1 / 2 at beta_c -+ 0.01
  (:147-165);
* Z2 = Trivial for the Ising tensor.
"""
import numpy as np
import pytest

from oracle import models as om
from oracle import observables as ob
from oracle import schemes as osch
from oracle import torus
from oracle.ztensor import ZT


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_untruncated_step_is_exact(seed):
    rng = np.random.default_rng(seed)
    T = rng.standard_normal((2, 2, 2, 2)) + 1.0
    n = osch.run(osch.ATRG(ZT(T)), 16, 1)
    assert abs(n[0] ** 4 * n[1] / abs(torus.z_torus(T, 2)) - 1) < 1e-12


def test_planted_leg_swap_fails_the_torus_pin(monkeypatch):
    """Q = AX[3, -3, 4, -4] DY[-2, 4, -1, 3] with labels 3 and 4 swapped on
    DY keeps every shape but joins A to D instead of Y: the exact 2x2 torus
    must break.  (Choosing the other bipartition for the first SVD would
    not: that only stacks the two tensors the other way round on a periodic
    torus.)"""
    rng = np.random.default_rng(4)
    T = rng.standard_normal((2, 2, 2, 2)) + 1.0
    orig = osch.contract
    hits = []

    def wrong(out, *ops):
        if len(ops) == 2 and list(ops[0][1]) == [3, -3, 4, -4] and list(ops[1][1]) == [-2, 4, -1, 3]:
            hits.append(1)
            ops = (ops[0], (ops[1][0], [-2, 3, -1, 4], False))
        return orig(out, *ops)
    monkeypatch.setattr(osch, "contract", wrong)
    n = osch.run(osch.ATRG(ZT(T)), 16, 1)
    assert hits
    assert abs(n[0] ** 4 * n[1] / abs(torus.z_torus(T, 2)) - 1) > 1e-6


def test_reference_onsager_gate():
    sc = osch.ATRG(om.classical_ising_z2(om.ising_betac_mp))
    d = osch.run(sc, 24, 25)
    f = osch.free_energy(d, om.ising_betac, scalefactor=4.0)
    assert abs(f / om.f_onsager - 1) < 3e-6


def test_reference_cft_gsd_gu_wen_gates():
    sc = osch.ATRG(om.classical_ising_z2(om.ising_betac_mp))
    osch.run(sc, 24, 3)
    c = ob.cft_data(sc)[1:]
    assert abs(c[0] / (1 / 8) - 1) < 1e-2 and abs(c[1] - 1) < 1e-2
    for db, want in ((-0.01, 1.0), (0.01, 2.0)):
        sc = osch.ATRG(om.classical_ising_z2(om.ising_betac + db))
        osch.run(sc, 16, 20)
        x1, x2 = ob.gu_wen_ratio(sc)
        for v in (ob.ground_state_degeneracy(sc), x1, x2):
            assert abs(v / want - 1) < 1e-2


def test_z2_equals_trivial():
    a = osch.run(osch.ATRG(om.classical_ising_z2(0.43)), 8, 4)
    b = osch.run(osch.ATRG(om.classical_ising_trivial(0.43)), 8, 4)
    assert np.max(np.abs(np.array(a) / np.array(b) - 1)) < 1e-10
