"""The oracle's HOTRG_3D (oracle/schemes.py, src/schemes/hotrg3d.jl:48-141)
against every pin there is (CPU only):

* exactness: one z compression with chi above the x and y pair dimensions
  loses nothing, so the trace T[1 1; 2 3 2 3] of the result equals the
  brute-force two-site z torus sum_{..} A[a b; y x y x] A[b a; y' x' y' x']
  of a random tensor; exchanging the y and x legs of A1 in the step network
  fails it (exchanging x2 with x2' would not: the trace identifies them);
* the reference's test: free energy at beta_c, chi 8, 25 steps, within
  1e-3 of -3.507 (scalefactor 8, test/schemes.jl:376-383);
* Z2 = Trivial for the 3D Ising tensor while nothing is truncated.
"""
import numpy as np
import pytest

from oracle import models as om
from oracle import schemes as osch
from oracle.ztensor import ZT


def _zpair(A):
    return np.einsum("abcdcd,baefef->", A, A)


@pytest.mark.parametrize("seed", [1, 2])
def test_untruncated_z_compression_is_exact(seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((2,) * 6) + 0.5
    sc = osch.HOTRG_3D(ZT(A))
    sc.substep(8)
    assert abs(np.einsum("aabcbc->", sc.T.data) / _zpair(A) - 1) < 1e-13


def test_planted_leg_swap_fails_the_pin(monkeypatch):
    rng = np.random.default_rng(3)
    A = rng.standard_normal((2,) * 6) + 0.5
    orig = osch.contract
    hits = []

    def wrong(out, *ops):
        # B1 = A1[a z; y1' x1' y1 x1] Ux[x1' x2'; n4] with y1' and x1' exchanged
        if len(ops) == 2 and list(ops[0][1]) == [-1, -2, -3, 1, -4, -5]:
            hits.append(1)
            ops = ((ops[0][0], [-1, -2, 1, -3, -4, -5], False), ops[1])
        return orig(out, *ops)
    monkeypatch.setattr(osch, "contract", wrong)
    sc = osch.HOTRG_3D(ZT(A))
    sc.substep(8)
    assert hits
    assert abs(np.einsum("aabcbc->", sc.T.data) / _zpair(A) - 1) > 1e-6


def test_reference_free_energy_gate():
    sc = osch.HOTRG_3D(om.classical_ising_3d_z2(om.ising_betac_3d))
    d = osch.run(sc, 8, 25)
    f = osch.free_energy(d, om.ising_betac_3d, scalefactor=8.0)
    assert abs(f / om.f_benchmark3d - 1) < 1e-3


def test_z2_equals_trivial_before_truncation():
    a = osch.run(osch.HOTRG_3D(om.classical_ising_3d_z2(0.2)), 4, 1)
    b = osch.run(osch.HOTRG_3D(om.classical_ising_3d_trivial(0.2)), 4, 1)
    assert np.max(np.abs(np.array(a) / np.array(b) - 1)) < 1e-12
