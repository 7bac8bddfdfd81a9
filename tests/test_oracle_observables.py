"""The oracle's observables (oracle/observables.py) against every pin there
is (CPU only):

* the reference's own tests: cft_data against ising_cft_exact
  (test/schemes.jl:25-31, 64-70, 101-107), ground_state_degeneracy and
  gu_wen_ratio -> 1 / 2 at beta_c -+ 0.01 (test/schemes.jl:33-53, 72-88,
  109-125);
* closed forms on product tensors T[a,b,c,d] = P[a,d] A[b,c]: transfer
  matrix tr(P) A (unitcell 1) and tr(P^2) A (x) A (unitcell 2), X1 =
  tr(P)^2 / tr(P^2), X2 = tr(P)^2 tr(A)^2 / (tr(P^2) tr(A^2)), 2x2 torus
  tr(P^2)^2 tr(A^2)^2, central charge from |tr P| sigma_1(A); BTRG bond
  weights as diag(S2) P and diag(S1) A.  Each formula tells the traced leg
  pairs apart (a transposed pairing changes it);
* Z2 = Trivial for the Ising tensor.
"""
import numpy as np
import pytest

from oracle import models as om
from oracle import observables as ob
from oracle import schemes as osch
from oracle.ztensor import ZT

EXACT = [1 / 8, 1, 9 / 8]


@pytest.mark.parametrize("cls,chi,n,tol", [(osch.TRG, 24, 10, (2e-4, 1e-2)),
                                           (osch.BTRG, 24, 10, (3e-4, 2e-2)),
                                           (osch.HOTRG, 16, 4, (6e-4, 1e-2))])
def test_cft_data_reference_gates(cls, chi, n, tol):
    sc = cls(om.classical_ising_z2(om.ising_betac_mp))
    osch.run(sc, chi, n)
    c = ob.cft_data(sc)[1:]
    assert abs(c[0] / EXACT[0] - 1) < tol[0] and abs(c[1] / EXACT[1] - 1) < tol[1]


@pytest.mark.parametrize("cls,chi", [(osch.TRG, 16), (osch.BTRG, 16), (osch.HOTRG, 12)])
def test_gsd_and_gu_wen_reference_gates(cls, chi):
    for db, want in ((-0.01, 1.0), (0.01, 2.0)):
        sc = cls(om.classical_ising_z2(om.ising_betac + db))
        osch.run(sc, chi, 20)
        x1, x2 = ob.gu_wen_ratio(sc)
        for v in (ob.ground_state_degeneracy(sc), x1, x2):
            assert abs(v / want - 1) < 1e-2


def _product(rng, p, q):
    P = rng.uniform(0.2, 1.0, (p, p))
    A = rng.standard_normal((q, q)) + 2 * np.eye(q)
    return P, A, np.einsum("ad,bc->abcd", P, A)


def same_multiset(a, b, tol):
    """Every value of a matched to a distinct value of b within tol (spectra
    with exact modulus ties have no roundoff-free order)."""
    a, b = list(np.asarray(a, dtype=complex)), list(np.asarray(b, dtype=complex))
    if len(a) != len(b):
        return False
    for x in a:
        j = int(np.argmin([abs(x - y) for y in b]))
        if abs(x - b[j]) > tol * max(1.0, abs(x)):
            return False
        b.pop(j)
    return True


class _Plain:
    def __init__(self, T):
        self.T = ZT(np.array(T, copy=True))


def test_closed_forms_on_product_tensors():
    rng = np.random.default_rng(11)
    P, A, T = _product(rng, 3, 5)
    sc = _Plain(T)
    trP, trP2, trA, trA2 = np.trace(P), np.trace(P @ P), np.trace(A), np.trace(A @ A)
    ev = ob.sort_spectrum(trP * np.linalg.eigvals(A))
    assert same_multiset(ob.transfer_spectrum(sc, 1), ev, 1e-13)
    lam = np.linalg.eigvals(A)
    ev2 = ob.sort_spectrum(trP2 * np.outer(lam, lam).ravel())
    assert same_multiset(ob.transfer_spectrum(sc, 2), ev2, 1e-12)
    x1, x2 = ob.gu_wen_ratio(sc)
    assert abs(x1 / (trP ** 2 / trP2) - 1) < 1e-13
    assert abs(x2 / (trP ** 2 * trA ** 2 / (trP2 * trA2)) - 1) < 1e-13
    s1 = np.linalg.svd(A, compute_uv=False)[0]
    assert abs(ob.central_charge(sc, 1.7) - np.log(abs(trP) * s1 / 1.7) * 6 / np.pi) < 1e-13
    D = np.linalg.eigvals(trP * A)
    D = np.abs(D / D.sum())
    assert abs(ob.ground_state_degeneracy(sc) - np.exp(-np.sum(D * np.log(D)))) < 1e-13
    r = ob.finalize_two_by_two(sc)
    assert abs(r / np.sqrt(abs(trP2 * trA2)) - 1) < 1e-13
    assert np.allclose(sc.T.data * r, T, rtol=1e-15, atol=0)


def test_btrg_weights_enter_as_diagonal_factors():
    rng = np.random.default_rng(12)
    P, A, T = _product(rng, 4, 3)
    sc = osch.BTRG(ZT(np.array(T, copy=True)))
    s1, s2 = rng.uniform(0.5, 2.0, 3), rng.uniform(0.5, 2.0, 4)
    sc.S1 = ZT(np.diag(s1), sc.S1.q, 1)
    sc.S2 = ZT(np.diag(s2), sc.S2.q, 1)
    Pw, Aw = np.diag(s2) @ P, np.diag(s1) @ A      # T_unit = S2[a] S1[b] T[a,b,c,d]
    x1, x2 = ob.gu_wen_ratio(sc)
    assert abs(x1 / (np.trace(Pw) ** 2 / np.trace(Pw @ Pw)) - 1) < 1e-13
    assert abs(x2 / (np.trace(Pw) ** 2 * np.trace(Aw) ** 2 /
                     (np.trace(Pw @ Pw) * np.trace(Aw @ Aw))) - 1) < 1e-13
    assert same_multiset(ob.transfer_spectrum(sc), np.trace(Pw) * np.linalg.eigvals(Aw), 1e-13)
    # central charge: S2 on the traced pair, S1 on the top leg (cft.jl:263-266)
    M = np.trace(np.diag(s2) @ P) * A @ np.diag(s1)
    want = np.log(np.linalg.svd(M, compute_uv=False)[0] / 2.0) * 6 / np.pi
    assert abs(ob.central_charge(sc, 2.0) - want) < 1e-13
    # 2x2 torus (finalize.jl:29-38): four weights on eight bonds, one in
    # each loop of two P (S2) and of two A (S1): n = tr(S2 P P)^2 tr(S1 A A)^2
    want = np.sqrt(abs(np.trace(np.diag(s2) @ P @ P) * np.trace(np.diag(s1) @ A @ A)))
    assert abs(ob.finalize_two_by_two(sc) / want - 1) < 1e-13


def test_z2_equals_trivial():
    for cls in (osch.TRG, osch.HOTRG):
        a = cls(om.classical_ising_z2(0.42))
        b = cls(om.classical_ising_trivial(0.42))
        osch.run(a, 8, 4)
        osch.run(b, 8, 4)
        ca, cb = ob.cft_data(a), ob.cft_data(b)
        m = min(len(ca), len(cb), 8)
        assert np.allclose(ca[:m], cb[:m], rtol=1e-8, atol=1e-9)
        assert abs(ob.ground_state_degeneracy(a) / ob.ground_state_degeneracy(b) - 1) < 1e-9
        assert np.allclose(ob.gu_wen_ratio(a), ob.gu_wen_ratio(b), rtol=1e-10)
