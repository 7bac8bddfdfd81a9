"""Observables on the device (src/utility/cft.jl, finalize.jl:17-41) through
the C ABI: the eigenvalue kernel against LAPACK, closed forms on planted
product tensors, every observable against the oracle after real runs, and
the reference's own observable tests on the device."""
import ctypes as C

import numpy as np
import pytest

from conftest import to_sym
from oracle import models as om
from oracle import observables as ob
from oracle import schemes as osch
from oracle.ztensor import ZT
from test_oracle_observables import same_multiset

pytestmark = pytest.mark.gpu
D = C.POINTER(C.c_double)


def p(a):
    return a.ctypes.data_as(D)


def kernel_eig(tk, As):
    cnt, n = As.shape[0], As.shape[1]
    A = np.ascontiguousarray(np.stack([np.asfortranarray(a).ravel(order="F") for a in As]))
    wr = np.zeros(cnt * n)
    wi = np.zeros(cnt * n)
    st = tk.lib().tnr_kernel_eig(cnt, n, p(A), p(wr), p(wi))
    assert st == 0, tk.lib().tnr_last_error()
    return (wr + 1j * wi).reshape(cnt, n)


# above 128 the matrix lives in global memory (the same solver); the
# reference's eig_full (cft.jl:16) has no size limit
@pytest.mark.parametrize("n", [1, 2, 3, 17, 64, 96, 128, 129, 160, 300])
def test_eig_kernel_vs_lapack(tk, n):
    rng = np.random.default_rng(n)
    As = [rng.standard_normal((n, n))]
    Dg = np.diag(np.logspace(0, 5, n))
    As.append(Dg @ rng.standard_normal((n, n)) @ np.linalg.inv(Dg))      # badly balanced
    Z = rng.standard_normal((n, n))
    if n > 2:
        Z[0, :] = 0.0
        Z[:, 1] = 0.0
    As.append(Z)
    got = kernel_eig(tk, np.stack(As))
    for A, g in zip(As, got):
        ref = np.linalg.eigvals(A)
        assert same_multiset(g, ref, 1e-10 * max(1.0, np.max(np.abs(ref)))), (g, ref)


def test_eig_kernel_special_matrices(tk):
    cases = [np.zeros((4, 4)), np.eye(5), np.array([[0.0, -1.0], [1.0, 0.0]]),
             np.diag([3.0, 1.0, 2.0]), np.array([[2.0, 1.0], [0.0, 2.0]])]
    for A in cases:
        g = kernel_eig(tk, A[None])[0]
        ref = np.linalg.eigvals(A)
        tol = 1e-7 if A.shape == (2, 2) and A[1, 0] == 0 else 1e-14   # Jordan block: sqrt(eps)
        assert same_multiset(g, ref, tol)


def _product(rng, p_, q):
    P = rng.uniform(0.2, 1.0, (p_, p_))
    A = rng.standard_normal((q, q)) + 2 * np.eye(q)
    return P, A, np.einsum("ad,bc->abcd", P, A)


@pytest.mark.parametrize("kind", ["TRG", "BTRG", "HOTRG"])
def test_closed_forms_on_product_tensors(tk, kind):
    """T[a,b,c,d] = P[a,d] A[b,c] (bond weights start at 1): the transfer
    matrix is tr(P) A, twice tr(P^2) A (x) A, and the traces follow."""
    rng = np.random.default_rng(21)
    P, A, T = _product(rng, 3, 5)
    sc = getattr(tk, kind)(tk.trivial(T))
    trP, trP2, trA, trA2 = np.trace(P), np.trace(P @ P), np.trace(A), np.trace(A @ A)
    lam = np.linalg.eigvals(A)
    assert same_multiset(tk.transfer_spectrum(sc, 1), trP * lam, 1e-12)
    assert same_multiset(tk.transfer_spectrum(sc, 2), trP2 * np.outer(lam, lam).ravel(), 1e-11)
    ref = ob.cft_data(type("S", (), {"T": ZT(T)})(), is_real=False)
    got = tk.cft_data(sc, is_real=False)
    assert same_multiset(got, ref, 1e-11)
    x1, x2 = tk.gu_wen_ratio(sc)
    assert abs(x1 / (trP ** 2 / trP2) - 1) < 1e-12
    assert abs(x2 / (trP ** 2 * trA ** 2 / (trP2 * trA2)) - 1) < 1e-12
    s1 = np.linalg.svd(A, compute_uv=False)[0]
    assert abs(tk.central_charge(sc, 1.7) - np.log(abs(trP) * s1 / 1.7) * 6 / np.pi) < 1e-12
    Dv = np.linalg.eigvals(trP * A)
    Dv = np.abs(Dv / Dv.sum())
    assert abs(tk.ground_state_degeneracy(sc) - np.exp(-np.sum(Dv * np.log(Dv)))) < 1e-12
    r = tk.finalize_two_by_two(sc)
    assert abs(r / np.sqrt(abs(trP2 * trA2)) - 1) < 1e-12
    assert np.max(np.abs(sc.T.to_dense() * r - T)) < 1e-13 * np.max(np.abs(T))
    sc.close()


def _pair(tk, cls_name, sym, beta, chi, n):
    T = om.classical_ising_z2(beta) if sym == "Z2" else om.classical_ising_trivial(beta)
    o = getattr(osch, cls_name)(T)
    osch.run(o, chi, n)
    g = getattr(tk, cls_name)(tk.classical_ising(None if beta is om.ising_betac_mp else float(beta),
                                                  symmetry=sym))
    d = tk.run(g, tk.truncrank(chi), tk.maxiter(n), verbosity=0)
    return g, o, d


def _big(ev):
    """eigenvalues above 1e-6 of the largest (the rest is truncation noise,
    where two runs that agree to 1e-10 need not agree relatively)"""
    ev = np.asarray(ev)
    return ev[np.abs(ev) > 1e-6 * np.abs(ev[0])]


@pytest.mark.parametrize("cls_name", ["TRG", "BTRG", "HOTRG"])
@pytest.mark.parametrize("sym", ["Z2", "Trivial"])
def test_observables_vs_oracle_after_runs(tk, cls_name, sym):
    """Gauge-invariant observables of the same run: device vs oracle.  The
    run norms agree to 1e-10; eigenvalues to 1e-8 of the largest."""
    g, o, d = _pair(tk, cls_name, sym, om.ising_betac_mp, 16, 6)
    ev_g, ev_o = tk.transfer_spectrum(g), ob.transfer_spectrum(o)
    assert same_multiset(_big(ev_g), _big(ev_o), 1e-8)
    cg, co = tk.cft_data(g), ob.cft_data(o)
    assert np.max(np.abs(cg[:4] - co[:4])) < 1e-7
    assert abs(tk.ground_state_degeneracy(g) / ob.ground_state_degeneracy(o) - 1) < 1e-8
    assert np.allclose(tk.gu_wen_ratio(g), ob.gu_wen_ratio(o), rtol=1e-8)
    assert abs(tk.central_charge(g, d[-1]) - ob.central_charge(o, d[-1])) < 1e-8
    g.close()


def test_unitcell_two_spectrum_vs_oracle(tk):
    g, o, _ = _pair(tk, "TRG", "Z2", om.ising_betac_mp, 8, 4)
    ev_g, ev_o = tk.transfer_spectrum(g, 2), ob.transfer_spectrum(o, 2)
    assert same_multiset(_big(ev_g), _big(ev_o), 1e-8)
    assert abs(tk.ground_state_degeneracy(g, 2) / ob.ground_state_degeneracy(o, 2) - 1) < 1e-8
    g.close()


def test_unitcell_two_spectrum_above_128_vs_oracle(tk):
    """Trivial TRG at chi = 12: the two-site transfer matrix is one 144 x 144
    block, solved with the matrix in global memory (cft.jl:16, 296)."""
    g, o, _ = _pair(tk, "TRG", "Trivial", om.ising_betac_mp, 12, 4)
    ev_g, ev_o = tk.transfer_spectrum(g, 2), ob.transfer_spectrum(o, 2)
    assert len(ev_g) == 144
    assert same_multiset(_big(ev_g), _big(ev_o), 1e-8)
    cg, co = tk.cft_data(g, unitcell=2), ob.cft_data(o, unitcell=2)
    assert np.max(np.abs(cg[:4] - co[:4])) < 1e-7
    g.close()


@pytest.mark.parametrize("cls_name", ["TRG", "BTRG", "HOTRG"])
def test_two_by_two_finalizer_run_vs_oracle(tk, cls_name):
    """run!(scheme, truncrank(chi), maxiter(n), two_by_two_Finalizer): every
    finalize returns n^(1/4) of the 2x2 torus and rescales T by it."""
    chi, n = 12, 5
    o = getattr(osch, cls_name)(om.classical_ising_z2(om.ising_betac_mp))
    ref = [ob.finalize_two_by_two(o)]
    for _ in range(n):
        o.step(chi)
        ref.append(ob.finalize_two_by_two(o))
    g = getattr(tk, cls_name)(tk.classical_ising(symmetry="Z2"))
    got = tk.run(g, tk.truncrank(chi), tk.maxiter(n), tk.two_by_two_Finalizer, verbosity=0)
    g.close()
    assert np.max(np.abs(np.asarray(got) / np.asarray(ref) - 1)) < 1e-10


@pytest.mark.parametrize("cls_name,chi,n,tol", [("TRG", 24, 10, (2e-4, 1e-2)),
                                                ("BTRG", 24, 10, (3e-4, 2e-2)),
                                                ("HOTRG", 16, 4, (6e-4, 1e-2))])
def test_reference_cft_gates(tk, cls_name, chi, n, tol):
    """test/schemes.jl:25-31, 64-70, 101-107 on the device."""
    g = getattr(tk, cls_name)(tk.classical_ising(symmetry="Z2"))
    tk.run(g, tk.truncrank(chi), tk.maxiter(n), verbosity=0)
    c = tk.cft_data(g)[1:]
    g.close()
    assert abs(c[0] / tk.ising_cft_exact[0] - 1) < tol[0]
    assert abs(c[1] / tk.ising_cft_exact[1] - 1) < tol[1]


@pytest.mark.parametrize("cls_name,chi", [("TRG", 16), ("BTRG", 16), ("HOTRG", 12)])
def test_reference_gsd_gu_wen_gates(tk, cls_name, chi):
    """test/schemes.jl:33-53, 72-88, 109-125 on the device."""
    for db, want in ((-0.01, 1.0), (0.01, 2.0)):
        g = getattr(tk, cls_name)(tk.classical_ising(tk.ising_betac + db, symmetry="Z2"))
        tk.run(g, tk.truncrank(chi), tk.maxiter(20), verbosity=0)
        x1, x2 = tk.gu_wen_ratio(g)
        gsd = tk.ground_state_degeneracy(g)
        g.close()
        for v in (gsd, x1, x2):
            assert abs(v / want - 1) < 1e-2


@pytest.mark.parametrize("cls_name", ["TRG", "BTRG", "HOTRG"])
def test_observables_z3_potts_vs_oracle(tk, cls_name):
    """Z3 sectors are not self-dual: the traces pair a leg with the dual of
    another (the torus through charge-conjugated permutes)."""
    beta, chi, n = 1.0, 12, 4
    o = getattr(osch, cls_name)(om.classical_potts_zn(3, beta))
    osch.run(o, chi, n)
    g = getattr(tk, cls_name)(tk.classical_potts(3, beta))
    d = tk.run(g, tk.truncrank(chi), tk.maxiter(n), verbosity=0)
    assert same_multiset(_big(tk.transfer_spectrum(g)), _big(ob.transfer_spectrum(o)), 1e-8)
    assert abs(tk.ground_state_degeneracy(g) / ob.ground_state_degeneracy(o) - 1) < 1e-8
    assert np.allclose(tk.gu_wen_ratio(g), ob.gu_wen_ratio(o), rtol=1e-8)
    assert abs(tk.central_charge(g, d[-1]) - ob.central_charge(o, d[-1])) < 1e-8
    assert abs(tk.finalize_two_by_two(g) / ob.finalize_two_by_two(o) - 1) < 1e-8
    g.close()
