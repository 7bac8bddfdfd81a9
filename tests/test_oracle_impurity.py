"""The impurity schemes of the oracle against every pin available (CPU only).

* test/schemes.jl:399-410, 441-450: ImpurityHOTRG (chi 16) and ImpurityTRG
  (chi 24) free energies vs Onsager at 6e-7 and 2e-6.
* test/schemes.jl:412-438, 452-475: magnetisation limits, <m^2> -> 0 / 1
  (ImpurityHOTRG) and <m> -> 0 / 1 (ImpurityTRG, h = 1e-6 at low T).
* Exact spin sums: before any step the tensors are single Ising spins, so
  npure = Z(1-site torus) and nimp = Z(2x2 torus with sigma_1) / npure^4,
  enumerated over all spin configurations (independent of the networks).
* With every impurity tensor equal to T the impurity schemes reduce to TRG
  and to HOTRG with the y projector first.
* The pivoted-row SVD gauge (ztensor.pivoted_gauge) used by the device
  ImpurityTRG: a function of each cluster's span (scrambled or sign-flipped
  input bases give the same vectors), Z2 parities kept apart, the same
  impurity norms from two LAPACK drivers, the reference's gates passed, and
  <m> = 0 to roundoff at h = 0, where LAPACK's basis leaks up to 1e-3.
"""
import itertools
import math

import numpy as np
import pytest

from oracle import models as om
from oracle import schemes as osch


def _mag(beta, h=0.0):
    return om.classical_ising_trivial(beta, h), om.classical_ising_impurity(beta, h)


def test_impurity_trg_free_energy_gate():
    T, Ti = _mag(om.ising_betac_mp)
    d = osch.run(osch.ImpurityTRG(T, Ti, T, T, T), 24, 25)
    f = osch.free_energy([x[0] for x in d], om.ising_betac)
    assert abs(f / om.f_onsager - 1) < 2e-6          # test/schemes.jl:449


def test_impurity_hotrg_free_energy_gate():
    T, Ti = _mag(om.ising_betac_mp)
    d = osch.run(osch.ImpurityHOTRG(T, Ti, Ti, T), 16, 25)
    f = osch.free_energy([x[0] for x in d], om.ising_betac, scalefactor=4.0)
    assert abs(f / om.f_onsager - 1) < 6e-7          # test/schemes.jl:409


def test_impurity_trg_magnetisation_limits():
    T, Ti = _mag(0.1)
    d = osch.run(osch.ImpurityTRG(T, Ti, T, T, T), 16, 25)
    assert abs(d[-1][1] / d[-1][0]) < 1e-4           # test/schemes.jl:463
    T, Ti = _mag(2.0, 1e-6)
    d = osch.run(osch.ImpurityTRG(T, Ti, T, T, T), 16, 25)
    assert abs(d[-1][1] / d[-1][0] - 1) < 1e-4       # test/schemes.jl:474


def test_impurity_hotrg_m2_limits():
    T, Ti = _mag(0.2)
    d = osch.run(osch.ImpurityHOTRG(T, Ti, Ti, T), 8, 25)
    assert abs(d[-1][3] / d[-1][0]) < 1e-14          # test/schemes.jl:424
    T, Ti = _mag(1.0)
    d = osch.run(osch.ImpurityHOTRG(T, Ti, Ti, T), 8, 25)
    assert abs(d[-1][3] / d[-1][0] - 1) < 1e-2       # test/schemes.jl:437


@pytest.mark.parametrize("beta,h", [(0.3, 0.0), (0.44, 0.2), (0.7, -0.5)])
def test_first_finalize_matches_spin_sums(beta, h):
    T, Ti = _mag(beta, h)
    # one spin: two self bonds (e^{2b}) and the field
    z1 = sum(math.exp(2 * beta) * math.exp(h * beta * s) for s in (1, -1))
    # 2x2 torus, sites 1 2 / 4 3, every neighbour pair joined by two bonds,
    # sigma_1 inserted at site 1
    zimp = 0.0
    for s in itertools.product((1, -1), repeat=4):
        e = 2 * beta * (s[0] * s[1] + s[1] * s[2] + s[2] * s[3] + s[3] * s[0])
        e += h * beta * sum(s)
        zimp += s[0] * math.exp(e)
    npure, nimp = osch.ImpurityTRG(T, Ti, T, T, T).finalize()
    assert abs(npure / z1 - 1) < 1e-14
    assert abs(nimp - abs(zimp) / z1 ** 4) <= 1e-14 * (abs(zimp) / z1 ** 4) + 1e-300
    n, n11, n12, n2 = osch.ImpurityHOTRG(T, Ti, Ti, T).finalize()
    z1_sigma = sum(s * math.exp(2 * beta) * math.exp(h * beta * s) for s in (1, -1))
    assert abs(n / z1 - 1) < 1e-14 and abs(n2 / z1 - 1) < 1e-14
    assert abs(n11 - abs(z1_sigma)) <= 1e-14 * z1


def test_impurity_schemes_reduce_to_pure_schemes():
    T = om.classical_ising_trivial(0.42)
    d = osch.run(osch.ImpurityTRG(T, T, T, T, T), 8, 6)
    ref = osch.run(osch.TRG(T), 8, 6)
    assert np.max(np.abs(np.array([x[0] for x in d]) / np.array(ref) - 1)) < 1e-13
    d = osch.run(osch.ImpurityHOTRG(T, T, T, T), 8, 5)
    for row in d:
        assert max(abs(v / row[0] - 1) for v in row) < 1e-12


def _ilt(beta, h, chi, n, gauge):
    T, Ti = _mag(beta, h)
    return np.array(osch.run(osch.ImpurityTRG(T, Ti, T, T, T, gauge=gauge), chi, n))


def test_pivoted_gauge_is_a_function_of_the_span():
    from oracle.ztensor import pivoted_gauge
    rng = np.random.default_rng(3)
    m, k = 30, 8
    s = np.array([1.0, 0.7, 0.7, 0.7, 0.4, 0.2, 0.2, 1e-17])
    U, _ = np.linalg.qr(rng.standard_normal((m, k)))
    Vt = np.linalg.qr(rng.standard_normal((m, k)))[0].T
    a_u, a_v = pivoted_gauge(U, s, Vt)
    R = np.eye(k)
    for i, j in ((1, 4), (5, 7)):                 # rotate inside each cluster
        R[i:j, i:j] = np.linalg.qr(rng.standard_normal((j - i, j - i)))[0]
    R[0, 0] = R[4, 4] = -1.0                      # and flip single vectors
    b_u, b_v = pivoted_gauge(U @ R, s, R.T @ Vt)
    assert np.max(np.abs(a_u - b_u)) < 1e-13 and np.max(np.abs(a_v - b_v)) < 1e-13
    # still an SVD of the same matrix (null triplet dropped), orthonormal
    M = (U[:, :7] * s[:7]) @ Vt[:7]
    assert np.max(np.abs((a_u * s) @ a_v - M)) < 1e-14
    assert np.max(np.abs(a_u[:, :7].T @ a_u[:, :7] - np.eye(7))) < 1e-13
    assert np.all(a_u[:, 7] == 0.0)
    # single values: positive at the largest entry; a cluster's first vector
    # is positive at the row where the cluster's span has the largest
    # weight, and the cluster's other vectors vanish at that row
    for j in (0, 4):
        assert a_u[np.argmax(np.abs(a_u[:, j])), j] > 0
    for i, j in ((1, 4), (5, 7)):
        r = np.argmax(np.sum(U[:, i:j] ** 2, axis=1))
        assert a_u[r, i] > 0 and np.max(np.abs(a_u[r, i + 1:j])) < 1e-14


def test_pivoted_gauge_keeps_z2_parities_apart():
    from oracle.ztensor import pivoted_gauge
    rng = np.random.default_rng(5)
    E = np.linalg.qr(rng.standard_normal((10, 3)))[0]
    O = np.linalg.qr(rng.standard_normal((10, 3)))[0]
    U = np.zeros((20, 6))
    U[:10, :3] = E
    U[10:, 3:] = O
    order = [0, 3, 1, 4, 2, 5]
    U = U[:, order]
    s = np.array([1.0, 1.0, 0.5, 0.5, 0.25, 0.25])
    R = np.eye(6)
    for i in (0, 2, 4):
        R[i:i + 2, i:i + 2] = np.linalg.qr(rng.standard_normal((2, 2)))[0]
    u, _ = pivoted_gauge(U @ R, s, np.eye(6))
    for j in range(6):
        assert min(np.linalg.norm(u[:10, j]), np.linalg.norm(u[10:, j])) < 1e-14


def test_pivoted_impurity_trg_gates_and_symmetry():
    d = _ilt(0.1, 0.0, 16, 25, "pivoted")
    assert abs(d[-1][1] / d[-1][0]) < 1e-4              # test/schemes.jl:463
    d = _ilt(2.0, 1e-6, 16, 25, "pivoted")
    assert abs(d[-1][1] / d[-1][0] - 1) < 1e-4          # test/schemes.jl:474
    T, Ti = _mag(om.ising_betac_mp)
    d = osch.run(osch.ImpurityTRG(T, Ti, T, T, T, gauge="pivoted"), 24, 25)
    f = osch.free_energy([x[0] for x in d], om.ising_betac)
    assert abs(f / om.f_onsager - 1) < 2e-6             # test/schemes.jl:449
    # h = 0: <m> = 0 by the Z2 symmetry; the pivoted gauge keeps it at
    # roundoff for the first steps, LAPACK's basis does not
    piv = _ilt(0.44, 0.0, 16, 7, "pivoted")
    lap = _ilt(0.44, 0.0, 16, 7, "lapack")
    assert np.max(np.abs(piv[:, 1] / piv[:, 0])) < 1e-14
    assert np.max(np.abs(lap[:, 1] / lap[:, 0])) > 1e-6
    assert np.max(np.abs(piv[:, 0] / lap[:, 0] - 1)) < 1e-13


def test_pivoted_impurity_norms_do_not_depend_on_the_lapack_driver(monkeypatch):
    import scipy.linalg as sla
    import oracle.ztensor as zt
    for beta, h, chi, n in ((0.44, 0.1, 8, 8), (0.6, 0.05, 16, 6)):
        a = _ilt(beta, h, chi, n, "pivoted")
        with monkeypatch.context() as mp:
            mp.setattr(zt.np.linalg, "svd", lambda M, full_matrices=False:
                       sla.svd(M, full_matrices=False, lapack_driver="gesvd"))
            b = _ilt(beta, h, chi, n, "pivoted")
            c = _ilt(beta, h, chi, n, "lapack")
        assert np.max(np.abs(a / b - 1)) < 1e-12
        # LAPACK's own basis is not pinned by the tensors: the two drivers
        # give different impurity norms (the npure column agrees)
        ref = _ilt(beta, h, chi, n, "lapack")
        assert np.max(np.abs(c[:, 0] / ref[:, 0] - 1)) < 1e-13
