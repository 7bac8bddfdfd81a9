"""The FP64 oracle against every pin the reference offers (CPU only).

* README.md:87 - BTRG, Z2 Ising at beta_c, truncrank(16), maxiter(25):
  f = -2.1096504926141826902647832 (the only published number).
* test/schemes.jl:18-23, 57-62, 94-99 - TRG / BTRG / HOTRG vs Onsager.
* src/models/ising.jl:2 - the exact f_onsager constant vs the closed form.
* Exact small-torus partition functions for the un-truncated first steps of
  all three schemes, on random (symmetry-free) tensors, so a transposed
  operand or a wrong index in any network changes the numbers.
* Z2 vs Trivial Ising, Z_N vs Trivial Potts / clock (same lattice model).
* test/models.jl:9-16 - Potts q=3 and clock q=3/q=4 free energies.
"""
import mpmath
import numpy as np
import pytest

from oracle import models as om
from oracle import onsager, schemes, torus
from oracle.ztensor import ZT, contract

from conftest import rel

README_F = float(mpmath.mpf("-2.1096504926141826902647832"))


def test_onsager_closed_form_matches_reference_constant():
    assert abs(onsager.f_onsager(om.ising_betac) / om.f_onsager - 1) < 1e-15


def test_readme_btrg_free_energy():
    sc = schemes.BTRG(om.classical_ising_z2(om.ising_betac_mp))
    data = schemes.run(sc, 16, 25)
    assert len(data) == 26
    f = schemes.free_energy(data, om.ising_betac)
    assert abs(f / README_F - 1) < 1e-13
    # README.md:93: |f - f_onsager| / |f_onsager| = 3.1e-07
    assert abs(abs(f / om.f_onsager - 1) - 3.09e-7) < 0.01e-7


@pytest.mark.parametrize("cls,chi,rtol,sf", [
    (schemes.TRG, 24, 2e-6, 2.0),      # test/schemes.jl:20-23
    (schemes.BTRG, 24, 6e-8, 2.0),     # test/schemes.jl:59-62
    (schemes.HOTRG, 16, 6e-7, 4.0),    # test/schemes.jl:96-99
])
def test_reference_onsager_tolerances(cls, chi, rtol, sf):
    sc = cls(om.classical_ising_z2(om.ising_betac_mp))
    data = schemes.run(sc, chi, 25)
    f = schemes.free_energy(data, om.ising_betac, scalefactor=sf)
    assert abs(f / om.f_onsager - 1) < rtol


@pytest.mark.parametrize("seed", [1, 2])
def test_exact_first_steps_on_random_tensors(seed):
    rng = np.random.default_rng(seed)
    T = rng.standard_normal((2, 2, 2, 2)) + 1.0
    z1, zt2 = torus.z_1site(T), torus.z_tilted2(T)
    z22, z44 = torus.z_torus(T, 2), torus.z_torus(T, 4)
    for cls in (schemes.TRG, schemes.BTRG):
        n = schemes.run(cls(ZT(T)), 16, 2)
        assert abs(n[0] / abs(z1) - 1) < 1e-13
        assert abs(n[0] ** 2 * n[1] / abs(zt2) - 1) < 1e-12
        assert abs(n[0] ** 4 * n[1] ** 2 * n[2] / abs(z22) - 1) < 1e-12
    n = schemes.run(schemes.HOTRG(ZT(T)), 16, 2)
    assert abs(n[0] ** 4 * n[1] / abs(z22) - 1) < 1e-12
    assert abs(n[0] ** 16 * n[1] ** 4 * n[2] / abs(z44) - 1) < 1e-11


def test_torus_pin_catches_a_planted_transpose():
    """A TRG whose second factorization uses the wrong bipartition must fail
    the exact 2-site torus check (the pin is sensitive to index mistakes)."""
    rng = np.random.default_rng(3)
    T = rng.standard_normal((2, 2, 2, 2)) + 1.0

    class BadTRG(schemes.TRG):
        def step(self, chi):
            from oracle.ztensor import permute
            A, B = schemes.svd12(self.T, chi)
            Tp = permute(self.T, (3, 1, 0, 2))         # planted: legs 2 and 4 swapped
            C, D = schemes.svd12(Tp, chi)
            X = contract([-1, -2, -3, -4], (D, [-2, 1, -3], False), (B, [-1, -4, 1], False))
            Y = contract([-1, -2, -3, -4], (C, [-1, 3, -3], False), (A, [3, -2, -4], False))
            self.T = contract([-1, -2, -3, -4], (X, [-1, -2, 2, 4], False), (Y, [4, 2, -3, -4], False))
            return self

    n = schemes.run(BadTRG(ZT(T)), 16, 1)
    assert abs(n[0] ** 2 * n[1] / abs(torus.z_tilted2(T)) - 1) > 1e-6


def test_z2_equals_trivial_ising():
    a = schemes.run(schemes.BTRG(om.classical_ising_z2(om.ising_betac_mp)), 16, 12)
    b = schemes.run(schemes.BTRG(om.classical_ising_trivial(om.ising_betac_mp)), 16, 12)
    assert rel(a, b) < 1e-11


def test_spin_delta_gauge_gives_same_norms():
    b = om.ising_betac
    a = schemes.run(schemes.TRG(om.ising_spin_delta(b)), 8, 8)
    c = schemes.run(schemes.TRG(om.classical_ising_trivial(b)), 8, 8)
    assert rel(a, c) < 1e-11


@pytest.mark.parametrize("make_zn,make_tr,q,beta,ref", [
    (om.classical_potts_zn, om.classical_potts_trivial, 3, om.potts_betac(3), -4.119552029995684),
    (om.classical_clock_zn, om.classical_clock_trivial, 3, 2 * np.log(np.sqrt(3) + 1) / 3, -4.17924244901635),
    (om.classical_clock_zn, om.classical_clock_trivial, 4, np.log(np.sqrt(2) + 1), 2 * om.f_onsager),
])
def test_zn_models_match_trivial_and_reference(make_zn, make_tr, q, beta, ref):
    """test/models.jl:9-16 with the reference's rtol 1e-3 (TRG, chi=16, 25 steps)."""
    tz, tt = make_zn(q, beta), make_tr(q, beta)
    assert tz.check_symmetric() == 0.0
    a = schemes.run(schemes.TRG(tz), 16, 25)
    b = schemes.run(schemes.TRG(tt), 16, 25)
    assert abs(schemes.free_energy(a, beta) / ref - 1) < 1e-3
    # the Z_N and Trivial runs are the same lattice model: the first steps
    # are un-truncated and agree to roundoff
    assert rel(a[:3], b[:3]) < 1e-12


def test_charge_flow_check_rejects_wrong_pairing():
    t = om.classical_potts_zn(3, 1.0)
    with pytest.raises(ValueError):
        contract([-1, -2, -3, -4, -5, -6], (t, [-1, 1, -2, -3], False), (t, [-4, 1, -5, -6], False))


def test_high_temperature_limit():
    sc = schemes.TRG(om.classical_ising_trivial(1e-12))
    d = schemes.run(sc, 4, 6)
    assert abs(-schemes.free_energy(d, 1.0) - np.log(2)) < 1e-9
