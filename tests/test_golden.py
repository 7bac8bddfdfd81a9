"""The oracle against the committed golden fixtures (CPU only).

* onsager_table.json: Onsager's closed form at 30 digits (mpmath) on a beta
  grid pins oracle/onsager.py away from beta_c as well.
* reference_values.json: the reference's published numbers (README.md:87,
  src/models/ising.jl:2) pin the oracle's whole BTRG pipeline and constants.
* oracle_norms.json: stored oracle norm vectors; the oracle must reproduce
  them (a change of the oracle shows up here before it reaches the GPU
  parity tests, which compare against the same vectors).
"""
import json
import os

import numpy as np
import pytest

from oracle import models as om
from oracle import onsager
from oracle import schemes as osch

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def test_onsager_table():
    t = load("onsager_table.json")
    for b, f in zip(t["beta"], t["f"]):
        assert abs(onsager.f_onsager(b) / f - 1) < 1e-13, b


def test_reference_values():
    r = load("reference_values.json")
    assert abs(float(r["f_onsager_betac"]["value"]) / om.f_onsager - 1) < 1e-16
    sc = osch.BTRG(om.classical_ising_z2(om.ising_betac_mp))
    f = osch.free_energy(osch.run(sc, 16, 25), om.ising_betac)
    assert abs(f / float(r["readme_btrg_z2_chi16_maxiter25_f"]["value"]) - 1) < 1e-13


def _model(spec_name, beta):
    if spec_name == "ising_z2":
        return om.classical_ising_z2(om.ising_betac_mp if beta is None else beta)
    if spec_name == "ising_trivial":
        return om.classical_ising_trivial(beta)
    if spec_name == "potts_z3":
        return om.classical_potts_zn(3, om.potts_betac(3) if beta is None else beta)
    if spec_name == "clock_z6":
        return om.classical_clock_zn(6, beta)
    raise ValueError(spec_name)


@pytest.mark.parametrize("name", sorted(load("oracle_norms.json")))
def test_oracle_reproduces_stored_norms(name):
    case = load("oracle_norms.json")[name]
    cls = {"TRG": osch.TRG, "BTRG": osch.BTRG, "HOTRG": osch.HOTRG}[case["scheme"]]
    d = osch.run(cls(_model(case["model"], case["beta"])), case["chi"], case["maxiter"])
    ref = np.array(case["norms"])
    assert len(d) == len(ref)
    # the oracle is deterministic up to the host BLAS's threading
    assert np.max(np.abs(np.array(d) - ref) / np.abs(ref)) < 1e-12


# ------------------------------------------------- schemes beyond TRG/BTRG/HOTRG
def test_oracle_reproduces_stored_extensions():
    """oracle_extensions.json: ATRG and HOTRG_3D norms, ImpurityTRG (pivoted
    gauge) and ImpurityHOTRG rows, observables after fixed runs."""
    from oracle import observables as ob
    g = load("oracle_extensions.json")
    c = g["atrg_z2_ising_betac_chi12_8"]
    d = osch.run(osch.ATRG(om.classical_ising_z2(om.ising_betac_mp)), c["chi"], c["maxiter"])
    assert np.max(np.abs(np.array(d) / np.array(c["norms"]) - 1)) < 1e-12
    c = g["hotrg3d_z2_ising3d_betac_chi6_3"]
    d = osch.run(osch.HOTRG_3D(om.classical_ising_3d_z2(om.ising_betac_3d)), c["chi"], c["maxiter"])
    assert np.max(np.abs(np.array(d) / np.array(c["norms"]) - 1)) < 1e-12
    c = g["impurity_trg_b0.44_h0.1_chi8_6"]
    T, Ti = om.classical_ising_trivial(c["beta"], c["h"]), om.classical_ising_impurity(c["beta"], c["h"])
    d = osch.run(osch.ImpurityTRG(T, Ti, T, T, T, gauge="pivoted"), c["chi"], c["maxiter"])
    assert np.max(np.abs(np.array(d) / np.array(c["rows"]) - 1)) < 1e-12
    c = g["impurity_hotrg_b0.5_h0.1_chi8_4"]
    T, Ti = om.classical_ising_trivial(c["beta"], c["h"]), om.classical_ising_impurity(c["beta"], c["h"])
    d = osch.run(osch.ImpurityHOTRG(T, Ti, Ti, T), c["chi"], c["maxiter"])
    assert np.max(np.abs(np.array(d) / np.array(c["rows"]) - 1)) < 1e-12
    for kind in ("TRG", "BTRG"):
        c = g[f"observables_{kind.lower()}_z2_ising_betac_chi16_6"]
        sc = {"TRG": osch.TRG, "BTRG": osch.BTRG}[kind](om.classical_ising_z2(om.ising_betac_mp))
        d = osch.run(sc, c["chi"], c["maxiter"])
        assert np.max(np.abs(ob.cft_data(sc)[:6] - np.array(c["cft_data_first6"]))) < 1e-10
        assert abs(ob.ground_state_degeneracy(sc) / c["ground_state_degeneracy"] - 1) < 1e-10
        assert np.allclose(ob.gu_wen_ratio(sc), c["gu_wen_ratio"], rtol=1e-12)
        assert abs(ob.central_charge(sc, d[-1]) - c["central_charge_n_last"]) < 1e-12
