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
