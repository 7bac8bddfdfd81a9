"""Package model constructors vs the oracle's (CPU only)."""
import numpy as np
import pytest

from oracle import models as om
from conftest import to_sym


def same(tk, a, zt):
    b = to_sym(tk, zt)
    assert a.N == b.N
    assert np.array_equal(a.leg_dims, b.leg_dims)
    assert np.max(np.abs(a.data - b.data)) <= 4e-16 * np.max(np.abs(b.data))


def test_ising(tk):
    same(tk, tk.classical_ising(symmetry="Z2"), om.classical_ising_z2(om.ising_betac_mp))
    same(tk, tk.classical_ising(symmetry="Trivial"), om.classical_ising_trivial(om.ising_betac_mp))
    same(tk, tk.classical_ising(0.4, symmetry="Z2"), om.classical_ising_z2(0.4))
    same(tk, tk.classical_ising(0.4, symmetry="Trivial"), om.classical_ising_trivial(0.4))


@pytest.mark.parametrize("q", [2, 3, 4])
def test_potts(tk, q):
    b = om.potts_betac(q)
    same(tk, tk.classical_potts(q, b), om.classical_potts_zn(q, b))
    same(tk, tk.classical_potts(q, b, symmetry="Trivial"), om.classical_potts_trivial(q, b))


@pytest.mark.parametrize("q,beta", [(3, 0.8), (6, 1.1)])
def test_clock(tk, q, beta):
    same(tk, tk.classical_clock(q, beta), om.classical_clock_zn(q, beta))
    same(tk, tk.classical_clock(q, beta, symmetry="Trivial"), om.classical_clock_trivial(q, beta))


def test_block_roundtrip(tk):
    t = om.classical_clock_zn(6, 1.1)
    s = to_sym(tk, t)
    dense = s.to_dense()
    assert np.array_equal(dense, t.data)
