"""Exact free energy of the 2D square-lattice Ising model (Onsager).

-beta f = ln(2)/2 + (1/2pi) int_0^pi ln[cosh^2(2b) + (1/k) sqrt(1 + k^2 - 2k cos 2t)] dt,
k = 1/sinh^2(2b).  Pinned against the reference constant f_onsager at
beta_c (src/models/ising.jl:2), reproduced to ~1e-16 in
tests/test_oracle_pins.py, and against a 30-digit mpmath evaluation of the
same closed form on a beta grid (tests/golden/onsager_table.json).
Test infrastructure: only tests, the smoke check and bench.py's CPU baseline
use this package.
"""
from __future__ import annotations

import warnings

import numpy as np
from scipy.integrate import quad


def f_onsager(beta: float) -> float:
    b = float(beta)
    k = 1.0 / np.sinh(2 * b) ** 2
    c2 = np.cosh(2 * b) ** 2

    def integrand(t):
        return np.log(c2 + np.sqrt(1 + k * k - 2 * k * np.cos(2 * t)) / k)

    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        val, _ = quad(integrand, 0.0, np.pi, epsabs=1e-15, epsrel=1e-15, limit=400)
    return -(0.5 * np.log(2.0) + val / (2 * np.pi)) / b
