"""Initial tensors of the 2D classical models, in the engine's format.

src/models/ising.jl:1-67, src/models/potts.jl:6-81, src/models/clock.jl:1-49.
Leg order (left, bottom, top, right) = T[1 2; 3 4]; index i of a Z_N leg
carries sector i.
"""
from __future__ import annotations

import math
from decimal import Decimal, getcontext

import numpy as np

from .tensors import NTensor, SymTensor, trivial

getcontext().prec = 50
#: src/models/ising.jl:1 (beta_c = ln(1 + sqrt 2) / 2)
ising_betac = math.log(1.0 + math.sqrt(2.0)) / 2.0
#: src/models/ising.jl:2
f_onsager = -2.1096511446082074596677792835110847808254932754354
# src/models/ising.jl:3-7: the exact Ising CFT scaling dimensions
ising_cft_exact = [1 / 8, 1, 9 / 8, 9 / 8, 2, 2, 2, 2, 17 / 8, 17 / 8, 17 / 8, 3, 3, 3, 3, 3,
                   25 / 8, 25 / 8, 25 / 8, 25 / 8, 25 / 8, 25 / 8]


def _bond(beta, exact_betac):
    """(sqrt(cosh b), sqrt(sinh b)); at the reference's BigFloat beta_c the
    entries are correctly rounded (src/models/ising.jl:10-15)."""
    if exact_betac:
        b = (Decimal(1) + Decimal(2).sqrt()).ln() / 2
        e, ei = b.exp(), (-b).exp()
        x, y = (e + ei) / 2, (e - ei) / 2
        return float(x.sqrt()), float(y.sqrt()), float(x), float(y)
    x, y = float(np.cosh(beta)), float(np.sinh(beta))
    return float(np.sqrt(x)), float(np.sqrt(y)), x, y


def _field_factors(b, exact, h):
    """cosh(h b), sinh(h b); at the exact beta_c evaluated at high precision
    first, as the reference does in BigFloat."""
    if h == 0.0:
        return 1.0, 0.0
    if exact:
        bd = (Decimal(1) + Decimal(2).sqrt()).ln() / 2 * Decimal(float(h))
        e, ei = bd.exp(), (-bd).exp()
        return float((e + ei) / 2), float((e - ei) / 2)
    return float(np.cosh(h * b)), float(np.sinh(h * b))


def _ising_dense(beta, h, impurity):
    exact = beta is None
    b = ising_betac if exact else float(beta)
    sx, sy, x, y = _bond(b, exact)
    ch, sh = _field_factors(b, exact, float(h))
    even, odd = (sh, ch) if impurity else (ch, sh)
    w = np.array([sx, sy])
    T = np.zeros((2, 2, 2, 2))
    for i in range(2):
        for j in range(2):
            for k in range(2):
                for l in range(2):
                    init = even if (i + j + k + l) % 2 == 0 else odd
                    T[i, j, k, l] = 2.0 * init * w[i] * w[j] * w[k] * w[l]
    return trivial(T)


def classical_ising(beta=None, symmetry="Z2", h=0.0):
    """2D Ising partition-function tensor.

    ``symmetry`` "Trivial" follows src/models/ising.jl:44-55 (with field
    ``h``: cosh(h b) on even index sums, sinh(h b) on odd ones), "Z2" follows
    :56-67 (h = 0 only, ising.jl:57).  ``beta=None`` means the reference's
    beta_c."""
    if symmetry == "Trivial":
        return _ising_dense(beta, h, impurity=False)
    if h != 0.0:
        raise ValueError("an external field is not compatible with Z2 symmetry (ising.jl:57)")
    exact = beta is None
    b = ising_betac if exact else float(beta)
    sx, sy, x, y = _bond(b, exact)
    if symmetry == "Z2":
        if exact:
            bd = (Decimal(1) + Decimal(2).sqrt()).ln() / 2
            e, ei = bd.exp(), (-bd).exp()
            xd, yd = (e + ei) / 2, (e - ei) / 2
            xx, xy, yy = float(2 * xd * xd), float(2 * xd * yd), float(2 * yd * yd)
        else:
            xx, xy, yy = 2 * x * x, 2 * x * y, 2 * y * y
        # blocks: c=0 rows (0,0),(1,1) -> [2x^2 2xy; 2xy 2y^2]; c=1 -> 2xy
        data = np.array([xx, xy, xy, yy, xy, xy, xy, xy])
        return SymTensor(2, np.ones((4, 2), dtype=np.int32), data)
    raise ValueError("symmetry must be 'Trivial' or 'Z2'")


def classical_ising_impurity(beta=None, h=0.0):
    """src/models/ising.jl:70-102: the magnetisation impurity (Trivial only),
    the Ising tensor with the parities of init swapped."""
    return _ising_dense(beta, h, impurity=True)


def potts_betac(q):
    """src/models/potts.jl:39"""
    return math.log(1.0 + math.sqrt(q))


def _potts_dense(q, beta):
    A = np.zeros((q, q, q, q))
    for i in range(q):
        for j in range(q):
            for k in range(q):
                for l in range(q):
                    E = -((i == j) + (j == l) + (k == l) + (k == i))
                    A[i, j, k, l] = math.exp(-beta * E)
    return A


def _clock_dense(q, beta):
    def clock(a, b):
        return -math.cos(2 * math.pi / q * (a - b))
    A = np.zeros((q, q, q, q))
    for i in range(q):
        for j in range(q):
            for k in range(q):
                for l in range(q):
                    E = clock(i, j) + clock(j, l) + clock(l, k) + clock(k, i)
                    A[i, j, k, l] = math.exp(-beta * E)
    return A


def _fourier(A, q):
    """conj(F) on codomain legs, F on domain legs, F[x, a] = w^{xa}/sqrt(q);
    real part of the Z_q-allowed entries."""
    idx = np.arange(q)
    F = np.exp(2j * np.pi * np.outer(idx, idx) / q) / math.sqrt(q)
    U = np.einsum("wxyz,wa,xb,yc,zd->abcd", A.astype(complex), F.conj(), F.conj(), F, F)
    lab = [idx, idx, idx, idx]
    allowed = ((idx[:, None, None, None] + idx[None, :, None, None]
                - idx[None, None, :, None] - idx[None, None, None, :]) % q) == 0
    return SymTensor.from_dense(np.where(allowed, U.real, 0.0), lab, q)


def classical_potts(q, beta=None, symmetry="ZN"):
    """q-state Potts tensor (src/models/potts.jl:22-30,63-81)."""
    b = potts_betac(q) if beta is None else float(beta)
    A = _potts_dense(q, b)
    if symmetry == "Trivial":
        return trivial(A)
    return _fourier(A, q)


def classical_clock(q, beta, symmetry="ZN"):
    """q-state clock tensor (src/models/clock.jl:1-12,26-49)."""
    A = _clock_dense(q, float(beta))
    if symmetry == "Trivial":
        return trivial(A)
    return _fourier(A, q)


# ------------------------------------------------------------ 3D Ising
#: src/models/ising.jl:8 (beta_c of the 3D Ising model)
ising_betac_3D = 1.0 / 4.51152469
#: test/schemes.jl:10 (Fig. 5 of PRB 102, 054432): the free energy at beta_c
f_benchmark3D = -3.507


def classical_ising_3D(beta=None, symmetry="Z2"):
    """3D Ising tensor T[z z'; Y X y x] (2 + 4 legs) for HOTRG_3D.

    "Z2" follows src/models/ising.jl:149-165 (t = sum_a prod W[a, i_k], W =
    [[sqrt x, sqrt y], [sqrt x, -sqrt y]], permuted to ((1, 4), (5, 6, 2, 3));
    Z2 charge = index on every leg), "Trivial" :130-148 (the spin delta with
    the square root of the bond Boltzmann matrix on each leg)."""
    b = ising_betac_3D if beta is None else float(beta)
    if symmetry == "Z2":
        x, y = np.cosh(b), np.sinh(b)
        W = np.array([[np.sqrt(x), np.sqrt(y)], [np.sqrt(x), -np.sqrt(y)]])
        t = np.einsum("ai,aj,ak,al,am,an->ijklmn", W, W, W, W, W, W)
        t = np.transpose(t, (0, 3, 4, 5, 1, 2))
        return NTensor.from_dense(t, [[0, 1]] * 6, 2, 2)
    if symmetry == "Trivial":
        t = np.array([[np.exp(b), np.exp(-b)], [np.exp(-b), np.exp(b)]])
        w, v = np.linalg.eigh(t)
        q = (v * np.sqrt(w)) @ v.T
        O = np.zeros((2,) * 6)
        O[(0,) * 6] = 1.0
        O[(1,) * 6] = 1.0
        o = np.einsum("abcdef,Aa,Bb,Cc,Dd,Ee,Ff->ABCDEF", O, q, q, q, q, q, q)
        return NTensor(1, np.full((6, 1), 2, dtype=np.int32), o.ravel(order="F"), 2)
    raise ValueError("symmetry must be 'Z2' or 'Trivial'")
