"""Initial tensors of the 2D classical models (oracle).

Follows src/models/ising.jl:1-67, src/models/potts.jl:6-81 and
src/models/clock.jl:1-49.  Leg order is (left, bottom, top, right) =
T[1 2; 3 4] (docs/src/index.md:107-119); legs 1, 2 are the codomain and
legs 3, 4 the domain.  Index value i of a Z_N leg carries sector i.

Pins: Z2 Ising == Trivial Ising entry by entry (ising.jl:44-55 vs 56-67);
the spin-delta form gives identical norms; Z_N Potts / clock == their
Trivial forms under the per-leg Fourier gauge (identical norms); the
reference's own physics tolerances (test/models.jl:5-40) on free energies.
"""
from __future__ import annotations

import mpmath
import numpy as np

from .ztensor import ZT

mpmath.mp.dps = 60

#: src/models/ising.jl:1 (BigFloat in the reference)
ising_betac_mp = mpmath.log(1 + mpmath.sqrt(2)) / 2
ising_betac = float(ising_betac_mp)
#: src/models/ising.jl:2, first 40 digits of the 100-digit constant
f_onsager_str = "-2.109651144608207459667779283511084780825"
f_onsager = float(mpmath.mpf(f_onsager_str))


def potts_betac(q):
    """src/models/potts.jl:39"""
    return float(np.log(1.0 + np.sqrt(q)))


def _cosh_sinh(beta):
    if isinstance(beta, mpmath.mpf):
        return mpmath.cosh(beta), mpmath.sinh(beta)
    b = float(beta)
    return np.cosh(b), np.sinh(b)


def _ising_trivial(beta, h, impurity):
    x, y = _cosh_sinh(beta)
    if isinstance(beta, mpmath.mpf):
        w = np.array([float(mpmath.sqrt(x)), float(mpmath.sqrt(y))])
        ch, sh = float(mpmath.cosh(h * beta)), float(mpmath.sinh(h * beta))
    else:
        w = np.array([np.sqrt(x), np.sqrt(y)])
        ch, sh = np.cosh(h * float(beta)), np.sinh(h * float(beta))
    even, odd = (sh, ch) if impurity else (ch, sh)
    init = np.zeros((2, 2, 2, 2))
    for i in range(2):
        for j in range(2):
            for k in range(2):
                for l in range(2):
                    init[i, j, k, l] = even if (i + j + k + l) % 2 == 0 else odd
    t = 2.0 * init * w[:, None, None, None] * w[None, :, None, None] \
        * w[None, None, :, None] * w[None, None, None, :]
    return ZT(t)


def classical_ising_trivial(beta, h=0.0) -> ZT:
    """src/models/ising.jl:10-15,44-55: 2 * init * bond^4, init = cosh(h b)
    on even-parity index sums and sinh(h b) on odd ones, bond =
    diag(sqrt(cosh b), sqrt(sinh b)); BigFloat beta is honoured by
    evaluating the entries in high precision before rounding."""
    return _ising_trivial(beta, h, impurity=False)


def classical_ising_impurity(beta, h=0.0) -> ZT:
    """src/models/ising.jl:70-102: the magnetisation impurity, init with the
    parities swapped (sinh(h b) on even sums, cosh(h b) on odd)."""
    return _ising_trivial(beta, h, impurity=True)


def classical_ising_z2(beta) -> ZT:
    """src/models/ising.jl:56-67: blocks c=0 -> [2x^2 2xy; 2xy 2y^2] (rows
    (0,0),(1,1)), c=1 -> 2xy everywhere; embedded densely with leg charges
    [0, 1] (codomain) and -[0, 1] (domain)."""
    x, y = _cosh_sinh(beta)
    xx, xy, yy = float(2 * x * x), float(2 * x * y), float(2 * y * y)
    t = np.zeros((2, 2, 2, 2))
    t[0, 0, 0, 0] = xx
    t[0, 0, 1, 1] = xy
    t[1, 1, 0, 0] = xy
    t[1, 1, 1, 1] = yy
    for a1, a2 in ((1, 0), (0, 1)):
        for a3, a4 in ((1, 0), (0, 1)):
            t[a1, a2, a3, a4] = xy
    qa = np.array([0, 1])
    return ZT(t, [qa, qa, -qa, -qa], 2)


def ising_spin_delta(beta) -> ZT:
    """The same Ising tensor in the spin basis: sqrt(W) contracted onto each
    leg of the spin delta, W = [[e^b, e^-b], [e^-b, e^b]].  Differs from
    ``classical_ising_trivial`` by a Hadamard gauge on every bond."""
    b = float(beta)
    W = np.array([[np.exp(b), np.exp(-b)], [np.exp(-b), np.exp(b)]])
    w, v = np.linalg.eigh(W)
    sW = (v * np.sqrt(w)) @ v.T
    t = np.einsum("si,sj,sk,sl->ijkl", sW, sW, sW, sW)
    return ZT(t)


def potts_tensor(q, beta) -> np.ndarray:
    """src/models/potts.jl:22-30"""
    A = np.zeros((q, q, q, q))
    for i in range(q):
        for j in range(q):
            for k in range(q):
                for l in range(q):
                    E = -((i == j) + (j == l) + (k == l) + (k == i))
                    A[i, j, k, l] = np.exp(-beta * E)
    return A


def clock_tensor(q, beta) -> np.ndarray:
    """src/models/clock.jl:1-12"""
    def clock(a, b):
        return -np.cos(2 * np.pi / q * (a - b))
    A = np.zeros((q, q, q, q))
    for i in range(q):
        for j in range(q):
            for k in range(q):
                for l in range(q):
                    E = clock(i, j) + clock(j, l) + clock(l, k) + clock(k, i)
                    A[i, j, k, l] = np.exp(-beta * E)
    return A


def _fourier_zn(A, q) -> ZT:
    """Conjugate every leg by the q-point DFT F[x, a] = w^(x a)/sqrt(q):
    codomain legs get conj(F), domain legs F (src/models/potts.jl:75-77,
    src/models/clock.jl:38-45); keep the real part of the Z_q-allowed
    entries (TensorMap projects onto the symmetric blocks, then ``real``)."""
    idx = np.arange(q)
    F = np.exp(2j * np.pi * np.outer(idx, idx) / q) / np.sqrt(q)
    U = np.einsum("wxyz,wa,xb,yc,zd->abcd", A.astype(complex), F.conj(), F.conj(), F, F)
    qa = idx
    t = ZT(np.zeros((q, q, q, q)), [qa, qa, -qa, -qa], q)
    tot = (idx[:, None, None, None] + idx[None, :, None, None]
           - idx[None, None, :, None] - idx[None, None, None, :]) % q
    t.data[...] = np.where(tot == 0, U.real, 0.0)
    return t


def classical_potts_trivial(q, beta) -> ZT:
    return ZT(potts_tensor(q, beta))


def classical_potts_zn(q, beta) -> ZT:
    """src/models/potts.jl:71-81"""
    return _fourier_zn(potts_tensor(q, beta), q)


def classical_clock_trivial(q, beta) -> ZT:
    return ZT(clock_tensor(q, beta))


def classical_clock_zn(q, beta) -> ZT:
    """src/models/clock.jl:32-49"""
    return _fourier_zn(clock_tensor(q, beta), q)


# ------------------------------------------------------------ 3D Ising
#: src/models/ising.jl:8 and test/schemes.jl:10 (Fig. 5 of PRB 102, 054432)
ising_betac_3d = 1.0 / 4.51152469
f_benchmark3d = -3.507


def classical_ising_3d_z2(beta) -> ZT:
    """src/models/ising.jl:149-165: t[i..n] = sum_a prod_k W[a, i_k] with
    W = [[sqrt x, sqrt y], [sqrt x, -sqrt y]] (x = cosh b, y = sinh b) on
    S^3 <- S^3, permuted to ((1, 4), (5, 6, 2, 3)); Z2 charges = index
    parity on every leg (self-dual)."""
    b = float(beta)
    x, y = np.cosh(b), np.sinh(b)
    W = np.array([[np.sqrt(x), np.sqrt(y)], [np.sqrt(x), -np.sqrt(y)]])
    t = np.einsum("ai,aj,ak,al,am,an->ijklmn", W, W, W, W, W, W)
    t = np.transpose(t, (0, 3, 4, 5, 1, 2))
    qa = np.array([0, 1])
    return ZT(t, [qa, qa, qa, qa, qa, qa], 2)


def classical_ising_3d_trivial(beta) -> ZT:
    """src/models/ising.jl:130-148: the spin delta O with q = sqrt of the
    bond Boltzmann matrix [[e^K, e^-K], [e^-K, e^K]] on each of its six legs."""
    K = float(beta)
    t = np.array([[np.exp(K), np.exp(-K)], [np.exp(-K), np.exp(K)]])
    w, v = np.linalg.eigh(t)
    q = (v * np.sqrt(w)) @ v.T
    O = np.zeros((2,) * 6)
    O[(0,) * 6] = 1.0
    O[(1,) * 6] = 1.0
    o = np.einsum("abcdef,Aa,Bb,Cc,Dd,Ee,Ff->ABCDEF", O, q, q, q, q, q, q)
    return ZT(o)
