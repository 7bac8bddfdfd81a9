"""Observables on a scheme's tensor (oracle, CPU, FP64 numpy; test
infrastructure only).

Restates src/utility/cft.jl (cft_data, central_charge,
ground_state_degeneracy, gu_wen_ratio) and src/utility/finalize.jl:17-41
(finalize_two_by_two!) for TRG / HOTRG (the plain tensor) and BTRG (S1, S2
absorbed).  Dense einsum throughout; eigenvalues and singular values from
LAPACK (numpy.linalg.eigvals / svd).  Pins: the reference's tests
(test/schemes.jl:25-31, 33-53, 64-88, 101-125), closed forms on product
tensors T = P (x) A, and Z2 = Trivial.
"""
from __future__ import annotations

import numpy as np

from .schemes import BTRG
from .ztensor import ZT, matrix_blocks


def unit_tensor(scheme) -> np.ndarray:
    """T_unit[-1 -2; -3 -4] := T[1 2; -3 -4] S1[-2; 2] S2[-1; 1]
    (src/utility/cft.jl:44-45) for BTRG, T otherwise."""
    T = scheme.T.data
    if isinstance(scheme, BTRG):
        return np.einsum("abcd,Bb,Aa->ABcd", T, scheme.S1.data, scheme.S2.data)
    return T


def transfer_matrix(scheme, unitcell=1) -> np.ndarray:
    """ncon of `unitcell` tensors in a periodic row, (bottoms | tops)
    (src/utility/cft.jl:7-15, 280-295)."""
    T = unit_tensor(scheme)
    if unitcell == 1:
        return np.einsum("abta->bt", T)
    if unitcell == 2:
        d = T.shape[1]
        M = np.einsum("xbty,yBTx->bBtT", T, T)
        return M.reshape(d * d, d * d)
    raise ValueError("unitcell 1 or 2")


def sort_spectrum(ev):
    """Decreasing modulus, ties by decreasing real part, then imaginary part
    (the reference's stable sort by |x| keeps LAPACK's order for ties; this
    total order makes the result a function of the eigenvalues alone)."""
    ev = np.asarray(ev, dtype=np.complex128)
    order = np.lexsort((-ev.imag, -ev.real, -np.abs(ev)))
    return ev[order]


def transfer_spectrum(scheme, unitcell=1):
    """eig_full of the transfer matrix (src/utility/cft.jl:16)."""
    return sort_spectrum(np.linalg.eigvals(transfer_matrix(scheme, unitcell)))


def cft_data(scheme, v=1.0, unitcell=1, is_real=True):
    """src/utility/cft.jl:5-36 (and :39-71 for BTRG)."""
    data = transfer_spectrum(scheme, unitcell)
    data = data[data.real > 0]
    data = data[np.abs(data) > 1e-12]
    if is_real:
        data = data.real
    return unitcell * (1.0 / (2.0 * np.pi * v)) * np.log(data[0] / data)


def ground_state_degeneracy(scheme, unitcell=1):
    """src/utility/cft.jl:278-308 (and :311-339 for BTRG): exp of the
    Shannon entropy of |eigenvalues / trace|."""
    D = np.linalg.eigvals(transfer_matrix(scheme, unitcell))
    D = D / np.sum(D)
    vals = np.abs(D)
    vals = vals[vals != 0]
    return float(np.exp(-np.sum(vals * np.log(vals))))


def gu_wen_ratio(scheme):
    """src/utility/cft.jl:374-396: X1 = tr(T)^2 / T[1 2; 2 3] T[3 4; 4 1],
    X2 = tr(T)^2 / T[1 2; 3 4] T[4 3; 2 1]."""
    T = unit_tensor(scheme)
    one = abs(np.einsum("abba->", T))
    two1 = abs(np.einsum("abbc,cdda->", T, T))
    two2 = abs(np.einsum("abcd,dcba->", T, T))
    return one * one / two1, one * one / two2


def central_charge(scheme, n):
    """src/utility/cft.jl:256-268: M = (T/n)[1 -1; -2 1] (BTRG:
    T[1 -1; 3 2] S1[3; -2] S2[2; 1] / n); log(S.data[1]) 6 / pi with
    S.data[1] the largest singular value of M's first non-empty block."""
    T = scheme.T
    if isinstance(scheme, BTRG):
        M = np.einsum("axcd,cy,da->xy", T.data, scheme.S1.data, scheme.S2.data) / n
    else:
        M = np.einsum("axya->xy", T.data) / n
    Mz = ZT(M, [T.q[1], T.q[2]], T.N)
    for c, r, k in matrix_blocks(Mz, 1):
        if len(r) and len(k):
            s = np.linalg.svd(M[np.ix_(r, k)], compute_uv=False)
            return float(np.log(s[0]) * 6.0 / np.pi)
    raise ValueError("empty transfer matrix")


def finalize_two_by_two(scheme):
    """src/utility/finalize.jl:17-41: n = |2x2 torus|, T /= n^(1/4),
    returns n^(1/4)."""
    T = scheme.T.data
    if isinstance(scheme, BTRG):
        S1, S2 = scheme.S1.data, scheme.S2.data
        # T[11 1; 9 8] S2[8; 2] T[2 6; 10 11] S1[3; 6] T[7 10; 3 12] S2[4; 7]
        # T[12 9; 5 4] S1[5; 1]
        n = np.einsum("kaih,hb,bfjk,cf,gjcl,dg,lied,ea->", T, S2, T, S1, T, S2, T, S1,
                      optimize="optimal")
    else:
        # T[7 1; 5 4] T[4 2; 6 7] T[3 6; 2 8] T[8 5; 1 3]
        n = np.einsum("gaed,dbfg,cfbh,heac->", T, T, T, T, optimize="optimal")
    r = abs(float(n)) ** 0.25
    scheme.T.data /= r
    return r
