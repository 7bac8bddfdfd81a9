"""Exact partition functions of small periodic lattices of one tensor.

Independent of the coarse-graining code: built from lattice geometry only.
Leg order is (left, bottom, top, right).  Used to pin the un-truncated first
steps: for TRG/BTRG the 1-site torus of the lattice after one step is the
tilted 2-site torus with periods (1,1), (1,-1), after two steps the 2x2
torus; for HOTRG one step gives the 2x2 torus and two steps the 4x4 torus.
Hence (norms n_0, n_1, n_2 of an exact run)

    TRG/BTRG:  n0 = Z1,  n0^2 n1 = Z_tilted2,  n0^4 n1^2 n2 = Z_2x2
    HOTRG:     n0 = Z1,  n0^4 n1 = Z_2x2,      n0^16 n1^4 n2 = Z_4x4
"""
from __future__ import annotations

import numpy as np


def z_1site(T: np.ndarray) -> float:
    """left<->right, bottom<->top of one tensor."""
    return float(np.einsum("abba->", T))


def z_tilted2(T: np.ndarray) -> float:
    """Sites A=(0,0), B=(1,0) with periods (1,1), (1,-1): every neighbour of
    A is B, so B.left=A.right, B.bottom=A.top, B.right=A.left, B.top=A.bottom."""
    return float(np.einsum("lbtr,rtbl->", T, T))


def z_torus(T: np.ndarray, L: int) -> float:
    """L x L torus by a row transfer matrix.  Row y: sites x=0..L-1 with
    horizontal bonds h_x between x and x+1 (periodic); the row maps the
    vertical bonds below (v_x) to those above (w_x)."""
    D = T.shape[0]
    if T.shape != (D, D, D, D):
        raise ValueError("square tensor with equal leg dims required")
    R = T  # R[h_left, VB, VA, h_right], VB/VA = combined vertical bonds
    for _ in range(1, L):
        R = _append_site(R, T)
    n = D ** L
    M = np.einsum("avwa->vw", R.reshape(D, n, n, D))
    Z = np.trace(np.linalg.matrix_power(M, L))
    return float(Z)


def _append_site(R, T):
    D = T.shape[0]
    a, nb, na, h = R.shape[0], R.shape[1], R.shape[2], R.shape[3]
    out = np.einsum("aBCh,hbcr->aBbCcr", R, T)
    return out.reshape(a, nb * D, na * D, D)
