"""TRG / BTRG / HOTRG coarse-graining, finalize, run loop, free energy (oracle).

Restates, function by function:

* ``TRG.step``   - src/schemes/trg.jl:35-41 (with ``svd12`` =
  src/utility/projectors.jl:211-217)
* ``BTRG``       - src/schemes/btrg.jl:26-94 (``pseudopow`` :50-58)
* ``HOTRG``      - src/schemes/hotrg.jl:42-160
* ``finalize``   - src/utility/finalize.jl:1-14
* ``run``        - src/schemes/tnrscheme.jl:33-59 with ``maxiter``
  (src/utility/stopping.jl:3-5,18)
* ``free_energy``- src/utility/free_energy.jl:13-20
* ``ImpurityTRG``  - src/schemes/impuritytrg.jl:63-91, finalize.jl:74-88
* ``ImpurityHOTRG``- src/schemes/impurityhotrg.jl:48-84, finalize.jl:90-101

Contractions are pairwise in a cost-optimal order (TensorOperations'
order only changes roundoff).  Each factorization appends its truncation
diagnostics (smallest kept / largest dropped value) to ``scheme.log``.

Pins (tests/test_oracle_pins.py): the published BTRG free energy
(README.md:87), the reference's Onsager tolerances (test/schemes.jl:18-23,
57-62, 94-99), exact small-torus partition functions for the un-truncated
first steps of all three schemes on random tensors (oracle.torus), Z2 vs
Trivial Ising equality, Z_N vs Trivial Potts/clock equality.
"""
from __future__ import annotations

import numpy as np

from .ztensor import ZT, contract, permute, svd_trunc, eigh_trunc, \
    project_hermitian, scale_leg

EPS = np.finfo(np.float64).eps


def _diag(s, q_cod, N) -> ZT:
    """Dense (1,1) tensor diag(s) on a bond whose codomain leg has charges q."""
    return ZT(np.diag(s), [q_cod, (-q_cod) % N], N)


def svd12(T: ZT, chi: int, log=None, gauge="lapack"):
    """src/utility/projectors.jl:211-217: (U*sqrt(s), sqrt(s)*V); ``gauge``
    as in svd_trunc."""
    U, s, Vt, info = svd_trunc(T, 2, chi, gauge)
    if log is not None:
        log.append(info)
    sq = np.sqrt(s)
    return scale_leg(U, 2, sq), scale_leg(Vt, 0, sq)


def pseudopow(s, a, tol=EPS ** 0.75):
    """src/schemes/btrg.jl:50-58: s**a where s >= tol, else s unchanged."""
    s = np.asarray(s, dtype=np.float64)
    out = s.copy()
    m = ~(s < tol)
    out[m] = s[m] ** a
    return out


class TRG:
    """src/schemes/trg.jl:26-41"""
    scalefactor = 2.0

    def __init__(self, T: ZT):
        self.T = T.copy()
        self.log = []

    def step(self, chi):
        T = self.T
        A, B = svd12(T, chi, self.log)
        Tp = permute(T, (1, 3, 0, 2))            # transpose(T, ((2,4),(1,3)))
        C, D = svd12(Tp, chi, self.log)
        # T[-1 -2; -3 -4] := D[-2; 1 2] * B[-1; 4 1] * C[4 3; -3] * A[3 2; -4]
        X = contract([-1, -2, -3, -4], (D, [-2, 1, -3], False), (B, [-1, -4, 1], False))  # X[p,q,2,4]
        Y = contract([-1, -2, -3, -4], (C, [-1, 3, -3], False), (A, [3, -2, -4], False))  # Y[4,2,r,s]
        self.T = contract([-1, -2, -3, -4], (X, [-1, -2, 2, 4], False), (Y, [4, 2, -3, -4], False))
        return self

    def finalize(self):
        """src/utility/finalize.jl:4-8: n = |tr T[1 2; 2 1]|, T /= n."""
        n = abs(float(np.einsum("abba->", self.T.data)))
        self.T.data /= n
        return n


class HOTRG(TRG):
    """src/schemes/hotrg.jl:27-160"""
    scalefactor = 4.0

    def step(self, chi):
        Ux, _ = get_hotrg_xproj(self.T, self.T, chi, self.log)
        self.T = step_hotrg_y(self.T, self.T, Ux)
        Uy, _ = get_hotrg_yproj(self.T, self.T, chi, self.log)
        self.T = step_hotrg_x(self.T, self.T, Uy)
        return self


class BTRG:
    """src/schemes/btrg.jl:26-94 (S1 on vertical bonds, S2 on horizontal)."""
    scalefactor = 2.0

    def __init__(self, T: ZT, k=-0.5):
        self.T = T.copy()
        self.k = float(k)
        N = T.N
        self.S1 = _diag(np.ones(T.shape[1]), T.q[1], N)   # id(space(T, 2))
        self.S2 = _diag(np.ones(T.shape[0]), T.q[0], N)   # id(space(T, 1))
        self.log = []

    def step(self, chi):
        T, k = self.T, self.k
        U, s, V, info = svd_trunc(T, 2, chi)
        self.log.append(info)
        sa, sb = pseudopow(s, (1 - k) / 2), pseudopow(s, k)
        A = scale_leg(U, 2, sa)                  # A[-1 -2; -3] := U * S_a
        B = scale_leg(V, 0, sa)                  # B[-1; -2 -3] := S_a * V
        S1n = _diag(sb, V.q[0], T.N)             # S1'[-1; -2] := S_b
        U, s, V, info = svd_trunc(permute(T, (2, 0, 3, 1)), 2, chi)
        self.log.append(info)
        sa, sb = pseudopow(s, (1 - k) / 2), pseudopow(s, k)
        C = scale_leg(U, 2, sa)
        D = scale_leg(V, 0, sa)
        S2n = _diag(sb, V.q[0], T.N)
        S1, S2 = self.S1, self.S2
        # T[-1 -2; -3 -4] := D[-1; 4 7] * S1[1; 7] * B[-2; 1 3] * S2[3; 2]
        #                    * C[8 2; -4] * S1[8; 5] * A[6 5; -3] * S2[4; 6]
        # bond matrices are absorbed pairwise (same network, cost-optimal order)
        D1 = contract([-1, -2, -3], (D, [-1, -2, 7], False), (S1, [-3, 7], False))   # D[-1;4 7]S1[1;7]
        B1 = contract([-1, -2, -3], (B, [-1, -2, 3], False), (S2, [3, -3], False))   # B[-2;1 3]S2[3;2]
        C1 = contract([-1, -2, -3], (C, [8, -2, -3], False), (S1, [8, -1], False))   # C[8 2;-4]S1[8;5]
        A1 = contract([-1, -2, -3], (A, [6, -2, -3], False), (S2, [-1, 6], False))   # A[6 5;-3]S2[4;6]
        # D1[p, 4, 1] B1[q, 1, 2] -> X[p, q, 4, 2];  C1[5, 2, s] A1[4, 5, r] -> Y[2, 4, r, s]
        X = contract([-1, -2, -3, -4], (D1, [-1, -3, 1], False), (B1, [-2, 1, -4], False))
        Y = contract([-1, -2, -3, -4], (C1, [5, -1, -4], False), (A1, [-2, 5, -3], False))
        self.T = contract([-1, -2, -3, -4], (X, [-1, -2, 4, 2], False), (Y, [2, 4, -3, -4], False))
        self.S1, self.S2 = S1n, S2n
        return self

    def finalize(self):
        """src/utility/finalize.jl:10-14: n = |T[1 2; 4 3] S1[4; 2] S2[3; 1]|."""
        n = abs(float(np.einsum("abdc,db,ca->", self.T.data, self.S1.data, self.S2.data)))
        self.T.data /= n
        return n


class ATRG(TRG):
    """src/schemes/atrg.jl:35-76 (anisotropic TRG): two half-steps, each
    three truncated SVDs; the tensor is rotated between and after them."""
    scalefactor = 4.0

    def _half(self, chi):
        T = self.T
        # A, S, B = svd_trunc(permute(T, ((1, 3), (2, 4)))); C, D = copies
        A, s, D, info = svd_trunc(permute(T, (0, 2, 1, 3)), 2, chi)
        self.log.append(info)
        B = scale_leg(D, 0, s)                   # B[-1; -2 -3] = S[-1; 1] B[1; -2 -3]
        C = scale_leg(A, 2, s)                   # C[-1 -2; -3] = C[-1 -2; 1] S[1; -3]
        # M[-1 -2; -3 -4] := B[-3; 1 -4] * C[-1 1; -2]
        M = contract([-1, -2, -3, -4], (B, [-3, 1, -4], False), (C, [-1, 1, -2], False))
        X, s2, Y, info = svd_trunc(permute(M, (0, 2, 1, 3)), 2, chi)
        self.log.append(info)
        sq = np.sqrt(s2)
        X, Y = scale_leg(X, 2, sq), scale_leg(Y, 0, sq)
        # Q[-1 -2; -3 -4] := A[3 -3; 2] * D[1; -2 4] * X[4 2; -4] * Y[-1 1; 3]
        AX = contract([-1, -2, -3, -4], (A, [-1, -2, 1], False), (X, [-3, 1, -4], False))
        DY = contract([-1, -2, -3, -4], (D, [1, -1, -2], False), (Y, [-3, 1, -4], False))
        Q = contract([-1, -2, -3, -4], (AX, [3, -3, 4, -4], False), (DY, [-2, 4, -1, 3], False))
        H, s3, G, info = svd_trunc(Q, 2, chi)
        self.log.append(info)
        sq = np.sqrt(s3)
        H, G = scale_leg(H, 2, sq), scale_leg(G, 0, sq)
        # T[-1 -2; -3 -4] := G[-1; -3 1] * H[1 -2; -4]
        self.T = contract([-1, -2, -3, -4], (G, [-1, -3, 1], False), (H, [1, -2, -4], False))

    def step(self, chi):
        self._half(chi)
        self.T = permute(self.T, (1, 3, 0, 2))   # permute(T, ((2, 4), (1, 3)))
        self._half(chi)
        self.T = permute(self.T, (2, 0, 3, 1))   # permute(T, ((3, 1), (4, 2)))
        return self


# ---------------------------------------------------------------- HOTRG parts

def _pick(A, B, chi, log):
    MM = project_hermitian(A, 2)
    _, U, eps, info = eigh_trunc(MM, 2, chi)
    MM2 = project_hermitian(B, 2)
    _, U2, eps2, info2 = eigh_trunc(MM2, 2, chi)
    if log is not None:
        log.append(info)
        log.append(info2)
    if eps > eps2:                                  # src/schemes/hotrg.jl:112,148
        return U2, eps2
    return U, eps


def get_hotrg_xproj(A1: ZT, A2: ZT, chi, log=None):
    """src/schemes/hotrg.jl:82-116"""
    # MM[-1 -2; -3 -4] := A2[-1 5; 1 2] A1[-2 3; 5 4] conj(A2[-3 6; 1 2]) conj(A1[-4 3; 6 4])
    E2 = contract([-1, -2, -3, -4], (A2, [-1, -2, 1, 2], False), (A2, [-3, -4, 1, 2], True))
    E1 = contract([-1, -2, -3, -4], (A1, [-1, 3, -2, 4], False), (A1, [-3, 3, -4, 4], True))
    MM = contract([-1, -2, -3, -4], (E2, [-1, 5, -3, 6], False), (E1, [-2, 5, -4, 6], False))
    # MM[-1 -2; -3 -4] := conj(A2[2 5; 1 -1]) conj(A1[4 3; 5 -2]) A2[2 6; 1 -3] A1[4 3; 6 -4]
    F2 = contract([-1, -2, -3, -4], (A2, [2, -1, 1, -2], True), (A2, [2, -3, 1, -4], False))
    F1 = contract([-1, -2, -3, -4], (A1, [4, 3, -1, -2], True), (A1, [4, 3, -3, -4], False))
    MMp = contract([-1, -2, -3, -4], (F2, [5, -1, 6, -3], False), (F1, [5, -2, 6, -4], False))
    return _pick(MM, MMp, chi, log)


def get_hotrg_yproj(A1: ZT, A2: ZT, chi, log=None):
    """src/schemes/hotrg.jl:118-152"""
    # MM[-1 -2; -3 -4] := A1[1 -1; 2 5] A2[5 -2; 4 3] conj(A1[1 -3; 2 6]) conj(A2[6 -4; 4 3])
    G1 = contract([-1, -2, -3, -4], (A1, [1, -1, 2, -2], False), (A1, [1, -3, 2, -4], True))
    G2 = contract([-1, -2, -3, -4], (A2, [-1, -2, 4, 3], False), (A2, [-3, -4, 4, 3], True))
    MM = contract([-1, -2, -3, -4], (G1, [-1, 5, -3, 6], False), (G2, [5, -2, 6, -4], False))
    # MM[-1 -2; -3 -4] := conj(A1[1 2; -1 5]) conj(A2[5 4; -2 3]) A1[1 2; -3 6] A2[6 4; -4 3]
    H1 = contract([-1, -2, -3, -4], (A1, [1, 2, -1, -2], True), (A1, [1, 2, -3, -4], False))
    H2 = contract([-1, -2, -3, -4], (A2, [-1, 4, -2, 3], True), (A2, [-3, 4, -4, 3], False))
    MMp = contract([-1, -2, -3, -4], (H1, [-1, 5, -3, 6], False), (H2, [5, -2, 6, -4], False))
    return _pick(MM, MMp, chi, log)


def step_hotrg_y(A1: ZT, A2: ZT, Ux: ZT) -> ZT:
    """src/schemes/hotrg.jl:42-60:
    T[-1 -2; -3 -4] := conj(Ux[1 2; -1]) Ux[3 4; -4] A2[1 5; -3 3] A1[2 -2; 5 4]"""
    P = contract([-1, -2, -3, -4, -5], (Ux, [1, -1, -2], True), (A2, [1, -3, -4, -5], False))
    Q = contract([-1, -2, -3, -4, -5], (P, [2, -1, 5, -2, -3], False), (A1, [2, -4, 5, -5], False))
    return contract([-1, -2, -3, -4], (Q, [-1, -3, 3, -2, 4], False), (Ux, [3, 4, -4], False))


def step_hotrg_x(A1: ZT, A2: ZT, Uy: ZT) -> ZT:
    """src/schemes/hotrg.jl:62-80:
    T[-1 -2; -3 -4] := A1[-1 1; 3 5] A2[5 2; 4 -4] conj(Uy[1 2; -2]) Uy[3 4; -3]"""
    R = contract([-1, -2, -3, -4, -5], (A1, [-1, 1, -2, -3], False), (Uy, [1, -4, -5], True))
    S = contract([-1, -2, -3, -4, -5], (R, [-1, -2, 5, 2, -3], False), (A2, [5, 2, -4, -5], False))
    return contract([-1, -2, -3, -4], (S, [-1, 3, -2, 4, -4], False), (Uy, [3, 4, -3], False))


# ---------------------------------------------------------------- HOTRG_3D
# tensors T[z_up z_down; Y X y x] (2|4); bosonic sectors, so the fermionic
# twists of src/schemes/hotrg3d.jl:36-46 are identities

def _mmdag_3d(A1: ZT, A2: ZT) -> ZT:
    """src/schemes/hotrg3d.jl:48-65:
    MM[x2 z z' x2'] := A2[z z2; Y2 X2 y2 x2] conj(A2[z' z2; Y2 X2 y2 x2'])
    MM[x1 x2; x1' x2'] := MM[x2 z z' x2'] A1[z1 z; Y1 X1 y1 x1] conj(A1[z1 z'; Y1 X1 y1 x1'])"""
    E2 = contract([-1, -2, -3, -4], (A2, [-2, 1, 2, 3, 4, -1], False), (A2, [-3, 1, 2, 3, 4, -4], True))
    E1 = contract([-1, -2, -3, -4], (A1, [1, -1, 2, 3, 4, -2], False), (A1, [1, -3, 2, 3, 4, -4], True))
    MM = contract([-1, -2, -3, -4], (E2, [-2, 5, 6, -4], False), (E1, [5, -1, 6, -3], False))
    return project_hermitian(MM, 2)


def _mdagm_3d(A1: ZT, A2: ZT) -> ZT:
    """src/schemes/hotrg3d.jl:67-84:
    MM[x2 z z' x2'] := conj(A2[z z2; Y2 x2 y2 X2]) A2[z' z2; Y2 x2' y2 X2]
    MM[x1 x2; x1' x2'] := MM[x2 z z' x2'] conj(A1[z1 z; Y1 x1 y1 X1]) A1[z1 z'; Y1 x1' y1 X1]"""
    E2 = contract([-1, -2, -3, -4], (A2, [-2, 1, 2, -1, 3, 4], True), (A2, [-3, 1, 2, -4, 3, 4], False))
    E1 = contract([-1, -2, -3, -4], (A1, [1, -1, 2, -2, 3, 4], True), (A1, [1, -3, 2, -4, 3, 4], False))
    MM = contract([-1, -2, -3, -4], (E2, [-2, 5, 6, -4], False), (E1, [5, -1, 6, -3], False))
    return project_hermitian(MM, 2)


def get_hotrg3d_xproj(A1: ZT, A2: ZT, chi, log=None):
    """src/schemes/hotrg3d.jl:86-99: U of MM unless eps > eps' (then U')."""
    MM = _mmdag_3d(A1, A2)
    _, U, eps, info = eigh_trunc(MM, 2, chi)
    MM2 = _mdagm_3d(A1, A2)
    _, U2, eps2, info2 = eigh_trunc(MM2, 2, chi)
    if log is not None:
        log.append(info)
        log.append(info2)
    if eps > eps2:
        return U2, eps2
    return U, eps


def get_hotrg3d_yproj(A1: ZT, A2: ZT, chi, log=None):
    """src/schemes/hotrg3d.jl:101-107: the x projector of ((1, 2), (4, 3, 6, 5))."""
    perm = (0, 1, 3, 2, 5, 4)
    return get_hotrg3d_xproj(permute(A1, perm), permute(A2, perm), chi, log)


def step_hotrg3d(A1: ZT, A2: ZT, Ux: ZT, Uy: ZT) -> ZT:
    """src/schemes/hotrg3d.jl:109-119:
    T[-1 -2; -3 -4 -5 -6] := conj(Ux[x1 x2; -6]) Ux[x1' x2'; -4] conj(Uy[y1 y2; -5])
                             Uy[y1' y2'; -3] A1[-1 z; y1' x1' y1 x1] A2[z -2; y2' x2' y2 x2]
    pairwise, at most eight legs at a time."""
    # B1[a z y1' y1 x1 x2' n4] = A1[a z; y1' x1' y1 x1] Ux[x1' x2'; n4]
    B1 = contract([-1, -2, -3, -4, -5, -6, -7], (A1, [-1, -2, -3, 1, -4, -5], False),
                  (Ux, [1, -6, -7], False))
    # B2[a z y1' y1 x2' n4 x2 n6] = B1 conj(Ux[x1 x2; n6])
    B2 = contract([-1, -2, -3, -4, -5, -6, -7, -8], (B1, [-1, -2, -3, -4, 1, -5, -6], False),
                  (Ux, [1, -7, -8], True))
    # B3[a y1' y1 n4 n6 b y2' y2] = B2 A2[z b; y2' x2' y2 x2]
    B3 = contract([-1, -2, -3, -4, -5, -6, -7, -8], (B2, [-1, 1, -2, -3, 2, -4, 3, -5], False),
                  (A2, [1, -6, -7, 2, -8, 3], False))
    # B4[a y1 n4 n6 b y2 n3] = B3 Uy[y1' y2'; n3]
    B4 = contract([-1, -2, -3, -4, -5, -6, -7], (B3, [-1, 1, -2, -3, -4, -5, 2, -6], False),
                  (Uy, [1, 2, -7], False))
    # T[a b; n3 n4 n5 n6] = B4 conj(Uy[y1 y2; n5])
    return contract([-1, -2, -3, -4, -5, -6], (B4, [-1, 1, -4, -6, -2, 2, -3], False),
                    (Uy, [1, 2, -5], True))


class HOTRG_3D:
    """src/schemes/hotrg3d.jl:27-141: three sub-steps per step, each
    compressing the z direction, the tensor rotated by ((6, 4), (2, 3, 1, 5))
    after each."""
    scalefactor = 8.0

    def __init__(self, T: ZT):
        self.T = T.copy()
        self.log = []

    def substep(self, chi):
        Ux, _ = get_hotrg3d_xproj(self.T, self.T, chi, self.log)
        Uy, _ = get_hotrg3d_yproj(self.T, self.T, chi, self.log)
        self.T = step_hotrg3d(self.T, self.T, Ux, Uy)

    def step(self, chi):
        for _ in range(3):
            self.substep(chi)
            self.T = permute(self.T, (5, 3, 1, 2, 0, 4))
        return self

    def finalize(self):
        """src/utility/finalize.jl:62-66: n = |T[1 1; 2 3 2 3]|, T /= n."""
        n = abs(float(np.einsum("aabcbc->", self.T.data)))
        self.T.data /= n
        return n


# ---------------------------------------------------------------- impurities

def _trg_net(D, B, C, A) -> ZT:
    """T[-1 -2; -3 -4] := D[-2; 1 2] * B[-1; 4 1] * C[4 3; -3] * A[3 2; -4]"""
    X = contract([-1, -2, -3, -4], (D, [-2, 1, -3], False), (B, [-1, -4, 1], False))
    Y = contract([-1, -2, -3, -4], (C, [-1, 3, -3], False), (A, [3, -2, -4], False))
    return contract([-1, -2, -3, -4], (X, [-1, -2, 2, 4], False), (Y, [4, 2, -3, -4], False))


class ImpurityTRG:
    """src/schemes/impuritytrg.jl:40-91: the pure tensor T and four
    impurity tensors on a 2x2 plaquette (1 2 / 4 3).

    The impurity norm of this network contracts B1 with Ap and A1 with Bp
    (factors of different SVDs), so it depends on the singular-vector basis
    of every SVD, not only on the tensors: ``gauge="lapack"`` is the
    reference (LAPACK's basis, whatever it is), ``gauge="pivoted"`` fixes
    the basis with ztensor.pivoted_gauge."""
    scalefactor = 2.0

    def __init__(self, T: ZT, T1: ZT, T2: ZT, T3: ZT, T4: ZT, gauge="lapack"):
        self.T, self.T1, self.T2, self.T3, self.T4 = (x.copy() for x in (T, T1, T2, T3, T4))
        self.log = []
        self.gauge = gauge

    def step(self, chi):
        tp = (1, 3, 0, 2)                          # transpose(., ((2, 4), (1, 3)))
        g = self.gauge
        A1, B1 = svd12(self.T1, chi, self.log, g)
        C2, D2 = svd12(permute(self.T2, tp), chi, self.log, g)
        A3, B3 = svd12(self.T3, chi, self.log, g)
        C4, D4 = svd12(permute(self.T4, tp), chi, self.log, g)
        Ap, Bp = svd12(self.T, chi, self.log, g)
        Cp, Dp = svd12(permute(self.T, tp), chi, self.log, g)
        self.T = _trg_net(Dp, Bp, Cp, Ap)
        self.T1 = _trg_net(Dp, B1, C2, Ap)
        self.T2 = _trg_net(D2, B3, Cp, Ap)
        self.T3 = _trg_net(D4, Bp, Cp, A3)
        self.T4 = _trg_net(Dp, Bp, C4, A1)
        return self

    def finalize(self):
        """finalize.jl:74-88: npure = |tr T[1 2; 2 1]|, every tensor /= npure;
        nimp = |T1[5 4; 6 1] T2[1 2; 7 5] T3[3 7; 2 8] T4[8 6; 4 3]|."""
        npure = abs(float(np.einsum("abba->", self.T.data)))
        for t in (self.T, self.T1, self.T2, self.T3, self.T4):
            t.data /= npure
        nimp = abs(float(np.einsum("edfa,abge,cgbh,hfdc->", self.T1.data, self.T2.data,
                                   self.T3.data, self.T4.data, optimize="optimal")))
        return (npure, nimp)


class ImpurityHOTRG:
    """src/schemes/impurityhotrg.jl:27-84: T, the two first-order impurity
    tensors and the second-order one; y projector then x step, x projector
    then y step, impurity tensors averaged over their placements."""
    scalefactor = 4.0

    def __init__(self, T: ZT, T11: ZT, T12: ZT, T2: ZT):
        self.T, self.T11, self.T12, self.T2 = (x.copy() for x in (T, T11, T12, T2))
        self.log = []

    def _half(self, proj, apply, chi):
        U, _ = proj(self.T, self.T, chi, self.log)
        f = lambda a, b: apply(a, b, U).data
        T, T11, T12, T2 = self.T, self.T11, self.T12, self.T2
        n_T = f(T, T)
        n_11 = 0.5 * (f(T11, T) + f(T, T11))
        n_12 = 0.5 * (f(T12, T) + f(T, T12))
        n_2 = 0.25 * (f(T2, T) + f(T, T2) + f(T11, T12) + f(T12, T11))
        q = apply(T, T, U).q
        self.T = ZT(n_T, q, T.N)
        self.T11 = ZT(n_11, q, T.N)
        self.T12 = ZT(n_12, q, T.N)
        self.T2 = ZT(n_2, q, T.N)

    def step(self, chi):
        self._half(get_hotrg_yproj, step_hotrg_x, chi)
        self._half(get_hotrg_xproj, step_hotrg_y, chi)
        return self

    def finalize(self):
        """finalize.jl:90-101: traces of all four, each divided by |tr T|."""
        ns = [abs(float(np.einsum("abba->", t.data)))
              for t in (self.T, self.T11, self.T12, self.T2)]
        for t in (self.T, self.T11, self.T12, self.T2):
            t.data /= ns[0]
        return tuple(ns)


# ---------------------------------------------------------------- driver

def run(scheme, chi, maxiter, finalize_beginning=True):
    """src/schemes/tnrscheme.jl:33-57 with ``maxiter(n)`` (stopping.jl:18):
    returns the list of norms (maxiter + 1 entries by default)."""
    data = []
    if finalize_beginning:
        data.append(scheme.finalize())
    steps = 0
    crit = True
    while crit:
        scheme.step(chi)
        data.append(scheme.finalize())
        steps += 1
        crit = steps < maxiter
    return data


def free_energy(data, beta, scalefactor=2.0, initial_size=1.0):
    """src/utility/free_energy.jl:13-20"""
    lnz = 0.0
    x = 1.0 - np.log(initial_size) / np.log(scalefactor)
    for i, z in enumerate(data, start=1):
        lnz += np.log(z) * scalefactor ** (x - i)
    return -lnz / beta
