"""Charge-labelled dense tensors and truncated factorizations (oracle).

A Z_N-symmetric tensor is stored densely: ``data`` is an ordinary numpy
array and ``q[leg]`` gives, for every index value of that leg, its Z_N
charge in the *outgoing* convention (a codomain leg carrying sector ``a``
has q = a, a domain leg carrying sector ``a`` has q = -a).  Entries that
violate sum(q) == 0 mod N are exactly zero and stay exactly zero under
contraction.  Trivial symmetry is N = 1 (all charges 0, one block).

This reproduces, for abelian Z_N with trivial F/R symbols (TensorKitSectors
``ZNIrrep``, ``Z2Irrep``), the arithmetic TensorKit performs block-sparsely:

* ``permute``/``transpose`` only reorder legs (no phases for Z_N);
* ``svd_trunc(t; trunc=truncrank(chi))`` (MatrixAlgebraKit, used at
  src/schemes/btrg.jl:61,72 and src/utility/projectors.jl:215) = LAPACK SVD
  per coupled-charge block, then the chi largest singular values across all
  blocks (stable sort, descending);
* ``eigh_trunc!(project_hermitian!(M); trunc)`` (src/schemes/hotrg.jl:103,
  110,139,146) = symmetrise, LAPACK eigh per block, keep the chi largest
  |lambda| across blocks; epsilon = 2-norm of the discarded eigenvalues.

Pins: block SVD/eigh are checked against a dense LAPACK factorization of
the embedded matrix (same singular values / eigenvalues, same projector when
the cut has a gap) in tests/test_oracle_ztensor.py; the contraction helper
checks that every contracted pair of legs carries opposite charges, so a
wrongly paired index fails loudly for N >= 3.
"""
from __future__ import annotations

import numpy as np


class ZT:
    """Dense tensor with per-leg outgoing Z_N charges."""

    __slots__ = ("data", "q", "N")

    def __init__(self, data, q=None, N=1):
        self.data = np.asarray(data, dtype=np.float64)
        self.N = int(N)
        if q is None:
            q = [np.zeros(d, dtype=np.int64) for d in self.data.shape]
        self.q = [np.asarray(x, dtype=np.int64) % self.N for x in q]
        if len(self.q) != self.data.ndim:
            raise ValueError("one charge vector per leg required")
        for d, x in zip(self.data.shape, self.q):
            if len(x) != d:
                raise ValueError("charge vector length must match leg dimension")

    @property
    def shape(self):
        return self.data.shape

    def copy(self):
        return ZT(self.data.copy(), [x.copy() for x in self.q], self.N)

    def check_symmetric(self, atol=0.0):
        """Max |entry| that violates charge conservation (should be 0)."""
        tot = np.zeros(self.data.shape, dtype=np.int64)
        for leg, x in enumerate(self.q):
            shp = [1] * self.data.ndim
            shp[leg] = len(x)
            tot = tot + x.reshape(shp)
        bad = (tot % self.N) != 0
        return float(np.max(np.abs(self.data[bad]))) if bad.any() else 0.0


def permute(t: ZT, perm) -> ZT:
    """TensorKit ``permute``/``transpose`` for bosonic abelian sectors."""
    perm = tuple(perm)
    return ZT(np.transpose(t.data, perm).copy(), [t.q[p] for p in perm], t.N)


def contract(out_labels, *operands):
    """einsum over charge-labelled tensors.

    ``operands`` are ``(ZT, labels, conj)`` triples; labels are ints, negative
    ones are open and ordered by ``out_labels``.  ``conj`` flips charges
    (the data is real).  Every contracted label must join legs with opposite
    charges, which is the charge-flow check of a TensorKit contraction.
    """
    N = operands[0][0].N
    where = {}
    for k, (t, labels, conj) in enumerate(operands):
        if t.N != N:
            raise ValueError("mixed symmetry groups")
        if len(labels) != t.data.ndim:
            raise ValueError("label count does not match tensor rank")
        for leg, lab in enumerate(labels):
            qv = (-t.q[leg]) % N if conj else t.q[leg]
            where.setdefault(lab, []).append((k, leg, qv, t.data.shape[leg]))
    for lab, occ in where.items():
        if lab < 0:
            if len(occ) != 1:
                raise ValueError(f"open label {lab} used twice")
            continue
        if len(occ) != 2:
            raise ValueError(f"contracted label {lab} must appear twice")
        (_, _, qa, da), (_, _, qb, db) = occ
        if da != db:
            raise ValueError(f"dimension mismatch on label {lab}: {da} vs {db}")
        if np.any((qa + qb) % N != 0):
            raise ValueError(f"charge-flow mismatch on label {lab}")
    args = []
    for t, labels, _ in operands:
        args.append(t.data)
        args.append([_lbl(x) for x in labels])
    out = [_lbl(x) for x in out_labels]
    data = np.einsum(*args, out, optimize="optimal")
    q = [where[lab][0][2] for lab in out_labels]
    return ZT(data, q, N)


def _lbl(x):
    # einsum sublist labels must be non-negative ints
    return 2 * x if x >= 0 else -2 * x - 1


def _row_charges(t: ZT, legs):
    tot = np.zeros(1, dtype=np.int64)
    for leg in legs:
        tot = (tot[:, None] + t.q[leg][None, :]).reshape(-1)
    return tot % t.N


def matrix_blocks(t: ZT, ncod: int):
    """Yield (c, row_idx, col_idx) of the coupled-charge blocks of the
    (legs[:ncod] | legs[ncod:]) matrix view, rows in C order of the legs."""
    nl = t.data.ndim
    rq = _row_charges(t, range(ncod))
    cq = (-_row_charges(t, range(ncod, nl))) % t.N
    for c in range(t.N):
        r = np.nonzero(rq == c)[0]
        k = np.nonzero(cq == c)[0]
        yield c, r, k


def block_row_rank(t: ZT, ncod: int, rows):
    """Position of each dense row (C-order index over legs[:ncod]) inside its
    coupled-charge block in TensorKit's block layout: sector tuples
    (labels = outgoing charges) in colexicographic order, first leg fastest,
    and inside a tuple the elements column-major over the legs.  The
    pivoted-row gauge breaks ties between rows in this order."""
    rows = np.asarray(rows)
    shape = t.data.shape[:ncod]
    idx = np.unravel_index(rows, shape)
    tup = np.zeros(len(rows), dtype=np.int64)
    inner = np.zeros(len(rows), dtype=np.int64)
    stride = np.ones(len(rows), dtype=np.int64)
    for l in range(ncod):
        q = np.asarray(t.q[l]) % t.N
        pos = np.zeros(len(q), dtype=np.int64)
        dim = np.zeros(t.N, dtype=np.int64)
        for i, a in enumerate(q):
            pos[i] = dim[a]
            dim[a] += 1
        a = q[idx[l]]
        tup += a * t.N ** l
        inner += pos[idx[l]] * stride
        stride = stride * dim[a]
    order = np.lexsort((inner, tup))
    rank = np.empty(len(rows), dtype=np.int64)
    rank[order] = np.arange(len(rows))
    return rank


class Trunc:
    """Diagnostics of one truncated factorization (the sigma gap log)."""

    __slots__ = ("kept", "first_dropped", "err", "sectors", "n", "top")

    def __init__(self, kept, first_dropped, err, sectors, n, top=0.0):
        self.kept = kept            # smallest kept value
        self.first_dropped = first_dropped  # largest dropped value (0 if none)
        self.err = err              # 2-norm of the dropped values
        self.sectors = sectors      # kept count per coupled charge
        self.n = n                  # matrix dimension (rows)
        self.top = top              # largest value over all blocks

    def rel_gap(self):
        if self.first_dropped == 0.0 or self.kept == 0.0:
            return np.inf
        return (self.kept - self.first_dropped) / self.kept


def _global_cut(vals_per_block, chi, by_abs):
    """Stable descending selection of the chi largest values across blocks.

    Returns a per-block list of kept local indices (in block order)."""
    allv, owner = [], []
    for b, v in enumerate(vals_per_block):
        allv.append(np.abs(v) if by_abs else v)
        owner.extend((b, i) for i in range(len(v)))
    if allv:
        flat = np.concatenate(allv)
    else:
        flat = np.zeros(0)
    order = np.argsort(-flat, kind="stable")
    keep = order[: min(chi, len(flat))]
    drop = order[min(chi, len(flat)):]
    kept = [[] for _ in vals_per_block]
    for j in keep:
        b, i = owner[j]
        kept[b].append(i)
    kept = [sorted(x) for x in kept]
    return kept, flat, keep, drop


# Pivoted-row gauge of a truncated SVD (used by the impurity TRG, whose
# impurity norm contracts factors of different SVDs with each other and so
# depends on the basis each SVD returns; not part of the reference, which
# takes whatever LAPACK returns).  Per block: singular values within
# GAUGE_CLUSTER * s_max of their neighbour form a cluster, values at or
# below GAUGE_NULL * s_max are null (their vectors are set to zero); inside
# each cluster the basis is rebuilt as u_t = P_t e_r / |P_t e_r|, with P_t
# the projector on the part of the cluster's span orthogonal to u_0..u_{t-1}
# and r the first row whose residual norm e_r^T P_t e_r is within GAUGE_TIE
# of the largest.  Every u_t is then a function of the span alone (any
# rotation or sign of the input basis gives the same result; a lone vector
# only has its sign set), is positive at its pivot row, and a symmetry that acts diagonally on the row
# coordinates (the hidden Z2 of a Trivial Ising tensor) maps each u_t to
# +-u_t, so the kept vectors never mix even and odd sectors.
GAUGE_CLUSTER = 1e-9
GAUGE_NULL = 1e-14
GAUGE_TIE = 1e-10


def pivoted_gauge(u, s, vt, row_rank=None):
    """u (rows x k), s (k, descending), vt (k x cols) of one block: the
    same factorization in the pivoted-row gauge; returns (u, vt).  Ties
    between rows go to the smallest ``row_rank`` (default: row index)."""
    u = np.array(u, dtype=np.float64, copy=True)
    vt = np.array(vt, dtype=np.float64, copy=True)
    k = len(s)
    if k == 0:
        return u, vt
    null = s <= GAUGE_NULL * s[0]
    u[:, null] = 0.0
    vt[null, :] = 0.0
    i = 0
    while i < k:
        if null[i]:
            break
        j = i + 1
        while j < k and not null[j] and s[j - 1] - s[j] <= GAUGE_CLUSTER * s[0]:
            j += 1
        d = j - i
        Uc = u[:, i:j]
        Cq = np.zeros((d, d))
        rn0 = np.sum(Uc * Uc, axis=1)
        for t in range(d):
            proj = Uc @ Cq[:, :t]
            rn = rn0 - np.sum(proj * proj, axis=1)
            cand = np.flatnonzero(rn >= (1.0 - GAUGE_TIE) * rn.max())
            r = int(cand[0] if row_rank is None else cand[np.argmin(row_rank[cand])])
            if d == 1:                     # a lone vector: only its sign
                Cq[0, 0] = 1.0 if Uc[r, 0] >= 0.0 else -1.0
                break
            c = Uc[r, :] - Cq[:, :t] @ proj[r, :]
            Cq[:, t] = c / np.sqrt(rn[r])
        u[:, i:j] = Uc @ Cq
        vt[i:j, :] = Cq.T @ vt[i:j, :]
        i = j
    return u, vt


def svd_trunc(t: ZT, ncod: int, chi: int, gauge: str = "lapack"):
    """Truncated SVD of the (legs[:ncod] | legs[ncod:]) matrix view.

    Returns (U, s, Vt, info): U has legs (cod..., new), Vt (new, dom...),
    s the kept singular values in new-leg order (by sector, descending
    within a sector).  ``gauge`` "lapack" keeps LAPACK's singular vectors
    (the reference), "pivoted" applies pivoted_gauge to each block's kept
    triplets."""
    nl = t.data.ndim
    rshape = t.data.shape[:ncod]
    cshape = t.data.shape[ncod:]
    M = t.data.reshape(int(np.prod(rshape)), int(np.prod(cshape)))
    blocks = []
    for c, r, k in matrix_blocks(t, ncod):
        if len(r) == 0 or len(k) == 0:
            continue
        u, s, vt = np.linalg.svd(M[np.ix_(r, k)], full_matrices=False)
        blocks.append((c, r, k, u, s, vt))
    kept, flat, keep, drop = _global_cut([b[4] for b in blocks], chi, False)
    newdim = sum(len(x) for x in kept)
    U = np.zeros((M.shape[0], newdim))
    Vt = np.zeros((newdim, M.shape[1]))
    s_out = np.zeros(newdim)
    qnew = np.zeros(newdim, dtype=np.int64)
    sectors = {}
    pos = 0
    for (c, r, k, u, s, vt), idx in zip(blocks, kept):
        if not idx:
            continue
        m = len(idx)
        ub, vb = u[:, idx], vt[idx, :]
        if gauge == "pivoted":
            ub, vb = pivoted_gauge(ub, s[idx], vb, block_row_rank(t, ncod, r))
        elif gauge != "lapack":
            raise ValueError("gauge must be 'lapack' or 'pivoted'")
        U[r, pos:pos + m] = ub
        Vt[pos:pos + m, k] = vb
        s_out[pos:pos + m] = s[idx]
        qnew[pos:pos + m] = c
        sectors[c] = m
        pos += m
    info = Trunc(float(flat[keep].min()) if len(keep) else 0.0,
                 float(flat[drop].max()) if len(drop) else 0.0,
                 float(np.linalg.norm(flat[drop])) if len(drop) else 0.0,
                 sectors, M.shape[0], float(flat.max()) if len(flat) else 0.0)
    Ut = ZT(U.reshape(*rshape, newdim), [t.q[i] for i in range(ncod)] + [(-qnew) % t.N], t.N)
    Vtt = ZT(Vt.reshape(newdim, *cshape), [qnew] + [t.q[i] for i in range(ncod, nl)], t.N)
    return Ut, s_out, Vtt, info


def project_hermitian(t: ZT, ncod: int) -> ZT:
    """(M + M^H)/2 of the matrix view (MatrixAlgebraKit ``project_hermitian!``)."""
    rshape = t.data.shape[:ncod]
    n = int(np.prod(rshape))
    M = t.data.reshape(n, -1)
    if M.shape[1] != n:
        raise ValueError("project_hermitian needs a square matrix view")
    return ZT(((M + M.T) / 2).reshape(t.data.shape), t.q, t.N)


def eigh_trunc(t: ZT, ncod: int, chi: int):
    """Truncated Hermitian eigendecomposition of the matrix view.

    Returns (lam, U, eps, info): U has legs (cod..., new); eps is the 2-norm
    of the discarded eigenvalues (the truncation error compared at
    src/schemes/hotrg.jl:112,148)."""
    rshape = t.data.shape[:ncod]
    n = int(np.prod(rshape))
    M = t.data.reshape(n, n)
    blocks = []
    for c, r, k in matrix_blocks(t, ncod):
        if len(r) == 0:
            continue
        if not np.array_equal(r, k):
            raise ValueError("eigh needs matching row/column sectors")
        w, v = np.linalg.eigh(M[np.ix_(r, r)])
        blocks.append((c, r, w, v))
    kept, flat, keep, drop = _global_cut([b[2] for b in blocks], chi, True)
    newdim = sum(len(x) for x in kept)
    U = np.zeros((n, newdim))
    lam = np.zeros(newdim)
    qnew = np.zeros(newdim, dtype=np.int64)
    sectors = {}
    pos = 0
    for (c, r, w, v), idx in zip(blocks, kept):
        if not idx:
            continue
        m = len(idx)
        U[r, pos:pos + m] = v[:, idx]
        lam[pos:pos + m] = w[idx]
        qnew[pos:pos + m] = c
        sectors[c] = m
        pos += m
    eps = float(np.linalg.norm(flat[drop])) if len(drop) else 0.0
    info = Trunc(float(flat[keep].min()) if len(keep) else 0.0,
                 float(flat[drop].max()) if len(drop) else 0.0,
                 eps, sectors, n, float(np.abs(flat).max()) if len(flat) else 0.0)
    Ut = ZT(U.reshape(*rshape, newdim), [t.q[i] for i in range(ncod)] + [(-qnew) % t.N], t.N)
    return lam, Ut, eps, info


def scale_leg(t: ZT, leg: int, w) -> ZT:
    """Multiply leg ``leg`` by the diagonal weights ``w`` (a DiagonalTensorMap
    contracted onto that leg)."""
    shp = [1] * t.data.ndim
    shp[leg] = len(w)
    return ZT(t.data * np.asarray(w).reshape(shp), t.q, t.N)
