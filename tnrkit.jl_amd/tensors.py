"""Host tensors in the engine's ABI format.

A ``SymTensor`` is a 4-leg map V1 (x) V2 <- V3 (x) V4 with Z_N sectors
(N = 1 is Trivial).  ``leg_dims[i, a]`` is the dimension of sector ``a`` in
V_{i+1}; ``data`` holds, for coupled charge c = 0..N-1, the column-major
block matrix at c with rows over codomain sector pairs (a1, a2), a1+a2 = c,
in colexicographic order (a1 fastest), columns over (a3, a4) likewise.  For
N = 1 ``data`` is the column-major dense array T[i1, i2, i3, i4], i.e.
TensorKit's ``t.data`` of a Trivial TensorMap.
"""
from __future__ import annotations

import numpy as np


class SymTensor:
    __slots__ = ("N", "leg_dims", "data")

    def __init__(self, N, leg_dims, data):
        self.N = int(N)
        self.leg_dims = np.ascontiguousarray(np.asarray(leg_dims, dtype=np.int32).reshape(4, self.N))
        self.data = np.ascontiguousarray(np.asarray(data, dtype=np.float64).ravel())
        if self.data.size != block_size(self.N, self.leg_dims):
            raise ValueError("data size does not match the sector dimensions")

    def copy(self):
        return SymTensor(self.N, self.leg_dims.copy(), self.data.copy())

    def to_dense(self):
        """Dense 4D array (index order per leg: sector by sector)."""
        return _blocks_to_dense(self.N, self.leg_dims, self.data)

    @staticmethod
    def from_dense(T, leg_labels, N):
        """Build from a dense 4D array whose index i of leg k carries space
        label leg_labels[k][i]; entries violating a1+a2 = a3+a4 must be 0."""
        return _dense_to_blocks(np.asarray(T, dtype=np.float64), leg_labels, N)


def _tuples(N, d_a, d_b, c):
    """Colex sector pairs (a, b) with a + b = c mod N and their sizes."""
    out = []
    for t in range(N * N):
        a, b = t % N, t // N
        if (a + b) % N == c:
            out.append((a, b, d_a[a] * d_b[b]))
    return out


def block_size(N, leg_dims):
    n = 0
    for c in range(N):
        R = sum(s for _, _, s in _tuples(N, leg_dims[0], leg_dims[1], c))
        Cc = sum(s for _, _, s in _tuples(N, leg_dims[2], leg_dims[3], c))
        n += R * Cc
    return n


def _sector_index(dims):
    """Index offsets of each sector along a leg (sectors in label order)."""
    off = np.concatenate([[0], np.cumsum(dims)]).astype(int)
    return off


def _dense_to_blocks(T, leg_labels, N):
    labels = [np.asarray(x, dtype=int) % N for x in leg_labels]
    dims = np.zeros((4, N), dtype=np.int32)
    order = []
    for k in range(4):
        for a in range(N):
            dims[k, a] = int(np.sum(labels[k] == a))
        # indices of leg k grouped by sector, original order inside a sector
        order.append([np.nonzero(labels[k] == a)[0] for a in range(N)])
    chunks = []
    for c in range(N):
        rows = []
        for a1, a2, _ in _tuples(N, dims[0], dims[1], c):
            i1, i2 = order[0][a1], order[1][a2]
            rows.append((i1[:, None] + 0 * i2[None, :], i2[None, :] + 0 * i1[:, None]))
        cols = []
        for a3, a4, _ in _tuples(N, dims[2], dims[3], c):
            i3, i4 = order[2][a3], order[3][a4]
            cols.append((i3[:, None] + 0 * i4[None, :], i4[None, :] + 0 * i3[:, None]))
        if not rows or not cols:
            continue
        r1 = np.concatenate([x.ravel(order="F") for x, _ in rows])
        r2 = np.concatenate([y.ravel(order="F") for _, y in rows])
        c3 = np.concatenate([x.ravel(order="F") for x, _ in cols])
        c4 = np.concatenate([y.ravel(order="F") for _, y in cols])
        B = T[r1[:, None], r2[:, None], c3[None, :], c4[None, :]]
        chunks.append(B.ravel(order="F"))
    data = np.concatenate(chunks) if chunks else np.zeros(0)
    return SymTensor(N, dims, data)


def _blocks_to_dense(N, dims, data):
    tot = [int(dims[k].sum()) for k in range(4)]
    off = [_sector_index(dims[k]) for k in range(4)]
    T = np.zeros(tot)
    pos = 0
    for c in range(N):
        rows = _tuples(N, dims[0], dims[1], c)
        cols = _tuples(N, dims[2], dims[3], c)
        R = sum(s for _, _, s in rows)
        Cc = sum(s for _, _, s in cols)
        if R == 0 or Cc == 0:
            continue
        B = data[pos:pos + R * Cc].reshape(R, Cc, order="F")
        pos += R * Cc
        r = 0
        for a1, a2, sr in rows:
            cc = 0
            for a3, a4, sc in cols:
                sub = B[r:r + sr, cc:cc + sc].reshape(dims[0][a1], dims[1][a2], dims[2][a3],
                                                     dims[3][a4], order="F")
                T[off[0][a1]:off[0][a1 + 1], off[1][a2]:off[1][a2 + 1],
                  off[2][a3]:off[2][a3 + 1], off[3][a4]:off[3][a4 + 1]] = sub
                cc += sc
            r += sr
    return T


def trivial(T):
    """SymTensor of a dense (Trivial) 4D array."""
    T = np.asarray(T, dtype=np.float64)
    return SymTensor(1, np.array(T.shape, dtype=np.int32).reshape(4, 1), T.ravel(order="F"))


# ------------------------------------------------------------ n-leg tensors
def _tuples_k(N, dims, c):
    """Colex sector tuples (first leg fastest) over the legs whose per-sector
    dimensions are ``dims`` (a list of arrays), labels summing to c mod N,
    with their element counts."""
    k = len(dims)
    out = []
    for t in range(N ** k):
        labels, x = [], t
        for _ in range(k):
            labels.append(x % N)
            x //= N
        if sum(labels) % N == c:
            size = 1
            for j, a in enumerate(labels):
                size *= int(dims[j][a])
            out.append((tuple(labels), size))
    return out


class NTensor:
    """A map V1 (x) .. (x) V_ncod <- V_{ncod+1} (x) .. (x) V_nleg in the ABI
    block format (the 3D tensors of HOTRG_3D: 2 + 4 legs).  ``leg_dims[i, a]``
    is the dimension of sector a in V_{i+1}; ``data`` holds, for coupled
    charge c = 0..N-1, the column-major block with rows over codomain sector
    tuples and columns over domain sector tuples summing to c, both
    colexicographic, elements column-major inside a tuple."""
    __slots__ = ("N", "leg_dims", "data", "ncod")

    def __init__(self, N, leg_dims, data, ncod):
        self.N = int(N)
        self.leg_dims = np.ascontiguousarray(np.asarray(leg_dims, dtype=np.int32).reshape(-1, self.N))
        self.ncod = int(ncod)
        self.data = np.ascontiguousarray(np.asarray(data, dtype=np.float64).ravel())
        if self.data.size != self._size():
            raise ValueError("data size does not match the sector dimensions")

    @property
    def nleg(self):
        return self.leg_dims.shape[0]

    def _size(self):
        cod, dom = list(self.leg_dims[:self.ncod]), list(self.leg_dims[self.ncod:])
        n = 0
        for c in range(self.N):
            R = sum(s for _, s in _tuples_k(self.N, cod, c))
            Cc = sum(s for _, s in _tuples_k(self.N, dom, c))
            n += R * Cc
        return n

    def copy(self):
        return NTensor(self.N, self.leg_dims.copy(), self.data.copy(), self.ncod)

    @staticmethod
    def from_dense(T, leg_labels, N, ncod):
        """From a dense array whose index i of leg k carries space label
        leg_labels[k][i] (entries that violate charge conservation must be 0)."""
        T = np.asarray(T, dtype=np.float64)
        nleg = T.ndim
        labels = [np.asarray(x, dtype=int) % N for x in leg_labels]
        dims = np.zeros((nleg, N), dtype=np.int32)
        order = []
        for k in range(nleg):
            for a in range(N):
                dims[k, a] = int(np.sum(labels[k] == a))
            order.append([np.nonzero(labels[k] == a)[0] for a in range(N)])

        def index_lists(legs, tup):
            grids = np.meshgrid(*[order[l][a] for l, a in zip(legs, tup)], indexing="ij")
            return [g.ravel(order="F") for g in grids]

        chunks = []
        cod, dom = list(range(ncod)), list(range(ncod, nleg))
        for c in range(N):
            rows = [index_lists(cod, t) for t, _ in _tuples_k(N, [dims[l] for l in cod], c)]
            cols = [index_lists(dom, t) for t, _ in _tuples_k(N, [dims[l] for l in dom], c)]
            if not rows or not cols:
                continue
            r = [np.concatenate([x[j] for x in rows]) for j in range(ncod)]
            cc = [np.concatenate([x[j] for x in cols]) for j in range(nleg - ncod)]
            idx = tuple([ri[:, None] for ri in r] + [ci[None, :] for ci in cc])
            chunks.append(T[idx].ravel(order="F"))
        data = np.concatenate(chunks) if chunks else np.zeros(0)
        return NTensor(N, dims, data, ncod)

    def to_dense(self):
        N, ncod, nleg = self.N, self.ncod, self.nleg
        off = [_sector_index(self.leg_dims[k]) for k in range(nleg)]
        T = np.zeros([int(self.leg_dims[k].sum()) for k in range(nleg)])
        cod, dom = list(range(ncod)), list(range(ncod, nleg))
        pos = 0
        for c in range(N):
            rows = _tuples_k(N, [self.leg_dims[l] for l in cod], c)
            cols = _tuples_k(N, [self.leg_dims[l] for l in dom], c)
            R = sum(s for _, s in rows)
            Cc = sum(s for _, s in cols)
            if R == 0 or Cc == 0:
                continue
            B = self.data[pos:pos + R * Cc].reshape(R, Cc, order="F")
            pos += R * Cc
            r = 0
            for rt, sr in rows:
                cc = 0
                for ct, sc in cols:
                    shape = [self.leg_dims[l][a] for l, a in zip(cod, rt)] + \
                            [self.leg_dims[l][a] for l, a in zip(dom, ct)]
                    sub = B[r:r + sr, cc:cc + sc].reshape(shape, order="F")
                    sl = tuple(slice(off[l][a], off[l][a + 1]) for l, a in
                               zip(cod + dom, rt + ct))
                    T[sl] = sub
                    cc += sc
                r += sr
        return T
