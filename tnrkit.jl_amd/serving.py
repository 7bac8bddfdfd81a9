"""Batched runs and the continuous-batching serving loop (host side).

``run_batch`` drives many independent ``run!`` calls through grouped device
launches; ``Server`` admits requests as HBM and slots free up and advances
every resident scheme one coarse-graining step per ``pump``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib

KINDS = {"TRG": _lib.TNR_TRG, "BTRG": _lib.TNR_BTRG, "HOTRG": _lib.TNR_HOTRG,
         "ATRG": _lib.TNR_ATRG, "HOTRG_3D": _lib.TNR_HOTRG_3D,
         "ImpurityTRG": _lib.TNR_IMPURITY_TRG, "ImpurityHOTRG": _lib.TNR_IMPURITY_HOTRG}


class Request:
    """One run: scheme kind, initial tensor(s), chi, maxiter.  ``T`` is a
    SymTensor (an NTensor with 2 + 4 legs for HOTRG_3D), or for the impurity
    methods the list of their tensors in constructor order."""

    def __init__(self, kind, T, chi, maxiter, k=-0.5, finalize_beginning=True):
        self.kind = KINDS[kind] if isinstance(kind, str) else int(kind)
        tensors = list(T) if isinstance(T, (list, tuple)) else [T]
        self.T = tensors[0]
        self._extra = [np.ascontiguousarray(t.data, dtype=np.float64) for t in tensors[1:]]
        self._extra_ptrs = (C.POINTER(C.c_double) * max(1, len(self._extra)))(
            *[e.ctypes.data_as(C.POINTER(C.c_double)) for e in self._extra])
        self.chi = int(chi)
        self.maxiter = int(maxiter)
        self.k = float(k)
        self.finalize_beginning = bool(finalize_beginning)
        self.ncomp = lib().tnr_ncomponents(self.kind)
        if self.ncomp < 1:
            check(self.ncomp)
        self._dims = np.ascontiguousarray(self.T.leg_dims, dtype=np.int32).ravel()

    @property
    def n_norms(self):
        """Rows of the run's norm output (one per finalize!)."""
        return max(self.maxiter, 1) + (1 if self.finalize_beginning else 0)

    def desc(self):
        imp = C.cast(self._extra_ptrs, C.POINTER(C.POINTER(C.c_double))) if self._extra else None
        return _lib.RunDesc(self.kind, self.T.N, self._dims.ctypes.data_as(C.POINTER(C.c_int32)),
                            self.T.data.ctypes.data_as(C.POINTER(C.c_double)), self.T.data.size,
                            self.k, self.chi, self.maxiter, int(self.finalize_beginning), imp)

    def shape(self, flat):
        """The run's norms: a vector, or rows of ncomp values for the impurity methods."""
        return flat.copy() if self.ncomp == 1 else flat.reshape(-1, self.ncomp).copy()


def run_batch(requests, precision=_lib.TNR_PREC_FP64):
    """Norms of every request, computed together on the device."""
    descs = (_lib.RunDesc * len(requests))(*[r.desc() for r in requests])
    out = np.zeros(sum(r.n_norms * r.ncomp for r in requests))
    check(lib().tnr_run_batch(len(requests), descs, precision,
                              out.ctypes.data_as(C.POINTER(C.c_double))))
    res, o = [], 0
    for r in requests:
        n = r.n_norms * r.ncomp
        res.append(r.shape(out[o:o + n]))
        o += n
    return res


class Server:
    def __init__(self, hbm_budget_bytes=0, max_slots=1024, precision=_lib.TNR_PREC_FP64):
        self._h = C.c_void_p()
        check(lib().tnr_server_create(C.byref(self._h), int(hbm_budget_bytes), int(max_slots),
                                      int(precision)))
        self._reqs = {}

    def submit(self, req: Request) -> int:
        i = C.c_int64()
        d = req.desc()
        check(lib().tnr_server_submit(self._h, C.byref(d), C.byref(i)))
        self._reqs[i.value] = req
        return i.value

    def pump(self) -> int:
        a = C.c_int()
        check(lib().tnr_server_pump(self._h, C.byref(a)))
        return a.value

    def poll(self, rid):
        """(norms, t_submit, t_done) or None while the request is in flight."""
        req = self._reqs[rid]
        out = np.zeros(req.n_norms * req.ncomp)
        n = C.c_int()
        ts, td = C.c_double(), C.c_double()
        st = lib().tnr_server_poll(self._h, rid, out.ctypes.data_as(C.POINTER(C.c_double)),
                                   C.byref(n), C.byref(ts), C.byref(td))
        if st == _lib.TNR_ERR_BUSY:
            return None
        check(st)
        return req.shape(out[:n.value * req.ncomp]), ts.value, td.value

    def stats(self):
        m = C.c_size_t()
        r, q = C.c_int(), C.c_int()
        check(lib().tnr_server_stats(self._h, C.byref(m), C.byref(r), C.byref(q)))
        return {"hbm_in_use": m.value, "resident": r.value, "queued": q.value}

    def close(self):
        if self._h:
            lib().tnr_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
