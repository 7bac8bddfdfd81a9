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
from .tensors import SymTensor

KINDS = {"TRG": _lib.TNR_TRG, "BTRG": _lib.TNR_BTRG, "HOTRG": _lib.TNR_HOTRG,
         "ATRG": _lib.TNR_ATRG}


class Request:
    """One run: scheme kind, initial tensor, chi, maxiter."""

    def __init__(self, kind, T: SymTensor, chi, maxiter, k=-0.5, finalize_beginning=True):
        self.kind = KINDS[kind] if isinstance(kind, str) else int(kind)
        self.T = T
        self.chi = int(chi)
        self.maxiter = int(maxiter)
        self.k = float(k)
        self.finalize_beginning = bool(finalize_beginning)
        self._dims = np.ascontiguousarray(T.leg_dims, dtype=np.int32).ravel()

    @property
    def n_norms(self):
        return max(self.maxiter, 1) + (1 if self.finalize_beginning else 0)

    def desc(self):
        return _lib.RunDesc(self.kind, self.T.N, self._dims.ctypes.data_as(C.POINTER(C.c_int32)),
                            self.T.data.ctypes.data_as(C.POINTER(C.c_double)), self.T.data.size,
                            self.k, self.chi, self.maxiter, int(self.finalize_beginning))


def run_batch(requests, precision=_lib.TNR_PREC_FP64):
    """Norm vectors of every request, computed together on the device."""
    descs = (_lib.RunDesc * len(requests))(*[r.desc() for r in requests])
    out = np.zeros(sum(r.n_norms for r in requests))
    check(lib().tnr_run_batch(len(requests), descs, precision,
                              out.ctypes.data_as(C.POINTER(C.c_double))))
    res, o = [], 0
    for r in requests:
        res.append(out[o:o + r.n_norms].copy())
        o += r.n_norms
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
        out = np.zeros(req.n_norms)
        n = C.c_int()
        ts, td = C.c_double(), C.c_double()
        st = lib().tnr_server_poll(self._h, rid, out.ctypes.data_as(C.POINTER(C.c_double)),
                                   C.byref(n), C.byref(ts), C.byref(td))
        if st == _lib.TNR_ERR_BUSY:
            return None
        check(st)
        return out[:n.value], ts.value, td.value

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
