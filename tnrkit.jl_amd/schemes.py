"""TNRKit's scheme interface over the HIP engine.

Mirrors src/schemes/tnrscheme.jl:16-59 (``run!`` -> ``run``, ``Finalizer``),
src/utility/stopping.jl (``maxiter``, ``convcrit``, ``&``), the scheme
constructors ``TRG(T)`` (src/schemes/trg.jl:30), ``BTRG(T, k=-0.5)``
(src/schemes/btrg.jl:39-48) and ``HOTRG(T)`` (src/schemes/hotrg.jl:31), and
``free_energy`` (src/utility/free_energy.jl:13-20).  Every step and finalize
runs in libtnr_hip.so on the GPU; there is no host fallback.
"""
from __future__ import annotations

import ctypes as C
import logging

import numpy as np

from . import _lib
from ._lib import check, lib
from .tensors import NTensor, SymTensor

log = logging.getLogger("tnrkit_amd")

FP64 = _lib.TNR_PREC_FP64
SPLIT32 = _lib.TNR_PREC_SPLIT32


# ----------------------------------------------------------------- truncation
class truncrank:
    """MatrixAlgebraKit ``truncrank(howmany)``: keep the howmany largest values."""

    def __init__(self, howmany):
        self.howmany = int(howmany)

    def __repr__(self):
        return f"truncrank({self.howmany})"


# ------------------------------------------------------------------ stopping
class stopcrit:
    def __and__(self, other):
        return MultipleCrit([self, other])


class maxiter(stopcrit):
    """src/utility/stopping.jl:3-5,18: continue while steps < n."""

    def __init__(self, n):
        self.n = int(n)

    def __call__(self, steps, data):
        return steps < self.n

    def info(self, steps, data):
        return f"Maximum amount of iterations reached: {steps}"


class convcrit(stopcrit):
    """src/utility/stopping.jl:7-10,19: continue while delta < f(steps, data)."""

    def __init__(self, delta, f):
        self.delta = float(delta)
        self.f = f

    def __call__(self, steps, data):
        return self.delta < self.f(steps, data)

    def info(self, steps, data):
        return f"Convergence criterion reached: {self.f(steps, data):.3e} <= {self.delta:.3e}"


class MultipleCrit(stopcrit):
    """src/utility/stopping.jl:12-16,22: every criterion must hold."""

    def __init__(self, crits):
        self.crits = list(crits)

    def __call__(self, steps, data):
        return all(c(steps, data) for c in self.crits)

    def info(self, steps, data):
        for c in self.crits:
            if not c(steps, data):
                return c.info(steps, data)
        return ""


def trivial_convcrit(delta):
    return convcrit(delta, lambda steps, data: data[-1])


# ------------------------------------------------------------------ schemes
class TNRScheme:
    """One device-resident scheme (src/schemes/tnrscheme.jl:31)."""

    kind = None
    scalefactor = 2.0

    def __init__(self, T: SymTensor, k: float = -0.5, precision: int = FP64):
        if not isinstance(T, SymTensor):
            raise TypeError("T must be a SymTensor")
        self._T0 = T
        self.k = float(k)
        self.precision = int(precision)
        self._h = C.c_void_p()
        dims = np.ascontiguousarray(T.leg_dims, dtype=np.int32).ravel()
        check(lib().tnr_create(C.byref(self._h), self.kind, T.N,
                               dims.ctypes.data_as(C.POINTER(C.c_int32)),
                               T.data.ctypes.data_as(C.POINTER(C.c_double)), T.data.size,
                               self.k, self.precision))

    def close(self):
        if self._h:
            lib().tnr_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state
    @property
    def T(self) -> SymTensor:
        N = self._T0.N
        dims = np.zeros(4 * N, dtype=np.int32)
        n = C.c_size_t()
        check(lib().tnr_get_dims(self._h, dims.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(n)))
        data = np.zeros(n.value)
        check(lib().tnr_get_tensor(self._h, data.ctypes.data_as(C.POINTER(C.c_double)), n.value))
        return SymTensor(N, dims.reshape(4, N), data)

    @T.setter
    def T(self, t: SymTensor):
        dims = np.ascontiguousarray(t.leg_dims, dtype=np.int32).ravel()
        check(lib().tnr_set_tensor(self._h, dims.ctypes.data_as(C.POINTER(C.c_int32)),
                                   t.data.ctypes.data_as(C.POINTER(C.c_double)), t.data.size))

    def trunc_log(self):
        n = C.c_int()
        check(lib().tnr_get_trunc_log(self._h, None, 0, C.byref(n)))
        out = np.zeros(4 * max(n.value, 1))
        check(lib().tnr_get_trunc_log(self._h, out.ctypes.data_as(C.POINTER(C.c_double)),
                                      n.value, C.byref(n)))
        return out[:4 * n.value].reshape(-1, 4)

    def __repr__(self):
        return f"{type(self).__name__} - T: {self._T0.N}-sector tensor on the GPU"


class TRG(TNRScheme):
    """src/schemes/trg.jl:26-41 (each step rescales the area by 2)."""
    kind = _lib.TNR_TRG

    def __init__(self, T, precision=FP64):
        super().__init__(T, 0.0, precision)


class BTRG(TNRScheme):
    """src/schemes/btrg.jl:26-94 (k = -1/2 by default)."""
    kind = _lib.TNR_BTRG

    def __init__(self, T, k=-0.5, precision=FP64):
        super().__init__(T, k, precision)

    def bond_weights(self, which):
        T = self.T
        d = T.leg_dims[1] if which == 1 else T.leg_dims[0]
        w = np.zeros(int(np.sum(d)))
        check(lib().tnr_get_bond_weights(self._h, which, w.ctypes.data_as(C.POINTER(C.c_double)),
                                         w.size))
        return w


class HOTRG(TNRScheme):
    """src/schemes/hotrg.jl:27-160 (each step rescales the area by 4)."""
    kind = _lib.TNR_HOTRG
    scalefactor = 4.0

    def __init__(self, T, precision=FP64):
        super().__init__(T, 0.0, precision)


class ATRG(TNRScheme):
    """ATRG(T) (src/schemes/atrg.jl:26-33): anisotropic TRG, one step = two
    half-steps of three truncated SVDs (free energies with scalefactor 4,
    as in test/schemes.jl:131-136)."""
    kind = _lib.TNR_ATRG
    scalefactor = 4.0

    def __init__(self, T, precision=FP64):
        super().__init__(T, 0.0, precision)


class HOTRG_3D(TNRScheme):
    """HOTRG_3D(T) (src/schemes/hotrg3d.jl:27-141) for a 6-leg tensor
    T[z z'; Y X y x] (``NTensor`` with 2 + 4 legs); each step compresses z,
    rotates, three times (free energies with scalefactor 8)."""
    kind = _lib.TNR_HOTRG_3D
    scalefactor = 8.0

    def __init__(self, T: NTensor, precision: int = FP64):
        if not isinstance(T, NTensor) or T.nleg != 6 or T.ncod != 2:
            raise TypeError("T must be a 2 + 4 leg NTensor")
        self._T0 = T
        self.k = 0.0
        self.precision = int(precision)
        self._h = C.c_void_p()
        dims = np.ascontiguousarray(T.leg_dims, dtype=np.int32).ravel()
        check(lib().tnr_create_3d(C.byref(self._h), T.N, dims.ctypes.data_as(C.POINTER(C.c_int32)),
                                  T.data.ctypes.data_as(C.POINTER(C.c_double)), T.data.size,
                                  self.precision))

    @property
    def T(self) -> NTensor:
        N = self._T0.N
        dims = np.zeros(6 * N, dtype=np.int32)
        n = C.c_size_t()
        check(lib().tnr_get_dims(self._h, dims.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(n)))
        data = np.zeros(n.value)
        check(lib().tnr_get_tensor(self._h, data.ctypes.data_as(C.POINTER(C.c_double)), n.value))
        return NTensor(N, dims.reshape(6, N), data, 2)

    @T.setter
    def T(self, t: NTensor):
        dims = np.ascontiguousarray(t.leg_dims, dtype=np.int32).ravel()
        check(lib().tnr_set_tensor(self._h, dims.ctypes.data_as(C.POINTER(C.c_int32)),
                                   t.data.ctypes.data_as(C.POINTER(C.c_double)), t.data.size))


class _MultiScheme(TNRScheme):
    """A scheme holding the pure tensor and impurity tensors on its spaces."""
    ntensors = 0

    def __init__(self, *tensors, precision=FP64):
        if len(tensors) != self.ntensors or not all(isinstance(t, SymTensor) for t in tensors):
            raise TypeError(f"{type(self).__name__} takes {self.ntensors} SymTensors")
        T = tensors[0]
        self._T0 = T
        self.k = 0.0
        self.precision = int(precision)
        self._h = C.c_void_p()
        dims = np.ascontiguousarray(T.leg_dims, dtype=np.int32).ravel()
        datas = [np.ascontiguousarray(t.data, dtype=np.float64) for t in tensors]
        ptrs = (C.POINTER(C.c_double) * len(datas))(
            *[d.ctypes.data_as(C.POINTER(C.c_double)) for d in datas])
        sizes = (C.c_size_t * len(datas))(*[d.size for d in datas])
        check(lib().tnr_create_multi(C.byref(self._h), self.kind, T.N,
                                     dims.ctypes.data_as(C.POINTER(C.c_int32)), ptrs, sizes,
                                     len(datas), self.precision))

    def tensor(self, index):
        """Tensor ``index`` (0 = T, then the impurity tensors in constructor order)."""
        N = self._T0.N
        dims = np.zeros(4 * N, dtype=np.int32)
        n = C.c_size_t()
        check(lib().tnr_get_dims(self._h, dims.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(n)))
        data = np.zeros(n.value)
        check(lib().tnr_get_tensor_at(self._h, int(index), data.ctypes.data_as(C.POINTER(C.c_double)),
                                      n.value))
        return SymTensor(N, dims.reshape(4, N), data)


class ImpurityTRG(_MultiScheme):
    """src/schemes/impuritytrg.jl:40-103: ImpurityTRG(T, T_imp1, T_imp2,
    T_imp3, T_imp4); finalize! returns (npure, nimp) (finalize.jl:74-88)."""
    kind = _lib.TNR_IMPURITY_TRG
    ntensors = 5


class ImpurityHOTRG(_MultiScheme):
    """src/schemes/impurityhotrg.jl:27-95: ImpurityHOTRG(T, T_imp_order1_1,
    T_imp_order1_2, T_imp_order2); finalize! returns (n, n_11, n_12, n_2)
    (finalize.jl:90-101); each step rescales the area by 4."""
    kind = _lib.TNR_IMPURITY_HOTRG
    ntensors = 4
    scalefactor = 4.0


def _ncomp(scheme):
    return lib().tnr_ncomponents(scheme.kind)


# ---------------------------------------------------------------- driver
def step(scheme: TNRScheme, trunc: truncrank):
    """step!(scheme, trunc)"""
    check(lib().tnr_step(scheme._h, trunc.howmany))
    return scheme


def finalize(scheme: TNRScheme):
    """finalize!(scheme) (src/utility/finalize.jl:4-14); a tuple for the
    impurity schemes (finalize.jl:74-101)."""
    nc = _ncomp(scheme)
    if nc == 1:
        n = C.c_double()
        check(lib().tnr_finalize(scheme._h, C.byref(n)))
        return n.value
    out = np.zeros(nc)
    check(lib().tnr_finalize_multi(scheme._h, out.ctypes.data_as(C.POINTER(C.c_double))))
    return tuple(float(x) for x in out)


class Finalizer:
    """src/schemes/tnrscheme.jl:16-24: only the default norm finalizer runs
    on the device."""

    def __init__(self, f=finalize, E=float):
        self.f = f
        self.E = E


default_Finalizer = Finalizer()


def run(scheme: TNRScheme, trscheme: truncrank, criterion: stopcrit,
        finalizer: Finalizer = default_Finalizer, finalize_beginning=True, verbosity=1):
    """run!(scheme, trscheme, criterion[, finalizer]; finalize_beginning,
    verbosity) (src/schemes/tnrscheme.jl:33-59).  A pure ``maxiter`` run is a
    single device call (tnr_run); other criteria are checked on the host
    after every step, exactly as the reference does."""
    if verbosity >= 1:
        log.info("Starting simulation\n %r\n", scheme)
    if isinstance(criterion, maxiter) and finalizer is default_Finalizer:
        n_steps = max(criterion.n, 1)
        nc = _ncomp(scheme)
        out = np.zeros((n_steps + 1) * nc)
        n = C.c_int()
        check(lib().tnr_run(scheme._h, trscheme.howmany, criterion.n, int(bool(finalize_beginning)),
                            out.ctypes.data_as(C.POINTER(C.c_double)), C.byref(n)))
        if nc == 1:
            data = list(out[:n.value])
        else:
            data = [tuple(float(x) for x in out[i * nc:(i + 1) * nc]) for i in range(n.value)]
        steps = n_steps
    else:
        data = []
        if finalize_beginning:
            data.append(finalizer.f(scheme))
        steps = 0
        crit = True
        while crit:
            if verbosity >= 2:
                log.info("Step %d, data[end]: %s", steps + 1, data[-1] if data else "empty")
            step(scheme, trscheme)
            data.append(finalizer.f(scheme))
            steps += 1
            crit = criterion(steps, data)
    if verbosity >= 1:
        log.info("Simulation finished\n %s\n Iterations: %d", criterion.info(steps, data), steps)
    return data


def solver_config(spec: str = ""):
    """Limits and thresholds of the device's truncated SVD / eigensolver
    ("key=value,..."; "" restores the defaults and TNR_SOLVER_CFG).  A solve
    that reaches ``max_iters`` or a Jacobi sweep cap (``sweeps``,
    ``sweeps_iter``) unconverged makes its step raise TNRError with status
    TNR_ERR_NOCONV, as the reference's LAPACK-backed factorizations raise."""
    check(lib().tnr_solver_config(spec.encode()))


def free_energy(data, beta, scalefactor=2.0, initial_size=1.0):
    """src/utility/free_energy.jl:13-20"""
    d = np.ascontiguousarray(np.asarray(data, dtype=np.float64))
    f = C.c_double()
    check(lib().tnr_free_energy(d.ctypes.data_as(C.POINTER(C.c_double)), d.size, float(beta),
                                float(scalefactor), float(initial_size), C.byref(f)))
    return f.value
