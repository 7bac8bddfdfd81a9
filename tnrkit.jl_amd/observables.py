"""Observables on a scheme's tensor, the names of src/utility/cft.jl and
src/utility/finalize.jl: cft_data, central_charge, ground_state_degeneracy,
gu_wen_ratio, finalize_two_by_two and the finalizers built on them.  All
arithmetic runs in libtnr_hip.so (contractions, traces, SVD and the
transfer-matrix eigenvalues on the device)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib
from .schemes import Finalizer, TNRScheme, finalize

_DP = C.POINTER(C.c_double)


def _spectrum_call(fn, scheme, unitcell, *args):
    """Run an entry point that fills (re, im, cap, count), sized for the
    whole transfer matrix: (dimension of leg 2)^unitcell values."""
    N = scheme._T0.N
    dims = np.zeros(4 * N, dtype=np.int32)
    nd = C.c_size_t()
    check(lib().tnr_get_dims(scheme._h, dims.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(nd)))
    cap = int(dims.reshape(4, N)[1].sum()) ** max(int(unitcell), 1)
    re = np.zeros(max(cap, 1))
    im = np.zeros(max(cap, 1))
    n = C.c_size_t()
    check(fn(scheme._h, int(unitcell), *args, re.ctypes.data_as(_DP), im.ctypes.data_as(_DP),
             cap, C.byref(n)))
    return re[:n.value], im[:n.value]


def transfer_spectrum(scheme: TNRScheme, unitcell: int = 1):
    """All eigenvalues of the unitcell-site transfer matrix (eig_full at
    src/utility/cft.jl:16), by decreasing modulus (ties: real part, then
    imaginary part, decreasing)."""
    re, im = _spectrum_call(lib().tnr_transfer_spectrum, scheme, unitcell)
    return re + 1j * im


def cft_data(scheme: TNRScheme, v=1, unitcell: int = 1, is_real: bool = True):
    """cft_data(scheme; v, unitcell, is_real) (src/utility/cft.jl:5-71)."""
    re, im = _spectrum_call(lib().tnr_cft_data, scheme, unitcell, C.c_double(float(v)),
                            int(bool(is_real)))
    return re if is_real else re + 1j * im


def ground_state_degeneracy(scheme: TNRScheme, unitcell: int = 1) -> float:
    """ground_state_degeneracy(scheme; unitcell) (src/utility/cft.jl:278-339)."""
    g = C.c_double()
    check(lib().tnr_ground_state_degeneracy(scheme._h, int(unitcell), C.byref(g)))
    return g.value


def gu_wen_ratio(scheme: TNRScheme):
    """gu_wen_ratio(scheme) -> (X1, X2) (src/utility/cft.jl:374-396)."""
    x1, x2 = C.c_double(), C.c_double()
    check(lib().tnr_gu_wen_ratio(scheme._h, C.byref(x1), C.byref(x2)))
    return x1.value, x2.value


def central_charge(scheme: TNRScheme, n: float) -> float:
    """central_charge(scheme, n) (src/utility/cft.jl:256-268)."""
    c = C.c_double()
    check(lib().tnr_central_charge(scheme._h, float(n), C.byref(c)))
    return c.value


def finalize_two_by_two(scheme: TNRScheme) -> float:
    """finalize_two_by_two!(scheme) (src/utility/finalize.jl:17-41)."""
    n = C.c_double()
    check(lib().tnr_finalize_two_by_two(scheme._h, C.byref(n)))
    return n.value


def finalize_central_charge(scheme: TNRScheme) -> float:
    """finalize_central_charge!(scheme) (src/utility/finalize.jl:145-148)."""
    return central_charge(scheme, finalize(scheme))


def finalize_groundstatedegeneracy(scheme: TNRScheme) -> float:
    """finalize_groundstatedegeneracy!(scheme) (src/utility/finalize.jl:155-163)."""
    finalize(scheme)
    return ground_state_degeneracy(scheme, unitcell=1)


def finalize_gu_wen_ratio(scheme: TNRScheme):
    """finalize_gu_wen_ratio!(scheme) (src/utility/finalize.jl:173-186)."""
    finalize(scheme)
    return gu_wen_ratio(scheme)


two_by_two_Finalizer = Finalizer(finalize_two_by_two, float)
GSDegeneracy_Finalizer = Finalizer(finalize_groundstatedegeneracy, float)
guwenratio_Finalizer = Finalizer(finalize_gu_wen_ratio, tuple)
