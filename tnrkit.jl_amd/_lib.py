"""ctypes binding of libtnr_hip.so (declarations mirror include/tnr.h)."""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtnr_hip.so")

TNR_OK = 0
TNR_ERR_ARG = -1
TNR_ERR_HIP = -2
TNR_ERR_NOMEM = -3
TNR_ERR_NONFINITE = -4
TNR_ERR_NOCONV = -5
TNR_ERR_STATE = -6
TNR_ERR_NODEVICE = -7
TNR_ERR_INTERNAL = -8
TNR_ERR_BUDGET = -9
TNR_ERR_BUSY = -10

TNR_TRG, TNR_BTRG, TNR_HOTRG = 0, 1, 2
TNR_IMPURITY_TRG, TNR_IMPURITY_HOTRG = 3, 4
TNR_ATRG = 5
TNR_HOTRG_3D = 6
TNR_PREC_FP64, TNR_PREC_SPLIT32 = 0, 1


class RunDesc(C.Structure):
    _fields_ = [
        ("kind", C.c_int),
        ("zn", C.c_int),
        ("leg_dims", C.POINTER(C.c_int32)),
        ("data", C.POINTER(C.c_double)),
        ("ndata", C.c_size_t),
        ("k", C.c_double),
        ("chi", C.c_int),
        ("maxiter", C.c_int),
        ("finalize_beginning", C.c_int),
        ("impurity", C.POINTER(C.POINTER(C.c_double))),
    ]


_P = C.c_void_p
_PP = C.POINTER(C.c_void_p)
_I32P = C.POINTER(C.c_int32)
_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int)
_SZP = C.POINTER(C.c_size_t)
_I64P = C.POINTER(C.c_int64)

# name -> (restype, argtypes)
SIGNATURES = {
    "tnr_create": (C.c_int, [_PP, C.c_int, C.c_int, _I32P, _DP, C.c_size_t, C.c_double, C.c_int]),
    "tnr_step": (C.c_int, [_P, C.c_int]),
    "tnr_finalize": (C.c_int, [_P, _DP]),
    "tnr_run": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, _DP, _IP]),
    "tnr_free_energy": (C.c_int, [_DP, C.c_int, C.c_double, C.c_double, C.c_double, _DP]),
    "tnr_get_dims": (C.c_int, [_P, _I32P, _SZP]),
    "tnr_get_tensor": (C.c_int, [_P, _DP, C.c_size_t]),
    "tnr_create_multi": (C.c_int, [_PP, C.c_int, C.c_int, _I32P, C.POINTER(_DP), _SZP, C.c_int,
                                   C.c_int]),
    "tnr_ncomponents": (C.c_int, [C.c_int]),
    "tnr_finalize_multi": (C.c_int, [_P, _DP]),
    "tnr_get_tensor_at": (C.c_int, [_P, C.c_int, _DP, C.c_size_t]),
    "tnr_set_tensor": (C.c_int, [_P, _I32P, _DP, C.c_size_t]),
    "tnr_get_bond_weights": (C.c_int, [_P, C.c_int, _DP, C.c_size_t]),
    "tnr_get_trunc_log": (C.c_int, [_P, _DP, C.c_int, _IP]),
    "tnr_destroy": (C.c_int, [_P]),
    "tnr_strerror": (C.c_char_p, [C.c_int]),
    "tnr_last_error": (C.c_char_p, []),
    "tnr_run_batch": (C.c_int, [C.c_int, C.POINTER(RunDesc), C.c_int, _DP]),
    "tnr_server_create": (C.c_int, [_PP, C.c_size_t, C.c_int, C.c_int]),
    "tnr_server_submit": (C.c_int, [_P, C.POINTER(RunDesc), _I64P]),
    "tnr_server_pump": (C.c_int, [_P, _IP]),
    "tnr_server_poll": (C.c_int, [_P, C.c_int64, _DP, _IP, _DP, _DP]),
    "tnr_server_stats": (C.c_int, [_P, _SZP, _IP, _IP]),
    "tnr_server_destroy": (C.c_int, [_P]),
    # include/tnr_kernels.h
    "tnr_kernel_gemm": (C.c_int, [C.c_int, C.c_int, C.c_int, _DP, C.c_int, C.c_int, _DP, C.c_int,
                                  C.c_int, _DP, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                                  _DP]),
    "tnr_kernel_svd": (C.c_int, [C.c_int, C.c_int, _DP, C.c_int, _DP, _DP, _DP, _IP]),
    "tnr_kernel_pivoted_gauge": (C.c_int, [C.c_int, C.c_int, _DP, _DP, _DP]),
    "tnr_kernel_eig": (C.c_int, [C.c_int, C.c_int, _DP, _DP, _DP]),
    "tnr_create_3d": (C.c_int, [_PP, C.c_int, C.POINTER(C.c_int32), _DP, C.c_size_t, C.c_int]),
    "tnr_get_nlegs": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "tnr_transfer_spectrum": (C.c_int, [C.c_void_p, C.c_int, _DP, _DP, C.c_size_t,
                                        C.POINTER(C.c_size_t)]),
    "tnr_cft_data": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_int, _DP, _DP, C.c_size_t,
                               C.POINTER(C.c_size_t)]),
    "tnr_ground_state_degeneracy": (C.c_int, [C.c_void_p, C.c_int, _DP]),
    "tnr_gu_wen_ratio": (C.c_int, [C.c_void_p, _DP, _DP]),
    "tnr_central_charge": (C.c_int, [C.c_void_p, C.c_double, _DP]),
    "tnr_finalize_two_by_two": (C.c_int, [C.c_void_p, _DP]),
    "tnr_kernel_eigh": (C.c_int, [C.c_int, _DP, C.c_int, _DP, _DP, _IP]),
    "tnr_kernel_jacobi": (C.c_int, [C.c_int, C.c_int, _DP, _DP, _DP, _IP, _DP]),
    "tnr_profile_enable": (C.c_int, [C.c_int]),
    "tnr_autotune_enable": (C.c_int, [C.c_int]),
    "tnr_solver_config": (C.c_int, [C.c_char_p]),
    "tnr_kernel_jacobi_noalloc": (C.c_int, [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                            _IP, C.c_void_p]),
    "tnr_device_sync": (C.c_int, []),
    "tnr_profile_read": (C.c_int, [C.c_int, _I64P, _DP, _DP, _DP, _DP, _DP]),
}

_lib = None


class TNRError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{msg} (status {status})")
        self.status = status


def lib():
    """Load the HIP engine; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != TNR_OK:
        L = lib()
        raise TNRError(status, L.tnr_last_error().decode() or L.tnr_strerror(status).decode())
    return status
