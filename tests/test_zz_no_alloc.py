"""tnr_kernel_jacobi_noalloc (include/tnr_kernels.h) against
tnr_kernel_jacobi: the same eigensolver on caller-owned device buffers with
its working memory in the library's static scratch array.  Two calls in a
row, the second smaller, must give the existing entry point's eigenvalues,
eigenvectors and sweep counts bit for bit; a count beyond the scratch
array's capacity is refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)


def _symmetric(rng, count, n):
    A = rng.standard_normal((count, n, n))
    return np.ascontiguousarray(A + A.transpose(0, 2, 1))


def _host(tk, G):
    count, n, _ = G.shape
    W = np.zeros_like(G)
    lam = np.zeros((count, n))
    sw = np.zeros(count, dtype=np.int32)
    st = tk.lib().tnr_kernel_jacobi(count, n, G.ctypes.data_as(DP), W.ctypes.data_as(DP),
                                    lam.ctypes.data_as(DP), sw.ctypes.data_as(IP), None)
    assert st == 0, tk.lib().tnr_last_error()
    return W, lam, sw


def _hip():
    """The HIP runtime the library itself runs on (device buffers for the
    test: the caller owns them, the entry point allocates nothing)."""
    h = C.CDLL("/opt/rocm/lib/libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


def _noalloc(tk, G):
    h = _hip()
    count, n, _ = G.shape
    # the matrices are stored column-major per problem: a C-order (count, n, n)
    # array of symmetric matrices is the same bytes
    bufs = []
    for nbytes in (G.nbytes, G.nbytes, count * n * 8):
        p = C.c_void_p()
        assert h.hipMalloc(C.byref(p), nbytes) == 0
        bufs.append(p)
    dG, dW, dL = bufs
    try:
        assert h.hipMemcpy(dG, G.ctypes.data, G.nbytes, H2D) == 0
        sw = np.zeros(count, dtype=np.int32)
        st = tk.lib().tnr_kernel_jacobi_noalloc(count, n, dG, dW, dL, sw.ctypes.data_as(IP), None)
        assert st == 0, tk.lib().tnr_last_error()
        W = np.zeros_like(G)
        lam = np.zeros((count, n))
        assert h.hipMemcpy(W.ctypes.data, dW, W.nbytes, D2H) == 0
        assert h.hipMemcpy(lam.ctypes.data, dL, lam.nbytes, D2H) == 0
    finally:
        for p in bufs:
            h.hipFree(p)
    return W, lam, sw


def test_noalloc_jacobi_equals_the_allocating_entry_point(tk):
    rng = np.random.default_rng(11)
    for count, n in ((48, 64), (5, 24)):          # the second call is smaller
        G = _symmetric(rng, count, n)
        W0, L0, S0 = _host(tk, G)
        W1, L1, S1 = _noalloc(tk, G)
        assert np.array_equal(L0, L1) and np.array_equal(W0, W1) and np.array_equal(S0, S1)
        # and they are eigen-decompositions: G W = W diag(lam)
        i = count - 1
        Wt = W1[i].T                                 # column-major per problem
        assert np.allclose(G[i] @ Wt, Wt * L1[i], atol=1e-11 * np.abs(L1[i]).max())


def test_noalloc_jacobi_refuses_counts_beyond_its_scratch(tk):
    st = tk.lib().tnr_kernel_jacobi_noalloc(13798, 2, 1, 1, 1, None, None)
    assert st == -1 and "scratch" in tk.lib().tnr_last_error().decode()
