"""Error paths of the C ABI on the device build.

Every case here is rejected on the host before a kernel would see the bad
input: the entry point returns its status code and names the problem in
tnr_last_error, and the engine stays usable afterwards (the last test runs
the README case in the same process after a run that failed)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

I32P = C.POINTER(C.c_int32)
DP = C.POINTER(C.c_double)


def _ptr(a, t):
    return a.ctypes.data_as(t)


def _create(tk, kind, N, dims, data):
    h = C.c_void_p()
    dims = np.ascontiguousarray(dims, dtype=np.int32).ravel()
    data = np.ascontiguousarray(data, dtype=np.float64)
    st = tk.lib().tnr_create(C.byref(h), kind, N, _ptr(dims, I32P), _ptr(data, DP), data.size,
                             -0.5, 0)
    return st, h


def _last_error(tk):
    return tk.lib().tnr_last_error().decode()


def test_create_rejects_bad_kind_charge_count_and_spaces(tk):
    lib = tk.lib()
    T = tk.classical_ising(symmetry="Trivial")
    dims, data = np.asarray(T.leg_dims), T.data
    st, _ = _create(tk, 99, 1, dims, data)
    assert st == -1 and "kind" in _last_error(tk)
    st, _ = _create(tk, tk.BTRG.kind, 9, np.ones(36, dtype=np.int32), np.ones(1))
    assert st == -1 and "zn" in _last_error(tk)
    # legs 2 and 3 must carry the same space: 2 x 3 x 2 x 2 does not
    st, _ = _create(tk, tk.TRG.kind, 1, [2, 3, 2, 2], np.ones(24))
    assert st == -1 and "legs" in _last_error(tk)
    # the 4-leg constructor does not take the 3D scheme
    st, _ = _create(tk, tk.HOTRG_3D.kind, 1, dims, data)
    assert st == -1
    # data length must match the dims
    st, _ = _create(tk, tk.TRG.kind, 1, dims, np.ones(data.size + 1))
    assert st == -1
    assert lib.tnr_strerror(-1).decode() == "invalid argument"


def test_run_rejects_chi_below_one_and_wrong_state_calls(tk):
    lib = tk.lib()
    T = tk.classical_ising(symmetry="Z2")
    s = tk.TRG(T)
    out = np.zeros(8)
    n = C.c_int()
    assert lib.tnr_run(s._h, 0, 3, 1, _ptr(out, DP), C.byref(n)) == -1
    assert "chi" in _last_error(tk)
    w = np.zeros(4)
    assert lib.tnr_get_bond_weights(s._h, 1, _ptr(w, DP), w.size) == -6   # TRG has none
    b = tk.BTRG(T)
    assert lib.tnr_get_bond_weights(b._h, 3, _ptr(w, DP), w.size) == -1
    s.close()
    b.close()


def test_eigensolver_rejects_blocks_above_its_limit(tk):
    n = 129
    A = np.eye(n)
    wr = np.zeros(n)
    wi = np.zeros(n)
    assert tk.lib().tnr_kernel_eig(1, n, _ptr(A, DP), _ptr(wr, DP), _ptr(wi, DP)) == -1


def test_chi_beyond_the_solver_limit_is_an_error_and_the_engine_recovers(tk):
    """A Trivial 10^4 tensor gives 100 x 100 SVD blocks; above 96 rows the
    block iteration keeps at most 88 values, so chi = 90 is refused (the
    check precedes any launch for that solve).  The same process then runs
    the README case to the published free energy."""
    rng = np.random.default_rng(5)
    D = 10
    dense = rng.standard_normal((D, D, D, D))
    dense = dense + dense.transpose(3, 2, 1, 0)
    T = tk.SymTensor.from_dense(dense, [np.zeros(D, dtype=int)] * 4, 1)
    s = tk.TRG(T)
    with pytest.raises(tk.TNRError) as e:
        tk.run(s, tk.truncrank(90), tk.maxiter(1), verbosity=0)
    assert e.value.status == -1 and "88" in str(e.value)
    s.close()
    # recovery: README.md:87 (BTRG, Z2 Ising at beta_c, chi 16, 25 steps)
    b = tk.BTRG(tk.classical_ising(tk.ising_betac, symmetry="Z2"))
    data = tk.run(b, tk.truncrank(16), tk.maxiter(25), verbosity=0)
    f = tk.free_energy(data, tk.ising_betac)
    assert abs(f / -2.1096504926141826902647832 - 1) < 1e-12
