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
    n = 513
    A = np.eye(n)
    wr = np.zeros(n)
    wi = np.zeros(n)
    assert tk.lib().tnr_kernel_eig(1, n, _ptr(A, DP), _ptr(wr, DP), _ptr(wi, DP)) == -1


def test_chi_above_the_rayleigh_ritz_width_runs_the_wide_solver(tk):
    """A Trivial 10^4 tensor gives 100 x 100 SVD blocks; keeping chi = 90 of
    them needs a Rayleigh-Ritz step wider than 96, which the wide-block
    solver (block one-sided Jacobi) takes: TRG and HOTRG first steps match
    the oracle.  The same process then runs the README case to the published
    free energy."""
    from oracle import schemes as osch
    from oracle.ztensor import ZT
    rng = np.random.default_rng(5)
    D = 10
    dense = rng.standard_normal((D, D, D, D))
    dense = dense + dense.transpose(3, 2, 1, 0)
    T = tk.SymTensor.from_dense(dense, [np.zeros(D, dtype=int)] * 4, 1)
    for name in ("TRG", "HOTRG"):
        s = getattr(tk, name)(T)
        d = np.array(tk.run(s, tk.truncrank(90), tk.maxiter(1), verbosity=0))
        s.close()
        ref = np.array(osch.run(getattr(osch, name)(ZT(dense)), 90, 1))
        assert np.max(np.abs(d / ref - 1)) < 1e-10, (name, d, ref)
    # recovery: README.md:87 (BTRG, Z2 Ising at beta_c, chi 16, 25 steps)
    b = tk.BTRG(tk.classical_ising(tk.ising_betac, symmetry="Z2"))
    data = tk.run(b, tk.truncrank(16), tk.maxiter(25), verbosity=0)
    f = tk.free_energy(data, tk.ising_betac)
    assert abs(f / -2.1096504926141826902647832 - 1) < 1e-12


# --------------------------------------------------------------------------
# Failures the reference raises from its LAPACK-backed factorizations
# (src/schemes/btrg.jl:61,72; src/schemes/hotrg.jl:103,110), forced from the
# host through the solver's own caps: every kernel still exits normally.

@pytest.fixture
def solver_caps(tk):
    yield tk.solver_config
    tk.solver_config("")


def test_block_iteration_cap_is_reported(tk, solver_caps):
    """max_iters=2: the chi=32 SVDs of 1024 x 1024 matrices (block
    iteration) stop unconverged once the spectrum has full rank (from step 5
    on; before, the rank is below 32 and two iterations converge); the step
    fails with TNR_ERR_NOCONV, the handle keeps failing until its tensor is
    replaced, and with the defaults restored the same handle runs to the
    oracle's norms."""
    from oracle import models as om, schemes as osch
    T = tk.classical_ising(0.44, symmetry="Trivial")
    s = tk.BTRG(T)
    solver_caps("max_iters=2")
    with pytest.raises(tk.TNRError) as e:
        tk.run(s, tk.truncrank(32), tk.maxiter(8), verbosity=0)
    assert e.value.status == -5 and "converge" in str(e.value)
    solver_caps("")
    with pytest.raises(tk.TNRError) as e:          # sticky until set_tensor
        tk.run(s, tk.truncrank(32), tk.maxiter(4), verbosity=0)
    assert e.value.status == -5
    dims = np.ascontiguousarray(T.leg_dims, dtype=np.int32).ravel()
    assert tk.lib().tnr_set_tensor(s._h, _ptr(dims, I32P), _ptr(T.data, DP), T.data.size) == 0
    d = np.array(tk.run(s, tk.truncrank(32), tk.maxiter(4), verbosity=0))
    ref = np.array(osch.run(osch.BTRG(om.classical_ising_trivial(0.44)), 32, 4))
    assert np.max(np.abs(d / ref - 1)) < 1e-10
    s.close()


def test_jacobi_sweep_cap_is_reported(tk, solver_caps):
    """sweeps=1: the final Jacobi solves of the direct (small-block) SVDs
    stop after one sweep with rotations still due -> TNR_ERR_NOCONV."""
    solver_caps("sweeps=1")
    s = tk.TRG(tk.classical_ising(symmetry="Z2"))
    with pytest.raises(tk.TNRError) as e:
        tk.run(s, tk.truncrank(8), tk.maxiter(3), verbosity=0)
    assert e.value.status == -5
    s.close()


def test_failed_run_in_a_batch_leaves_the_others_untouched(tk, solver_caps):
    """In one server a chi=32 Trivial run (block iteration) fails under
    max_iters=2 while a chi=8 Z2 run (direct solves only) in the same grouped
    launches finishes with the same norms as alone; tnr_run_batch returns the
    failed run's status and names it."""
    small = tk.Request("BTRG", tk.classical_ising(0.41, symmetry="Z2"), 8, 6)
    big = tk.Request("BTRG", tk.classical_ising(0.41, symmetry="Trivial"), 32, 6)
    alone = tk.run_batch([small])[0]
    solver_caps("max_iters=2")
    srv = tk.Server(max_slots=4)
    i_small, i_big = srv.submit(small), srv.submit(big)
    while srv.pump() > 0:
        pass
    res = srv.poll(i_small)
    assert res is not None and np.array_equal(res[0], alone)
    with pytest.raises(tk.TNRError) as e:
        srv.poll(i_big)
    assert e.value.status == -5
    srv.close()
    with pytest.raises(tk.TNRError) as e:
        tk.run_batch([small, big])
    assert e.value.status == -5 and "run 1" in str(e.value)


def test_nonfinite_tensors_are_refused(tk):
    """NaN / Inf entries: tnr_create and tnr_server_submit return
    TNR_ERR_NONFINITE; a server keeps serving its other requests."""
    T = tk.classical_ising(symmetry="Trivial")
    bad = T.data.copy()
    bad[3] = np.nan
    st, _ = _create(tk, tk.BTRG.kind, 1, np.asarray(T.leg_dims), bad)
    assert st == -4 and "finite" in _last_error(tk)
    bad[3] = np.inf
    st, _ = _create(tk, tk.TRG.kind, 1, np.asarray(T.leg_dims), bad)
    assert st == -4
    srv = tk.Server(max_slots=2)
    good = tk.Request("BTRG", T, 8, 4)
    Tb = tk.trivial(bad.reshape(2, 2, 2, 2, order="F"))
    with pytest.raises(tk.TNRError) as e:
        srv.submit(tk.Request("BTRG", Tb, 8, 4))
    assert e.value.status == -4
    i = srv.submit(good)
    while srv.pump() > 0:
        pass
    assert np.array_equal(srv.poll(i)[0], tk.run_batch([good])[0])
    srv.close()
