"""The oracle's block-sparse truncated factorizations against dense LAPACK.

A Z_N-symmetric matrix is block diagonal after a permutation of its rows
and columns, so the union of its blocks' singular values (eigenvalues) is
the dense matrix's spectrum and the global cut of ``truncrank(chi)``
(MatrixAlgebraKit, used at src/schemes/btrg.jl:61,72 and
src/schemes/hotrg.jl:103,110) keeps the chi largest values of that spectrum.
These tests factor random charge-conserving tensors both ways:

* kept singular values / |eigenvalues| = the chi largest of the dense
  spectrum, and the truncation error = the 2-norm of the rest;
* the kept factors span the dense top-chi subspace (equal projectors
  when the cut has a gap) and reproduce the dense best rank-chi
  approximation;
* the factors conserve charge and are isometries.

Each check is chosen so that a plausible slip fails it: a per-block cut
instead of the global one, sorting by value instead of |value| for eigh,
an error taken over the kept values, or rows of a block mapped to the
wrong dense rows.
"""
import numpy as np
import pytest

from oracle.ztensor import (ZT, _global_cut, eigh_trunc, pivoted_gauge,
                            project_hermitian, svd_trunc)


def random_zt(rng, dims, N, ncod, scales=None):
    """Random tensor whose entries off charge conservation are zero.  With
    ``scales`` the block of coupled charge c is multiplied by scales[c], so
    the blocks' spectra have very different magnitudes."""
    q = [rng.integers(0, N, size=d) for d in dims]
    data = rng.standard_normal(dims)
    tot = np.zeros(dims, dtype=np.int64)
    for leg, x in enumerate(q):
        shp = [1] * len(dims)
        shp[leg] = len(x)
        tot = tot + x.reshape(shp)
    data[(tot % N) != 0] = 0.0
    if scales is not None:
        rq = np.zeros(1, dtype=np.int64)
        for leg in range(ncod):
            rq = (rq[:, None] + q[leg][None, :]).reshape(-1)
        rq = rq % N
        M = data.reshape(len(rq), -1)
        M *= np.asarray(scales)[rq][:, None]
        data = M.reshape(dims)
    return ZT(data, q, N)


def dense(t, ncod):
    r = int(np.prod(t.data.shape[:ncod]))
    return t.data.reshape(r, -1)


CASES = [
    # dims, N, ncod, chi
    ((3, 4, 3, 4), 1, 2, 5),
    ((4, 3, 4, 3), 2, 2, 7),
    ((3, 3, 3, 3), 3, 2, 6),
    ((2, 3, 4, 5), 3, 2, 9),
    ((4, 4, 4), 2, 1, 3),
    ((2, 3, 2, 3, 2), 3, 3, 8),
]


@pytest.mark.parametrize("dims,N,ncod,chi", CASES)
def test_block_svd_matches_dense_lapack(dims, N, ncod, chi):
    rng = np.random.default_rng(11 + 7 * N + sum(dims))
    t = random_zt(rng, dims, N, ncod, scales=[1.0, 30.0, 0.05][:N])
    M = dense(t, ncod)
    U, s, Vt, info = svd_trunc(t, ncod, chi)
    sd = np.linalg.svd(M, compute_uv=False)
    k = min(chi, min(M.shape))
    # kept values: the k largest of the dense spectrum (order inside the new
    # leg is by sector, so compare sorted)
    np.testing.assert_allclose(np.sort(s)[::-1], sd[:k], rtol=1e-12, atol=1e-13 * sd[0])
    assert info.err == pytest.approx(np.linalg.norm(sd[k:]), rel=1e-10, abs=1e-13 * sd[0])
    assert info.kept == pytest.approx(sd[k - 1], rel=1e-12)
    assert info.first_dropped == pytest.approx(sd[k] if k < len(sd) else 0.0, rel=1e-12, abs=1e-300)
    assert info.top == pytest.approx(sd[0], rel=1e-12)
    Ud = dense(U, ncod)
    Vd = Vt.data.reshape(len(s), -1)
    # isometries and charge conservation of the factors
    np.testing.assert_allclose(Ud.T @ Ud, np.eye(len(s)), atol=1e-12)
    np.testing.assert_allclose(Vd @ Vd.T, np.eye(len(s)), atol=1e-12)
    assert U.check_symmetric() == 0.0 and Vt.check_symmetric() == 0.0
    # the best rank-k approximation and its residual
    A = (Ud * s) @ Vd
    ud, sdd, vdt = np.linalg.svd(M, full_matrices=False)
    Ad = (ud[:, :k] * sdd[:k]) @ vdt[:k]
    assert sd[k - 1] - (sd[k] if k < len(sd) else 0.0) > 1e-6 * sd[0]
    np.testing.assert_allclose(A, Ad, atol=1e-11 * sd[0])
    assert np.linalg.norm(M - A) == pytest.approx(info.err, rel=1e-9, abs=1e-12 * sd[0])
    # the same column space as the dense top-k singular vectors
    np.testing.assert_allclose(Ud @ Ud.T, ud[:, :k] @ ud[:, :k].T, atol=1e-10)


@pytest.mark.parametrize("dims,N,ncod,chi", [c for c in CASES if c[2] * 2 == len(c[0])
                                              and c[0][:c[2]] == c[0][c[2]:]]
                         + [((3, 4, 3, 4), 3, 2, 10), ((4, 2, 4, 2), 2, 2, 5)])
def test_block_eigh_matches_dense_lapack(dims, N, ncod, chi):
    rng = np.random.default_rng(5 + N + sum(dims))
    half = len(dims) // 2
    # a Hermitian matrix view: legs (a, b | a', b') with a' = a, b' = b and
    # domain charges opposite to the codomain ones
    q_cod = [rng.integers(0, N, size=d) for d in dims[:half]]
    n = int(np.prod(dims[:half]))
    rq = np.zeros(1, dtype=np.int64)
    for x in q_cod:
        rq = (rq[:, None] + x[None, :]).reshape(-1)
    rq = rq % N
    B = rng.standard_normal((n, n))
    B[rq[:, None] != rq[None, :]] = 0.0
    B = B + B.T
    # one block scaled by -40 so that the |lambda| order differs from the
    # signed order
    B[np.ix_(rq == 0, rq == 0)] *= -40.0
    t = ZT(B.reshape(dims), q_cod + [(-x) % N for x in q_cod], N)
    assert t.check_symmetric() == 0.0
    lam, U, eps, info = eigh_trunc(t, ncod, chi)
    wd, vd = np.linalg.eigh(B)
    order = np.argsort(-np.abs(wd), kind="stable")
    k = min(chi, n)
    np.testing.assert_allclose(np.sort(np.abs(lam))[::-1], np.abs(wd[order[:k]]), rtol=1e-12)
    np.testing.assert_allclose(np.sort(lam), np.sort(wd[order[:k]]), rtol=1e-12,
                               atol=1e-13 * np.abs(wd).max())
    assert eps == pytest.approx(np.linalg.norm(wd[order[k:]]), rel=1e-10)
    Ud = dense(U, ncod)
    np.testing.assert_allclose(Ud.T @ Ud, np.eye(len(lam)), atol=1e-12)
    np.testing.assert_allclose(Ud.T @ B @ Ud, np.diag(lam), atol=1e-11 * np.abs(wd).max())
    assert U.check_symmetric() == 0.0
    gap = np.abs(wd[order[k - 1]]) - (np.abs(wd[order[k]]) if k < n else 0.0)
    assert gap > 1e-6 * np.abs(wd).max()
    P = vd[:, order[:k]] @ vd[:, order[:k]].T
    np.testing.assert_allclose(Ud @ Ud.T, P, atol=1e-10)
    # eps is the residual of the truncated decomposition
    assert np.linalg.norm(B - (Ud * lam) @ Ud.T) == pytest.approx(eps, rel=1e-9)


def test_global_cut_is_not_a_per_block_cut():
    # two blocks, the second 100x larger: truncrank(3) keeps all three values
    # of the large block and none of the small one
    kept, flat, keep, drop = _global_cut([np.array([3.0, 2.0, 1.0]),
                                          np.array([300.0, 200.0, 100.0])], 3, False)
    assert kept == [[], [0, 1, 2]]
    # eigh cuts by |lambda|
    kept, *_ = _global_cut([np.array([-5.0, 1.0]), np.array([4.0, -0.5])], 2, True)
    assert kept == [[0], [0]]
    # ties: stable order (earlier block first)
    kept, *_ = _global_cut([np.array([1.0, 0.5]), np.array([1.0, 0.5])], 3, False)
    assert kept == [[0, 1], [0]]


def test_project_hermitian_closed_form():
    rng = np.random.default_rng(3)
    A = rng.standard_normal((6, 6))
    t = project_hermitian(ZT(A.reshape(2, 3, 2, 3)), 2)
    np.testing.assert_array_equal(t.data.reshape(6, 6), (A + A.T) / 2)
    with pytest.raises(ValueError):
        project_hermitian(ZT(rng.standard_normal((2, 3, 3))), 2)


def test_pivoted_gauge_depends_on_the_span_only():
    # a degenerate cluster of three singular values: any rotation of its
    # basis (and any signs) gives the same pivoted factors
    rng = np.random.default_rng(8)
    m, n = 9, 7
    u0, _ = np.linalg.qr(rng.standard_normal((m, 5)))
    v0, _ = np.linalg.qr(rng.standard_normal((n, 5)))
    s = np.array([4.0, 2.0, 2.0, 2.0, 0.5])
    Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    u1, v1 = u0.copy(), v0.copy()
    u1[:, 1:4] = u0[:, 1:4] @ Q
    v1[:, 1:4] = v0[:, 1:4] @ Q
    u1[:, 0] *= -1
    v1[:, 0] *= -1
    a_u, a_vt = pivoted_gauge(u0, s, v0.T)
    b_u, b_vt = pivoted_gauge(u1, s, v1.T)
    np.testing.assert_allclose(a_u, b_u, atol=1e-13)
    np.testing.assert_allclose(a_vt, b_vt, atol=1e-13)
    # still a factorization of the same matrix
    np.testing.assert_allclose((a_u * s) @ a_vt, (u0 * s) @ v0.T, atol=1e-13)
