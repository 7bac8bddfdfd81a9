"""Kernel-level parity on the GPU: grouped FP64 MFMA GEMM, FP32 split GEMM,
grouped Jacobi eigensolver, truncated SVD and eigensolver, each against
numpy/LAPACK FP64 on seeded inputs.  Tolerances are stated per test."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

D = C.POINTER(C.c_double)


def p(a):
    return a.ctypes.data_as(D)


def gemm(tk, A, B, Cm, ta, tb, alpha, beta, split=0):
    M, N = Cm.shape
    K = A.shape[0] if ta else A.shape[1]
    A = np.asfortranarray(A); B = np.asfortranarray(B); Co = np.asfortranarray(Cm.copy())
    st = tk.lib().tnr_kernel_gemm(M, N, K, p(A), max(1, A.shape[0]), ta, p(B), max(1, B.shape[0]), tb, p(Co),
                                  M, alpha, beta, split, 1, None)
    assert st == 0, tk.lib().tnr_last_error()
    return Co


@pytest.mark.parametrize("M,N,K,ta,tb", [
    (1, 1, 1, 0, 0), (64, 64, 16, 0, 0), (100, 37, 53, 0, 0), (100, 37, 53, 1, 0),
    (100, 37, 53, 0, 1), (100, 37, 53, 1, 1), (1024, 64, 1024, 0, 0), (64, 64, 4096, 1, 0),
    (300, 200, 0, 0, 0), (513, 65, 129, 1, 1)])
def test_gemm_f64(tk, M, N, K, ta, tb):
    rng = np.random.default_rng(M * 7 + N * 3 + K)
    A = rng.standard_normal((K, M) if ta else (M, K))
    B = rng.standard_normal((N, K) if tb else (K, N))
    Cm = rng.standard_normal((M, N))
    got = gemm(tk, A, B, Cm, ta, tb, 1.3, 0.7)
    ref = 1.3 * ((A.T if ta else A) @ (B.T if tb else B)) + 0.7 * Cm
    scale = 1.3 * np.sqrt(max(K, 1)) + 1
    assert np.max(np.abs(got - ref)) <= 1e-14 * scale * 8


@pytest.mark.parametrize("M,N,K,ta,tb", [
    (1, 1, 1, 0, 0), (100, 37, 53, 1, 1), (300, 200, 70, 0, 1), (129, 257, 33, 1, 0),
    (1024, 64, 1024, 0, 0), (64, 64, 4096, 1, 0), (1024, 1024, 32, 0, 0)])
def test_gemm_tile_shapes_agree_bitwise(tk, M, N, K, ta, tb):
    """Each output element is summed in the same k order under every tile
    shape (64x64, 128x64, 128x128, 64x64 with 32-deep k-tiles), so the
    kernels agree bit for bit;
    each also matches numpy to the FP64 GEMM bound."""
    rng = np.random.default_rng(M + 11 * N + 101 * K)
    A = rng.standard_normal((K, M) if ta else (M, K))
    B = rng.standard_normal((N, K) if tb else (K, N))
    Cm = rng.standard_normal((M, N))
    outs = [gemm(tk, A, B, Cm, ta, tb, 0.9, -0.4, split=mode) for mode in (2, 3, 4, 5, 0)]
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    ref = 0.9 * ((A.T if ta else A) @ (B.T if tb else B)) - 0.4 * Cm
    assert np.max(np.abs(outs[0] - ref)) <= 1e-14 * (np.sqrt(K) + 1) * 8


@pytest.mark.parametrize("M,N,K,ta,tb", [
    (64, 64, 1024, 1, 0), (1024, 64, 1024, 0, 0), (100, 37, 777, 1, 1), (300, 130, 2049, 0, 1)])
def test_gemm_split_k_does_not_change_bits(tk, M, N, K, ta, tb):
    """K is summed in 256-deep chunks added in order whether one workgroup
    runs all chunks or each chunk has its own workgroup (then k_gemm_reduce
    adds them): split and unsplit launches agree bit for bit, for every tile
    shape."""
    rng = np.random.default_rng(M + 7 * N + 13 * K)
    A = rng.standard_normal((K, M) if ta else (M, K))
    B = rng.standard_normal((N, K) if tb else (K, N))
    Cm = rng.standard_normal((M, N))
    outs = [gemm(tk, A, B, Cm, ta, tb, 1.1, 0.3, split=shape | flag)
            for shape in (2, 3, 4, 5) for flag in (8, 16)]
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    ref = 1.1 * ((A.T if ta else A) @ (B.T if tb else B)) + 0.3 * Cm
    assert np.max(np.abs(outs[0] - ref)) <= 1e-14 * (np.sqrt(K) + 1) * 8


def gemm_padded(tk, A, B, Cm, ta, tb, pad, mode):
    """tnr_kernel_gemm on operands stored with leading dimensions `pad`
    larger than their rows; the padding holds NaN, which must not reach
    the product."""
    M, N = Cm.shape
    K = A.shape[0] if ta else A.shape[1]
    As = np.full((A.shape[0] + pad, A.shape[1]), np.nan, order="F")
    As[:A.shape[0]] = A
    Bs = np.full((B.shape[0] + pad, B.shape[1]), np.nan, order="F")
    Bs[:B.shape[0]] = B
    Co = np.asfortranarray(Cm.copy())
    st = tk.lib().tnr_kernel_gemm(M, N, K, p(As), As.shape[0], ta, p(Bs), Bs.shape[0], tb, p(Co),
                                  M, 1.0, 0.5, mode, 1, None)
    assert st == 0, tk.lib().tnr_last_error()
    return Co


@pytest.mark.parametrize("M,N,K,ta,tb,pad", [
    (1024, 64, 1024, 0, 0, 0), (1024, 64, 1024, 1, 0, 0), (1024, 64, 1024, 0, 1, 0),
    (1024, 64, 1024, 1, 1, 0), (300, 130, 2049, 0, 1, 0), (129, 257, 33, 1, 0, 3),
    (200, 100, 77, 0, 0, 5), (200, 100, 77, 1, 1, 5), (100, 37, 53, 1, 0, 1),
    (257, 129, 4, 0, 0, 2), (128, 64, 16, 0, 0, 0), (5, 3, 1, 1, 1, 1)])
def test_gemm_lds_dma_matches_register_loads_bitwise(tk, M, N, K, ta, tb, pad):
    """The 16-deep tile shapes load operands through LDS-DMA (three LDS
    stages) by default and through registers with mode | 32: the MFMA
    sequence is the same, so the bits are.  Covered: all four storage
    orders, K, M and N not multiples of the tile, split and unsplit K, and
    leading dimensions above the operand's rows with NaN in the padding
    (DMA reads past a row or past K must contribute exact zeros)."""
    rng = np.random.default_rng(M + 3 * N + 7 * K + ta + 2 * tb)
    A = rng.standard_normal((K, M) if ta else (M, K))
    B = rng.standard_normal((N, K) if tb else (K, N))
    Cm = rng.standard_normal((M, N))
    ref = (A.T if ta else A) @ (B.T if tb else B) + 0.5 * Cm
    for shape in (2, 3, 4):
        for flag in (8, 16):
            if flag == 16 and K < 512:
                continue
            dma = gemm_padded(tk, A, B, Cm, ta, tb, pad, shape | flag)
            reg = gemm_padded(tk, A, B, Cm, ta, tb, pad, shape | flag | 32)
            assert np.array_equal(dma, reg), (shape, flag)
            assert np.all(np.isfinite(dma))
            assert np.max(np.abs(dma - ref)) <= 1e-14 * (np.sqrt(K) + 1) * 8


@pytest.mark.parametrize("M,N,K", [(128, 128, 256), (200, 100, 1000)])
def test_gemm_split32(tk, M, N, K):
    rng = np.random.default_rng(K)
    A = rng.standard_normal((M, K)); B = rng.standard_normal((K, N)); Cm = np.zeros((M, N))
    got = gemm(tk, A, B, Cm, 0, 0, 1.0, 0.0, split=1)
    ref = A @ B
    # hi/lo split products accumulated on FP32 MFMA within 16-deep k-tiles:
    # error ~ 2^-24 * sqrt(K) * |a||b| per entry; stated bound 1e-6 relative
    # to the entry scale sqrt(K)
    assert np.max(np.abs(got - ref)) / np.sqrt(K) < 1e-6


def jacobi(tk, G):
    cnt, n = G.shape[0], G.shape[1]
    Gf = np.ascontiguousarray(np.stack([np.asfortranarray(g).ravel(order="F") for g in G]))
    W = np.zeros((cnt, n * n)); lam = np.zeros((cnt, n)); sw = np.zeros(cnt, dtype=np.int32)
    st = tk.lib().tnr_kernel_jacobi(cnt, n, p(Gf), p(W), p(lam),
                                    sw.ctypes.data_as(C.POINTER(C.c_int)), None)
    assert st == 0, tk.lib().tnr_last_error()
    return lam, np.stack([w.reshape(n, n, order="F") for w in W]), int(sw.max())


@pytest.mark.parametrize("n", [1, 2, 5, 64, 95, 96])
def test_jacobi_random_symmetric(tk, n):
    rng = np.random.default_rng(n)
    G = rng.standard_normal((3, n, n)); G = G + G.transpose(0, 2, 1)
    lam, W, sw = jacobi(tk, G)
    for g, l, w in zip(G, lam, W):
        ref = np.linalg.eigvalsh(g)
        ref = ref[np.argsort(-np.abs(ref), kind="stable")]
        s = np.max(np.abs(ref))
        assert np.max(np.abs(l - ref)) < 1e-13 * s * n
        assert np.max(np.abs(g @ w - w * l)) < 1e-13 * s * n
        assert np.max(np.abs(w.T @ w - np.eye(n))) < 1e-13 * n


def test_jacobi_graded_gram_relative_accuracy(tk):
    rng = np.random.default_rng(9)
    A = rng.standard_normal((200, 64)) * np.logspace(0, -12, 64)
    G = (A.T @ A)[None]
    lam, W, sw = jacobi(tk, G)
    ref = np.linalg.svd(A, compute_uv=False) ** 2
    # relative accuracy of every eigenvalue, down to 1e-24 of the largest
    assert np.max(np.abs(lam[0] / ref - 1)) < 1e-12


def svd(tk, M, k):
    m, n = M.shape
    Mf = np.asfortranarray(M)
    U = np.zeros((m, k), order="F"); s = np.zeros(k); V = np.zeros((n, k), order="F")
    it = C.c_int()
    st = tk.lib().tnr_kernel_svd(m, n, p(Mf), k, p(U), p(s), p(V), C.byref(it))
    assert st == 0, tk.lib().tnr_last_error()
    return U, s, V, it.value


def decaying(rng, m, n, rate):
    U, _ = np.linalg.qr(rng.standard_normal((m, min(m, n))))
    V, _ = np.linalg.qr(rng.standard_normal((n, min(m, n))))
    s = np.exp(-rate * np.arange(min(m, n)))
    return (U * s) @ V.T


# k above 88 on blocks above 96 takes the wide-block solver (block
# one-sided Jacobi as the Rayleigh-Ritz step)
@pytest.mark.parametrize("m,n,k,rate", [
    (4, 4, 4, 0.5), (16, 16, 16, 0.3), (50, 200, 20, 0.1), (200, 50, 20, 0.1),
    (256, 256, 16, 0.05), (1024, 1024, 32, 0.02), (400, 300, 40, 0.05), (300, 500, 48, 0.05),
    (300, 400, 120, 0.02), (1024, 1024, 160, 0.01)])
def test_svd_truncated(tk, m, n, k, rate):
    rng = np.random.default_rng(m + n + k)
    M = decaying(rng, m, n, rate)
    U, s, V, it = svd(tk, M, k)
    Ur, sr, Vtr = np.linalg.svd(M, full_matrices=False)
    assert np.max(np.abs(s - sr[:k])) < 1e-13 * sr[0]
    # truncated reconstruction equals LAPACK's (the cut has a gap here)
    Mk = (U * s) @ V.T
    Mr = (Ur[:, :k] * sr[:k]) @ Vtr[:k]
    assert np.max(np.abs(Mk - Mr)) < 1e-12 * sr[0]
    assert np.max(np.abs(U.T @ U - np.eye(k))) < 1e-12
    assert np.max(np.abs(V.T @ V - np.eye(k))) < 1e-12


# k = 40 iterates an 80-column block on a rank-70 matrix: the block's
# Cholesky QR (65-96 columns) meets null directions and drops them; the kept
# triplets still match LAPACK's
def test_svd_block_wider_than_rank(tk):
    rng = np.random.default_rng(70)
    m, n, r, k = 400, 300, 70, 40
    Ua, _ = np.linalg.qr(rng.standard_normal((m, r)))
    Va, _ = np.linalg.qr(rng.standard_normal((n, r)))
    sa = np.exp(-0.05 * np.arange(r))
    M = (Ua * sa) @ Va.T
    U, s, V, it = svd(tk, M, k)
    assert it > 0   # the block iteration, not a direct solve
    Ur, sr, Vtr = np.linalg.svd(M, full_matrices=False)
    assert np.max(np.abs(s - sr[:k])) < 1e-13 * sr[0]
    Mk = (U * s) @ V.T
    Mr = (Ur[:, :k] * sr[:k]) @ Vtr[:k]
    assert np.max(np.abs(Mk - Mr)) < 1e-12 * sr[0]
    assert np.max(np.abs(U.T @ U - np.eye(k))) < 1e-12
    assert np.max(np.abs(V.T @ V - np.eye(k))) < 1e-12


@pytest.mark.parametrize("n,k", [(64, 16), (512, 32), (400, 120)])
def test_eigh_truncated(tk, n, k):
    rng = np.random.default_rng(n)
    X = decaying(rng, n, n, 0.04)
    M = X @ X.T
    Mf = np.asfortranarray(M)
    U = np.zeros((n, k), order="F"); lam = np.zeros(k); it = C.c_int()
    st = tk.lib().tnr_kernel_eigh(n, p(Mf), k, p(U), p(lam), C.byref(it))
    assert st == 0, tk.lib().tnr_last_error()
    w, v = np.linalg.eigh(M)
    o = np.argsort(-np.abs(w)); w, v = w[o], v[:, o]
    assert np.max(np.abs(lam - w[:k])) < 1e-13 * w[0]
    Pk = U @ U.T
    Pr = v[:, :k] @ v[:, :k].T
    assert np.max(np.abs(Pk - Pr)) < 1e-9


# A spectrum that decays too slowly for the block iteration (the top-k
# subspace contracts by sigma_l / sigma_k ~ 0.97 per iteration, so 80
# iterations leave it far from converged): such blocks, up to 384 on their
# smaller side, are factorised whole instead, and match LAPACK as usual
# (the reference's LAPACK-backed svd_trunc / eigh_trunc!, btrg.jl:61 and
# hotrg.jl:103, have no iteration to stall).
def _flat(rng, m, n, lo):
    U, _ = np.linalg.qr(rng.standard_normal((m, min(m, n))))
    V, _ = np.linalg.qr(rng.standard_normal((n, min(m, n))))
    s = np.linspace(1.0, lo, min(m, n))
    return (U * s) @ V.T, s


@pytest.mark.parametrize("m,n,k", [(200, 150, 10), (300, 384, 16)])
def test_svd_stalled_iteration_is_factorised_whole(tk, m, n, k):
    rng = np.random.default_rng(m * n + k)
    M, _ = _flat(rng, m, n, 0.3)
    U, s, V, it = svd(tk, M, k)
    Ur, sr, Vtr = np.linalg.svd(M, full_matrices=False)
    assert np.max(np.abs(s - sr[:k])) < 1e-13 * sr[0]
    Mk = (U * s) @ V.T
    Mr = (Ur[:, :k] * sr[:k]) @ Vtr[:k]
    assert np.max(np.abs(Mk - Mr)) < 1e-11 * sr[0]
    assert np.max(np.abs(U.T @ U - np.eye(k))) < 1e-12
    assert np.max(np.abs(V.T @ V - np.eye(k))) < 1e-12


@pytest.mark.parametrize("n,k", [(144, 12), (300, 20)])
def test_eigh_stalled_iteration_is_factorised_whole(tk, n, k):
    rng = np.random.default_rng(n + k)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    w0 = np.linspace(1.0, 0.2, n) * np.where(np.arange(n) % 3 == 2, -1.0, 1.0)
    M = (Q * w0) @ Q.T
    Mf = np.asfortranarray(M)
    U = np.zeros((n, k), order="F"); lam = np.zeros(k); it = C.c_int()
    st = tk.lib().tnr_kernel_eigh(n, p(Mf), k, p(U), p(lam), C.byref(it))
    assert st == 0, tk.lib().tnr_last_error()
    w, v = np.linalg.eigh(M)
    o = np.argsort(-np.abs(w), kind="stable"); w, v = w[o], v[:, o]
    assert np.max(np.abs(lam - w[:k])) < 1e-13
    Pk = U @ U.T
    Pr = v[:, :k] @ v[:, :k].T
    assert np.max(np.abs(Pk - Pr)) < 1e-9


def _z2_factor(rng, m, k, s):
    """Left singular vectors of a matrix whose rows split into even (first
    half) and odd coordinates, every value of s shared by one even and one
    odd vector, handed over scrambled inside each degenerate pair."""
    h = m // 2
    E, _ = np.linalg.qr(rng.standard_normal((h, k)))
    O, _ = np.linalg.qr(rng.standard_normal((m - h, k)))
    U = np.zeros((m, 2 * k))
    U[:h, 0::2] = E
    U[h:, 1::2] = O
    for j in range(k):
        q, _ = np.linalg.qr(rng.standard_normal((2, 2)))
        U[:, 2 * j:2 * j + 2] = U[:, 2 * j:2 * j + 2] @ q
    return U, np.repeat(s, 2)


@pytest.mark.parametrize("m,k,null", [(40, 6, 0), (300, 24, 4), (97, 40, 8)])
def test_pivoted_gauge_kernel_vs_oracle(tk, m, k, null):
    """k_canon against ztensor.pivoted_gauge: scrambled degenerate pairs
    come back as the same basis (to 1e-12), each vector purely even or
    purely odd, positive at its pivot; distinct values only change sign;
    null columns are left alone."""
    from oracle.ztensor import pivoted_gauge
    rng = np.random.default_rng(m + k)
    s = np.sort(rng.uniform(0.1, 1.0, k // 2))[::-1]
    U, sv = _z2_factor(rng, m, k // 2, s)
    if null:                      # whole degenerate pairs
        U[:, -null:] = 0.0
        sv[-null:] = 0.0
    Q = np.zeros((k, k), order="F")
    Uf = np.asfortranarray(U)
    st = tk.lib().tnr_kernel_pivoted_gauge(m, k, p(Uf), p(np.ascontiguousarray(sv)), p(Q))
    assert st == 0, tk.lib().tnr_last_error()
    got = U @ Q
    ref, _ = pivoted_gauge(U, sv, np.eye(k))
    assert np.max(np.abs(got - ref)) < 1e-12
    h = m // 2
    live = sv > 1e-14 * sv[0]
    for j in np.flatnonzero(live):
        even, odd = np.linalg.norm(got[:h, j]), np.linalg.norm(got[h:, j])
        assert min(even, odd) < 1e-13
    assert np.max(np.abs(got.T[live][:, :] @ got[:, live] - np.eye(int(live.sum())))) < 1e-12
    # distinct values: a sign per column, taken from the pivot row
    s2 = np.linspace(1.0, 0.2, k)
    V, _ = np.linalg.qr(rng.standard_normal((m, k)))
    Q2 = np.zeros((k, k), order="F")
    Vf = np.asfortranarray(V)
    assert tk.lib().tnr_kernel_pivoted_gauge(m, k, p(Vf), p(s2), p(Q2)) == 0
    assert np.array_equal(np.abs(Q2), np.eye(k))
    ref2, _ = pivoted_gauge(V, s2, np.eye(k))
    assert np.array_equal(V @ Q2, ref2)
