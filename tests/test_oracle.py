"""CPU checks that pin the oracle (oracle/slate_oracle.cc) before it is trusted.

Each test compares the oracle with something independent of it: LAPACK via
scipy (committed fixtures in tests/golden, made by make_golden.py), a
brute-force extended-precision GEPP, closed forms, or invariants.
"""
import os

import numpy as np
import pytest
import scipy.linalg as sl

import oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
EPS = np.finfo(np.float64).eps


def gepp_longdouble(A):
    """Textbook unblocked partial-pivot LU in extended precision (first max wins)."""
    A = np.array(A, dtype=np.longdouble)
    n = A.shape[0]
    ipiv = np.zeros(n, dtype=np.int64)
    for j in range(n):
        p = j + int(np.argmax(np.abs(A[j:, j])))
        ipiv[j] = p + 1
        if p != j:
            A[[j, p], :] = A[[p, j], :]
        if A[j, j] != 0:
            A[j + 1:, j] /= A[j, j]
            A[j + 1:, j + 1:] -= np.outer(A[j + 1:, j], A[j, j + 1:])
    return A, ipiv


# ---------------------------------------------------------------- matgen
def test_rands_distribution_and_range():
    A = oracle.rands(0, 300, 300)
    assert A.min() >= -1.0 and A.max() < 1.0
    assert abs(A.mean()) < 0.01
    assert abs(A.var() - 1.0 / 3.0) < 0.01


def test_randn_moments():
    A = oracle.randn(3, 300, 300)
    assert abs(A.mean()) < 0.01
    assert abs(A.var() - 1.0) < 0.02


def test_matgen_independent_of_tiling():
    """Entry (i, j) depends only on (seed, i, j) (matgen/random.cc:163-173)."""
    big = oracle.rands(5, 64, 48)
    sub = oracle.generate(oracle.DIST_RANDS, 5, 20, 10, ioff=7, joff=30)
    np.testing.assert_array_equal(sub, big[7:27, 30:40])


def test_matgen_seed_changes_values_and_same_seed_pitfall():
    """Same seed for A and B makes B(:,0) = A(:,0) (the tester's seedB rule)."""
    A = oracle.rands(0, 50, 50)
    b_same = oracle.rands(0, 50, 1)
    b_diff = oracle.rands(1, 50, 1)
    np.testing.assert_array_equal(A[:, 0], b_same[:, 0])
    assert not np.array_equal(A[:, 0], b_diff[:, 0])


def test_philox_rounds_change_every_bit_position():
    o = [oracle.philox(0, j, 0) for j in range(64)]
    assert len(set(o)) == 64
    hi_bits = np.array([x[0] >> 63 for x in o])
    assert 10 < hi_bits.sum() < 54


# ---------------------------------------------------------------- getrf
def test_getrf_pivots_match_lapack_n2000():
    A = oracle.rands(0, 2000, 2000)
    ref = np.load(os.path.join(GOLD, "ipiv_lapack_n2000.npy")).astype(np.int64)
    for nthreads in (1, 4):
        LU, ipiv, info = oracle.getrf(A, nb=256, ib=16, nthreads=nthreads)
        assert info == 0
        np.testing.assert_array_equal(ipiv, ref)
    lu, piv, _ = sl.lapack.dgetrf(A)
    assert np.abs(LU - lu).max() < 1e-10


@pytest.mark.parametrize("name", ["perm", "diag_dominant", "hilbert5", "singular_col"])
def test_getrf_tiny_fixtures(name):
    z = np.load(os.path.join(GOLD, "ipiv_lapack_tiny.npz"))
    A = z[name + "_A"]
    LU, ipiv, info = oracle.getrf(A, nb=2, ib=1)
    np.testing.assert_array_equal(ipiv, z[name + "_ipiv"])
    assert info == int(z[name + "_info"])
    np.testing.assert_allclose(LU, z[name + "_lu"], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("n,nb,ib", [(7, 3, 2), (12, 5, 2), (33, 8, 4), (64, 16, 16)])
def test_getrf_matches_bruteforce_gepp(n, nb, ib):
    A = oracle.rands(11 + n, n, n)
    LU, ipiv, info = oracle.getrf(A, nb=nb, ib=ib)
    LUr, ipr = gepp_longdouble(A)
    np.testing.assert_array_equal(ipiv, ipr)
    np.testing.assert_allclose(LU, np.asarray(LUr, dtype=np.float64), rtol=0, atol=1e-12)


def test_getrf_plu_invariant_and_threshold0():
    n = 150
    A = oracle.rands(21, n, n)
    LU, ipiv, info = oracle.getrf(A, nb=32, ib=8)
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    PA = A.copy()
    for i in range(n):
        p = ipiv[i] - 1
        PA[[i, p]] = PA[[p, i]]
    assert np.abs(PA - L @ U).max() / np.abs(A).max() < 50 * n * EPS
    # threshold 0: no pivoting (Tile_getrf.hh:283-289), LU of A itself
    D = A + n * np.eye(n)
    LU0, ip0, _ = oracle.getrf(D, nb=32, ib=8, threshold=0.0)
    np.testing.assert_array_equal(ip0, np.arange(1, n + 1))
    L0 = np.tril(LU0, -1) + np.eye(n)
    assert np.abs(D - L0 @ np.triu(LU0)).max() < 1e-10


def test_getrf_info_on_zero_column():
    n = 40
    A = oracle.rands(4, n, n)
    A[:, 13] = 0.0
    _, _, info = oracle.getrf(A, nb=8, ib=4)
    _, _, linfo = sl.lapack.dgetrf(A)
    assert info == linfo == 14


# ---------------------------------------------------------------- getrs / gesv
@pytest.mark.parametrize("trans", [False, True])
def test_getrs_matches_scipy(trans):
    n = 300
    A = oracle.rands(7, n, n)
    B = oracle.rands(8, n, 5)
    LU, ipiv, info = oracle.getrf(A, nb=64, ib=16)
    X = oracle.getrs(LU, ipiv, B, nb=64, trans=trans)
    Xr = sl.solve(A.T if trans else A, B)
    np.testing.assert_allclose(X, Xr, rtol=1e-9, atol=1e-11)


def test_gesv_scaled_residual_n2000():
    A = oracle.rands(0, 2000, 2000)
    b = oracle.rands(1, 2000, 1)
    _, _, X, info = oracle.gesv(A, b, nb=256)
    assert info == 0
    assert oracle.scaled_residual(A, X, b) <= 3.0


# ---------------------------------------------------------------- Cholesky
def test_potrf_matches_lapack_fixture():
    G = oracle.randn(2, 128, 128)
    S = G @ G.T / 128 + np.eye(128)
    ref = np.load(os.path.join(GOLD, "potrf_lapack_n128.npy"))
    for nb in (16, 50, 128):
        L, info = oracle.potrf(S, nb=nb)
        assert info == 0
        np.testing.assert_allclose(np.tril(L), ref, rtol=0, atol=1e-13)


def test_potrf_info_first_non_spd_minor():
    n = 60
    G = oracle.randn(9, n, n)
    S = G @ G.T / n + np.eye(n)
    S[37, 37] = -1.0
    _, info = oracle.potrf(S, nb=16)
    _, linfo = sl.lapack.dpotrf(S, lower=1)
    assert info == linfo and info > 0


def test_posv_matches_scipy():
    n = 200
    G = oracle.randn(2, n, n)
    S = G @ G.T / n + np.eye(n)
    B = oracle.rands(3, n, 3)
    _, X, info = oracle.posv(S, B, nb=48)
    assert info == 0
    np.testing.assert_allclose(X, sl.solve(S, B, assume_a="pos"), rtol=1e-10, atol=1e-12)


# ---------------------------------------------------------------- mixed precision
def test_gesv_mixed_converges_and_codes():
    n = 400
    A = oracle.rands(0, n, n)
    b = oracle.rands(1, n, 2)
    _, ipiv, X, info, it = oracle.gesv_mixed(A, b, nb=64)
    assert info == 0 and 1 <= it <= 10
    assert oracle.scaled_residual(A, X, b) <= 3.0
    # itermax = 0 forces the FP64 fallback: iter = -(0+1), result = plain gesv
    A2, ipiv2, X2, info2, it2 = oracle.gesv_mixed(A, b, nb=64, itermax=0)
    _, _, Xg, _ = oracle.gesv(A, b, nb=64)
    assert it2 == -1 and info2 == 0
    np.testing.assert_array_equal(X2, Xg)
    # without fallback the negative code stays and A is untouched
    A3, _, _, _, it3 = oracle.gesv_mixed(A, b, nb=64, itermax=0, use_fallback=False)
    assert it3 == -1
    np.testing.assert_array_equal(A3, A)


def test_gesv_mixed_singular_low_precision_factor():
    n = 64
    A = oracle.rands(5, n, n)
    A[:, 10] = 0.0
    b = oracle.rands(6, n, 1)
    _, _, _, info, it = oracle.gesv_mixed(A, b, nb=16, use_fallback=False)
    assert it == -3 and info == 11


def test_gmres_ir_converges():
    n = 300
    A = oracle.rands(0, n, n)
    b = oracle.rands(1, n, 1)
    _, _, x, info, it = oracle.gesv_mixed_gmres(A, b, nb=64)
    assert info == 0 and 0 <= it <= 10
    assert oracle.scaled_residual(A, x, b) <= 3.0


def test_posv_mixed_and_gmres_converge():
    n = 250
    G = oracle.randn(2, n, n)
    S = G @ G.T / n + np.eye(n)
    b = oracle.rands(3, n, 1)
    _, X, info, it = oracle.posv_mixed(S, b, nb=64)
    assert info == 0 and it >= 0 and oracle.scaled_residual(S, X, b) <= 3.0
    _, x, info, it = oracle.posv_mixed_gmres(S, b, nb=64)
    assert info == 0 and it >= 0 and oracle.scaled_residual(S, x, b) <= 3.0


@pytest.mark.parametrize("a,b", [(3.0, 4.0), (-2.0, 1.0), (0.0, 5.0), (7.0, 0.0), (1e-300, 1e-300)])
def test_rotg_closed_form(a, b):
    r, z, c, s = oracle.rotg(a, b)
    assert abs(c * a + s * b - r) <= 1e-14 * max(abs(r), 1e-300)
    assert abs(-s * a + c * b) <= 1e-14 * max(abs(r), 1e-300)
    assert abs(c * c + s * s - 1) < 1e-14


def test_scaled_residual_definition():
    """test/test_gesv.cc:330-377: ||B-AX||_1 / (n ||A||_1 ||X||_1), here in eps units."""
    A = np.diag([2.0, 4.0])
    X = np.array([[1.0], [1.0]])
    B = A @ X + np.array([[0.0], [8.0 * EPS * 2 * 4 * 1]])
    # ||R||_1 = 64 eps, n = 2, ||A||_1 = 4, ||X||_1 = 2  ->  4 eps
    assert abs(oracle.scaled_residual(A, X, B) - 4.0) < 1e-9


# ---- spectrum-controlled generator -----------------------------------------

def test_sigma_closed_forms():
    n, cond = 9, 1e6
    g = oracle.sigma(2, cond, n)
    assert g[0] == 1.0 and abs(g[-1] - 1 / cond) <= 1e-20
    np.testing.assert_allclose(g[1:] / g[:-1], cond ** (-1.0 / (n - 1)), rtol=1e-14)
    a = oracle.sigma(1, cond, n)
    assert a[0] == 1.0 and abs(a[-1] - 1 / cond) < 1e-16
    np.testing.assert_allclose(np.diff(a), -(1 - 1 / cond) / (n - 1), rtol=1e-12)
    c0 = oracle.sigma(3, cond, n)
    assert c0[0] == 1 and np.all(c0[1:] == 1 / cond)
    c1 = oracle.sigma(4, cond, n)
    assert np.all(c1[:-1] == 1) and c1[-1] == 1 / cond


def test_reflector_product_invariants():
    n = 40
    Q = oracle.reflector_product(5, n)
    np.testing.assert_allclose(Q.T @ Q, np.eye(n), atol=1e-13)
    # each reflector has determinant -1: a dropped or doubled one flips the sign
    assert np.sign(np.linalg.det(Q)) == (-1) ** n
    # the last reflector applied to its own vector: H_n v_n = -v_n, so
    # Q v_n = -(H_1 ... H_{n-1}) v_n; check with the product of the first n-1
    V = oracle.randn(5, n, n)
    Qm = np.eye(n)
    for i in range(n - 2, -1, -1):
        v = V[:, i]
        Qm -= np.outer(v, (2.0 / (v @ v)) * (v @ Qm))
    np.testing.assert_allclose(Q @ V[:, -1], -(Qm @ V[:, -1]), atol=1e-12)


@pytest.mark.parametrize("sym", [False, True])
@pytest.mark.parametrize("dist,cond", [(2, 1e6), (1, 1e2), (3, 1e10)])
def test_generate_svd_spectrum(sym, dist, cond):
    n = 60
    A = oracle.generate_svd(dist, cond, 3, n, sym=sym)
    sv = np.linalg.svd(A, compute_uv=False)
    np.testing.assert_allclose(sv, np.sort(oracle.sigma(dist, cond, n))[::-1], rtol=0, atol=1e-14)
    if sym:
        np.testing.assert_allclose(A, A.T, atol=1e-15)
        assert np.all(np.linalg.eigvalsh(A) > 0)
    else:
        assert not np.allclose(A, A.T)


# ---- selector MLP -------------------------------------------------------------

def test_mlp_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    P = rng.standard_normal(oracle.mlp_num_params()) * 0.3
    X = rng.standard_normal((7, oracle.MLP_F))
    y = rng.integers(0, oracle.MLP_K, 7)
    loss, g = oracle.mlp_loss_grad(P, X, y)
    idx = rng.choice(P.size, 60, replace=False)
    h = 1e-6
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        fd = (oracle.mlp_loss_grad(Pp, X, y)[0] - oracle.mlp_loss_grad(Pm, X, y)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])
    p = oracle.mlp_forward(P, X)
    np.testing.assert_allclose(p.sum(axis=1), 1.0, rtol=1e-12)
    assert abs(loss + np.mean(np.log(p[np.arange(7), y]))) < 1e-12


def test_adam_first_step_moves_by_lr():
    # at t = 1 the bias-corrected step is lr * g / (|g| + eps): magnitude lr
    g = np.array([0.5, -2.0, 1e-3])
    P, m, v = oracle.adam_step(np.zeros(3), g, np.zeros(3), np.zeros(3), 0.01, 1)
    np.testing.assert_allclose(P, -0.01 * np.sign(g), rtol=1e-4)
