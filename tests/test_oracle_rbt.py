"""Oracle checks for the random butterfly transforms and the RBT solver
(src/gerbt.cc, src/internal/Tile_gerbt.hh, src/gesv_rbt.cc).

The oracle's loops are checked against dense butterfly matrices assembled
here from the definition of one level (a different formulation: explicit
n x n factors multiplied together), against orthogonality of the unit-weight
transform, and against numpy's solver."""
import numpy as np
import pytest

import oracle as O


def level_matrix(w, h, n, trans):
    """Dense matrix of one butterfly level with weights w and half length h.

    trans: the U^T form, rows (i1, i2) -> (w1 (x1 + x2), w2 (x1 - x2)) / sqrt2;
    otherwise the V form, (w1 x1 + w2 x2, w1 x1 - w2 x2) / sqrt2.  An index
    whose partner is beyond n maps to w1 x1."""
    M = np.zeros((n, n))
    s = 1 / np.sqrt(2)
    for i1 in range(n):
        if i1 % (2 * h) >= h:
            continue
        i2 = i1 + h
        if i2 >= n:
            M[i1, i1] = w[i1]
            continue
        if trans:
            M[i1, i1] = M[i1, i2] = s * w[i1]
            M[i2, i1] = s * w[i2]
            M[i2, i2] = -s * w[i2]
        else:
            M[i1, i1] = M[i2, i1] = s * w[i1]
            M[i1, i2] = s * w[i2]
            M[i2, i2] = -s * w[i2]
    return M


def half(n, nb, d, k):
    nt = -(-n // nb)
    inner = -(-nt // 2 ** d)
    return 2 ** (d - k - 1) * inner * nb


def dense_transforms(u, v, n, nb, d):
    """(M_U, M_V): U^T = L_0 ... L_{d-1}, V = R_{d-1} ... R_0."""
    MU = np.eye(n)
    MV = np.eye(n)
    for k in range(d):
        h = half(n, nb, d, k)
        MU = MU @ level_matrix(u, h, n, True)
        MV = level_matrix(v, h, n, False) @ MV
    return MU, MV


CASES = [(64, 8, 2), (37, 4, 2), (37, 4, 3), (50, 16, 1), (45, 7, 2), (9, 4, 3), (20, 32, 2)]


@pytest.mark.parametrize("n,nb,d", CASES)
def test_half_lengths(n, nb, d):
    for k in range(d):
        assert O.rbt_half(n, nb, d, k) == half(n, nb, d, k)


def test_half_lengths_follow_tile_counts():
    # gerbt.cc:60-77: inner_len = ceil(nt / 2^d) tiles, half_len = 2^(d-k-1) inner_len
    assert O.rbt_half(65536, 256, 2, 0) == 32768
    assert O.rbt_half(65536, 256, 2, 1) == 16384
    assert O.rbt_half(1000, 96, 2, 1) == 3 * 96      # nt = 11 -> inner 3


@pytest.mark.parametrize("n,nb,d", CASES)
def test_gerbt_two_sided_matches_dense(n, nb, d):
    rng = np.random.default_rng(n * 100 + d)
    u = O.rbt_weights(42, n)
    v = O.rbt_weights(43, n)
    A = rng.standard_normal((n, n))
    MU, MV = dense_transforms(u, v, n, nb, d)
    np.testing.assert_allclose(O.gerbt(u, A, v, d, nb), MU @ A @ MV, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("n,nb,d", CASES)
def test_gerbt_rhs_matches_dense(n, nb, d):
    rng = np.random.default_rng(n + d)
    u = O.rbt_weights(7, n)
    B = rng.standard_normal((n, 3))
    MU, MV = dense_transforms(u, u, n, nb, d)
    np.testing.assert_allclose(O.gerbt_rhs(u, B, d, nb, True), MU @ B, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(O.gerbt_rhs(u, B, d, nb, False), MV @ B, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("n,nb,d", CASES)
def test_two_sided_equals_one_sided_on_both_sides(n, nb, d):
    # U^T A V computed by the two-sided pass equals the one-sided U^T applied
    # to the columns of A V, where A V = (V^T A^T)^T and V^T acts like U^T
    # with the V weights.
    rng = np.random.default_rng(3)
    u = O.rbt_weights(1, n)
    v = O.rbt_weights(2, n)
    A = rng.standard_normal((n, n))
    AV = O.gerbt_rhs(v, A.T, d, nb, True).T
    np.testing.assert_allclose(O.gerbt(u, A, v, d, nb), O.gerbt_rhs(u, AV, d, nb, True),
                               rtol=1e-13, atol=1e-13)


def test_unit_weights_orthogonal():
    n, nb, d = 64, 8, 3
    one = np.ones(n)
    MU = O.gerbt_rhs(one, np.eye(n), d, nb, True)
    MV = O.gerbt_rhs(one, np.eye(n), d, nb, False)
    np.testing.assert_allclose(MU @ MU.T, np.eye(n), atol=1e-14)
    np.testing.assert_allclose(MV, MU.T, atol=1e-15)
    # a single level is the 2x2 Haar butterfly on each pair
    M1 = O.gerbt_rhs(one, np.eye(4), 1, 1, True)
    s = 1 / np.sqrt(2)
    np.testing.assert_allclose(M1, s * np.array([[1, 0, 1, 0], [0, 1, 0, 1], [1, 0, -1, 0], [0, 1, 0, -1]]),
                               atol=1e-16)


def test_weights_closed_form():
    u = O.rbt_weights(42, 1000)
    r = O.rands(42, 1000, 1)[:, 0]
    np.testing.assert_allclose(u, np.exp(r / 20), rtol=1e-15)
    assert np.all(u > np.exp(-0.05)) and np.all(u < np.exp(0.05))


def test_depth_zero_is_identity():
    A = np.random.default_rng(0).standard_normal((10, 10))
    u = O.rbt_weights(1, 10)
    np.testing.assert_array_equal(O.gerbt(u, A, u, 0, 4), A)


@pytest.mark.parametrize("n,nb,nrhs", [(200, 32, 1), (157, 24, 3)])
def test_gesv_rbt_solves(n, nb, nrhs):
    A = O.rands(0, n, n)
    B = O.rands(1, n, nrhs)
    A_out, X, info, it = O.gesv_rbt(A, B, nb=nb)
    assert info == 0 and 0 <= it <= 30
    Xr = np.linalg.solve(A, B)
    np.testing.assert_allclose(X, Xr, rtol=1e-10, atol=1e-10 * np.abs(Xr).max())
    # A_out holds the unpivoted factors of U^T A V
    u = O.rbt_weights(42, n)
    v = O.rbt_weights(42 + -(-n // nb), n)
    T = O.gerbt(u, A, v, 2, nb)
    L = np.tril(A_out, -1) + np.eye(n)
    U = np.triu(A_out)
    assert np.abs(L @ U - T).max() <= n * np.finfo(np.float64).eps * (np.abs(L) @ np.abs(U)).max()


def test_gesv_rbt_no_refinement():
    n = 120
    A = O.rands(3, n, n)
    B = O.rands(4, n, 1)
    _, X, info, it = O.gesv_rbt(A, B, nb=16, itermax=0)
    assert info == 0 and it == 0
    np.testing.assert_allclose(X, np.linalg.solve(A, B), rtol=1e-8, atol=1e-8)


def test_gesv_rbt_fallback():
    n = 96
    A = O.rands(5, n, n)
    B = O.rands(6, n, 2)
    A_out, X, info, it = O.gesv_rbt(A, B, nb=16, itermax=3, tol=1e-300)
    assert info == 0 and it == -4
    np.testing.assert_array_equal(A_out, A)            # gesv_rbt.cc:222: A gets the copy back
    _, _, Xg, _ = O.gesv(A, B, nb=16)
    np.testing.assert_array_equal(X, Xg)
    _, X2, _, it2 = O.gesv_rbt(A, B, nb=16, itermax=3, tol=1e-300, use_fallback=False)
    assert it2 == -4
    assert np.abs(X2 - np.linalg.solve(A, B)).max() < 1e-8
