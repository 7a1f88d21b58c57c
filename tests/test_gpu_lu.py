"""GPU parity of the LU path (getrf/getrs/gesv, gemm, matgen) against the oracle.

Every call goes through the C ABI (include/slate_amd.h) via slate_amd.  Integer
and index results (pivots, info, matgen bits) must be identical; floating-point
factors agree with the oracle within the stated tolerance and the solutions
meet the tester's scaled-residual bound (<= 3, stricter than test_gesv.cc's 25).
"""
import os

import numpy as np
import pytest
import torch

import oracle
import slate_amd
from helpers import EPS, dev_rands, say, scaled_residual, to_dev, to_host

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_generate_bitwise_equals_oracle(gpu):
    say("generate rands/uniform 300x200 seed 3")
    for dist in (slate_amd.DIST_UNIFORM, slate_amd.DIST_RANDS):
        A = slate_amd.empty_colmajor(300, 200, torch.float64)
        slate_amd.generate(A, dist, 3, ioffset=5, joffset=7)
        ref = oracle.generate(dist, 3, 300, 200, ioff=5, joff=7)
        np.testing.assert_array_equal(to_host(A), ref)
    A = slate_amd.empty_colmajor(300, 200, torch.float64)
    slate_amd.generate(A, slate_amd.DIST_RANDN, 3)
    np.testing.assert_allclose(to_host(A), oracle.randn(3, 300, 200), rtol=1e-13, atol=1e-14)


# (32, 4097, 32) ... (48, 2000, 40): C has at most 64 rows and k <= 64, the
# narrow VALU kernel of the panel's lowest splits (16-byte vectors, odd
# leading dimensions on the element path, k tails after the 8-vector batches)
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (5, 7, 3), (128, 128, 16), (300, 257, 129), (1000, 513, 700),
                                   (32, 4097, 32), (64, 3001, 64), (33, 777, 17), (48, 2000, 40)])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_gemm_matches_reference(gpu, m, n, k, dtype):
    say("gemm %dx%dx%d %s" % (m, n, k, dtype))
    rng = np.random.default_rng(m * 7 + n + k)
    A, B, C = (rng.standard_normal(s) for s in ((m, k), (k, n), (m, n)))
    dA, dB, dC = to_dev(A, dtype), to_dev(B, dtype), to_dev(C, dtype)
    slate_amd.gemm(-1.0, dA, dB, 0.5, dC)
    ref = 0.5 * C - A @ B
    tol = 1e-13 if dtype == torch.float64 else 2e-5
    assert np.abs(to_host(dC) - ref).max() <= tol * max(1.0, k) ** 0.5 * 4


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_gemm_beta_zero_does_not_read_c(gpu, dtype):
    # beta = 0: C is written, never read (NaN in C must not reach the result),
    # on the narrow kernel (m = 32, k = 32) and the MFMA tile (m = 200)
    for m, n, k in ((32, 1500, 32), (200, 300, 48)):
        say("gemm beta=0 %dx%dx%d %s" % (m, n, k, dtype))
        rng = np.random.default_rng(m + n + k)
        A, B = rng.standard_normal((m, k)), rng.standard_normal((k, n))
        dA, dB = to_dev(A, dtype), to_dev(B, dtype)
        dC = to_dev(np.full((m, n), np.nan), dtype)
        slate_amd.gemm(2.0, dA, dB, 0.0, dC)
        tol = 1e-13 if dtype == torch.float64 else 2e-5
        assert np.abs(to_host(dC) - 2.0 * (A @ B)).max() <= tol * k ** 0.5 * 8


def test_getrf_n2000_pivots_equal_lapack_and_oracle(gpu):
    say("dgetrf n=2000 nb=256 rands seed 0")
    n = 2000
    A = dev_rands(0, n, n)
    info, ipiv = slate_amd.getrf(A, slate_amd.Options(block_size=256))
    assert info == 0
    ref = np.load(os.path.join(GOLD, "ipiv_lapack_n2000.npy")).astype(np.int64)
    np.testing.assert_array_equal(ipiv, ref)
    LU_o, ipiv_o, _ = oracle.getrf(oracle.rands(0, n, n), nb=256, ib=16)
    np.testing.assert_array_equal(ipiv, ipiv_o)
    # factors agree within the LU rounding bound, 4 n eps max|LU|
    assert np.abs(to_host(A) - LU_o).max() <= 4 * n * EPS * np.abs(LU_o).max()


@pytest.mark.parametrize("n", [1, 2, 17, 100, 257, 513, 1031])
@pytest.mark.parametrize("nb", [32, 128])
def test_getrf_sizes_pivots_equal_oracle(gpu, n, nb):
    say("dgetrf n=%d nb=%d" % (n, nb))
    A = dev_rands(100 + n, n, n)
    info, ipiv = slate_amd.getrf(A, slate_amd.Options(block_size=nb))
    LU_o, ipiv_o, info_o = oracle.getrf(oracle.rands(100 + n, n, n), nb=nb, ib=16)
    assert info == info_o == 0
    np.testing.assert_array_equal(ipiv, ipiv_o)
    assert np.abs(to_host(A) - LU_o).max() <= 4 * max(n, 16) * EPS * np.abs(LU_o).max()


def test_getrf_lookahead_is_bitwise_neutral(gpu):
    say("dgetrf n=1500 nb=128 lookahead 0 vs 1")
    n = 1500
    A0 = dev_rands(9, n, n)
    A1 = dev_rands(9, n, n)
    i0, p0 = slate_amd.getrf(A0, slate_amd.Options(block_size=128, lookahead=0))
    i1, p1 = slate_amd.getrf(A1, slate_amd.Options(block_size=128, lookahead=1))
    np.testing.assert_array_equal(p0, p1)
    assert torch.equal(A0, A1)


def test_getrf_zero_column_info(gpu):
    say("dgetrf singular n=300 zero column 77")
    n = 300
    H = oracle.rands(4, n, n)
    H[:, 77] = 0.0
    A = to_dev(H)
    info, ipiv = slate_amd.getrf(A, slate_amd.Options(block_size=64))
    _, ipiv_o, info_o = oracle.getrf(H, nb=64, ib=16)
    assert info == info_o == 78
    np.testing.assert_array_equal(ipiv, ipiv_o)


def test_getrf_threshold_zero_means_no_pivoting(gpu):
    say("dgetrf threshold 0 n=200")
    n = 200
    H = oracle.rands(21, n, n) + n * np.eye(n)
    A = to_dev(H)
    info, ipiv = slate_amd.getrf(A, slate_amd.Options(block_size=64, pivot_threshold=0.0))
    np.testing.assert_array_equal(ipiv, np.arange(1, n + 1))
    LU = to_host(A)
    L = np.tril(LU, -1) + np.eye(n)
    assert np.abs(H - L @ np.triu(LU)).max() < 1e-10


def _perm_from_ipiv(ipiv, n):
    """perm with (P^T b)[i] = b[perm[i]] for LAPACK's 1-based interchange sequence."""
    perm = np.arange(n)
    for i, p in enumerate(np.asarray(ipiv, dtype=np.int64) - 1):
        perm[i], perm[p] = perm[p], perm[i]
    return perm


@pytest.mark.parametrize("n", [129, 300, 1031])
@pytest.mark.parametrize("trans", [False, True])
def test_sgetrs_single_rhs_backward_error(gpu, n, trans):
    """FP32 getrs with one right-hand side runs the look-back triangular
    solves: a ragged last block, a narrow first tile in the upper sweep, the
    column-major forms when transposed.  Its normwise backward error
    ||P^T b - L U x|| / (||L|| ||U|| ||x||) against the float factors it used
    stays below 8 eps (measured ~1e-9 at these sizes); one wrong factor entry
    of size 1 would leave about 1 / (||L|| ||U||), orders of magnitude more."""
    say("sgetrs n=%d nrhs=1 trans=%s" % (n, trans))
    A = dev_rands(11, n, n, torch.float32)
    info, ipiv = slate_amd.getrf(A)
    assert info == 0
    B = dev_rands(12, n, 1, torch.float32)
    b = to_host(B)[:, 0]
    slate_amd.getrs(A, ipiv, B, trans=trans)
    x = to_host(B)[:, 0]
    LU = to_host(A)
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    perm = _perm_from_ipiv(ipiv, n)
    if trans:
        # A^T x = U^T L^T P^T x
        r = b - U.T @ (L.T @ x[perm])
        scale = np.abs(U.T).sum(1).max() * np.abs(L.T).sum(1).max() * np.abs(x).max()
    else:
        # P^T A = L U
        r = b[perm] - L @ (U @ x)
        scale = np.abs(L).sum(1).max() * np.abs(U).sum(1).max() * np.abs(x).max()
    berr = np.abs(r).max() / scale
    say("  backward error %.3e (bound 8 eps = %.3e)" % (berr, 8 * np.finfo(np.float32).eps))
    assert berr <= 8 * np.finfo(np.float32).eps


@pytest.mark.parametrize("nrhs", [1, 7])
@pytest.mark.parametrize("trans", [False, True])
def test_getrs_matches_oracle(gpu, nrhs, trans):
    say("dgetrs n=700 nrhs=%d trans=%s" % (nrhs, trans))
    n = 700
    A = dev_rands(7, n, n)
    info, ipiv = slate_amd.getrf(A, slate_amd.Options(block_size=128))
    B = dev_rands(8, n, nrhs)
    slate_amd.getrs(A, ipiv, B, trans=trans)
    H = oracle.rands(7, n, n)
    Bh = oracle.rands(8, n, nrhs)
    LU_o, ip_o, _ = oracle.getrf(H, nb=128, ib=16)
    X_o = oracle.getrs(LU_o, ip_o, Bh, nb=128, trans=trans)
    np.testing.assert_allclose(to_host(B), X_o, rtol=1e-8, atol=1e-10)
    opA = H.T if trans else H
    assert scaled_residual(opA, to_host(B), Bh) <= 3.0


def test_gesv_n2000_scaled_residual(gpu):
    say("dgesv n=2000 nb=256 nrhs=1 (config 1 shape)")
    n = 2000
    A = dev_rands(0, n, n)
    B = dev_rands(1, n, 1)
    info, ipiv = slate_amd.gesv(A, B, slate_amd.Options(block_size=256))
    assert info == 0
    H = oracle.rands(0, n, n)
    b = oracle.rands(1, n, 1)
    r = scaled_residual(H, to_host(B), b)
    say("  scaled residual %.3e" % r)
    assert r <= 3.0
    _, _, X_o, _ = oracle.gesv(H, b, nb=256)
    np.testing.assert_allclose(to_host(B), X_o, rtol=1e-8, atol=1e-10)


def test_sgetrf_pivots_equal_oracle(gpu):
    say("sgetrf n=400 nb=64")
    n = 400
    A = dev_rands(31, n, n, dtype=torch.float32)
    info, ipiv = slate_amd.getrf(A, slate_amd.Options(block_size=64))
    H = oracle.rands(31, n, n).astype(np.float32)
    LU_o, ip_o, info_o = oracle.getrf(H, nb=64, ib=16, dtype=np.float32)
    assert info == info_o == 0
    np.testing.assert_array_equal(ipiv, ip_o)
    eps32 = float(np.finfo(np.float32).eps)
    assert np.abs(to_host(A) - LU_o).max() <= 4 * n * eps32 * np.abs(LU_o).max()


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_getrf_odd_leading_dimension(gpu, dtype):
    # lda = n + 1: every kernel on the path (panel leaf, its interchanges of the
    # other panel columns, the one-launch triangular solves, the GEMMs) takes
    # its unaligned, scalar-access branch
    # float at n = 400 (as the aligned float test): larger float cases can
    # meet near-ties whose pivot depends on the summation order
    n, nb = (700, 128) if dtype == torch.float64 else (400, 64)
    say("getrf n=%d lda=%d nb=%d %s" % (n, n + 1, nb, dtype))
    H = oracle.rands(41, n, n)
    A = slate_amd.empty_colmajor(n, n, dtype, ld=n + 1)
    A.copy_(to_dev(H, dtype))
    info, ipiv = slate_amd.getrf(A, slate_amd.Options(block_size=nb))
    np_dtype = np.float64 if dtype == torch.float64 else np.float32
    LU_o, ip_o, info_o = oracle.getrf(H.astype(np_dtype), nb=nb, ib=16, dtype=np_dtype)
    assert info == info_o == 0
    np.testing.assert_array_equal(ipiv, ip_o)
    eps = float(np.finfo(np_dtype).eps)
    assert np.abs(to_host(A) - LU_o).max() <= 4 * n * eps * np.abs(LU_o).max()
