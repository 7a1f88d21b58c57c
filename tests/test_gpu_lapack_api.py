"""LAPACK-compatible entry points (include/slate_amd_lapack.h) on host arrays:
slate_dgesv_ pivots equal LAPACK dgetrf's on the n = 2000 matgen matrix
(committed fixture), solutions meet SLATE's residual bound, slate_dpotrf_
matches LAPACK dpotrf, and the mixed entries return LAPACK's iter codes."""
import ctypes
import os

import numpy as np
import pytest
import scipy.linalg as sl

import oracle
import slate_amd
from helpers import say, scaled_residual, spd

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))

I = ctypes.c_int


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _r(v):
    return ctypes.byref(v)


def test_dgesv_lapack_pivots_and_residual(gpu):
    n = 2000
    say("slate_dgesv_ n=%d" % n)
    A0 = np.asfortranarray(oracle.rands(0, n, n))
    B0 = np.asfortranarray(oracle.rands(1, n, 1))
    A, B = A0.copy(order="F"), B0.copy(order="F")
    ipiv = np.zeros(n, dtype=np.int32)
    info = I(0)
    slate_amd.lib().slate_dgesv_(_r(I(n)), _r(I(1)), _p(A), _r(I(n)), _p(ipiv), _p(B), _r(I(n)), _r(info))
    assert info.value == 0
    ref = np.load(os.path.join(HERE, "golden", "ipiv_lapack_n2000.npy"))
    np.testing.assert_array_equal(ipiv, ref)
    assert scaled_residual(A0, B, B0) <= 3.0


def test_dgetrf_dgetrs_trans(gpu):
    n = 700
    A0 = np.asfortranarray(oracle.rands(4, n, n))
    B0 = np.asfortranarray(oracle.rands(5, n, 3))
    A, B = A0.copy(order="F"), B0.copy(order="F")
    ipiv = np.zeros(n, dtype=np.int32)
    info = I(0)
    L = slate_amd.lib()
    L.slate_dgetrf_(_r(I(n)), _r(I(n)), _p(A), _r(I(n)), _p(ipiv), _r(info))
    assert info.value == 0
    lu, piv, _ = sl.lapack.dgetrf(A0)
    np.testing.assert_array_equal(ipiv, piv + 1)
    trans = ctypes.c_char(b"T")
    L.slate_dgetrs_(_r(trans), _r(I(n)), _r(I(3)), _p(A), _r(I(n)), _p(ipiv), _p(B), _r(I(n)), _r(info))
    assert info.value == 0
    assert scaled_residual(np.ascontiguousarray(A0.T), B, B0) <= 3.0


def test_dpotrf_dposv_dsposv(gpu):
    n = 600
    S0 = np.asfortranarray(spd(2, n))
    L = slate_amd.lib()
    S = S0.copy(order="F")
    info = I(0)
    uplo = ctypes.c_char(b"L")
    L.slate_dpotrf_(_r(uplo), _r(I(n)), _p(S), _r(I(n)), _r(info))
    assert info.value == 0
    c, _ = sl.lapack.dpotrf(S0, lower=1)
    np.testing.assert_allclose(np.tril(S), np.tril(c), rtol=0, atol=1e-12)
    B0 = np.asfortranarray(oracle.rands(3, n, 2))
    S, B = S0.copy(order="F"), B0.copy(order="F")
    L.slate_dposv_(_r(uplo), _r(I(n)), _r(I(2)), _p(S), _r(I(n)), _p(B), _r(I(n)), _r(info))
    assert info.value == 0
    assert scaled_residual(S0, B, B0) <= 3.0
    S, B = S0.copy(order="F"), B0.copy(order="F")
    X = np.zeros((n, 2), order="F")
    it = I(-99)
    L.slate_dsposv_(_r(uplo), _r(I(n)), _r(I(2)), _p(S), _r(I(n)), _p(B), _r(I(n)), _p(X), _r(I(n)),
                    None, None, _r(it), _r(info))
    assert info.value == 0 and it.value >= 0
    assert scaled_residual(S0, X, B0) <= 3.0


def test_dsgesv_iter_and_not_positive_definite(gpu):
    n = 800
    A0 = np.asfortranarray(oracle.rands(6, n, n))
    B0 = np.asfortranarray(oracle.rands(7, n, 1))
    A, B = A0.copy(order="F"), B0.copy(order="F")
    X = np.zeros((n, 1), order="F")
    ipiv = np.zeros(n, dtype=np.int32)
    info, it = I(0), I(-99)
    L = slate_amd.lib()
    L.slate_dsgesv_(_r(I(n)), _r(I(1)), _p(A), _r(I(n)), _p(ipiv), _p(B), _r(I(n)), _p(X), _r(I(n)),
                    None, None, _r(it), _r(info))
    assert info.value == 0 and it.value >= 0
    np.testing.assert_array_equal(A, A0)                 # A untouched after convergence
    assert scaled_residual(A0, X, B0) <= 3.0
    # an indefinite matrix: potrf reports the first non-positive minor
    M = np.asfortranarray(np.eye(5))
    M[3, 3] = -1.0
    uplo = ctypes.c_char(b"L")
    L.slate_dpotrf_(_r(uplo), _r(I(5)), _p(M), _r(I(5)), _r(info))
    assert info.value == 4
