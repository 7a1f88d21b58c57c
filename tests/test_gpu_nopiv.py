"""MethodLU::NoPiv (src/getrf_nopiv.cc): LU without row interchanges.  On a
diagonally dominant matrix the factors equal the oracle's no-pivot
restatement (Tile_getrf with threshold 0) and the unique L U = A
factorization; the pivot vector is the identity; gesv and the mixed driver
solve to SLATE's residual bound through it."""
import numpy as np
import pytest
import torch

import oracle
import slate_amd
from helpers import say, scaled_residual, to_dev, to_host

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,nb", [(200, 64), (1000, 256), (1537, 256)])
def test_getrf_nopiv_matches_oracle(gpu, n, nb):
    say("dgetrf nopiv n=%d nb=%d" % (n, nb))
    H = oracle.rands(21, n, n) + n * np.eye(n)
    A = to_dev(H)
    info, ipiv = slate_amd.getrf(A, slate_amd.Options(block_size=nb, method_lu=slate_amd.LU_NOPIV))
    assert info == 0
    np.testing.assert_array_equal(ipiv, np.arange(1, n + 1))
    LU = to_host(A)
    LU_o, ip_o, info_o = oracle.getrf(H, nb=nb, ib=16, threshold=0.0)
    assert info_o == 0
    tol = 4 * n * np.finfo(np.float64).eps * np.abs(LU_o).max()
    assert np.abs(LU - LU_o).max() <= tol
    L = np.tril(LU, -1) + np.eye(n)
    assert np.abs(H - L @ np.triu(LU)).max() <= tol * 4


def test_gesv_nopiv_and_mixed(gpu):
    n = 1200
    H = oracle.rands(23, n, n) + n * np.eye(n)
    b = oracle.rands(24, n, 1)
    opts = slate_amd.Options(block_size=256, method_lu=slate_amd.LU_NOPIV)
    A, B = to_dev(H), to_dev(b)
    info, ipiv = slate_amd.gesv(A, B, opts)
    assert info == 0
    assert scaled_residual(H, to_host(B), b) <= 3.0
    A, B = to_dev(H), to_dev(b)
    X = slate_amd.empty_colmajor(n, 1, torch.float64)
    info, it, _ = slate_amd.gesv_mixed(A, B, X, opts)
    assert info == 0 and 0 <= it <= 30
    assert scaled_residual(H, to_host(X), b) <= 3.0


def test_nopiv_zero_pivot_info(gpu):
    n = 300
    H = oracle.rands(25, n, n) + n * np.eye(n)
    H[0, 0] = 0.0          # no interchange allowed: U(1,1) = 0
    A = to_dev(H)
    info, _ = slate_amd.getrf(A, slate_amd.Options(block_size=128, method_lu=slate_amd.LU_NOPIV))
    assert info == 1
