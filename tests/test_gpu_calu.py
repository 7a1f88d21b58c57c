"""MethodLU::CALU (tournament pivoting, src/getrf_tntpiv.cc).  The pivot
sequence depends on the shape of the tournament tree (SLATE's leaves are
nb-row tiles, here the workgroups of the panel kernel), so exact pivots are
parity-unpinned; pinned instead: the interchanges form a valid sequence,
P A = L U to working precision, a bounded growth factor, and gesv / the
mixed driver meeting SLATE's residual bound."""
import numpy as np
import pytest
import torch

import oracle
import slate_amd
from helpers import say, scaled_residual, to_dev, to_host

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


def apply_ipiv(H, ipiv):
    P = H.copy()
    for i, p in enumerate(ipiv - 1):
        if p != i:
            P[[i, p]] = P[[p, i]]
    return P


@pytest.mark.parametrize("n,nb", [(300, 64), (1500, 256), (4500, 512)])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_getrf_calu_factorization(gpu, n, nb, dtype):
    say("getrf CALU n=%d nb=%d %s" % (n, nb, dtype))
    H = oracle.rands(51, n, n)
    A = to_dev(H, dtype)
    info, ipiv = slate_amd.getrf(A, slate_amd.Options(block_size=nb, method_lu=slate_amd.LU_CALU))
    assert info == 0
    assert np.all(ipiv >= np.arange(1, n + 1)) and np.all(ipiv <= n)
    LU = to_host(A)
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    growth = np.abs(U).max() / np.abs(H).max()
    eps = EPS if dtype == torch.float64 else np.finfo(np.float32).eps
    err = np.abs(apply_ipiv(H, ipiv) - L @ U).max() / (n * eps * np.abs(H).max() * max(growth, 1.0))
    say("  growth %.2f, scaled |PA - LU| %.3g, max|L| %.2f" % (growth, err, np.abs(L).max()))
    assert err <= 1.0
    assert growth <= 1e3


def test_gesv_calu_and_mixed(gpu):
    n = 3000
    H = oracle.rands(52, n, n)
    b = oracle.rands(53, n, 1)
    opts = slate_amd.Options(method_lu=slate_amd.LU_CALU)
    A, B = to_dev(H), to_dev(b)
    info, _ = slate_amd.gesv(A, B, opts)
    r = scaled_residual(H, to_host(B), b)
    say("gesv CALU n=%d residual %.3f" % (n, r))
    assert info == 0 and r <= 3.0
    A, B = to_dev(H), to_dev(b)
    X = slate_amd.empty_colmajor(n, 1, torch.float64)
    info, it, _ = slate_amd.gesv_mixed(A, B, X, opts)
    r = scaled_residual(H, to_host(X), b)
    say("gesv_mixed CALU iter %d residual %.3f" % (it, r))
    assert info == 0 and 0 <= it <= 30 and r <= 3.0


def test_calu_singular_column_info(gpu):
    n = 500
    H = oracle.rands(54, n, n)
    H[:, 7] = 0.0
    A = to_dev(H)
    info, _ = slate_amd.getrf(A, slate_amd.Options(block_size=128, method_lu=slate_amd.LU_CALU))
    assert info == 8
