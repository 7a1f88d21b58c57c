"""Full-size checks through size-independent properties (the oracle is too
slow at these sizes): SLATE's scaled residual ||B - A X||_1 /
(n ||A||_1 ||X||_1 eps) <= 3 computed on the GPU in FP64, valid pivot
sequences, and the mixed drivers leaving A untouched after convergence."""
import numpy as np
import pytest
import torch

import slate_amd
from helpers import dev_rands, say

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


def residual_dev(A, X, B):
    n = A.shape[0]
    R = B - A @ X
    return float(R.abs().sum(dim=0).max() / (n * A.abs().sum(dim=0).max() * X.abs().sum(dim=0).max()) / EPS)


@pytest.mark.parametrize("n,nrhs", [(16384, 1), (8192, 64)])
def test_gesv_full_size(gpu, n, nrhs):
    say("gesv n=%d nrhs=%d" % (n, nrhs))
    A = dev_rands(0, n, n)
    B = dev_rands(1, n, nrhs)
    A0, B0 = A.clone(), B.clone()
    info, ipiv = slate_amd.gesv(A, B)
    r = residual_dev(A0, B, B0)
    say("  info %d scaled residual %.3f (%.1f ms)" % (info, r, slate_amd.last_time_ms()))
    assert info == 0 and r <= 3.0
    assert np.all(ipiv >= np.arange(1, n + 1)) and np.all(ipiv <= n)


@pytest.mark.parametrize("prec", [slate_amd.PREC_FP32, slate_amd.PREC_BF16X3])
def test_gesv_mixed_full_size(gpu, prec):
    n = 16384
    say("gesv_mixed n=%d prec=%d" % (n, prec))
    A = dev_rands(0, n, n)
    B = dev_rands(1, n, 1)
    X = slate_amd.empty_colmajor(n, 1, torch.float64)
    A0 = A.clone()
    info, it, _ = slate_amd.gesv_mixed(A, B, X, slate_amd.Options(factor_precision=prec))
    r = residual_dev(A0, X, B)
    say("  info %d iter %d scaled residual %.3f (%.1f ms)" % (info, it, r, slate_amd.last_time_ms()))
    assert info == 0 and 0 <= it <= 30 and r <= 3.0
    assert torch.equal(A, A0)


def test_posv_full_size(gpu):
    n = 16384
    say("posv n=%d" % n)
    G = dev_rands(5, n, n)
    S = slate_amd.colmajor(G @ G.t() / n + torch.eye(n, dtype=torch.float64, device="cuda"))
    del G
    B = dev_rands(6, n, 1)
    S0, B0 = S.clone(), B.clone()
    info = slate_amd.posv(S, B)
    r = residual_dev(S0, B, B0)
    say("  info %d scaled residual %.3f (%.1f ms)" % (info, r, slate_amd.last_time_ms()))
    assert info == 0 and r <= 3.0


@pytest.mark.parametrize("nb", [768, 1024])
def test_gesv_large_blocks(gpu, nb):
    """The block sizes chosen for the largest matrices, on a smaller one."""
    n = 8192
    say("gesv n=%d nb=%d" % (n, nb))
    A = dev_rands(2, n, n)
    B = dev_rands(3, n, 2)
    A0, B0 = A.clone(), B.clone()
    info, ipiv = slate_amd.gesv(A, B, slate_amd.Options(block_size=nb))
    r = residual_dev(A0, B, B0)
    say("  info %d scaled residual %.3f (%.1f ms)" % (info, r, slate_amd.last_time_ms()))
    assert info == 0 and r <= 3.0


def test_gesv_mixed_benchmark_size(gpu):
    """The benchmark's configuration end to end: n = 65536, float LU with
    three-term split-BF16 updates (block 1024, CU-masked trailing stream),
    FP64 classic IR, checked by the scaled residual."""
    n = 65536
    say("gesv_mixed n=%d bf16x3" % n)
    A = dev_rands(0, n, n)
    B = dev_rands(1, n, 1)
    X = slate_amd.empty_colmajor(n, 1, torch.float64)
    info, it, _ = slate_amd.gesv_mixed(A, B, X, slate_amd.Options(factor_precision=slate_amd.PREC_BF16X3))
    ms = slate_amd.last_time_ms()
    r = residual_dev(A, X, B)
    say("  info %d iter %d scaled residual %.4f (%.1f ms)" % (info, it, r, ms))
    assert info == 0 and 0 <= it <= 30 and r <= 3.0


def test_getrf_block_above_one_pass_swap_limit(gpu):
    """nb = 1536 in FP64: the lookahead block's interchanges exceed the one-pass
    LDS staging and take the two-pass path."""
    n = 6144
    say("getrf fp64 n=%d nb=1536" % n)
    A = dev_rands(4, n, n)
    B = dev_rands(5, n, 1)
    A0, B0 = A.clone(), B.clone()
    info, ipiv = slate_amd.gesv(A, B, slate_amd.Options(block_size=1536))
    r = residual_dev(A0, B, B0)
    say("  info %d scaled residual %.3f" % (info, r))
    assert info == 0 and r <= 3.0
