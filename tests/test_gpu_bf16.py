"""Split-BF16 trailing updates of the float factor (factor_precision BF16,
BF16X2, BF16X3) on v_mfma_f32_16x16x32_bf16.

The kernel is checked against an FP64 product of the same float inputs with
the error bound each split implies (unit roundoff of one bf16 term 2^-9; a
t-term split leaves a remainder below 2^-9t relative, and the dropped cross
terms are of the same order), so a transposed operand, a dropped term or a
wrong plane stride fails it.  The factorization and the refinement drivers
are then checked end to end: PA = LU to the accuracy of the split, and the
mixed solvers reaching SLATE's FP64 residual bound.
"""
import numpy as np
import pytest
import torch

import slate_amd
from helpers import dev_rands, say, scaled_residual, to_host

pytestmark = pytest.mark.gpu

# max |C - C_ref| / max_ij sum_l |a_il||b_lj| allowed per number of terms
SPLIT_TOL = {1: 1.2e-2, 2: 8e-5, 3: 4e-6}


def _rand(shape, seed):
    g = torch.Generator().manual_seed(seed)
    return (torch.rand(shape, generator=g, dtype=torch.float64) * 2 - 1)


@pytest.mark.parametrize("terms", [1, 2, 3])
@pytest.mark.parametrize("shape", [(777, 650, 515), (256, 384, 64), (33, 17, 7)])
def test_split_gemm_error_bound(gpu, terms, shape):
    m, n, k = shape
    say("sgemm_bf16 terms=%d m=%d n=%d k=%d" % (terms, m, n, k))
    A64 = _rand((m, k), 1).float().double()
    B64 = _rand((k, n), 2).float().double()
    C64 = _rand((m, n), 3).float().double()
    A = slate_amd.colmajor(A64.float().cuda())
    B = slate_amd.colmajor(B64.float().cuda())
    C = slate_amd.colmajor(C64.float().cuda())
    alpha = -1.0
    slate_amd.gemm_bf16(terms, alpha, A, B, C)
    ref = C64 + alpha * (A64 @ B64)
    scale = (A64.abs() @ B64.abs()).max().item()
    err = (C.double().cpu() - ref).abs().max().item() / scale
    say("  relative error %.3g (bound %.1g)" % (err, SPLIT_TOL[terms]))
    assert err <= SPLIT_TOL[terms]
    if terms == 3:
        # float-accurate: within a small factor of the FP32-MFMA sgemm's error
        C2 = slate_amd.colmajor(C64.float().cuda())
        slate_amd.gemm(alpha, A, B, 1.0, C2)
        err32 = (C2.double().cpu() - ref).abs().max().item() / scale
        say("  fp32 sgemm error %.3g" % err32)
        assert err <= 4 * err32 + 1e-7


def test_split_terms_monotone(gpu):
    m = n = 512
    k = 512
    A64 = _rand((m, k), 5).float().double()
    B64 = _rand((k, n), 6).float().double()
    A = slate_amd.colmajor(A64.float().cuda())
    B = slate_amd.colmajor(B64.float().cuda())
    ref = A64 @ B64
    errs = []
    for t in (1, 2, 3):
        C = slate_amd.colmajor(torch.zeros(m, n, dtype=torch.float32, device="cuda"))
        slate_amd.gemm_bf16(t, 1.0, A, B, C)
        errs.append((C.double().cpu() - ref).abs().max().item())
    say("split errors %s" % errs)
    assert errs[0] > 50 * errs[1]
    assert errs[1] > 3 * errs[2]


def _lu_backward_error(A0, LU, ipiv):
    n = A0.shape[0]
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    PA = A0.copy()
    for i in range(n):
        p = ipiv[i] - 1
        if p != i:
            PA[[i, p]] = PA[[p, i]]
    return np.abs(PA - L @ U).max() / np.abs(A0).max()


def test_sgetrf_split_updates(gpu):
    n = 1200
    A0 = to_host(dev_rands(7, n, n, torch.float32))
    be = {}
    for prec in (slate_amd.PREC_FP32, slate_amd.PREC_BF16X3, slate_amd.PREC_BF16X2, slate_amd.PREC_BF16):
        say("sgetrf n=%d factor_precision=%d" % (n, prec))
        A = dev_rands(7, n, n, torch.float32)
        info, ipiv = slate_amd.getrf(A, slate_amd.Options(block_size=256, factor_precision=prec))
        assert info == 0
        assert np.all(ipiv >= np.arange(1, n + 1)) and np.all(ipiv <= n)
        be[prec] = _lu_backward_error(A0, to_host(A), ipiv)
        say("  max|PA - LU| / max|A| = %.3g" % be[prec])
    f32 = be[slate_amd.PREC_FP32]
    assert f32 <= 1e-3
    assert be[slate_amd.PREC_BF16X3] <= 3 * f32          # float-accurate update
    assert f32 < be[slate_amd.PREC_BF16X2] <= 2000 * f32  # ~16-bit products
    assert be[slate_amd.PREC_BF16X2] < be[slate_amd.PREC_BF16] <= 4.0   # one bf16 term: O(1) in max norm


@pytest.mark.parametrize("prec", [slate_amd.PREC_BF16X3, slate_amd.PREC_BF16X2, slate_amd.PREC_BF16])
@pytest.mark.parametrize("gmres", [False, True])
def test_gesv_mixed_split_reaches_fp64(gpu, prec, gmres):
    n = 2000
    say("dsgesv%s n=%d factor_precision=%d" % ("_gmres" if gmres else "", n, prec))
    A = dev_rands(0, n, n)
    B = dev_rands(1, n, 1)
    X = slate_amd.empty_colmajor(n, 1, torch.float64)
    A_in = to_host(A)
    f = slate_amd.gesv_mixed_gmres if gmres else slate_amd.gesv_mixed
    info, it, _ = f(A, B, X, slate_amd.Options(block_size=256, factor_precision=prec))
    say("  info %d iter %d" % (info, it))
    assert info == 0
    assert it >= 0 or it == -31                  # converged, or the FP64 fallback ran
    if prec == slate_amd.PREC_BF16X3:
        assert it >= 0                           # float-accurate factor: IR converges
    if prec == slate_amd.PREC_BF16X2 and gmres:
        assert 0 <= it <= 30                     # the benchmark's form converges without the fallback
    assert scaled_residual(A_in, to_host(X), to_host(B)) <= 3.0


def test_gesv_mixed_bf16_well_conditioned_converges(gpu):
    # diagonally dominant: kappa ~ 1, so even the one-term bf16 factor converges
    n = 1500
    A = dev_rands(9, n, n)
    A.diagonal().add_(float(n))
    B = dev_rands(10, n, 1)
    X = slate_amd.empty_colmajor(n, 1, torch.float64)
    A_in = to_host(A)
    info, it, _ = slate_amd.gesv_mixed(A, B, X, slate_amd.Options(block_size=256,
                                                                  factor_precision=slate_amd.PREC_BF16))
    say("dsgesv bf16 diag-dominant: info %d iter %d" % (info, it))
    assert info == 0 and 0 <= it <= 10
    assert scaled_residual(A_in, to_host(X), to_host(B)) <= 3.0
