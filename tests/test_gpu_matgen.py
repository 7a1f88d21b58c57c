"""Spectrum-controlled generator on the GPU against the oracle, and the
refinement drivers on the kappa-controlled inputs of the mixed-precision
configurations (A = U diag(sigma) V^T, sigma geometric from 1 to 1/kappa):
the solution must meet SLATE's FP64 residual bound for every kappa, through
refinement when kappa * u_factor is small enough and through the FP64
fallback otherwise (iteration code -(itermax+1))."""
import numpy as np
import pytest
import torch

import oracle
import slate_amd
from helpers import dev_rands, say, scaled_residual, to_host

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("sym", [False, True])
@pytest.mark.parametrize("n,dist,cond", [(300, 2, 1e6), (257, 1, 1e2), (64, 3, 1e10)])
def test_generate_svd_matches_oracle(gpu, sym, n, dist, cond):
    say("generate_svd n=%d dist=%d cond=%g sym=%d" % (n, dist, cond, sym))
    A = slate_amd.empty_colmajor(n, n, torch.float64)
    slate_amd.generate_svd(A, dist, cond, 11, sym=sym)
    H = to_host(A)
    ref = oracle.generate_svd(dist, cond, 11, n, sym=sym)
    err = np.abs(H - ref).max()
    say("  max |gpu - oracle| = %.3g" % err)
    assert err <= 50 * n * np.finfo(np.float64).eps
    sv = np.linalg.svd(H, compute_uv=False)
    np.testing.assert_allclose(sv, np.sort(oracle.sigma(dist, cond, n))[::-1], rtol=0,
                               atol=20 * n * np.finfo(np.float64).eps)


@pytest.mark.parametrize("cond", [1e2, 1e6, 1e10])
@pytest.mark.parametrize("prec", [slate_amd.PREC_FP32, slate_amd.PREC_BF16X3])
def test_gesv_mixed_kappa(gpu, cond, prec):
    n = 1024
    say("dsgesv kappa=%g n=%d prec=%d" % (cond, n, prec))
    A = slate_amd.empty_colmajor(n, n, torch.float64)
    slate_amd.generate_svd(A, slate_amd.SIGMA_GEO, cond, 21)
    B = dev_rands(22, n, 1)
    X = slate_amd.empty_colmajor(n, 1, torch.float64)
    A_in, B_in = to_host(A), to_host(B)
    info, it, _ = slate_amd.gesv_mixed(A, B, X, slate_amd.Options(block_size=256, factor_precision=prec))
    say("  info %d iter %d" % (info, it))
    assert info == 0
    if cond <= 1e6:
        assert 0 <= it <= 30                          # kappa * u_float < 1: IR converges
    assert it >= 0 or it == -31
    assert scaled_residual(A_in, to_host(X), B_in) <= 3.0


@pytest.mark.parametrize("cond", [1e2, 1e6])
def test_posv_mixed_kappa(gpu, cond):
    n = 1024
    say("dsposv kappa=%g n=%d" % (cond, n))
    A = slate_amd.empty_colmajor(n, n, torch.float64)
    slate_amd.generate_svd(A, slate_amd.SIGMA_GEO, cond, 23, sym=True)
    B = dev_rands(24, n, 1)
    X = slate_amd.empty_colmajor(n, 1, torch.float64)
    A_in, B_in = to_host(A), to_host(B)
    info, it = slate_amd.posv_mixed(A, B, X, opts=slate_amd.Options(block_size=256))
    say("  info %d iter %d" % (info, it))
    assert info == 0 and 0 <= it <= 30
    assert scaled_residual(A_in, to_host(X), B_in) <= 3.0


@pytest.mark.parametrize("cond", [1e6, 1e10])
def test_gesv_mixed_gmres_kappa(gpu, cond):
    n = 1024
    say("dsgesv_gmres kappa=%g n=%d" % (cond, n))
    A = slate_amd.empty_colmajor(n, n, torch.float64)
    slate_amd.generate_svd(A, slate_amd.SIGMA_GEO, cond, 25)
    B = dev_rands(26, n, 1)
    X = slate_amd.empty_colmajor(n, 1, torch.float64)
    A_in, B_in = to_host(A), to_host(B)
    info, it, _ = slate_amd.gesv_mixed_gmres(A, B, X, slate_amd.Options(block_size=256))
    say("  info %d iter %d" % (info, it))
    assert info == 0
    assert it >= 0 or it == -31
    assert scaled_residual(A_in, to_host(X), B_in) <= 3.0
