"""FP8 e4m3 factor precision on the GPU against independent restatements.

* The e4m3 update kernel (slate_amd_sgemm_fp8): quantization per 128-row
  tile of A and per 128-column tile of B with power-of-two scales, one or two
  planes, fp32 accumulation, against a torch/numpy restatement of the same
  rule (tests/helpers.quantized_update_ref, itself pinned on the CPU against
  torch's float8_e4m3fn and the OCP spec values); the difference must stay at
  fp32 accumulation level.  An exactly-representable case (every product and
  partial sum exact in fp32) must agree bitwise: it pins the MFMA's reading of
  the bytes as OCP E4M3FN and the operand lane layout.
* FP8 / FP8X2 refinement (gesv_mixed, gesv_mixed_gmres, posv_mixed_gmres) on
  kappa-controlled inputs against the oracle with the same update rounding
  (oracle.update_precision(UPD_FP8, terms)): same info, same iteration-code
  class, iterations within the bound of tests/test_gpu_c3.py.  Measured with
  the oracle at n = 512-2048: one plane never converges within 30 iterations
  even at kappa = 10 (it falls back to FP64), two planes converge up to
  kappa ~ 1e2.
"""
import numpy as np
import pytest
import torch

import oracle
import slate_amd as S
from helpers import EPS, UPD_FP8, dev_rands, quantized_update_ref, say, scaled_residual, to_dev, to_host

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("terms", [1, 2])
@pytest.mark.parametrize("m,n,k", [(300, 200, 96), (128, 128, 128), (513, 260, 1000), (64, 700, 300)])
def test_fp8_gemm_matches_restatement(gpu, terms, m, n, k):
    say("sgemm_fp8 terms=%d m=%d n=%d k=%d" % (terms, m, n, k))
    rng = np.random.default_rng(m + n + k)
    A = (rng.standard_normal((m, k)) * np.exp(rng.uniform(-4, 4, (m, 1)))).astype(np.float32)
    B = (rng.uniform(-1, 1, (k, n)) * np.exp(rng.uniform(-3, 3, (1, n)))).astype(np.float32)
    C = rng.uniform(-1, 1, (m, n)).astype(np.float32)
    Cd = to_dev(C, torch.float32)
    S.gemm_fp8(terms, -1.0, to_dev(A, torch.float32), to_dev(B, torch.float32), Cd)
    got = to_host(Cd).astype(np.float32)
    want = quantized_update_ref(A, B, C, UPD_FP8, terms)
    # sum of |products| bounds the fp32 accumulation error of either side
    mag = np.abs(A).astype(np.float64) @ np.abs(B).astype(np.float64) + np.abs(C)
    d = np.abs(got.astype(np.float64) - want) / mag
    say("  max |gpu - restatement| / sum|terms| = %.3g" % d.max())
    assert d.max() <= 4 * k * np.finfo(np.float32).eps


@pytest.mark.parametrize("terms", [1, 2])
def test_fp8_gemm_exact_case(gpu, terms):
    """Operands whose quantization is exact and whose products and sums are
    exact in fp32: GPU and restatement must agree bitwise."""
    say("sgemm_fp8 exact terms=%d" % terms)
    rng = np.random.default_rng(5)
    m, n, k = 256, 256, 256
    # e4m3 values with at most 3 significant bits, scaled so the tile max is 448 / 2^s
    vals = np.array([0.0, 0.5, 0.625, 0.75, 0.875, 1.0, 1.25, 1.5, 1.75])
    A = (rng.choice(vals, (m, k)) * rng.choice([-1, 1], (m, k))).astype(np.float32)
    B = (rng.choice(vals, (k, n)) * rng.choice([-1, 1], (k, n))).astype(np.float32)
    A[:, 0] = 1.75
    B[0, :] = 1.75
    C = np.zeros((m, n), dtype=np.float32)
    Cd = to_dev(C, torch.float32)
    S.gemm_fp8(terms, -1.0, to_dev(A, torch.float32), to_dev(B, torch.float32), Cd)
    got = to_host(Cd).astype(np.float32)
    want = quantized_update_ref(A, B, C, UPD_FP8, terms).astype(np.float32)
    exact = -(A.astype(np.float64) @ B.astype(np.float64))
    assert np.array_equal(want.astype(np.float64), exact)
    assert np.array_equal(got, want)


def _case(n, cond, seed, sym=False):
    A = S.empty_colmajor(n, n, torch.float64)
    S.generate_svd(A, S.SIGMA_GEO, cond, seed, sym=sym)
    B = dev_rands(seed + 1, n, 1)
    return A, B, to_host(A), to_host(B)


def _compare(it, it_o, info, info_o, x, x_o, A_h, B_h, cond, n, itermax=30):
    say("  gpu info %d iter %d | oracle info %d iter %d" % (info, it, info_o, it_o))
    assert info == info_o
    cls = lambda v: "conv" if v >= 0 else ("fallback" if v == -(itermax + 1) else v)
    if not (it_o >= itermax - 3 or it >= itermax - 3):
        assert cls(it) == cls(it_o)
    if it >= 0 and it_o >= 0:
        assert abs(it - it_o) <= max(2, it_o // 4)
    x = np.asarray(x).reshape(-1)
    x_o = np.asarray(x_o).reshape(-1)
    r = scaled_residual(A_h, x, B_h)
    d = np.abs(x - x_o).max() / np.abs(x_o).max()
    say("  scaled residual %.3g, |x - x_oracle| / |x_oracle| = %.3g" % (r, d))
    assert r <= 3.0
    assert d <= 20 * cond * n * EPS


@pytest.mark.parametrize("cond", [1e1, 1e2, 1e4])
@pytest.mark.parametrize("prec,terms", [(S.PREC_FP8, 1), (S.PREC_FP8X2, 2)])
@pytest.mark.parametrize("gmres", [False, True])
def test_gesv_mixed_fp8_vs_oracle(gpu, cond, prec, terms, gmres):
    n, nb = 2048, 256
    say("fp8 %s terms=%d kappa=%g n=%d" % ("dsgesv_gmres" if gmres else "dsgesv", terms, cond, n))
    A, B, A_h, B_h = _case(n, cond, 71)
    X = S.empty_colmajor(n, 1, torch.float64)
    f = S.gesv_mixed_gmres if gmres else S.gesv_mixed
    info, it, _ = f(A, B, X, S.Options(block_size=nb, factor_precision=prec))
    with oracle.tile_blas(), oracle.update_precision(oracle.UPD_FP8, terms):
        g = oracle.gesv_mixed_gmres if gmres else oracle.gesv_mixed
        _, _, X_o, info_o, it_o = g(A_h, B_h, nb=nb)
    _compare(it, it_o, info, info_o, to_host(X), X_o, A_h, B_h, cond, n)


@pytest.mark.parametrize("cond", [1e1, 1e3])
def test_posv_mixed_gmres_fp8x2_vs_oracle(gpu, cond):
    n, nb = 1024, 128
    say("fp8x2 dsposv_gmres kappa=%g n=%d" % (cond, n))
    A, B, A_h, B_h = _case(n, cond, 81, sym=True)
    X = S.empty_colmajor(n, 1, torch.float64)
    info, it = S.posv_mixed_gmres(A, B, X, opts=S.Options(block_size=nb, factor_precision=S.PREC_FP8X2))
    with oracle.tile_blas(), oracle.update_precision(oracle.UPD_FP8, 2):
        _, X_o, info_o, it_o = oracle.posv_mixed_gmres(np.tril(A_h), B_h, nb=nb)
    _compare(it, it_o, info, info_o, to_host(X), X_o, A_h, B_h, cond, n)


def test_fp8_block_size_rule(gpu):
    """FP8 forms need a block size that is a multiple of 128 (the scale tile)."""
    n = 512
    A = dev_rands(0, n, n)
    B = dev_rands(1, n, 1)
    X = S.empty_colmajor(n, 1, torch.float64)
    with pytest.raises(S.SlateError):
        S.gesv_mixed(A, B, X, S.Options(block_size=192, factor_precision=S.PREC_FP8X2))
