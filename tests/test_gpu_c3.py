"""Mixed-precision refinement on kappa-controlled inputs against the oracle.

A = U diag(sigma) V^T (U, V products of n Householder reflectors; SPD:
V diag(sigma) V^T) with sigma geometric, arithmetic or clustered from 1 to
1/kappa (matgen/generate_sigma.hh:61-111), generated on the GPU and copied to
the host, so the GPU driver and the oracle see bit-identical FP64 inputs.
For every refinement driver (src/gesv_mixed.cc, src/gesv_mixed_gmres.cc,
src/posv_mixed.cc, src/posv_mixed_gmres.cc) and factor precision (float, and
the float factor with one-, two- and three-term split-BF16 trailing updates,
which the oracle emulates with the same operand rounding and FP32
accumulation, its other tile gemm/trsm on a BLAS as SLATE's HostTask tiles
run them), checked:

* the same info;
* the same iteration-code class (converged >= 0, -(itermax+1) no convergence
  with FP64 fallback, -3 singular factor), except when the oracle converges
  within 3 iterations of the cap, where rounding may tip either way;
* when both converge, iter <= iter_oracle + max(2, iter_oracle / 4): the
  GPU needs no more iterations than the oracle beyond that margin (it may
  need fewer: its float factor applies L11^{-1} with a refinement step and
  accumulates updates in a different order, and refinement counts react
  strongly to small differences in the factor's accuracy; exact counts are
  rounding-dependent, parity unpinned beyond this bound);
* the solution within kappa-scaled distance of the oracle's:
  ||x - x_o||_inf / ||x_o||_inf <= 20 kappa n eps, and the tester's scaled
  residual <= 3 (test/test_gesv.cc:330-380).
"""
import numpy as np
import pytest
import torch

import oracle
import slate_amd as S
from helpers import EPS, dev_rands, say, scaled_residual, to_host

pytestmark = pytest.mark.gpu

UPD = {  # GPU factor precision -> oracle update emulation (kind, terms)
    S.PREC_FP32: (oracle.UPD_FLOAT, 0),
    S.PREC_BF16: (oracle.UPD_BF16, 1),
    S.PREC_BF16X2: (oracle.UPD_BF16, 2),
    S.PREC_BF16X3: (oracle.UPD_BF16, 3),
}
PNAME = {S.PREC_FP32: "fp32", S.PREC_BF16: "bf16", S.PREC_BF16X2: "bf16x2", S.PREC_BF16X3: "bf16x3"}


def _case(n, dist, cond, seed, sym):
    A = S.empty_colmajor(n, n, torch.float64)
    S.generate_svd(A, dist, cond, seed, sym=sym)
    B = dev_rands(seed + 1, n, 1)
    return A, B, to_host(A), to_host(B)


def _compare(it, it_o, info, info_o, x, x_o, A_h, B_h, cond, n, itermax=30):
    say("  gpu info %d iter %d | oracle info %d iter %d" % (info, it, info_o, it_o))
    assert info == info_o
    cls = lambda v: "conv" if v >= 0 else ("fallback" if v == -(itermax + 1) else v)
    if not (it_o >= itermax - 3 or it >= itermax - 3):
        assert cls(it) == cls(it_o)
    if it >= 0 and it_o >= 0:
        assert it <= it_o + max(2, it_o // 4)
    x = np.asarray(x).reshape(-1)
    x_o = np.asarray(x_o).reshape(-1)
    r = scaled_residual(A_h, x, B_h)
    d = np.abs(x - x_o).max() / np.abs(x_o).max()
    say("  scaled residual %.3g, |x - x_oracle| / |x_oracle| = %.3g" % (r, d))
    assert r <= 3.0
    assert d <= 20 * cond * n * EPS


@pytest.mark.parametrize("cond", [1e2, 1e6, 1e10])
@pytest.mark.parametrize("prec", [S.PREC_FP32, S.PREC_BF16, S.PREC_BF16X2, S.PREC_BF16X3])
@pytest.mark.parametrize("gmres", [False, True])
def test_gesv_mixed_kappa_vs_oracle(gpu, cond, prec, gmres):
    n, nb = 1024, 256
    say("c3 %s %s kappa=%g n=%d" % ("dsgesv_gmres" if gmres else "dsgesv", PNAME[prec], cond, n))
    A, B, A_h, B_h = _case(n, S.SIGMA_GEO, cond, 31, False)
    X = S.empty_colmajor(n, 1, torch.float64)
    opts = S.Options(block_size=nb, factor_precision=prec)
    f = S.gesv_mixed_gmres if gmres else S.gesv_mixed
    info, it, _ = f(A, B, X, opts)
    with oracle.tile_blas(), oracle.update_precision(*UPD[prec]):
        g = oracle.gesv_mixed_gmres if gmres else oracle.gesv_mixed
        _, _, X_o, info_o, it_o = g(A_h, B_h, nb=nb)
    _compare(it, it_o, info, info_o, to_host(X), X_o, A_h, B_h, cond, n)


@pytest.mark.parametrize("cond", [1e2, 1e6, 1e10])
@pytest.mark.parametrize("prec", [S.PREC_FP32, S.PREC_BF16, S.PREC_BF16X3])
@pytest.mark.parametrize("gmres", [False, True])
def test_posv_mixed_kappa_vs_oracle(gpu, cond, prec, gmres):
    n, nb = 1024, 256
    say("c3 %s %s kappa=%g n=%d" % ("dsposv_gmres" if gmres else "dsposv", PNAME[prec], cond, n))
    A, B, A_h, B_h = _case(n, S.SIGMA_GEO, cond, 41, True)
    X = S.empty_colmajor(n, 1, torch.float64)
    opts = S.Options(block_size=nb, factor_precision=prec)
    f = S.posv_mixed_gmres if gmres else S.posv_mixed
    info, it = f(A, B, X, opts=opts)
    with oracle.tile_blas(), oracle.update_precision(*UPD[prec]):
        g = oracle.posv_mixed_gmres if gmres else oracle.posv_mixed
        _, X_o, info_o, it_o = g(np.tril(A_h), B_h, nb=nb)
    _compare(it, it_o, info, info_o, to_host(X), X_o, A_h, B_h, cond, n)


@pytest.mark.parametrize("dist", [S.SIGMA_ARITH, S.SIGMA_CLUSTER0, S.SIGMA_CLUSTER1])
def test_gesv_mixed_gmres_spectra(gpu, dist):
    n, nb, cond = 1024, 256, 1e6
    say("c3 dsgesv_gmres sigma dist %d kappa=%g" % (dist, cond))
    A, B, A_h, B_h = _case(n, dist, cond, 51, False)
    X = S.empty_colmajor(n, 1, torch.float64)
    info, it, _ = S.gesv_mixed_gmres(A, B, X, S.Options(block_size=nb))
    with oracle.tile_blas():
        _, _, X_o, info_o, it_o = oracle.gesv_mixed_gmres(A_h, B_h, nb=nb)
    _compare(it, it_o, info, info_o, to_host(X), X_o, A_h, B_h, cond, n)


@pytest.mark.parametrize("prec", [S.PREC_BF16, S.PREC_BF16X2, S.PREC_BF16X3])
@pytest.mark.parametrize("nb", [256, 64])
@pytest.mark.parametrize("gmres", [False, True])
def test_posv_mixed_split_bf16_ragged(gpu, prec, nb, gmres):
    """Split-BF16 Cholesky updates (triangular herk on the upper-tile
    enumeration, two-stream lookahead) on a ragged n with more than two
    steps; lookahead on and off give the same result bitwise.  Each run
    gets its own copy of A: the FP64 fallback factors A in place."""
    n, cond = 1100, 1e4
    say("dsposv%s %s n=%d nb=%d" % ("_gmres" if gmres else "", PNAME[prec], n, nb))
    A, B, A_h, B_h = _case(n, S.SIGMA_GEO, cond, 61, True)
    f = S.posv_mixed_gmres if gmres else S.posv_mixed
    xs = []
    for la in (1, 0):
        X = S.empty_colmajor(n, 1, torch.float64)
        info, it = f(A.clone(), B, X, opts=S.Options(block_size=nb, factor_precision=prec, lookahead=la))
        say("  lookahead %d: info %d iter %d" % (la, info, it))
        xs.append((info, it, X.clone()))
    assert xs[0][0] == xs[1][0] and xs[0][1] == xs[1][1]
    assert torch.equal(xs[0][2], xs[1][2])
    info, it, X = xs[0]
    with oracle.tile_blas(), oracle.update_precision(*UPD[prec]):
        g = oracle.posv_mixed_gmres if gmres else oracle.posv_mixed
        _, X_o, info_o, it_o = g(np.tril(A_h), B_h, nb=nb)
    _compare(it, it_o, info, info_o, to_host(X), X_o, A_h, B_h, cond, n)


def test_gesv_mixed_gmres_bf16x2_converges(gpu):
    """The benchmark's default form (BF16X2 + GMRES-IR) must converge, not fall
    back, on the n = 2000 C1 input."""
    n = 2000
    say("dsgesv_gmres bf16x2 n=%d" % n)
    A = dev_rands(0, n, n)
    B = dev_rands(1, n, 1)
    X = S.empty_colmajor(n, 1, torch.float64)
    A_h, B_h = to_host(A), to_host(B)
    info, it, _ = S.gesv_mixed_gmres(A, B, X, S.Options(block_size=256, factor_precision=S.PREC_BF16X2))
    with oracle.tile_blas(), oracle.update_precision(oracle.UPD_BF16, 2):
        _, _, X_o, info_o, it_o = oracle.gesv_mixed_gmres(A_h, B_h, nb=256)
    assert 0 <= it <= 30 and 0 <= it_o <= 30
    _compare(it, it_o, info, info_o, to_host(X), X_o, A_h, B_h, 1e6, n)
