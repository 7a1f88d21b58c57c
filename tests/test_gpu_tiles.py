"""SLATE tile storage at the boundary: matrices whose tiles are independent
device buffers with their own leading dimension (as
slate::Matrix::insertLocalTiles(Target::Devices) builds them,
include/slate/internal/MatrixStorage.hh:151), Hermitian ones with the lower
tiles only.  The tile entries must give bitwise the same factors, pivots,
solutions and iteration codes as the dense entries on the same input, and
leave the tiles a reader would use (lower only for Hermitian) holding them."""
import numpy as np
import pytest
import torch

import slate_amd as S
from helpers import dev_rands, say, spd, to_dev

pytestmark = pytest.mark.gpu

N, TB = 1100, 256          # ragged: 5 x 5 tiles, the last 76 wide


def _tiled(M, lower=False, nb=TB):
    return S.TiledMatrix.from_dense(M, nb, nb, lower=lower, pad=5)


@pytest.mark.parametrize("nrhs", [1, 3])
def test_getrf_gesv_getrs_tiles_equal_dense(gpu, nrhs):
    say("tiles gesv n=%d nb=%d nrhs=%d" % (N, TB, nrhs))
    A = dev_rands(11, N, N)
    B = dev_rands(12, N, nrhs)
    opts = S.Options(block_size=TB)
    At, Bt = _tiled(A), _tiled(B)
    info_t, ip_t = S.gesv_tiles(At, Bt, opts)
    Ad, Bd = A.clone(), B.clone()
    info_d, ip_d = S.gesv(Ad, Bd, opts)
    assert info_t == info_d == 0
    np.testing.assert_array_equal(ip_t, ip_d)
    assert torch.equal(At.to_dense(), Ad) and torch.equal(Bt.to_dense(), Bd)
    # getrf + getrs separately, transposed solve
    At2 = _tiled(A)
    info2, ip2 = S.getrf_tiles(At2, opts)
    np.testing.assert_array_equal(ip2, ip_d)
    assert torch.equal(At2.to_dense(), Ad)
    Bt2 = _tiled(B)
    S.getrs_tiles(At2, ip2, Bt2, trans=True, opts=opts)
    Bd2 = B.clone()
    S.getrs(Ad, ip_d, Bd2, trans=True, opts=opts)
    assert torch.equal(Bt2.to_dense(), Bd2)


def test_potrf_posv_tiles_lower_only(gpu):
    say("tiles posv n=%d" % N)
    A = to_dev(spd(13, N))
    B = dev_rands(14, N, 2)
    opts = S.Options(block_size=TB)
    At, Bt = _tiled(A, lower=True), _tiled(B)
    assert all(i >= j for (i, j) in At.tiles)            # upper tiles are not even allocated
    info_t = S.posv_tiles(At, Bt, opts)
    Ad, Bd = A.clone(), B.clone()
    info_d = S.posv(Ad, Bd, opts=opts)
    assert info_t == info_d == 0
    assert torch.equal(torch.tril(At.to_dense()), torch.tril(Ad))
    assert torch.equal(Bt.to_dense(), Bd)
    At2 = _tiled(A, lower=True)
    assert S.potrf_tiles(At2, opts) == 0
    Bt2 = _tiled(B)
    S.potrs_tiles(At2, Bt2, opts)
    assert torch.equal(Bt2.to_dense(), Bd)


@pytest.mark.parametrize("gmres", [False, True])
@pytest.mark.parametrize("prec", [S.PREC_FP32, S.PREC_BF16X2])
def test_mixed_tiles_equal_dense(gpu, gmres, prec):
    say("tiles dsgesv%s prec %d" % ("_gmres" if gmres else "", prec))
    A = dev_rands(15, N, N)
    B = dev_rands(16, N, 1)
    opts = S.Options(block_size=TB, factor_precision=prec)
    At, Bt = _tiled(A), _tiled(B)
    Xt = S.TiledMatrix(N, 1, TB, TB, pad=5)
    info_t, it_t, ip_t = S.gesv_mixed_tiles(At, Bt, Xt, gmres=gmres, opts=opts)
    X = S.empty_colmajor(N, 1, torch.float64)
    Ad = A.clone()
    f = S.gesv_mixed_gmres if gmres else S.gesv_mixed
    info_d, it_d, ip_d = f(Ad, B, X, opts)
    say("  iter tiles %d dense %d" % (it_t, it_d))
    assert info_t == info_d and it_t == it_d
    np.testing.assert_array_equal(ip_t, ip_d)
    assert torch.equal(Xt.to_dense(), X)
    assert torch.equal(At.to_dense(), Ad)                  # untouched on success, factors after a fallback


@pytest.mark.parametrize("gmres", [False, True])
def test_posv_mixed_tiles_equal_dense(gpu, gmres):
    A = to_dev(spd(17, N))
    B = dev_rands(18, N, 1)
    opts = S.Options(block_size=TB)
    At, Bt = _tiled(A, lower=True), _tiled(B)
    Xt = S.TiledMatrix(N, 1, TB, TB)
    info_t, it_t = S.posv_mixed_tiles(At, Bt, Xt, gmres=gmres, opts=opts)
    X = S.empty_colmajor(N, 1, torch.float64)
    f = S.posv_mixed_gmres if gmres else S.posv_mixed
    info_d, it_d = f(A.clone(), B, X, opts=opts)
    assert info_t == info_d == 0 and it_t == it_d
    assert torch.equal(Xt.to_dense(), X)


def test_tiles_fallback_writes_factors_back(gpu):
    """itermax = 0 forces the FP64 fallback: A's tiles then hold the FP64
    factors, as the dense entry leaves A."""
    n = 600
    A = dev_rands(19, n, n)
    B = dev_rands(20, n, 1)
    opts = S.Options(block_size=128, max_iterations=0)
    At, Bt = S.TiledMatrix.from_dense(A, 128, 128), S.TiledMatrix.from_dense(B, 128, 128)
    Xt = S.TiledMatrix(n, 1, 128, 128)
    info_t, it_t, ip_t = S.gesv_mixed_tiles(At, Bt, Xt, opts=opts)
    Ad = A.clone()
    X = S.empty_colmajor(n, 1, torch.float64)
    info_d, it_d, ip_d = S.gesv_mixed(Ad, B, X, opts)
    assert it_t == it_d == -1 and info_t == info_d == 0
    assert torch.equal(At.to_dense(), Ad) and torch.equal(Xt.to_dense(), X)


def test_tiles_reject_bad_descriptor(gpu):
    A = dev_rands(21, 300, 300)
    At = S.TiledMatrix.from_dense(A, 128, 128)
    At._ld[0] = 3                                          # below the tile's 128 rows
    with pytest.raises(S.SlateError):
        S.getrf_tiles(At)
