"""GPU parity of the Cholesky path (potrf/potrs/posv) against the oracle and LAPACK."""
import os

import numpy as np
import pytest
import scipy.linalg as sl
import torch

import oracle
import slate_amd
from helpers import say, scaled_residual, spd, to_dev, to_host

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_potrf_matches_lapack_fixture(gpu):
    say("dpotrf n=128 fixture")
    S = spd(2, 128)
    ref = np.load(os.path.join(GOLD, "potrf_lapack_n128.npy"))
    for nb in (32, 64, 128):
        A = to_dev(S)
        info = slate_amd.potrf(A, opts=slate_amd.Options(block_size=nb))
        assert info == 0
        np.testing.assert_allclose(np.tril(to_host(A)), ref, rtol=0, atol=1e-13)


@pytest.mark.parametrize("n,nb", [(1, 64), (65, 64), (500, 64), (1100, 256)])
def test_potrf_matches_oracle(gpu, n, nb):
    say("dpotrf n=%d nb=%d" % (n, nb))
    S = spd(5 + n, n)
    A = to_dev(S)
    info = slate_amd.potrf(A, opts=slate_amd.Options(block_size=nb))
    L_o, info_o = oracle.potrf(S, nb=nb)
    assert info == info_o == 0
    assert np.abs(np.tril(to_host(A)) - np.tril(L_o)).max() <= 1e-12
    # the unreferenced upper triangle is untouched
    np.testing.assert_array_equal(np.triu(to_host(A), 1), np.triu(S, 1))


def test_potrf_upper_storage(gpu):
    say("dpotrf upper n=300")
    n = 300
    S = spd(8, n)
    A = to_dev(S)
    info = slate_amd.potrf(A, uplo=slate_amd.UPPER, opts=slate_amd.Options(block_size=64))
    assert info == 0
    U = np.triu(to_host(A))
    np.testing.assert_allclose(U, np.linalg.cholesky(S).T, rtol=0, atol=1e-12)
    np.testing.assert_array_equal(np.tril(to_host(A), -1), np.tril(S, -1))


def test_potrf_info_non_spd(gpu):
    say("dpotrf non-SPD n=200, bad diagonal at 137")
    n = 200
    S = spd(9, n)
    S[137, 137] = -1.0
    A = to_dev(S)
    info = slate_amd.potrf(A, opts=slate_amd.Options(block_size=64))
    _, linfo = sl.lapack.dpotrf(S, lower=1)
    _, info_o = oracle.potrf(S, nb=64)
    assert info == linfo == info_o == 138


def test_spotrf_matches_oracle(gpu):
    say("spotrf n=300")
    n = 300
    S = spd(12, n)
    A = to_dev(S, torch.float32)
    info = slate_amd.potrf(A, opts=slate_amd.Options(block_size=64))
    L_o, info_o = oracle.potrf(S.astype(np.float32), nb=64, dtype=np.float32)
    assert info == info_o == 0
    assert np.abs(np.tril(to_host(A)) - np.tril(L_o)).max() <= 1e-4


@pytest.mark.parametrize("uplo", [slate_amd.LOWER, slate_amd.UPPER])
@pytest.mark.parametrize("nrhs", [1, 3])
def test_posv_residual(gpu, uplo, nrhs):
    say("dposv n=900 nrhs=%d uplo=%d" % (nrhs, uplo))
    n = 900
    S = spd(4, n)
    B = oracle.rands(5, n, nrhs)
    A = to_dev(S)
    dB = to_dev(B)
    info = slate_amd.posv(A, dB, uplo=uplo, opts=slate_amd.Options(block_size=128))
    assert info == 0
    X = to_host(dB)
    assert scaled_residual(S, X, B) <= 3.0
    np.testing.assert_allclose(X, sl.solve(S, B, assume_a="pos"), rtol=1e-9, atol=1e-11)
