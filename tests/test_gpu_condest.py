"""Condition estimation from the factors (slate::gecondest / pocondest):
the same Hager/Higham estimator as LAPACK dgecon / dpocon, so on the same
factors the estimates agree with scipy's LAPACK, and on matrices of known
spectrum the estimate brackets the true condition number (the estimator
returns a lower bound of ||A^{-1}||, usually within a factor of 3)."""
import numpy as np
import pytest
import scipy.linalg as sl
import torch

import oracle
import slate_amd
from helpers import dev_rands, say, spd, to_dev, to_host

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("norm", ["1", "I"])
def test_gecondest_matches_lapack_dgecon(gpu, norm):
    n = 500
    A0 = oracle.rands(31, n, n)
    A = to_dev(A0)
    info, ipiv = slate_amd.getrf(A)
    assert info == 0
    anorm = np.abs(A0).sum(axis=0).max() if norm == "1" else np.abs(A0).sum(axis=1).max()
    rc = slate_amd.gecondest(A, ipiv, anorm, norm)
    lu, piv, _ = sl.lapack.dgetrf(A0)
    rc_ref, _ = sl.lapack.dgecon(lu, anorm, norm=norm)
    say("gecondest %s: %.6g  LAPACK dgecon %.6g" % (norm, rc, rc_ref))
    assert abs(rc - rc_ref) <= 1e-6 * rc_ref


@pytest.mark.parametrize("cond", [1e3, 1e8])
def test_gecondest_brackets_known_kappa(gpu, cond):
    n = 400
    A = slate_amd.empty_colmajor(n, n, torch.float64)
    slate_amd.generate_svd(A, slate_amd.SIGMA_GEO, cond, 33)
    A0 = to_host(A)
    kappa1 = np.linalg.cond(A0, 1)
    info, ipiv = slate_amd.getrf(A)
    rc = slate_amd.gecondest(A, ipiv, np.abs(A0).sum(axis=0).max(), "1")
    say("kappa_1 true %.4g  estimated %.4g" % (kappa1, 1 / rc))
    assert kappa1 / 10 <= 1 / rc <= kappa1 * (1 + 1e-8)


def test_pocondest_matches_lapack_dpocon(gpu):
    n = 450
    S0 = spd(4, n)
    S = to_dev(S0)
    assert slate_amd.potrf(S) == 0
    anorm = np.abs(S0).sum(axis=0).max()
    rc = slate_amd.pocondest(S, anorm)
    c, _ = sl.lapack.dpotrf(S0, lower=1)
    rc_ref, _ = sl.lapack.dpocon(c, anorm, uplo="L")
    say("pocondest %.6g  LAPACK dpocon %.6g" % (rc, rc_ref))
    assert abs(rc - rc_ref) <= 1e-6 * rc_ref
