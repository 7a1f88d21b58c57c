"""Backward error of the low-precision LU factors against the oracle's.

The refinement drivers' iteration counts depend on how well the float
factor approximates A.  For the same float input, ||P A - L U||_F / ||A||_F of
the GPU's float LU (every factor precision) must stay within a small factor of
the oracle's tile LU with the same update rounding (src/getrf.cc:23-245 with
Tile_getrf.hh; oracle.update_precision for the split-BF16 and FP8 forms) and
its tile gemm/trsm on a BLAS, as SLATE's HostTask tiles run them
(oracle.tile_blas).
"""
import numpy as np
import pytest
import torch

import oracle
import slate_amd as S
from helpers import say, to_dev, to_host

pytestmark = pytest.mark.gpu


def lu_backward_error(A, LU, ipiv):
    """||P A - L U||_F / ||A||_F in float64 (ipiv 1-based LAPACK)."""
    A = np.asarray(A, dtype=np.float64)
    n = A.shape[0]
    LU = np.asarray(LU, dtype=np.float64)
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    PA = A.copy()
    for i, p in enumerate(np.asarray(ipiv) - 1):
        if p != i:
            PA[[i, p]] = PA[[p, i]]
    return np.linalg.norm(PA - L @ U) / np.linalg.norm(A)


UPD = {S.PREC_FP32: (0, 0), S.PREC_BF16: (1, 1), S.PREC_BF16X2: (1, 2), S.PREC_BF16X3: (1, 3),
       S.PREC_FP8: (2, 1), S.PREC_FP8X2: (2, 2)}


@pytest.mark.parametrize("prec", [S.PREC_FP32, S.PREC_BF16X3, S.PREC_BF16X2, S.PREC_FP8X2])
@pytest.mark.parametrize("kind", ["rands", "cluster0", "geo1e10"])
def test_float_lu_backward_error_vs_oracle(gpu, prec, kind):
    n, nb = 1024, 256
    if kind == "rands":
        H = oracle.rands(3, n, n)
    elif kind == "cluster0":
        H = oracle.generate_svd(S.SIGMA_CLUSTER0, 1e6, 51, n)
    else:
        H = oracle.generate_svd(S.SIGMA_GEO, 1e10, 31, n)
    H32 = H.astype(np.float32)
    A = to_dev(H32, torch.float32)
    info, ipiv = S.getrf(A, S.Options(block_size=nb, factor_precision=prec))
    e_gpu = lu_backward_error(H32, to_host(A), ipiv)
    with oracle.tile_blas(), oracle.update_precision(*UPD[prec]):
        LU_o, ipiv_o, info_o = oracle.getrf(H32, nb=nb, dtype=np.float32)
    e_o = lu_backward_error(H32, LU_o, ipiv_o)
    say("sgetrf %s prec %d: backward error gpu %.3e oracle %.3e (ratio %.2f)" % (kind, prec, e_gpu, e_o, e_gpu / e_o))
    assert info == info_o == 0
    assert e_gpu <= 4.0 * e_o
