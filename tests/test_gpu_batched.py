"""Batched uneven solves (one persistent launch): every instance against the
oracle's FP64 solve of the same matgen system, SLATE's scaled-residual bound
per instance, the refinement codes (0..30, -31 fallback, -3 singular float
factor) and the per-system FP64 mode."""
import numpy as np
import pytest
import torch

import oracle
import slate_amd
from helpers import say, scaled_residual

pytestmark = pytest.mark.gpu


def make_batch(ns, seed, spd_every=2):
    mats, rhs = [], []
    for i, n in enumerate(ns):
        A = oracle.rands(seed + 2 * i, n, n)
        if spd_every and i % spd_every == 1:
            A = A @ A.T / n + np.eye(n)
        mats.append(A)
        rhs.append(oracle.rands(seed + 2 * i + 1, n, 1)[:, 0])
    packA = np.concatenate([m.ravel(order="F") for m in mats])
    packB = np.concatenate(rhs)
    return mats, rhs, torch.tensor(packA, device="cuda"), torch.tensor(packB, device="cuda")


def unpack(X, ns):
    out, o = [], 0
    for n in ns:
        out.append(X[o:o + n])
        o += n
    return out


@pytest.mark.parametrize("prec", [slate_amd.PREC_FP32, slate_amd.PREC_FP64])
def test_batched_uneven_matches_oracle(gpu, prec):
    rng = np.random.default_rng(0)
    ns = [1, 2, 31, 32, 33, 63, 64, 65, 100, 257, 511, 1024] + list(rng.integers(3, 400, 36))
    mats, rhs, dA, dB = make_batch(ns, 100)
    dX = torch.zeros_like(dB)
    iters, info, berr = slate_amd.gesv_batched(ns, dA, dB, dX, prec)
    say("batched prec=%d: %d systems, %.1f ms, iters %s, max berr %.3g"
        % (prec, len(ns), slate_amd.last_time_ms(), np.bincount(iters[iters >= 0]), berr.max()))
    X = unpack(dX.cpu().numpy(), ns)
    for i, n in enumerate(ns):
        assert info[i] == 0
        r = scaled_residual(mats[i], X[i][:, None], rhs[i][:, None])
        assert r <= 3.0, (i, n, r)
        ref = np.linalg.solve(mats[i], rhs[i])
        rel = np.abs(X[i] - ref).max() / np.abs(ref).max()
        assert rel <= 1e-9 * np.linalg.cond(mats[i], np.inf), (i, n, rel)
        assert berr[i] <= 4 * n * np.finfo(np.float64).eps
        if prec == slate_amd.PREC_FP32:
            assert 0 <= iters[i] <= 30
        else:
            assert iters[i] == 0


def test_batched_modes_fallback_and_singular(gpu):
    ns = [40, 200, 150, 90]
    mats, rhs = [], []
    # 0: ill-conditioned (kappa 1e12): float IR cannot converge -> FP64 fallback
    A = slate_amd.empty_colmajor(40, 40, torch.float64)
    slate_amd.generate_svd(A, slate_amd.SIGMA_GEO, 1e12, 7)
    mats.append(A.cpu().numpy())
    # 1: forced to FP64 by mode; 2: well conditioned; 3: singular (zero matrix)
    mats.append(oracle.rands(8, 200, 200))
    mats.append(oracle.rands(9, 150, 150))
    mats.append(np.zeros((90, 90)))
    rhs = [oracle.rands(10 + i, n, 1)[:, 0] for i, n in enumerate(ns)]
    dA = torch.tensor(np.concatenate([m.ravel(order="F") for m in mats]), device="cuda")
    dB = torch.tensor(np.concatenate(rhs), device="cuda")
    dX = torch.zeros_like(dB)
    iters, info, berr = slate_amd.gesv_batched(ns, dA, dB, dX, slate_amd.PREC_FP32, mode=[0, 1, 0, 0])
    say("modes: iters %s info %s" % (iters, info))
    X = unpack(dX.cpu().numpy(), ns)
    assert iters[0] == -31 and info[0] == 0
    assert scaled_residual(mats[0], X[0][:, None], rhs[0][:, None]) <= 3.0
    assert iters[1] == 0 and info[1] == 0
    assert 0 <= iters[2] <= 30 and info[2] == 0
    assert iters[3] == -3 and info[3] == 1
