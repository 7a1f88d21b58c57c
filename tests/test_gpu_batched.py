"""Batched uneven solves (one persistent launch): every instance against the
oracle's FP64 solve of the same matgen system, SLATE's scaled-residual bound
per instance, the refinement codes (0..30, -31 fallback, -3 singular float
factor), the per-system FP64 mode, the BF16 factor, the learned choice
(forced and by a network, against the host restatement of its decision),
graph replay, every tile configuration, and the 4096-system workload
against the plain FP64 batched path."""
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


@pytest.mark.parametrize("prec", [slate_amd.PREC_FP32, slate_amd.PREC_FP64, slate_amd.PREC_BF16])
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
        elif prec == slate_amd.PREC_BF16:
            assert 0 <= iters[i] <= 30 or iters[i] == -31, iters[i]
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


def _check_solutions(ns, mats, rhs, dX, berr, tag):
    X = unpack(dX.cpu().numpy(), ns)
    for i, n in enumerate(ns):
        r = scaled_residual(mats[i], X[i][:, None], rhs[i][:, None])
        assert r <= 3.0, (tag, i, n, r)
        assert berr[i] <= 4 * n * np.finfo(np.float64).eps, (tag, i, n, berr[i])


@pytest.mark.parametrize("probe", [slate_amd.PREC_FP32, slate_amd.PREC_BF16])
@pytest.mark.parametrize("force", [0, 1, 2])
def test_batched_auto_forced_choice(gpu, force, probe):
    """Forced choices after an FP32 or BF16 probe factor: what ran (an FP32
    probe is kept for choices 0 and 1), the features, per-system times, and
    every instance against the oracle's system."""
    rng = np.random.default_rng(5)
    ns = [1, 7, 64, 129, 300, 1024] + list(rng.integers(3, 500, 18))
    mats, rhs, dA, dB = make_batch(ns, 300)
    dX = torch.zeros_like(dB)
    iters, info, berr, ch, feat, ms = slate_amd.gesv_batched_auto(
        ns, dA, dB, dX, force=force, opts=slate_amd.Options(factor_precision=probe), want_features=True,
        want_ms=True)
    say("probe %d forced %d: choices %s iters %s" % (probe, force, np.bincount(ch, minlength=3),
                                                    np.bincount(iters[iters >= 0])))
    _check_solutions(ns, mats, rhs, dX, berr, "force %d" % force)
    assert np.all(info == 0)
    expect = force if probe == slate_amd.PREC_BF16 else max(force, 1)
    assert np.all(ch == expect)
    assert np.all(ms > 0)
    assert np.all(np.isfinite(feat))
    np.testing.assert_allclose(feat[:, 0], 1.0)
    np.testing.assert_allclose(feat[:, 1], np.log2(ns) / 10, rtol=1e-6)
    assert np.all(feat[:, 7] == (1.0 if probe == slate_amd.PREC_BF16 else 0.0))
    assert np.all(feat[:, 8:] == 0)
    if force == 2:
        assert np.all(iters == 0)
    else:
        assert np.all((iters >= 0) & (iters <= 30) | (iters == -31))


def test_batched_auto_follows_network(gpu):
    """A random network: the kernel's per-system choice equals the host
    restatement of the forward pass on the features it reported."""
    from slate_amd import batched, selector
    rng = np.random.default_rng(9)
    ns = list(rng.integers(16, 400, 40))
    mats, rhs, dA, dB = make_batch(ns, 500)
    for seed, probe in ((1, slate_amd.PREC_FP32), (2, slate_amd.PREC_BF16), (3, slate_amd.PREC_BF16)):
        P = selector.init_params(seed) * 3.0
        dX = torch.zeros_like(dB)
        iters, info, berr, ch, feat, _ = slate_amd.gesv_batched_auto(
            ns, dA, dB, dX, params=P, opts=slate_amd.Options(factor_precision=probe), want_features=True)
        say("network seed %d: choices %s" % (seed, np.bincount(ch, minlength=3)))
        assert np.array_equal(ch, batched.predict(P, feat))
        _check_solutions(ns, mats, rhs, dX, berr, "network %d" % seed)


def test_batched_graph_replay_equals_eager(gpu):
    rng = np.random.default_rng(11)
    ns = list(rng.integers(8, 300, 30))
    mats, rhs, dA, dB = make_batch(ns, 700)
    X0 = torch.zeros_like(dB)
    ref = slate_amd.gesv_batched(ns, dA, dB, X0, slate_amd.PREC_FP32)
    b0, r0 = slate_amd.batched_graph_stats()
    o = slate_amd.Options(use_graph=1)
    X = torch.zeros_like(dB)                      # same buffers every call: the graph is replayed
    for rep in range(3):
        X.zero_()
        out = slate_amd.gesv_batched(ns, dA, dB, X, slate_amd.PREC_FP32, opts=o)
        torch.cuda.synchronize()
        assert torch.equal(X, X0), rep
        for a, b in zip(out, ref):
            assert np.array_equal(a, b), rep
        # eager work between replays
        Y = torch.zeros_like(dB)
        slate_amd.gesv_batched(ns, dA, dB, Y, slate_amd.PREC_FP32)
        assert torch.equal(Y, X0)
    b1, r1 = slate_amd.batched_graph_stats()
    say("batched graph: builds %d -> %d, replays %d -> %d" % (b0, b1, r0, r1))
    assert b1 - b0 == 1 and r1 - r0 == 2


def test_batched_tile_configurations(gpu):
    saved = slate_amd.tune_get_batched()
    rng = np.random.default_rng(13)
    ns = [1, 33, 1024] + list(rng.integers(2, 600, 13))
    mats, rhs, dA, dB = make_batch(ns, 900)
    try:
        for nbf, bt in ((16, 256), (32, 256), (16, 512)):
            for wgs in (1, 2):
                slate_amd.tune_set_batched(nbf, bt, 1)
                mx = slate_amd.tune_get_batched()[3]
                if wgs > mx:
                    continue
                slate_amd.tune_set_batched(nbf, bt, wgs)
                assert slate_amd.tune_get_batched()[:3] == (nbf, bt, wgs)
                for prec in (slate_amd.PREC_FP32, slate_amd.PREC_FP64, slate_amd.PREC_BF16):
                    dX = torch.zeros_like(dB)
                    iters, info, berr = slate_amd.gesv_batched(ns, dA, dB, dX, prec)
                    say("config nbf %d bt %d wgs %d (max %d) prec %d: %.1f ms" % (nbf, bt, wgs, mx, prec,
                                                                                 slate_amd.last_time_ms()))
                    _check_solutions(ns, mats, rhs, dX, berr, (nbf, bt, wgs, prec))
        with pytest.raises(slate_amd.SlateError):
            slate_amd.tune_set_batched(32, 512, 1)
    finally:
        slate_amd.tune_set_batched(*saved[:3])


def test_batched_4096_systems_vs_fp64_batched(gpu):
    """The batched workload at its full shape: 4096 systems, n uniform in
    [32, 1024], half general and half SPD, kappa log-uniform in [1e1, 1e8]
    (Householder generator); the learned choice per system (the committed
    network) against the plain FP64 batched path on the same systems."""
    from slate_amd import batched
    b = batched.make_batch(4096, 2024, log=say)
    ns, A, B = b["ns"], b["A"], b["B"]
    X64 = torch.zeros_like(B)
    it64, info64, berr64 = slate_amd.gesv_batched(ns, A, B, X64, slate_amd.PREC_FP64)
    t64 = slate_amd.last_time_ms()
    params = slate_amd.batched_params()
    X = torch.zeros_like(B)
    if params is not None:
        iters, info, berr, ch, _, _ = slate_amd.gesv_batched_auto(ns, A, B, X, params=params)
    else:
        iters, info, berr, ch, _, _ = slate_amd.gesv_batched_auto(ns, A, B, X, force=1)
    t = slate_amd.last_time_ms()
    say("4096 systems: FP64 %.1f ms, learned %.1f ms (choices %s, fallbacks %d)"
        % (t64, t, np.bincount(ch, minlength=3), int(np.sum(iters == -31))))
    eps = np.finfo(np.float64).eps
    assert np.all(info64 == 0) and np.all(info == 0)
    assert np.all(berr64 <= 4 * ns * eps)
    assert np.all(berr <= 4 * ns * eps), np.max(berr / (ns * eps))
    # per instance: scaled residual on the GPU and forward error against FP64
    xs, x64s = X.cpu().numpy(), X64.cpu().numpy()
    for i in range(len(ns)):
        n = int(ns[i])
        Ai = A[b["offA"][i]:b["offA"][i + 1]].view(n, n).t()
        bi = B[b["offB"][i]:b["offB"][i + 1]]
        xi = X[b["offB"][i]:b["offB"][i + 1]]
        r = bi - Ai @ xi
        sr = float(r.abs().sum() / (n * Ai.abs().sum(dim=0).max() * xi.abs().sum()) / eps)
        assert sr <= 3.0, (i, n, sr)
        x, x64 = xs[b["offB"][i]:b["offB"][i + 1]], x64s[b["offB"][i]:b["offB"][i + 1]]
        fe = np.abs(x - x64).max() / np.abs(x64).max()
        assert fe <= 16 * b["kappa"][i] * n * eps, (i, n, b["kappa"][i], fe)
