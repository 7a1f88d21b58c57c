"""Learned selector: the MLP's HIP training and inference kernels against
the CPU oracle (gradients, Adam step, probabilities), the features, the
solve driven by the selector, and what a trained network gains over SLATE's
default (classic IR with FP64 fallback) on held-out systems."""
import json

import numpy as np
import pytest
import torch

import oracle
import slate_amd
from helpers import dev_rands, say, scaled_residual, to_dev, to_host
from slate_amd import selector

pytestmark = pytest.mark.gpu


def _data(B, seed):
    rng = np.random.default_rng(seed)
    P = (rng.standard_normal(oracle.mlp_num_params()) * 0.2).astype(np.float32)
    X = rng.standard_normal((B, oracle.MLP_F)).astype(np.float32)
    y = rng.integers(0, oracle.MLP_K, B).astype(np.int32)
    return P, X, y


@pytest.mark.parametrize("B", [1, 100, 700])
def test_mlp_gradient_matches_oracle(gpu, B):
    P, X, y = _data(B, B)
    dP = torch.tensor(P, device="cuda")
    G, M, V = (torch.zeros_like(dP) for _ in range(3))
    loss = slate_amd.mlp_train_step(dP, G, M, V, torch.tensor(X, device="cuda"), torch.tensor(y, device="cuda"),
                                    0.0, 1)
    loss_o, g_o = oracle.mlp_loss_grad(P, X, y)
    say("mlp grad B=%d loss gpu %.6f oracle %.6f" % (B, loss, loss_o))
    assert abs(loss - loss_o) <= 1e-4 * max(1.0, abs(loss_o))
    g = G.cpu().numpy()
    err = np.abs(g - g_o).max() / max(np.abs(g_o).max(), 1e-12)
    say("  max rel grad error %.3g" % err)
    assert err <= 2e-4
    assert torch.equal(dP.cpu(), torch.tensor(P))          # lr = 0: parameters unchanged


def test_mlp_adam_step_and_forward(gpu):
    P, X, y = _data(64, 5)
    dP = torch.tensor(P, device="cuda")
    G, M, V = (torch.zeros_like(dP) for _ in range(3))
    dX, dy = torch.tensor(X, device="cuda"), torch.tensor(y, device="cuda")
    slate_amd.mlp_train_step(dP, G, M, V, dX, dy, 1e-2, 1)
    _, g_o = oracle.mlp_loss_grad(P, X, y)
    P1, _, _ = oracle.adam_step(P.astype(np.float64), g_o, np.zeros_like(g_o), np.zeros_like(g_o), 1e-2, 1)
    # entries with |g| near zero flip sign under rounding: compare where the step is well defined
    ok = np.abs(g_o) > 1e-5
    np.testing.assert_allclose(dP.cpu().numpy()[ok], P1[ok], rtol=0, atol=2e-5)
    probs = slate_amd.mlp_forward(dP, dX).cpu().numpy()
    np.testing.assert_allclose(probs, oracle.mlp_forward(dP.cpu().numpy(), X), rtol=0, atol=2e-5)


def test_selector_features_track_kappa(gpu):
    """f1 = log10(||A|| ||x0|| / ||b||) / 16 from the BF16X2 probe's first
    solve x0: at most kappa_inf (up to the probe's own error) and, for a
    random right-hand side, within about sqrt(n) of it while kappa * u_probe
    stays below 1 (u ~ 2^-16 for the two-term split); beyond, it only has to
    say "large".  The probe solve is backward stable: its normwise backward
    error (f21) stays near the probe's precision whatever kappa is."""
    n = 512
    B = dev_rands(5, n, 1)
    berr = []
    for kappa in (1e2, 1e4, 1e6, 1e10):
        A = slate_amd.empty_colmajor(n, n, torch.float64)
        slate_amd.generate_svd(A, slate_amd.SIGMA_GEO, kappa, 3)
        f, info = slate_amd.selector_features(A, B)
        k_inf = np.linalg.cond(to_host(A), np.inf)
        say("kappa %.0e: estimate 10^%.2f, true kappa_inf 10^%.2f, probe backward error 10^%.2f"
            % (kappa, 16 * f[1], np.log10(k_inf), 16 * f[21]))
        assert info == 0
        if kappa <= 1e4:
            assert np.log10(k_inf) - np.log10(n) <= 16 * f[1] <= np.log10(k_inf) + 0.3
        else:
            assert 16 * f[1] >= 3.0
        assert abs(f[0] - np.log2(n) / 16) < 1e-6
        assert abs(f[4:20].sum() - 1.0) < 1e-5
        assert f[20] == 1.0 and f[23] in (0.0, 1.0)
        berr.append(f[21])
    assert all(-10.0 <= 16 * b <= -4.0 for b in berr), berr


@pytest.mark.parametrize("choice", list(range(len(slate_amd.CHOICES))))
def test_gesv_auto_follows_the_network(gpu, choice):
    """A network whose output bias forces one class runs that class, and the
    result is the same as the forced-choice run (the labelling path)."""
    n = 600
    P = np.zeros(oracle.mlp_num_params(), dtype=np.float32)
    P[-oracle.MLP_K + choice] = 10.0                    # output bias: always this class
    A = dev_rands(40, n, n)
    B = dev_rands(41, n, 1)
    A_in = to_host(A)
    outs = []
    for kw in ({"params": P}, {"force": choice}):
        A.copy_(to_dev(A_in))
        X = slate_amd.empty_colmajor(n, 1, torch.float64)
        info, it, ch, ipiv, feat = slate_amd.gesv_auto(A, B, X, **kw)
        say("gesv_auto %s: choice %d (%s) iter %d" % (list(kw)[0], ch, slate_amd.CHOICES[ch], it))
        assert info == 0 and ch == choice
        assert it >= 0 or it == -31
        assert scaled_residual(A_in, to_host(X), to_host(B)) <= 3.0
        if it >= 0 and slate_amd.CHOICE_PREC[choice] != slate_amd.PREC_FP64:
            assert np.array_equal(to_host(A), A_in)   # refinement converged: A untouched
        outs.append((it, X.clone(), feat))
    assert outs[0][0] == outs[1][0]
    assert torch.equal(outs[0][1], outs[1][1])
    assert np.array_equal(outs[0][2], outs[1][2])


def test_gesv_auto_default_is_slate_gesv_mixed(gpu):
    """Without parameters: SLATE's default, the float factor with classic IR
    and no probe, bitwise the same solve as gesv_mixed."""
    n = 500
    A = dev_rands(42, n, n)
    B = dev_rands(43, n, 1)
    X = slate_amd.empty_colmajor(n, 1, torch.float64)
    info, it, ch, _, feat = slate_amd.gesv_auto(A, B, X, None)
    assert info == 0 and ch == slate_amd.DEFAULT_CHOICE and it >= 0
    assert not feat.any()
    A2 = dev_rands(42, n, n)
    X2 = slate_amd.empty_colmajor(n, 1, torch.float64)
    info2, it2, _ = slate_amd.gesv_mixed(A2, B, X2)
    assert it2 == it and torch.equal(X, X2)


def test_selector_training_on_solver_labels(gpu):
    """Label a few small systems with the solver itself, train on the GPU with
    the library's kernels; the loss must fall and the GPU forward pass must
    agree with the host restatement used to evaluate committed parameters."""
    rows = selector.label_many(24, seed=3, nmin=192, nmax=512, log=say)
    P0 = selector.init_params(1)
    X = torch.tensor(np.array([r["features"] for r in rows], dtype=np.float32), device="cuda")
    y = np.array([r["label"] for r in rows], dtype=np.int32)
    loss0, _ = oracle.mlp_loss_grad(P0, X.cpu().numpy(), y)
    params, loss = selector.train(rows, steps=300, batch=24, lr=1e-2, seed=1)
    say("selector labels %s, loss %.3f -> %.3f" % (np.bincount(y, minlength=8).tolist(), loss0, loss))
    assert loss < 0.7 * loss0
    probs = slate_amd.mlp_forward(torch.tensor(params, device="cuda"), X).cpu().numpy()
    assert np.array_equal(probs.argmax(axis=1), selector.predict_numpy(params, X.cpu().numpy()))
    rep = selector.evaluate(params, rows)
    say("selector (training set) %s" % json.dumps(rep))
    assert rep["ms_mlp"] <= rep["ms_fixed_choice"][slate_amd.CHOICES[0]] * 1.5
