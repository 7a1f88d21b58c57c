"""Learned selector: the MLP's HIP training and inference kernels against
the CPU oracle (gradients, Adam step, probabilities), the features, the
solve driven by the selector, and what a trained network gains over SLATE's
default (classic IR with FP64 fallback) on held-out systems."""
import numpy as np
import pytest
import torch

import oracle
import slate_amd
from helpers import dev_rands, say, scaled_residual, to_host
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
    n = 512
    for kappa in (1e2, 1e6, 1e10):
        A = slate_amd.empty_colmajor(n, n, torch.float64)
        slate_amd.generate_svd(A, slate_amd.SIGMA_GEO, kappa, 3)
        f, info = slate_amd.selector_features(A)
        k_inf = np.linalg.cond(to_host(A), np.inf)
        say("kappa %.0e: estimate 10^%.2f, true kappa_inf 10^%.2f" % (kappa, 16 * f[1], np.log10(k_inf)))
        assert info == 0
        # within a decade while the float factor is meaningful (kappa well
        # below 1/eps_float ~ 1.7e7); beyond that the factor itself carries
        # relative errors ~ kappa * eps_float and the estimate only tracks the
        # order of magnitude, so its rounding-dependent error is bounded looser
        assert abs(16 * f[1] - np.log10(k_inf)) <= (1.0 if kappa < 1e7 else 1.5)
        assert abs(f[0] - np.log2(n) / 16) < 1e-6
        assert abs(f[4:20].sum() - 1.0) < 1e-5


@pytest.mark.parametrize("choice", [0, 1, 2])
def test_gesv_auto_follows_the_network(gpu, choice):
    n = 600
    P = np.zeros(oracle.mlp_num_params(), dtype=np.float32)
    P[-oracle.MLP_K + choice] = 10.0                    # output bias: always this class
    A = dev_rands(40, n, n)
    B = dev_rands(41, n, 1)
    X = slate_amd.empty_colmajor(n, 1, torch.float64)
    A_in = to_host(A)
    info, it, ch, ipiv, feat = slate_amd.gesv_auto(A, B, X, P)
    say("gesv_auto forced %d: choice %d iter %d" % (choice, ch, it))
    assert info == 0 and ch == choice and it >= 0
    assert scaled_residual(A_in, to_host(X), to_host(B)) <= 3.0
    if choice != 2:
        assert np.array_equal(to_host(A), A_in)       # refinement converged: A untouched


def test_gesv_auto_default_is_classic_ir(gpu):
    n = 500
    A = dev_rands(42, n, n)
    B = dev_rands(43, n, 1)
    X = slate_amd.empty_colmajor(n, 1, torch.float64)
    info, it, ch, _, _ = slate_amd.gesv_auto(A, B, X, None)
    assert info == 0 and ch == 0 and it >= 0


def test_trained_selector_beats_default_on_holdout(gpu):
    """Train on systems the solver labels itself; on held-out systems the
    network's choices must cost less than SLATE's default (always classic IR,
    FP64 fallback on a miss), and its misselection rate is reported."""
    rows = selector.build_dataset(160, seed=3, nmin=192, nmax=640, log=say)
    train, hold = rows[:120], rows[120:]
    params, loss = selector.train(train, steps=600, batch=120, lr=3e-3, seed=1)
    rep = selector.evaluate(params, hold)
    say("selector holdout: %s (final loss %.3f)" % (rep, loss))
    assert rep["ms_mlp"] < rep["ms_default_ir"]
    assert rep["misselection_rate"] <= 0.5
