"""Learned choice of refinement scheme for the mixed-precision solve.

The selector sees, per system, the features of ``slate_amd.selector_features``
(log2 n, the condition estimate from the float factor, per-tile max |a|
spread, an exponent histogram) and picks one of ``slate_amd.CHOICES``:
classic IR, GMRES-IR or plain FP64.  Its training data comes from the solver
itself: each generated system is solved all three ways and timed, and the
label is the fastest way that reached SLATE's FP64 tolerance (a refinement
that misses it pays for the FP64 fallback, as in src/gesv_mixed.cc).  The
network is trained on the GPU by the library's own kernels
(``slate_amd.mlp_train_step``).

Usage (on a GPU):  python -m slate_amd.selector train --systems 600 --out PATH
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_PARAMS = os.path.join(HERE, "selector_params.npy")
SHAPES = (1, 2, 3, 4)          # sigma: arith, geo, cluster0, cluster1


def init_params(seed=0):
    """He-initialised parameters in the layout of slate_amd/csrc/mlp.hip."""
    import slate_amd as S
    rng = np.random.default_rng(seed)
    F, H, K = S.SELECTOR_FEATURES, 64, len(S.CHOICES)
    parts = [rng.standard_normal((F, H)) * math.sqrt(2.0 / F), np.zeros(H),
             rng.standard_normal((H, H)) * math.sqrt(2.0 / H), np.zeros(H),
             rng.standard_normal((H, K)) * math.sqrt(1.0 / H), np.zeros(K)]
    P = np.concatenate([p.ravel() for p in parts]).astype(np.float32)
    assert P.size == S.mlp_num_params()
    return P


NMIN, NMAX = 512, 4096


def sample_system(rng, nmin=NMIN, nmax=NMAX):
    """(n, kappa, sigma shape, seed) of one training system: n log-uniform in
    [nmin, nmax], kappa log-uniform in [1e1, 1e14]."""
    n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
    kappa = 10.0 ** rng.uniform(1.0, 14.0)
    return n, kappa, int(rng.choice(SHAPES)), int(rng.integers(1, 2 ** 31))


def label_system(n, kappa, shape, seed, opts=None):
    """Solve one generated system three ways; returns a dict with the
    features, the cost of each choice (ms), iterations and the label."""
    import torch
    import slate_amd as S
    A0 = S.empty_colmajor(n, n, torch.float64)
    S.generate_svd(A0, shape, kappa, seed)
    B = S.empty_colmajor(n, 1, torch.float64)
    S.generate(B, S.DIST_RANDS, seed + 7)
    feat, finfo = S.selector_features(A0, opts)
    # the float factor the auto path always computes (features + factor)
    A32 = S.colmajor(A0.float())
    S.getrf(A32, opts)
    t_factor = S.last_time_ms()
    del A32
    X = S.empty_colmajor(n, 1, torch.float64)
    A = A0.clone()
    _, it_ir, _ = S.gesv_mixed(A, B, X, opts)
    t_ir = S.last_time_ms()
    A = A0.clone()
    _, it_gm, _ = S.gesv_mixed_gmres(A, B, X, opts)
    t_gm = S.last_time_ms()
    A = A0.clone()
    B2 = B.clone()
    S.gesv(A, B2, opts)
    t_64 = S.last_time_ms()
    cost = [t_ir, t_gm, t_factor + t_64]
    return {"n": n, "kappa": kappa, "shape": shape, "seed": seed, "features": feat.tolist(),
            "cost": cost, "iters": [it_ir, it_gm], "label": int(np.argmin(cost))}


def build_dataset(count, seed=1, nmin=NMIN, nmax=NMAX, log=None):
    rng = np.random.default_rng(seed)
    rows = []
    t0 = time.time()
    for i in range(count):
        rows.append(label_system(*sample_system(rng, nmin, nmax)))
        if log and (i + 1) % 25 == 0:
            log("selector: %d/%d systems labelled (%.0f s)" % (i + 1, count, time.time() - t0))
    return rows


def train(rows, steps=1500, batch=256, lr=3e-3, seed=0, log=None):
    """Train on the GPU with the library's kernels; returns numpy parameters."""
    import torch
    import slate_amd as S
    X = torch.tensor(np.array([r["features"] for r in rows], dtype=np.float32), device="cuda")
    y = torch.tensor(np.array([r["label"] for r in rows], dtype=np.int32), device="cuda")
    P = torch.tensor(init_params(seed), device="cuda")
    G = torch.zeros_like(P)
    M = torch.zeros_like(P)
    V = torch.zeros_like(P)
    g = torch.Generator(device="cpu").manual_seed(seed)
    loss = float("nan")
    for t in range(1, steps + 1):
        idx = torch.randint(0, X.shape[0], (min(batch, X.shape[0]),), generator=g).cuda()
        Xb, yb = X[idx].contiguous(), y[idx].contiguous()
        loss = S.mlp_train_step(P, G, M, V, Xb, yb, lr, t)
        if log and t % 250 == 0:
            log("selector: step %d loss %.4f" % (t, loss))
    return P.cpu().numpy(), loss


def predict(params, rows):
    import torch
    import slate_amd as S
    X = torch.tensor(np.array([r["features"] for r in rows], dtype=np.float32), device="cuda")
    P = torch.tensor(np.asarray(params, dtype=np.float32), device="cuda")
    return S.mlp_forward(P, X).argmax(dim=1).cpu().numpy()


def evaluate(params, rows):
    """Misselection rate and total cost of the MLP's choices against SLATE's
    default (always classic IR, FP64 fallback on a miss) and the best choice."""
    pred = predict(params, rows)
    cost = np.array([r["cost"] for r in rows])
    label = np.array([r["label"] for r in rows])
    idx = np.arange(len(rows))
    return {
        "systems": len(rows),
        "misselection_rate": float(np.mean(pred != label)),
        "ms_mlp": float(cost[idx, pred].sum()),
        "ms_default_ir": float(cost[:, 0].sum()),
        "ms_always_fp64": float(cost[:, 2].sum()),
        "ms_best": float(cost[idx, label].sum()),
        "choices": [int(np.sum(pred == k)) for k in range(cost.shape[1])],
        "labels": [int(np.sum(label == k)) for k in range(cost.shape[1])],
    }


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("cmd", choices=["train"])
    ap.add_argument("--systems", type=int, default=600)
    ap.add_argument("--holdout", type=float, default=0.25)
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--out", default=DEFAULT_PARAMS)
    ap.add_argument("--data", default=None, help="also write the labelled systems (JSON) here")
    args = ap.parse_args(argv)

    def log(m):
        print(m, file=sys.stderr, flush=True)

    rows = build_dataset(args.systems, log=log)
    k = int(len(rows) * (1 - args.holdout))
    params, loss = train(rows[:k], steps=args.steps, log=log)
    np.save(args.out, params)
    report = {"train": evaluate(params, rows[:k]), "holdout": evaluate(params, rows[k:]), "final_loss": loss,
              "n_range": [NMIN, NMAX], "kappa_range": [1e1, 1e14]}
    if args.data:
        with open(args.data, "w") as f:
            json.dump(rows, f)
    print(json.dumps(report), flush=True)


if __name__ == "__main__":
    main()
