"""Learned choice of factorization precision and refinement scheme.

A probe factor in BF16X2 (slate_amd.SELECTOR_PROBE) gives the features of
``slate_amd.selector_features``: log2 n, the condition estimate from the probe
factor, per-tile max |a| spread and an exponent histogram.  The MLP picks one
of ``slate_amd.CHOICES``: keep the probe factor with classic IR or GMRES-IR,
factor again in BF16X3, FP32 or FP8X2 with IR / GMRES-IR, or solve in plain
FP64.  Its training data comes from the solver itself: every generated system
is solved with each choice (the probe-factor choices through
``slate_amd.gesv_auto(force=c)``, the others as the probe's time plus their own
solve), and the cost of a choice is its measured device time, probe included,
with a refinement that misses SLATE's tolerance paying for the FP64 fallback
(as in src/gesv_mixed.cc); the label is the cheapest choice.  Training runs on the
GPU with the library's own kernels (``slate_amd.mlp_train_step``).

Systems (the mixed-precision configurations): n log-uniform in
[512, 8192], kappa log-uniform in [1e1, 1e14], singular values geometric,
arithmetic or clustered (matgen/generate_sigma.hh:61-111) of
A = U diag(sigma) V^T with U, V products of n Householder reflectors.

Usage (on a GPU):
  python -m slate_amd.selector label --count 10000 --seed 1 --out part1.npz
  python -m slate_amd.selector train --data part1.npz part2.npz --out selector_params.npy
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
NMIN, NMAX = 512, 8192
KMIN, KMAX = 1e1, 1e14


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


def sample_system(rng, nmin=NMIN, nmax=NMAX):
    """(n, kappa, sigma shape, seed) of one system: n log-uniform in
    [nmin, nmax], kappa log-uniform in [1e1, 1e14]."""
    n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
    kappa = 10.0 ** rng.uniform(math.log10(KMIN), math.log10(KMAX))
    return n, kappa, int(rng.choice(SHAPES)), int(rng.integers(1, 2 ** 31))


def label_system(n, kappa, shape, seed, choices=None):
    """Solve one generated system with every choice; returns the features,
    per-choice cost (device ms, probe included), iteration codes and info,
    the probe's own time, and the label (cheapest choice).

    Choices 0 and 1 keep the probe factor and run through the selector entry
    itself (slate_amd.gesv_auto(force=c)).  Every other choice costs the probe
    (which the selector always pays) plus that choice's own solve, which is
    timed by running the solve directly (gesv_mixed / gesv_mixed_gmres with its
    factor precision, or the plain FP64 gesv) instead of repeating the probe
    inside gesv_auto.

    choices: the choices to time (default all); the others get cost NaN and
    cannot be the label."""
    import torch
    import slate_amd as S
    A0 = S.empty_colmajor(n, n, torch.float64)
    S.generate_svd(A0, shape, kappa, seed)
    B = S.empty_colmajor(n, 1, torch.float64)
    S.generate(B, S.DIST_RANDS, seed + 7)
    X = S.empty_colmajor(n, 1, torch.float64)
    K = len(S.CHOICES)
    cost = np.full(K, np.nan)
    iters = np.zeros(K, dtype=np.int32)
    infos = np.zeros(K, dtype=np.int32)
    todo = set(range(K) if choices is None else choices) | {0}      # choice 0 reports the features
    feat = None
    A = torch.empty_like(A0)
    A.copy_(A0)
    S.selector_features(A, B)
    t_probe = S.last_time_ms()
    for c in range(K):
        if c not in todo:
            continue
        A.copy_(A0)
        if c in (0, 1):
            info, it, ch, _, f = S.gesv_auto(A, B, X, force=c)
            assert ch == c
            cost[c] = S.last_time_ms()
            feat = f
        elif S.CHOICE_PREC[c] == S.PREC_FP64:
            X.copy_(B)
            info, _ = S.gesv(A, X)
            it = 0
            cost[c] = t_probe + S.last_time_ms()
        else:
            f_ = S.gesv_mixed_gmres if S.CHOICE_GMRES[c] else S.gesv_mixed
            info, it, _ = f_(A, B, X, S.Options(factor_precision=S.CHOICE_PREC[c]))
            cost[c] = t_probe + S.last_time_ms()
        iters[c], infos[c] = it, info
    return {"n": n, "kappa": kappa, "shape": shape, "seed": seed, "features": feat, "cost": cost,
            "iters": iters, "info": infos, "t_probe": t_probe, "label": int(np.nanargmin(cost))}


def label_many(count, seed, nmin=NMIN, nmax=NMAX, log=None, max_seconds=None, out=None, every=500,
               choices=None, full_every=0):
    """Label up to count systems drawn with this seed; stops early after
    max_seconds, and (with out) rewrites the .npz every `every` systems so a
    run cut short keeps what it labelled.  choices: time only these (every
    full_every-th system times all)."""
    rng = np.random.default_rng(seed)
    rows = []
    t0 = t_log = time.time()
    for i in range(count):
        ch = None if (choices is None or (full_every and i % full_every == 0)) else choices
        rows.append(label_system(*sample_system(rng, nmin, nmax), choices=ch))
        if log and ((i + 1) % 100 == 0 or time.time() - t_log > 30):
            t_log = time.time()
            log("selector: %d/%d systems labelled (%.0f s, last n = %d)" % (i + 1, count, t_log - t0, rows[-1]["n"]))
        if out and (i + 1) % every == 0:
            save_rows(out, rows)
        if max_seconds is not None and time.time() - t0 > max_seconds:
            if log:
                log("selector: time budget reached after %d systems" % len(rows))
            break
    return rows


FIELDS = ("n", "kappa", "shape", "seed", "features", "cost", "iters", "info", "t_probe", "label")


def refeature(rows, log=None, out=None, every=500):
    """Recompute the features of labelled systems (regenerated from their
    n, kappa, spectrum shape and seed) with the current feature definition.
    Every choice's cost includes the probe, so each cost moves by the change
    in the probe's time; the label is recomputed (for the choices that keep
    the probe factor this keeps the first solve and residual the new probe
    computes in their cost, so they are, if anything, overcharged)."""
    import torch
    import slate_amd as S
    out_rows = []
    t0 = t_log = time.time()
    for i, r in enumerate(rows):
        n, kappa, shape, seed = int(r["n"]), float(r["kappa"]), int(r["shape"]), int(r["seed"])
        A = S.empty_colmajor(n, n, torch.float64)
        S.generate_svd(A, shape, kappa, seed)
        B = S.empty_colmajor(n, 1, torch.float64)
        S.generate(B, S.DIST_RANDS, seed + 7)
        feat, _ = S.selector_features(A, B)
        t_probe = S.last_time_ms()
        cost = np.asarray(r["cost"], dtype=np.float64) - float(r["t_probe"]) + t_probe
        nr = dict(r)
        nr.update(features=feat, cost=cost, t_probe=t_probe, label=int(np.nanargmin(cost)))
        out_rows.append(nr)
        del A, B
        if log and ((i + 1) % 100 == 0 or time.time() - t_log > 30):
            t_log = time.time()
            log("selector: %d/%d systems re-featured (%.0f s, last n = %d)" % (i + 1, len(rows), t_log - t0, n))
        if out and (i + 1) % every == 0:
            save_rows(out, out_rows)
    return out_rows


def save_rows(path, rows):
    np.savez_compressed(path, **{k: np.array([r[k] for r in rows]) for k in FIELDS})


def load_rows(*paths):
    """Labelled systems from one or more .npz files (concatenated)."""
    cols = {k: [] for k in FIELDS}
    for p in paths:
        with np.load(p, allow_pickle=False) as z:
            for k in FIELDS:
                cols[k].append(z[k])
    cols = {k: np.concatenate(v) for k, v in cols.items()}
    return [{k: cols[k][i] for k in FIELDS} for i in range(len(cols["n"]))]


def train(rows, steps=3000, batch=512, lr=3e-3, seed=0, log=None):
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
        step_lr = lr * (0.1 if t > 0.8 * steps else 1.0)
        loss = S.mlp_train_step(P, G, M, V, Xb, yb, step_lr, t)
        if log and t % 500 == 0:
            log("selector: step %d loss %.4f" % (t, loss))
    return P.cpu().numpy(), loss


def _logits(params, features):
    """Network outputs on the host (the layout of slate_amd/csrc/mlp.hip)."""
    import slate_amd as S
    F, H, K = S.SELECTOR_FEATURES, 64, len(S.CHOICES)
    P = np.asarray(params, dtype=np.float64)
    o = 0
    mats = []
    for shape in ((F, H), (H,), (H, H), (H,), (H, K), (K,)):
        sz = int(np.prod(shape))
        mats.append(P[o:o + sz].reshape(shape))
        o += sz
    W1, b1, W2, b2, W3, b3 = mats
    X = np.asarray(features, dtype=np.float64)
    h1 = np.maximum(X @ W1 + b1, 0)
    h2 = np.maximum(h1 @ W2 + b2, 0)
    return h2 @ W3 + b3


def predict_numpy(params, features):
    """Forward pass on the host: the choice per row.  Used for evaluating
    committed parameters without a GPU."""
    return np.argmax(_logits(params, features), axis=1)


def evaluate(params, rows, pred=None):
    """Misselection rate and total cost of the MLP's choices against the best
    choice per system and against fixed policies: SLATE's default (float
    factor + classic IR without the probe), and always the same choice."""
    import slate_amd as S
    cost = np.array([r["cost"] for r in rows], dtype=np.float64)
    full = ~np.isnan(cost).any(axis=1)
    # a choice that was not timed for a system costs as much as its slowest timed one
    cost = np.where(np.isnan(cost), np.nanmax(cost, axis=1, keepdims=True), cost)
    label = np.array([r["label"] for r in rows])
    t_probe = np.array([r["t_probe"] for r in rows])
    if pred is None:
        pred = predict_numpy(params, np.array([r["features"] for r in rows]))
    idx = np.arange(len(rows))
    best = cost[idx, label].sum()
    mlp = cost[idx, pred].sum()
    fixed = {S.CHOICES[k]: float(cost[:, k].sum()) for k in range(cost.shape[1])}
    default = float(np.maximum(cost[:, S.DEFAULT_CHOICE] - t_probe, 0).sum())
    near = np.mean(cost[idx, pred] <= 1.1 * cost[idx, label])
    return {
        "systems": len(rows), "systems_all_choices_timed": int(full.sum()),
        "misselection_rate": float(np.mean(pred != label)),
        "within_10pct_of_best_rate": float(near),
        "ms_mlp": float(mlp), "ms_best": float(best), "mlp_over_best": float(mlp / best),
        "ms_slate_default_no_probe": default, "mlp_over_default": float(mlp / default),
        "ms_fixed_choice": fixed,
        "best_fixed_choice": min(fixed, key=fixed.get),
        "mlp_over_best_fixed": float(mlp / min(fixed.values())),
        "choices": [int(np.sum(pred == k)) for k in range(cost.shape[1])],
        "labels": [int(np.sum(label == k)) for k in range(cost.shape[1])],
    }


def split(rows, holdout=0.2, seed=0):
    rng = np.random.default_rng(seed)
    perm = rng.permutation(len(rows))
    k = int(len(rows) * (1 - holdout))
    return [rows[i] for i in perm[:k]], [rows[i] for i in perm[k:]]


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("cmd", choices=["label", "train", "refeature"])
    ap.add_argument("--count", type=int, default=1000)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--nmin", type=int, default=NMIN)
    ap.add_argument("--nmax", type=int, default=NMAX)
    ap.add_argument("--data", nargs="*", default=[])
    ap.add_argument("--holdout", type=float, default=0.2)
    ap.add_argument("--holdout-out", default=None, help="write the held-out systems here (.npz)")
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--out", default=None)
    ap.add_argument("--max-seconds", type=float, default=None, help="label: stop after this long")
    ap.add_argument("--choices", default=None, help="label: comma-separated choices to time (default all)")
    ap.add_argument("--full-every", type=int, default=0, help="label: every k-th system times all choices")
    args = ap.parse_args(argv)

    def log(m):
        print(m, file=sys.stderr, flush=True)

    if args.cmd == "label":
        rows = label_many(args.count, args.seed, args.nmin, args.nmax, log=log, max_seconds=args.max_seconds,
                          out=args.out, choices=None if not args.choices else [int(c) for c in args.choices.split(",")],
                          full_every=args.full_every)
        save_rows(args.out, rows)
        log("selector: wrote %d systems to %s" % (len(rows), args.out))
        return
    rows = load_rows(*args.data)
    if args.cmd == "refeature":
        rows = refeature(rows, log=log, out=args.out)
        save_rows(args.out, rows)
        log("selector: wrote %d re-featured systems to %s" % (len(rows), args.out))
        return
    tr, ho = split(rows, args.holdout)
    params, loss = train(tr, steps=args.steps, log=log)
    np.save(args.out or DEFAULT_PARAMS, params)
    if args.holdout_out:
        save_rows(args.holdout_out, ho)
    report = {"train": evaluate(params, tr), "holdout": evaluate(params, ho), "final_loss": loss,
              "n_range": [args.nmin, args.nmax], "kappa_range": [KMIN, KMAX], "systems": len(rows)}
    print(json.dumps(report), flush=True)


if __name__ == "__main__":
    main()
