"""Batched uneven solves: workload generator and the batched selector.

The workload is many independent systems of uneven size solved in one
persistent launch (slate_amd.gesv_batched / gesv_batched_auto).  Its standard
shape: n uniform in [32, 1024], half general and half SPD, kappa log-uniform
in [1e1, 1e8], A = U diag(sigma) V^T (SPD: V diag(sigma) V^T) with U, V
products of n Householder reflectors and sigma geometric
(matgen/generate_sigma.hh:61-111), one right-hand side each.

The batched selector picks per system between BF16 + IR (choice 0), FP32 +
IR (1) and FP64 (2) from the normwise quantities of a probe factor (FP32 by
default, or BF16; features computed inside the kernel), keeping the probe
factor when it has the chosen precision.  Labels come from the solver: the
same batch is solved with each choice forced, every system's own time on its
workgroup is measured, and the label is the cheapest choice (a refinement
that misses pays for its FP64 fallback).  With an FP32 probe, choices 0 and 1
run the same path, so only 1 and 2 are timed.

Usage (on a GPU):
  python -m slate_amd.batched label --count 4096 --seed 1 --out b1.npz
  python -m slate_amd.batched train --data b1.npz b2.npz --out slate_amd/selector_batched_params.npy
"""
import argparse
import json
import math
import sys
import time

import numpy as np

NMIN, NMAX = 32, 1024
KMIN, KMAX = 1e1, 1e8


def make_batch(count, seed, nmin=NMIN, nmax=NMAX, kmin=KMIN, kmax=KMAX, shapes=None, log=None):
    """Generate `count` systems on the GPU; returns a dict with ns, kappa,
    sym (bool per system), shape, seeds and the packed float64 CUDA buffers
    A (column-major systems back to back) and B (right-hand sides)."""
    import torch
    import slate_amd as S
    rng = np.random.default_rng(seed)
    ns = rng.integers(nmin, nmax + 1, count).astype(np.int32)
    kappa = 10.0 ** rng.uniform(math.log10(kmin), math.log10(kmax), count)
    sym = (np.arange(count) % 2) == 1
    shapes = shapes or (S.SIGMA_GEO,)
    shape = np.array([shapes[int(rng.integers(0, len(shapes)))] for _ in range(count)], dtype=np.int32)
    seeds = rng.integers(1, 2 ** 31, count)
    offs = np.concatenate([[0], np.cumsum(ns.astype(np.int64) ** 2)])
    A = torch.empty(int(offs[-1]), dtype=torch.float64, device="cuda")
    t0 = time.time()
    for i in range(count):
        n = int(ns[i])
        Ai = A[offs[i]:offs[i + 1]].view(n, n).t()              # column-major view
        S.generate_svd(Ai, int(shape[i]), float(kappa[i]), int(seeds[i]), sym=bool(sym[i]))
        if log and (i + 1) % 1024 == 0:
            log("batched: generated %d/%d systems (%.0f s)" % (i + 1, count, time.time() - t0))
    B = torch.empty(int(ns.sum()), dtype=torch.float64, device="cuda")
    S.generate(B, S.DIST_RANDS, int(seed) + 7)
    return {"ns": ns, "kappa": kappa, "sym": sym, "shape": shape, "seeds": seeds, "A": A, "B": B,
            "offA": offs, "offB": np.concatenate([[0], np.cumsum(ns.astype(np.int64))])}


def label(batch, probe="fp32", log=None):
    """Solve the batch with each choice forced; per-system features, costs
    (ms on its workgroup; NaN for a choice not timed), iteration codes, and
    the label."""
    import torch
    import slate_amd as S
    K = len(S.BATCHED_CHOICES)
    cnt = len(batch["ns"])
    X = torch.zeros_like(batch["B"])
    cost = np.full((cnt, K), np.nan)
    iters = np.zeros((cnt, K), dtype=np.int32)
    feat = None
    opts = S.Options(factor_precision=S.PREC_BF16 if probe == "bf16" else S.PREC_FP32)
    for c in ((0, 1, 2) if probe == "bf16" else (1, 2)):
        X.zero_()
        it, info, berr, ch, f, ms = S.gesv_batched_auto(batch["ns"], batch["A"], batch["B"], X, force=c, opts=opts,
                                                       want_features=True, want_ms=True)
        cost[:, c] = ms
        iters[:, c] = it
        feat = f
        if log:
            log("batched: forced %s: batch %.1f ms, sum of system times %.1f ms, max berr %.3g"
                % (S.BATCHED_CHOICES[c], S.last_time_ms(), ms.sum(), berr.max()))
    rows = []
    for i in range(cnt):
        rows.append({"n": int(batch["ns"][i]), "kappa": float(batch["kappa"][i]), "sym": int(batch["sym"][i]),
                     "features": feat[i], "cost": cost[i], "iters": iters[i],
                     "label": int(np.nanargmin(cost[i]))})
    return rows


FIELDS = ("n", "kappa", "sym", "features", "cost", "iters", "label")


def save_rows(path, rows):
    np.savez_compressed(path, **{k: np.array([r[k] for r in rows]) for k in FIELDS})


def load_rows(*paths):
    cols = {k: [] for k in FIELDS}
    for p in paths:
        with np.load(p, allow_pickle=False) as z:
            for k in FIELDS:
                cols[k].append(z[k])
    cols = {k: np.concatenate(v) for k, v in cols.items()}
    return [{k: cols[k][i] for k in FIELDS} for i in range(len(cols["n"]))]


def predict(params, features):
    """Host restatement of the kernel's choice: argmax over the first three
    network outputs; a probe flagged unusable (feature 4) cannot be kept
    (BF16 probe, feature 7 = 1: FP32 instead; FP32 probe: FP64), and with an
    FP32 probe choices 0 and 1 both keep it (FP32 + IR)."""
    import slate_amd as S
    from .selector import _logits
    f = np.asarray(features)
    z = _logits(params, f)[:, :len(S.BATCHED_CHOICES)]
    c = np.argmax(z, axis=1)
    bad = f[:, 4] != 0
    bf = f[:, 7] != 0
    c[bf & (c == 0) & bad] = 1
    low = ~bf & (c <= 1)
    c[low] = np.where(bad[low], 2, 1)
    return c


def evaluate(params, rows):
    """Misselection rate and summed system time of the network's choices
    against the best choice per system and against each fixed choice."""
    import slate_amd as S
    cost = np.array([r["cost"] for r in rows], dtype=np.float64)
    cost = np.where(np.isnan(cost), np.nanmax(cost, axis=1, keepdims=True), cost)
    lab = np.array([r["label"] for r in rows])
    pred = predict(params, np.array([r["features"] for r in rows]))
    idx = np.arange(len(rows))
    best = cost[idx, lab].sum()
    mlp = cost[idx, pred].sum()
    fixed = {S.BATCHED_CHOICES[k]: float(cost[:, k].sum()) for k in range(cost.shape[1])}
    return {
        "systems": len(rows),
        "misselection_rate": float(np.mean(pred != lab)),
        "within_10pct_of_best_rate": float(np.mean(cost[idx, pred] <= 1.1 * cost[idx, lab])),
        "ms_mlp": float(mlp), "ms_best": float(best), "mlp_over_best": float(mlp / best),
        "ms_fixed_choice": fixed, "best_fixed_choice": min(fixed, key=fixed.get),
        "mlp_over_best_fixed": float(mlp / min(fixed.values())),
        "choices": [int(np.sum(pred == k)) for k in range(cost.shape[1])],
        "labels": [int(np.sum(lab == k)) for k in range(cost.shape[1])],
    }


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("cmd", choices=["label", "train"])
    ap.add_argument("--count", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--data", nargs="*", default=[])
    ap.add_argument("--holdout", type=float, default=0.2)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--out", default=None)
    ap.add_argument("--holdout-out", default=None)
    ap.add_argument("--probe", default="fp32", choices=["fp32", "bf16"])
    args = ap.parse_args(argv)

    def log(m):
        print(m, file=sys.stderr, flush=True)

    if args.cmd == "label":
        batch = make_batch(args.count, args.seed, log=log)
        rows = label(batch, probe=args.probe, log=log)
        save_rows(args.out, rows)
        log("batched: wrote %d labelled systems to %s (labels %s)"
            % (len(rows), args.out, np.bincount([r["label"] for r in rows], minlength=3).tolist()))
        return
    from . import selector
    rows = load_rows(*args.data)
    tr, ho = selector.split(rows, args.holdout)
    params, loss = selector.train(tr, steps=args.steps, log=log)
    np.save(args.out, params)
    if args.holdout_out:
        save_rows(args.holdout_out, ho)
    print(json.dumps({"train": evaluate(params, tr), "holdout": evaluate(params, ho), "final_loss": loss,
                      "systems": len(rows)}), flush=True)


if __name__ == "__main__":
    main()
