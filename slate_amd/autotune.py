"""Block-size autotuner for the factorizations.

SLATE takes the tile size nb from the matrix (the tester sweeps it by hand,
test/test_gesv.cc:165-177).  Here nb is chosen per kind of factorization and
problem size by measurement: for each n of a grid and each candidate nb the
factorization is timed on a matgen matrix, and the fastest nb is installed in
the library (slate_amd_tune_set_nb) for every n up to the next grid point.
Results persist as JSON (tuning.json beside this file by default) and are
applied to each new context by slate_amd.context() when the file exists.

The batched solver's tile configuration (float panel width, workgroup
size, workgroups per CU) is chosen the same way: every valid configuration
solves one batch of the standard batched workload in each fixed precision
(BF16, FP32, FP64 factors), and the configuration with the smallest summed
time is installed (slate_amd_tune_set_batched) and stored under "batched".

Usage (on a GPU):  python -m slate_amd.autotune [--sizes 2048,...,65536] [--out path]
                   python -m slate_amd.autotune --batched [--batch 1024]   (adds the batched entry)
"""
import argparse
import ctypes
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_PATH = os.path.join(HERE, "tuning.json")
KINDS = {"lu_fp64": 0, "lu_fp32": 1, "lu_bf16": 2, "chol": 3}
CANDIDATES = (128, 256, 384, 512, 768, 1024)


def install(table, device=0):
    """table: {kind: [[n_upper, nb], ...]} -> the library context of `device`."""
    import slate_amd as S
    for kind, rows in table.items():
        rows = sorted(rows)
        nu = np.array([r[0] for r in rows], dtype=np.int64)
        nb = np.array([r[1] for r in rows], dtype=np.int64)
        S._check(S.lib().slate_amd_tune_set_nb(S.context(device), KINDS[kind], len(rows),
                                               nu.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                               nb.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))


def install_batched(cfg, device=0):
    import slate_amd as S
    S.tune_set_batched(int(cfg["nbf"]), int(cfg["bt"]), int(cfg["wgs"]), device=device)


def load(path=DEFAULT_PATH, device=0):
    if not os.path.exists(path):
        return None
    with open(path) as f:
        data = json.load(f)
    table = data["table"]
    install(table, device)
    if "batched" in data:
        install_batched(data["batched"], device)
    return table


BATCHED_CONFIGS = ((16, 256), (32, 256), (16, 512))


def tune_batched(count=1024, seed=77, log=None):
    """Time every valid batched configuration on one generated batch; returns
    the fastest {nbf, bt, wgs} and the timings (ms per precision)."""
    import torch
    import slate_amd as S
    from . import batched
    b = batched.make_batch(count, seed)
    X = torch.zeros_like(b["B"])
    saved = S.tune_get_batched()
    raw = {}
    try:
        for nbf, bt in BATCHED_CONFIGS:
            S.tune_set_batched(nbf, bt, 1)
            mx = S.tune_get_batched()[3]
            for wgs in range(1, min(mx, 2) + 1):
                S.tune_set_batched(nbf, bt, wgs)
                t = {}
                for name, prec in (("bf16", S.PREC_BF16), ("fp32", S.PREC_FP32), ("fp64", S.PREC_FP64)):
                    S.gesv_batched(b["ns"], b["A"], b["B"], X, prec)              # warm
                    best = float("inf")
                    for _ in range(2):
                        S.gesv_batched(b["ns"], b["A"], b["B"], X, prec)
                        best = min(best, S.last_time_ms())
                    t[name] = round(best, 3)
                raw["%d/%d/%d" % (nbf, bt, wgs)] = t
                if log:
                    log("autotune batched nbf %d bt %d wgs %d: %s" % (nbf, bt, wgs, t))
    finally:
        S.tune_set_batched(*saved[:3])
    key = min(raw, key=lambda k: sum(raw[k].values()))
    nbf, bt, wgs = (int(x) for x in key.split("/"))
    return {"nbf": nbf, "bt": bt, "wgs": wgs, "systems": count}, raw


def _source(kind, n):
    """The matrix a kind is timed on (kept on the device and copied per run):
    matgen rands seed 0, or G^T G / n + I for Cholesky."""
    import torch
    import slate_amd as S
    A = S.empty_colmajor(n, n, torch.float64)
    S.generate(A, S.DIST_RANDS, 0)
    if kind == "chol":
        A = S.colmajor(torch.addmm(torch.eye(n, dtype=torch.float64, device=A.device), A.t().contiguous(), A,
                                   alpha=1.0 / n))
    elif kind != "lu_fp64":
        A = S.colmajor(A.float())
    return A


def _time_one(kind, A0, nb, reps=2):
    import slate_amd as S
    best = float("inf")
    for _ in range(reps):
        A = A0.clone()
        if kind == "chol":
            S.potrf(A, opts=S.Options(block_size=nb))
        elif kind == "lu_fp64":
            S.getrf(A, S.Options(block_size=nb))
        else:
            prec = S.PREC_BF16X2 if kind == "lu_bf16" else S.PREC_FP32
            S.getrf(A, S.Options(block_size=nb, factor_precision=prec))
        best = min(best, S.last_time_ms())
        del A
    return best


def tune(sizes, kinds=("lu_fp64", "lu_fp32", "lu_bf16", "chol"), log=None, chol_max=32768):
    """Measure every (kind, n, nb); returns the table and the raw timings.
    The split-BF16 kind is timed with the benchmark's two-term form (BF16X2);
    sizes from 32768 on are timed once per nb, smaller ones twice (best);
    Cholesky stops at chol_max (its SPD source is built by an FP64 product)."""
    import torch
    table, raw = {}, {}
    for kind in kinds:
        rows = []
        for n in sorted(sizes):
            if kind == "chol" and n > chol_max:
                continue
            cands = [nb for nb in CANDIDATES if nb <= max(128, n // 4)]
            A0 = _source(kind, n)
            times = {nb: _time_one(kind, A0, nb, reps=1 if n >= 32768 else 2) for nb in cands}
            del A0
            torch.cuda.empty_cache()
            best = min(times, key=times.get)
            raw["%s/%d" % (kind, n)] = times
            rows.append([n, best])
            if log:
                log("autotune %s n=%d: %s -> nb %d" % (kind, n, {k: round(v, 2) for k, v in times.items()}, best))
        # each entry covers sizes up to the next grid point; the last one all larger n
        table[kind] = [[rows[i + 1][0] - 1 if i + 1 < len(rows) else 1 << 40, rows[i][1]]
                       for i in range(len(rows))]
    return table, raw


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--sizes", default="2048,4096,8192,16384,32768,65536")
    ap.add_argument("--kinds", default="lu_fp64,lu_fp32,lu_bf16,chol")
    ap.add_argument("--out", default=DEFAULT_PATH)
    ap.add_argument("--batched", action="store_true", help="tune the batched solver and add its entry")
    ap.add_argument("--batch", type=int, default=1024)
    args = ap.parse_args(argv)
    if args.batched:
        cfg, raw = tune_batched(args.batch, log=lambda m: print(m, file=sys.stderr, flush=True))
        data = {}
        if os.path.exists(args.out):
            with open(args.out) as f:
                data = json.load(f)
        data["batched"] = cfg
        data["batched_timings_ms"] = raw
        with open(args.out, "w") as f:
            json.dump(data, f, indent=1)
        install_batched(cfg)
        print(json.dumps({"batched": cfg}), flush=True)
        return
    sizes = [int(x) for x in args.sizes.split(",")]
    table, raw = tune(sizes, args.kinds.split(","), log=lambda m: print(m, file=sys.stderr, flush=True))
    data = {}
    if os.path.exists(args.out):
        with open(args.out) as f:
            data = json.load(f)
    data["table"] = table
    data["timings_ms"] = raw
    with open(args.out, "w") as f:
        json.dump(data, f, indent=1)
    install(table)
    print(json.dumps({"table": table}), flush=True)


if __name__ == "__main__":
    main()
