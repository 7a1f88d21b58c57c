"""Benchmark: FP64-accurate gesv TFLOP/s, mixed precision + iterative refinement
vs plain FP64 LU, n = 65536, one MI355X.

One step = one FP64-accurate solve of A x = b (A = matgen rands seed 0,
b = seed 1, both resident in HBM) in the MLP-selected form:
slate_amd.gesv_auto with the committed selector network
(slate_amd/selector_params.npy).  Inside the timed step it factors A in float
with two-term split-BF16 trailing updates (the probe, BF16X2), computes the
features (one pass over A, a sampled exponent histogram, the probe's first
solve x0 and its FP64 residual), runs the network, and continues with the
chosen factor precision and refinement scheme — on this matrix it keeps the
probe factor and runs FP64 GMRES-IR (src/gesv_mixed_gmres.cc), reusing x0 and
its residual as iteration 0.  --form fixed runs a fixed form instead
(--factor-precision, --classic).  TFLOP/s counts FP64-equivalent LU flops
(2 n^3 / 3), whatever precision executes them.  The same input is also solved
with the plain FP64 gesv for the "vs plain FP64 LU" comparison, and the
oracle's CPU restatement of the HostTask path is timed on a bounded sample.

Prints ONE JSON line (see the keys below).  --gpus > 1 is not supported: this
path is a single-GPU factor-and-solve.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FP64-accurate gesv TFLOP/s, mixed+IR vs plain FP64 LU, n=65536, 1 GPU"
PEAK = {"fp32": 157.3, "fp64": 78.6, "bf16": 2500.0, "fp8": 5000.0}
# factor precision -> (C-ABI value, low-precision MFMA products per update (0 = FP32 MFMA))
PRECS = {"fp32": (0, 0), "bf16": (1, 1), "bf16x2": (2, 3), "bf16x3": (4, 6), "fp8": (5, 1), "fp8x2": (6, 3)}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def measured_traffic(kernel_key, factor_precision):
    """HBM bytes per launch of the trailing-update kernel, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload in profiles/
    (the file whose `kernel` and `factor_precision` fields match; None if absent)."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*hbm_traffic*.txt"))):
        fields = {}
        with open(path) as fh:
            for line in fh:
                if line.startswith("#") or not line.strip():
                    continue
                k, _, v = line.strip().partition(" ")
                fields[k] = v.strip()
        if fields.get("kernel") != kernel_key or fields.get("factor_precision") != factor_precision:
            continue
        try:
            return float(fields["traffic_bytes_per_launch"]), os.path.relpath(path, ROOT)
        except (KeyError, ValueError):
            return None, None
    return None, None


def cpu_baseline(n_cpu, gmres, gpu_c1=None):
    """Oracle (CPU restatement of SLATE's HostTask algorithms, tile BLAS
    through scipy's bundled OpenBLAS when present) on bounded samples:

    * the headline workload's form (float tile LU + FP64 refinement) at
      n = n_cpu, timed once: cpu_baseline.value;
    * the plain FP64 gesv at n = 2000, nb = 256 (the reference's CPU-runnable
      case), median of 3, with its pivots compared with the GPU's (gpu_c1)."""
    import numpy as np
    import oracle
    threads = oracle.lib().oracle_num_threads()
    blas = oracle.openblas_path()
    oracle.set_blas(blas)
    nthreads = max(1, threads // 2)                 # panel threads max(cores/2, 1), getrf.cc:41-43
    # C1: n = 2000, nb = 256, ib = 32 (tester default), FP64 gesv
    n1 = 2000
    A1 = oracle.rands(0, n1, n1)
    b1 = oracle.rands(1, n1, 1)
    ts = []
    for _ in range(3):
        t = time.perf_counter()
        LU, ipiv, X1, info1 = oracle.gesv(A1, b1, nb=256, ib=32, nthreads=nthreads)
        ts.append(time.perf_counter() - t)
    t1 = float(np.median(ts))
    c1 = {"n": n1, "nb": 256, "ms": round(1e3 * t1, 2),
          "tflops": round((2.0 / 3.0) * n1 ** 3 / t1 / 1e12, 5), "info": int(info1),
          "scaled_residual": float("%.3e" % oracle.scaled_residual(A1, X1, b1))}
    if gpu_c1 is not None:
        c1["pivots_equal_gpu"] = bool(np.array_equal(np.asarray(ipiv), np.asarray(gpu_c1)))
    A = oracle.rands(0, n_cpu, n_cpu)
    b = oracle.rands(1, n_cpu, 1)
    solve = oracle.gesv_mixed_gmres if gmres else oracle.gesv_mixed
    t = time.perf_counter()
    _, _, X, info, it = solve(A, b, nb=256, ib=16, nthreads=nthreads)
    dt = time.perf_counter() - t
    tf = (2.0 / 3.0) * n_cpu ** 3 / dt / 1e12
    oracle.set_blas(None)
    return {
        "value": round(tf, 5), "unit": "TFLOP/s", "cores": int(threads), "kind": "port",
        "blas": os.path.basename(blas) if blas else "oracle loops",
        "sample": "oracle %s n=%d nrhs=1 (float tile LU nb=256 ib=16, %d panel threads, tile BLAS %s, "
                  "+ FP64 %s, %d iters), matgen rands seed 0/1, %.1f s"
                  % ("gesv_mixed_gmres" if gmres else "gesv_mixed", n_cpu, nthreads,
                     "OpenBLAS" if blas else "loops", "GMRES-IR" if gmres else "IR", it, dt),
        "c1_plain_fp64_gesv": c1,
    }


def c1_gpu_pivots():
    """Pivots of the GPU FP64 getrf at n = 2000, nb = 256 (matgen rands seed 0)."""
    import torch
    import slate_amd as S
    A = S.empty_colmajor(2000, 2000, torch.float64)
    S.generate(A, S.DIST_RANDS, 0)
    _, ipiv = S.getrf(A, S.Options(block_size=256))
    return ipiv


def errors_dev(A, x, b, x_ref):
    """Normwise backward error ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf),
    the tester's scaled residual ||b - A x||_1 / (n ||A||_1 ||x||_1 eps)
    (test/test_gesv.cc:330-377) and the forward error against x_ref,
    ||x - x_ref||_inf / ||x_ref||_inf; on the GPU (checker only)."""
    import torch
    n = A.shape[0]
    r = b - A @ x
    Aabs = A.abs()
    ainf = Aabs.sum(dim=1).max()
    a1 = Aabs.sum(dim=0).max()
    del Aabs
    eps = torch.finfo(torch.float64).eps
    out = {
        "backward_error": float(r.abs().max() / (ainf * x.abs().max() + b.abs().max())),
        "scaled_residual": float(r.abs().sum() / (n * a1 * x.abs().sum()) / eps),
    }
    if x_ref is not None:
        out["forward_error_vs_fp64"] = float((x - x_ref).abs().max() / x_ref.abs().max())
    return {k: float("%.3e" % v) for k, v in out.items()}


def bench_batched(args):
    """Batched uneven solves: args.batch systems, n uniform in [32, 1024], half
    general and half SPD, kappa log-uniform in [1e1, 1e8] (Householder
    generator), one right-hand side each, in one persistent launch.  Timed
    forms: the learned per-system choice (committed network; BF16 probe,
    keep / FP32 / FP64), fixed FP32 factor + FP64 IR (SLATE's gesv_mixed
    form), fixed BF16, and the plain FP64 batched path, which also serves as
    the reference solution for the per-instance forward error."""
    import numpy as np
    import torch
    import slate_amd as S
    from slate_amd import batched
    cnt = args.batch
    log("bench: generating %d systems (n uniform [32, 1024], half SPD, kappa log-uniform [1e1, 1e8])" % cnt)
    b = batched.make_batch(cnt, 2024, log=log)
    ns, A, B = b["ns"], b["A"], b["B"]
    params = S.batched_params()
    forms = [("plain_fp64", lambda X: S.gesv_batched(ns, A, B, X, S.PREC_FP64) + (None,)),
             ("fp32_ir", lambda X: S.gesv_batched(ns, A, B, X, S.PREC_FP32) + (None,)),
             ("bf16_ir", lambda X: S.gesv_batched(ns, A, B, X, S.PREC_BF16) + (None,))]
    if params is not None:
        forms.append(("learned", lambda X: S.gesv_batched_auto(ns, A, B, X, params=params)[:4]))
    res, xs = {}, {}
    for name, run in forms:
        X = torch.zeros_like(B)
        for _ in range(args.warmup):
            run(X)
        times = []
        for _ in range(args.steps):
            torch.cuda.synchronize()
            t = time.perf_counter()
            iters, info, berr, ch = run(X)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t)
        ms = 1e3 * float(np.median(times))
        xs[name] = X
        res[name] = {"ms": round(ms, 3), "systems_per_s": round(cnt / (ms * 1e-3), 1),
                     "max_berr": float("%.3e" % berr.max()), "median_berr": float("%.3e" % np.median(berr)),
                     "max_berr_over_n_eps": round(float(np.max(berr / (ns * np.finfo(np.float64).eps))), 3),
                     "info_nonzero": int(np.sum(info != 0)), "fallbacks": int(np.sum(iters == -31)),
                     "iters_hist": np.bincount(np.clip(iters, -1, 31) + 1).tolist()}
        if ch is not None:
            res[name]["choices"] = dict(zip(S.BATCHED_CHOICES, np.bincount(ch, minlength=3).tolist()))
        log("bench: batched %s %.1f ms, %.0f systems/s, max berr %.3g"
            % (name, ms, cnt / (ms * 1e-3), berr.max()))
    # per-instance forward error of every form against the plain FP64 batched solution
    x64 = xs["plain_fp64"].cpu().numpy()
    offB = b["offB"]
    for name in res:
        if name == "plain_fp64":
            continue
        x = xs[name].cpu().numpy()
        fe = np.array([np.abs(x[offB[i]:offB[i + 1]] - x64[offB[i]:offB[i + 1]]).max()
                       / np.abs(x64[offB[i]:offB[i + 1]]).max() for i in range(cnt)])
        res[name]["max_forward_error_vs_fp64"] = float("%.3e" % fe.max())
        res[name]["max_forward_error_over_kappa"] = float("%.3e" % np.max(fe / b["kappa"]))
    head = "learned" if "learned" in res else "fp32_ir"
    out = {
        "metric": "batched uneven solves, systems/s (FP64-accurate)", "value": res[head]["systems_per_s"],
        "unit": "systems/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": res[head]["ms"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16/f32 factor (per system) + f64 refinement" if head == "learned"
        else "f32 factor + f64 refinement", "data": "synthetic",
        "config": {"workload": "%d systems, n uniform [32,1024], half general / half SPD, kappa log-uniform "
                               "[1e1,1e8] (Householder generator), nrhs=1, one persistent launch; %s form"
                               % (cnt, head),
                   "tile_config": dict(zip(("nbf", "bt", "wgs_per_cu", "max_wgs_per_cu"), S.tune_get_batched())),
                   "forms": res,
                   "speedup_vs_plain_fp64_batched": round(res["plain_fp64"]["ms"] / res[head]["ms"], 3)},
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=65536)
    ap.add_argument("--nb", type=int, default=0)
    ap.add_argument("--cpu-n", type=int, default=12288)
    ap.add_argument("--factor-precision", default="bf16x2", choices=sorted(PRECS),
                    help="factor precision of the mixed solve (bf16x3 = float-accurate split-BF16)")
    ap.add_argument("--classic", action="store_true", help="classic IR instead of GMRES-IR")
    ap.add_argument("--method", default="partialpiv", choices=["partialpiv", "calu"],
                    help="MethodLU of the low-precision factor (CALU: tournament pivoting, src/getrf_tntpiv.cc)")
    sel_params = os.path.join(ROOT, "slate_amd", "selector_params.npy")
    ap.add_argument("--form", default="auto" if os.path.exists(sel_params) else "fixed", choices=["auto", "fixed"],
                    help="auto: the learned selector picks the factor precision and the refinement scheme "
                         "(slate_amd_dsgesv_auto with the committed parameters; probe factor and features "
                         "inside the timed step); fixed: --factor-precision with GMRES-IR or --classic")
    ap.add_argument("--matrix", default="rands", choices=["rands", "geo"],
                    help="rands: matgen uniform(-1,1) seed 0; geo: U diag(sigma) V^T with Householder U, V "
                         "and sigma geometric from 1 to 1/--cond")
    ap.add_argument("--cond", type=float, default=1e6, help="condition number of --matrix geo")
    ap.add_argument("--workload", default="gesv", choices=["gesv", "batched"],
                    help="batched: many uneven systems in one launch instead of the headline solve")
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 comparison solve")
    ap.add_argument("--fp64-runs", type=int, default=3, help="timed plain FP64 solves (after one warmup)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first run one plain elementwise copy (y = 2 x, float32) of this many MiB "
                         "each way, the calibration kernel for HBM byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this path runs on one GPU only (single-GPU factor-and-solve); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    import slate_amd as S

    if args.workload == "batched":
        return bench_batched(args)

    if args.copy_probe_mib > 0:
        cnt = args.copy_probe_mib * (1 << 20) // 4
        xp = torch.ones(cnt, dtype=torch.float32, device="cuda")
        yp = torch.empty_like(xp)
        torch.mul(xp, 2.0, out=yp)
        torch.cuda.synchronize()
        log("bench: copy probe %d MiB read + %d MiB written" % (args.copy_probe_mib, args.copy_probe_mib))
        del xp, yp

    n = args.n
    prec_code, products = PRECS[args.factor_precision]
    opts = S.Options(block_size=args.nb, factor_precision=prec_code,
                     method_lu=S.LU_CALU if args.method == "calu" else S.LU_PARTIALPIV)
    gmres = not args.classic
    choices = []
    if args.form == "auto":
        import numpy as np
        params = np.load(sel_params, allow_pickle=False).astype(np.float32)

        def solve(A_, B_, X_, o_):
            info_, it_, ch_, ipiv_, _ = S.gesv_auto(A_, B_, X_, params=params, opts=o_)
            choices.append(int(ch_))
            return info_, it_, ipiv_
    else:
        solve = S.gesv_mixed_gmres if gmres else S.gesv_mixed
    A = S.empty_colmajor(n, n, torch.float64)
    if args.matrix == "geo":
        mat_desc = "U diag(sigma) V^T, Householder U/V seed 11, sigma geometric 1..1/%g" % args.cond
        log("bench: generating A (%d x %d, %s) and b (rands seed 1) on the GPU" % (n, n, mat_desc))
        S.generate_svd(A, S.SIGMA_GEO, args.cond, 11, sym=False)
    else:
        mat_desc = "matgen rands seed 0"
        log("bench: generating A (%d x %d, %s) and b (seed 1) on the GPU" % (n, n, mat_desc))
        S.generate(A, S.DIST_RANDS, 0)
    B = S.empty_colmajor(n, 1, torch.float64)
    S.generate(B, S.DIST_RANDS, 1)
    X = S.empty_colmajor(n, 1, torch.float64)
    torch.cuda.synchronize()

    # A pristine copy restores A, outside the timed window, after a step that
    # overwrote it (the FP64 choice or an FP64 fallback factor A in place;
    # a converged mixed solve leaves it untouched).
    A0 = A.clone()
    touched = [False]

    def step():
        if touched[0]:
            A.copy_(A0)
            torch.cuda.synchronize()
        info_, it_, _ = solve(A, B, X, opts)
        fp64_form = args.form == "auto" and S.CHOICE_PREC[choices[-1]] == S.PREC_FP64
        touched[0] = it_ < 0 or fp64_form
        assert info_ == 0, "solve failed: info %d" % info_
        return it_

    iters = []
    for w in range(args.warmup):
        it = step()
        log("bench: warmup %d iter %d  %.1f ms%s" % (w, it, S.last_time_ms(), " (choice %d %s)"
            % (choices[-1], S.CHOICES[choices[-1]]) if choices else ""))

    S.profile_enable(True)
    wall = 0.0
    dev_ms = 0.0
    for k in range(args.steps):
        if touched[0]:
            A.copy_(A0)
            touched[0] = False
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        iters.append(step())
        torch.cuda.synchronize()
        wall += time.perf_counter() - t0
        dev_ms += S.last_time_ms()
    gemm_ms, gemm_flops, launches = S.profile_read()
    per_kernel = {k: S.profile_read_kernel(k) for k in (S.PROF_LT, S.PROF_SP, S.PROF_OTHER)}
    S.profile_enable(False)
    if touched[0]:
        A.copy_(A0)
    del A0
    torch.cuda.empty_cache()
    ms_per_step = 1e3 * wall / args.steps
    flops = 2.0 / 3.0 * n ** 3
    value = flops / (ms_per_step * 1e-3) / 1e12
    if args.form == "auto":
        # the form the selector picked in the timed steps (all the same here)
        ch = choices[-1]
        assert all(c == ch for c in choices[-args.steps:]), "selector choice changed between steps"
        pcode = S.CHOICE_PREC[ch]
        gmres = bool(S.CHOICE_GMRES[ch])
        pname = {v[0]: k for k, v in PRECS.items()}.get(pcode, "fp64")
        args.factor_precision = pname
        products = PRECS[pname][1] if pname in PRECS else 0
        log("bench: selector picked choice %d (%s)" % (ch, S.CHOICES[ch]))
    log("bench: %s %.1f ms/step (device %.1f ms), %.2f TFLOP/s, iters %s"
        % ("gesv_mixed_gmres" if gmres else "gesv_mixed", ms_per_step, dev_ms / args.steps, value, iters))

    fp64 = None
    x64 = None
    if not args.no_fp64:
        # plain FP64 gesv on the same input: 1 warmup + fp64_runs timed, median
        A64 = torch.empty_like(A)
        B64 = torch.empty_like(B)
        opts64 = S.Options(block_size=args.nb)
        ts = []
        for rep in range(1 + args.fp64_runs):
            A64.copy_(A)
            B64.copy_(B)
            torch.cuda.synchronize()
            t = time.perf_counter()
            info64, ipiv64 = S.gesv(A64, B64, opts64)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t
            assert info64 == 0, "plain FP64 gesv: info %d" % info64
            if rep > 0:
                ts.append(dt)
            log("bench: plain FP64 gesv %s %.1f ms" % ("warmup" if rep == 0 else "run %d" % rep, 1e3 * dt))
        t64 = sorted(ts)[len(ts) // 2]
        x64 = B64[:, 0].clone()
        # 1-norm condition estimate of A from the FP64 factor (outside the timed region)
        anorm1 = float(A.abs().sum(dim=0).max())
        rcond = S.gecondest(A64, ipiv64, anorm1, "1")
        log("bench: condition estimate kappa_1(A) ~ %.3g" % (1.0 / rcond if rcond > 0 else float("inf")))
        err64 = errors_dev(A, x64, B[:, 0], None)
        assert err64["scaled_residual"] <= 3.0, "plain FP64 gesv: scaled residual %g" % err64["scaled_residual"]
        fp64 = {"ms_median": round(1e3 * t64, 2), "ms_runs": [round(1e3 * v, 2) for v in ts],
                "tflops": round(flops / t64 / 1e12, 3), "info": info64, **err64,
                "kappa_1_estimate": float("%.3e" % (1.0 / rcond)) if rcond > 0 else None}
        log("bench: plain FP64 gesv median %.1f ms, %.2f TFLOP/s, scaled residual %.3e"
            % (1e3 * t64, fp64["tflops"], err64["scaled_residual"]))
        del A64, B64
        torch.cuda.empty_cache()
    err_mixed = errors_dev(A, X[:, 0], B[:, 0], x64)
    log("bench: mixed solution %s" % err_mixed)
    assert err_mixed["scaled_residual"] <= 3.0, "mixed solve: scaled residual %g" % err_mixed["scaled_residual"]

    # Dominant kernel: the trailing update.  roofline.achieved counts the
    # algorithmic 2mnk flops per launch of that kernel; the MFMA issue rate
    # (the bf16 products of the split form) is reported separately.  With one-
    # and two-term splits the large steps run the 256 x 256 plane kernel and
    # the small ones the 128 x 128 split-at-staging kernel: the roofline is
    # the plane kernel's, the other's share is listed beside it.
    mfma_products = products if products else 1
    nterms = {1: 1, 3: 2, 6: 3}.get(products, 0)
    names = {
        S.PROF_LT: "gemm_lt_kernel<%d,1> trailing update (256 x 256 tiles over pre-split bf16 planes, direct "
                   "global->LDS staging, v_mfma_f32_16x16x32_bf16, %d products per update)" % (nterms, products),
        S.PROF_SP: "gemm_sp_kernel<%d,1,3> trailing update (128 x 128 tiles, float operands split into bf16 "
                   "terms at LDS staging, v_mfma_f32_16x16x32_bf16, %d products per update)" % (nterms, products),
        S.PROF_OTHER: "gemm_kernel<float,0,0,128,128,32,2,2,2,1,1> trailing update (v_mfma_f32_16x16x4f32)",
    }
    fp8 = args.factor_precision.startswith("fp8")
    f64 = args.factor_precision == "fp64"
    if f64:
        names[S.PROF_OTHER] = "gemm_f64_kernel<1> trailing update (v_mfma_f64_16x16x4f64, FP64 LU)"
    if fp8:
        names[S.PROF_OTHER] = ("gemm_f8_kernel<%d,1> trailing update (e4m3 planes with a power-of-two scale per "
                               "tile, v_mfma_scale_f32_16x16x128_f8f6f4, %d products per update)"
                               % (nterms, products))
    dom = max(per_kernel, key=lambda k: per_kernel[k][0])
    k_ms, k_flops, k_launches = per_kernel[dom]
    peak = PEAK["fp64"] if f64 else (PEAK["fp8"] if fp8 else (PEAK["bf16"] if products else PEAK["fp32"]))
    kname = names[dom]
    achieved = k_flops / (k_ms * 1e-3) / 1e12 if k_ms > 0 else 0.0
    kernel_key = kname.split(" ")[0]
    traffic, traffic_src = measured_traffic(kernel_key, args.factor_precision) if n == 65536 else (None, None)
    roofline = {
        "bound": "mfma", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
        "frac": round(achieved / peak, 4), "traffic": traffic,
        "kernel": kname,
        "launches": int(k_launches), "avg_launch_ms": round(k_ms / max(k_launches, 1), 4),
        "algorithmic_flops_per_launch": round(k_flops / max(k_launches, 1), 1),
        "mfma_issue_tflops": round(achieved * mfma_products, 3),
        "mfma_issue_frac": round(achieved * mfma_products / peak, 4),
        "all_trailing_updates": {"ms": round(gemm_ms, 2), "launches": int(launches),
                                 "tflops": round(gemm_flops / (gemm_ms * 1e-3) / 1e12, 3) if gemm_ms > 0 else 0.0},
        "other_kernels": {names[k].split(" ")[0]: {"ms": round(v[0], 2), "launches": int(v[2])}
                          for k, v in per_kernel.items() if k != dom and v[2] > 0},
    }
    if traffic_src:
        roofline["traffic_source"] = traffic_src
    cpu = None if args.no_cpu else cpu_baseline(args.cpu_n, gmres, c1_gpu_pivots())
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "TFLOP/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if f64 else (("f32 factor" if not products else "f32 factor, %s updates"
                                       % args.factor_precision) + " + f64 refinement"),
        "data": "synthetic",
        "config": {
            "workload": ("%splain FP64 gesv (after the BF16X2 probe) n=%d nrhs=1, A = %s, b = rands seed 1, "
                         "already in HBM" % ("MLP-selected " if args.form == "auto" else "", n, mat_desc)) if f64
                        else "%s%s (float LU, %s updates, + FP64 %s) n=%d nrhs=1, A = %s, "
                        "b = rands seed 1, already in HBM"
                        % ("MLP-selected " if args.form == "auto" else "",
                           "gesv_mixed_gmres" if gmres else "gesv_mixed", args.factor_precision,
                           "GMRES-IR" if gmres else "classic IR", n, mat_desc),
            "n": n, "nrhs": 1, "nb": args.nb or "auto", "factor_precision": args.factor_precision,
            "method_lu": args.method,
            "refinement": "GMRES-IR" if gmres else "classic IR", "iters": iters,
            "form": ("selector (slate_amd_dsgesv_auto, probe + features in the timed step): choice %d %s"
                     % (choices[-1], S.CHOICES[choices[-1]])) if args.form == "auto" else "fixed",
            **err_mixed,
            "plain_fp64_gesv": fp64,
            "speedup_vs_plain_fp64": round(fp64["ms_median"] / ms_per_step, 3) if fp64 else None,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
