"""Shared helpers for the GPU parity tests: the same seeded matgen inputs on the
device (slate_amd.generate) and on the host (oracle.generate)."""
import sys

import numpy as np
import torch

import oracle
import slate_amd

EPS = np.finfo(np.float64).eps


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def dev_rands(seed, m, n, dtype=torch.float64):
    A = slate_amd.empty_colmajor(m, n, torch.float64)
    slate_amd.generate(A, slate_amd.DIST_RANDS, seed)
    return A if dtype == torch.float64 else slate_amd.colmajor(A.to(dtype))


def to_dev(M, dtype=torch.float64):
    t = torch.from_numpy(np.asfortranarray(M)).to(dtype)
    return slate_amd.colmajor(t.cuda())


def to_host(T):
    return np.asfortranarray(T.detach().cpu().double().numpy())


def spd(seed, n):
    """SPD test matrix G G^T / n + I with G matgen randn (the posv config)."""
    G = oracle.randn(seed, n, n)
    return G @ G.T / n + np.eye(n)


def scaled_residual(A, X, B):
    return oracle.scaled_residual(A, X, B)


UPD_BF16, UPD_FP8 = 1, 2


def torch_e4m3(x):
    """Round to OCP e4m3 (torch.float8_e4m3fn, round to nearest even) and back."""
    return torch.from_numpy(np.asarray(x, dtype=np.float32)).to(torch.float8_e4m3fn).to(torch.float64).numpy()


def quantized_update_ref(A, B, C, kind, terms, qrows=128):
    """Independent restatement: torch conversions, numpy products."""
    A = A.astype(np.float32).astype(np.float64)
    B = B.astype(np.float32).astype(np.float64)

    def planes(X):                       # X: rows x k, tiles of qrows rows
        if kind == UPD_BF16:
            out, r = [], X.astype(np.float32)
            for _ in range(terms):
                h = torch.from_numpy(r).to(torch.bfloat16).to(torch.float32).numpy()
                out.append(h.astype(np.float64))
                r = (r - h).astype(np.float32)
            return out, np.ones(X.shape[0])
        q0 = np.zeros_like(X)
        q1 = np.zeros_like(X)
        sc = np.ones(X.shape[0])
        for i0 in range(0, X.shape[0], qrows):
            blk = X[i0:i0 + qrows]
            amax = np.abs(blk).max()
            s = 0
            if amax > 0:
                s = int(np.floor(np.log2(448.0 / amax)))
                while amax * 2.0 ** s > 448.0:
                    s -= 1
                while amax * 2.0 ** (s + 1) <= 448.0:
                    s += 1
            xs = blk * 2.0 ** s
            q0[i0:i0 + qrows] = torch_e4m3(xs)
            q1[i0:i0 + qrows] = torch_e4m3((xs - q0[i0:i0 + qrows]) * 16.0)
            sc[i0:i0 + qrows] = 2.0 ** -s
        return ([q0, q1][:terms]), sc

    pa, sa = planes(A)
    pb, sb = planes(B.T)
    if kind == UPD_BF16:
        acc = sum(pa[i] @ pb[j].T for i in range(terms) for j in range(terms) if i + j < terms)
    else:
        acc = pa[0] @ pb[0].T
        if terms > 1:
            acc = acc + (pa[0] @ pb[1].T + pa[1] @ pb[0].T) / 16.0
    acc = acc * sa[:, None] * sb[None, :]
    return (C.astype(np.float32) - acc.astype(np.float32)).astype(np.float32)
