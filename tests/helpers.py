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
