"""Pins of the oracle's low-precision pieces (CPU only).

* FP8 e4m3 (OCP E4M3FN) decode/round/encode: every one of the 256 codes
  against torch's independent float8_e4m3fn conversion and against values the
  OCP 8-bit floating point specification fixes (max normal 448, min normal
  2^-6, min subnormal 2^-9, NaN codes), round-to-nearest-even tie cases, and
  brute-force nearest search.
* The per-tile power-of-two scale: brute force over the exponent.
* The bf16 split: bf16_rne against torch's bfloat16, three terms exact.
* The quantized update C -= A B: against an independent numpy/torch
  restatement of the same rule.
* GMRES-IR's first restart cycle (src/gesv_mixed_gmres.cc:296-352): the
  Arnoldi relation A W_j = V_{j+1} Hbar_j, orthonormal V, non-increasing
  |S(j+1)|; a planted one-pass classical Gram-Schmidt must fail the
  orthogonality check.
"""
import numpy as np
import pytest
import torch

import oracle
from helpers import quantized_update_ref, torch_e4m3

EPS = np.finfo(np.float64).eps


def _torch_e4m3(x):
    return torch_e4m3(x)


def test_e4m3_decode_all_codes():
    codes = torch.arange(256, dtype=torch.int32).to(torch.uint8)
    ref = codes.view(torch.float8_e4m3fn).to(torch.float64).numpy()
    got = np.array([oracle.e4m3_decode(b) for b in range(256)])
    assert np.array_equal(np.isnan(ref), np.isnan(got))
    ok = ~np.isnan(ref)
    assert np.array_equal(ref[ok], got[ok])
    # values fixed by the OCP FP8 spec (E4M3: bias 7, S.1111.111 = NaN, no inf)
    assert got[0x7E] == 448.0 and got[0xFE] == -448.0
    assert got[0x08] == 2.0 ** -6 and got[0x01] == 2.0 ** -9 and got[0x07] == 7 * 2.0 ** -9
    assert got[0x38] == 1.0 and got[0x40] == 2.0 and got[0x3C] == 1.5
    assert np.isnan(got[0x7F]) and np.isnan(got[0xFF])
    assert np.sum(np.isnan(got)) == 2
    assert got[0x00] == 0.0 and np.signbit(got[0x80])


def test_e4m3_encode_roundtrip():
    for b in range(256):
        v = oracle.e4m3_decode(b)
        if np.isnan(v):
            continue
        assert oracle.e4m3_encode(v) == b, hex(b)


def test_e4m3_round_ties_and_random():
    # ties go to the even mantissa
    cases = {
        1.0625: 1.0,                      # between 1.0 (m=0) and 1.125 (m=1)
        1.1875: 1.25,                     # between 1.125 (m=1) and 1.25 (m=2)
        -1.0625: -1.0,
        1.5 * 2.0 ** -9: 2 * 2.0 ** -9,   # subnormal tie: m=1 / m=2 -> 2
        2.5 * 2.0 ** -9: 2 * 2.0 ** -9,
        0.5 * 2.0 ** -9: 0.0,             # half the smallest subnormal -> 0 (even)
        0.75 * 2.0 ** -9: 2.0 ** -9,
        432.0: 448.0,                     # tie between 416 (m=5) and 448 (m=6)
        400.0: 384.0,                     # tie between 384 (m=4) and 416 (m=5)
    }
    for x, want in cases.items():
        assert oracle.e4m3_round(x) == want, (x, oracle.e4m3_round(x), want)
    rng = np.random.default_rng(7)
    x = rng.standard_normal(50000) * np.exp(rng.uniform(-14, 6, 50000))
    x = np.clip(x, -448, 448).astype(np.float32).astype(np.float64)
    got = np.array([oracle.e4m3_round(v) for v in x])
    assert np.array_equal(got, _torch_e4m3(x))
    # brute force: nearest finite code, ties to the even code
    vals = np.array([oracle.e4m3_decode(b) for b in range(128) if not np.isnan(oracle.e4m3_decode(b))])
    for v in x[:3000]:
        a = abs(v)
        d = np.abs(vals - a)
        best = np.flatnonzero(d == d.min())
        pick = vals[best[0]] if len(best) == 1 else vals[[b for b in best if (np.flatnonzero(vals == vals[b])[0] & 1) == 0][0]]
        assert abs(oracle.e4m3_round(v)) == pick


def test_e4m3_scale_exp_brute_force():
    rng = np.random.default_rng(3)
    amax = np.concatenate([np.exp(rng.uniform(-40, 40, 3000)), [448.0, 449.0, 447.0, 1.0, 0.875, 2.0 ** -30]])
    for a in amax:
        s = oracle.e4m3_scale_exp(a)
        assert a * 2.0 ** s <= 448.0 and a * 2.0 ** (s + 1) > 448.0, (a, s)
    assert oracle.e4m3_scale_exp(0.0) == 0


def test_bf16_split_exact():
    rng = np.random.default_rng(5)
    x = (rng.standard_normal(20000) * np.exp(rng.uniform(-20, 20, 20000))).astype(np.float32)
    h0 = np.array([oracle.bf16_rne(v) for v in x], dtype=np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(h0, ref)
    r1 = x - h0
    h1 = np.array([oracle.bf16_rne(v) for v in r1], dtype=np.float32)
    r2 = r1 - h1
    h2 = np.array([oracle.bf16_rne(v) for v in r2], dtype=np.float32)
    assert np.array_equal(h0.astype(np.float64) + h1 + h2, x.astype(np.float64))   # 3 x 8 bits hold 24


@pytest.mark.parametrize("kind,terms", [(1, 1), (1, 2), (1, 3), (2, 1), (2, 2)])
def test_quantized_update_matches_restatement(kind, terms):
    rng = np.random.default_rng(11)
    m, n, k = 300, 200, 96
    A = (rng.standard_normal((m, k)) * np.exp(rng.uniform(-3, 3, (m, 1)))).astype(np.float32)
    B = rng.uniform(-1, 1, (k, n)).astype(np.float32)
    C = rng.uniform(-1, 1, (m, n)).astype(np.float32)
    with oracle.update_precision(kind, terms):
        got = oracle.sgemm_update(A, B, C)
    want = quantized_update_ref(A, B, C, kind, terms)
    exact = C.astype(np.float64) - A.astype(np.float64) @ B.astype(np.float64)
    # accumulation order differs (double sums of exact float products): float rounding of the result
    assert np.abs(got - want).max() <= 4 * np.finfo(np.float32).eps * np.abs(exact).max() + 1e-30
    err = np.abs(got - exact).max() / (np.abs(A).max() * np.abs(B).max() * k)
    bound = {(1, 1): 2.0 ** -8, (1, 2): 2.0 ** -16, (1, 3): 2.0 ** -22, (2, 1): 2.0 ** -3, (2, 2): 2.0 ** -7}
    assert err <= bound[(kind, terms)], (err, bound[(kind, terms)])


def test_quantized_update_terms_order_accuracy():
    """Planted-mistake guard: dropping a product or a plane shows as a worse error."""
    rng = np.random.default_rng(12)
    A = rng.uniform(-1, 1, (256, 128)).astype(np.float32)
    B = rng.uniform(-1, 1, (128, 256)).astype(np.float32)
    C = np.zeros((256, 256), dtype=np.float32)
    exact = -(A.astype(np.float64) @ B.astype(np.float64))
    errs = {}
    for kind, terms in [(2, 1), (2, 2), (1, 1), (1, 2), (1, 3)]:
        with oracle.update_precision(kind, terms):
            errs[(kind, terms)] = np.abs(oracle.sgemm_update(A, B, C) - exact).max()
    assert errs[(2, 1)] > 4 * errs[(2, 2)]
    assert errs[(1, 1)] > 16 * errs[(1, 2)] > 16 * errs[(1, 3)]


ORTH_C = 10.0


def _svd_matrix(n, cond, seed, sym=False):
    return oracle.generate_svd(2, cond, seed, n, sym=sym)


@pytest.mark.parametrize("sym", [False, True])
def test_gmres_cycle_invariants(sym):
    n = 256
    A = _svd_matrix(n, 1e6, 21, sym)
    b = oracle.rands(22, n, 1)[:, 0]
    V, W, H, S = oracle.gmres_first_cycle(A, b, nb=64, sym=sym)
    j = W.shape[1]
    assert j >= 3
    Ae = np.tril(A) + np.tril(A, -1).T if sym else A
    # orthonormal basis: CGS2 keeps ||V^T V - I|| at the level of eps
    # ("twice is enough"); the bound 10 sqrt(n) eps leaves room for n
    orth = np.abs(V.T @ V - np.eye(j + 1)).max()
    assert orth <= ORTH_C * np.sqrt(n) * EPS, orth
    # Arnoldi relation with the right-preconditioned operator: A M^{-1} V_j = A W_j = V_{j+1} Hbar_j
    lhs = Ae @ W
    rel = np.abs(lhs - V @ H).max() / (np.abs(Ae).max() * np.abs(W).max() * n)
    assert rel <= 10 * EPS, rel
    # Hessenberg: zero below the first subdiagonal, positive subdiagonal
    assert np.all(np.tril(H, -2) == 0)
    assert np.all(np.diag(H, -1) > 0)
    # GMRES residual estimates never increase
    assert np.all(np.diff(S) <= 1e-12 * S[0])


@pytest.mark.parametrize("cond,kind", [(1e10, oracle.UPD_BF16), (1e14, oracle.UPD_FLOAT)])
def test_gmres_one_pass_cgs_fails_orthogonality(cond, kind):
    """The check above bites: one-pass classical Gram-Schmidt loses
    orthogonality beyond the CGS2 bound on an ill-conditioned preconditioned
    operator (kappa = 1e10 with a one-term bf16 factor, 1e14 with a float one),
    while the same cycle with CGS2 stays within it."""
    n = 256
    A = _svd_matrix(n, cond, 23)
    b = oracle.rands(24, n, 1)[:, 0]
    with oracle.update_precision(kind, 1 if kind else 0):
        V2, _, _, _ = oracle.gmres_first_cycle(A, b, nb=64, cgs_passes=2)
        V1, _, _, _ = oracle.gmres_first_cycle(A, b, nb=64, cgs_passes=1)
    o2 = np.abs(V2.T @ V2 - np.eye(V2.shape[1])).max()
    o1 = np.abs(V1.T @ V1 - np.eye(V1.shape[1])).max()
    assert o2 <= ORTH_C * np.sqrt(n) * EPS, o2
    assert o1 > ORTH_C * np.sqrt(n) * EPS, o1
