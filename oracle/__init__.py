"""CPU oracle for the factor-and-solve path — TEST INFRASTRUCTURE ONLY.

A ctypes wrapper over ``oracle/liboracle.so`` (built from ``slate_oracle.cc`` by
``oracle/Makefile``), a plain CPU restatement of SLATE's ``Target::HostTask``
single-rank algorithms.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker (or as the timed CPU baseline).  Nothing in ``slate_amd/`` uses it.

Parity status of each restated function (how each one is pinned):

================  ==========================================  =====================================
function          reference                                   pinned by
================  ==========================================  =====================================
generate          matgen/random.cc:53-127                     parity unpinned for exact bits (the
                                                              reference needs BLAS++ headers absent
                                                              here); distribution moments, tiling
                                                              independence and the committed
                                                              fixture guard regressions
dgetrf/sgetrf     src/getrf.cc:23-245,                        LAPACK dgetrf (scipy/OpenBLAS) pivots
                  src/internal/Tile_getrf.hh:159-447          (committed n=2000 fixture), brute-force
                                                              long-double GEPP on tiny n, PA=LU
                                                              invariant, info on exact-zero columns
dgetrs            src/getrs.cc:66-117                         scipy lu_solve, residual bound, trans
dpotrf/spotrf     src/potrf.cc:23-211                         LAPACK dpotrf (elementwise), info =
                                                              first non-SPD minor
dpotrs/dposv      src/potrs.cc:55-77, src/posv.cc:71-94       scipy cho_solve, residual bound
dsgesv/dsposv     src/gesv_mixed.cc:106-290,                  backward error <= bound, iteration
                  src/posv_mixed.cc:105-292                   codes (-3 singular, -(itermax+1)),
                                                              fallback result equals plain solve
dsgesv_gmres      src/gesv_mixed_gmres.cc:111-410             same, plus Arnoldi invariants via the
                                                              GPU path; exact iteration counts are
                                                              parity unpinned (rounding-dependent)
rotg              blas::rotg (reference BLAS drotg)           closed form: c*a+s*b=r, -s*a+c*b=0
sigma             matgen/generate_sigma.hh:61-111             closed forms (endpoints 1 and 1/cond,
                                                              geometric ratio, arithmetic step)
mlp_*             (this library's selector network; no        finite differences of the loss
                  reference counterpart)                      (tests/test_oracle.py), softmax sums
generate_svd      (this library's generator; the reference's  invariants: singular values = sigma,
                  svd kind, generate_type_svd.hh:90-157,      orthogonality, det = (-1)^n of each
                  uses QR of a Gaussian instead)              n-reflector product, symmetry for sym
norms/residual    src/norm.cc, src/colNorms.cc                numpy closed forms
rbt_weights       src/internal/internal_rbt_generate.cc:26-140 closed form exp(rands/20); the
                                                              LAPACK larnv stream of the reference
                                                              is parity unpinned (not available)
gerbt/gerbt_rhs   src/gerbt.cc, src/internal/Tile_gerbt.hh    explicit dense butterfly matrices
                                                              built from the definition, products
                                                              of levels, orthogonality with unit
                                                              weights, U^T A V = (U^T)(A)(V)
dgesv_rbt         src/gesv_rbt.cc:94-236                      numpy solve, L U = U^T A V for the
                                                              unpivoted factor, iteration codes,
                                                              fallback result equals plain gesv
================  ==========================================  =====================================
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

c_i64 = ctypes.c_int64
c_dbl = ctypes.c_double
c_int = ctypes.c_int
c_p = ctypes.c_void_p


def build():
    """Compile liboracle.so with the committed Makefile."""
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        sigs = {
            "oracle_generate": (None, [c_int, c_i64, c_i64, c_i64, c_i64, c_i64, c_p, c_i64]),
            "oracle_philox": (None, [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, c_p]),
            "oracle_dgetrf": (c_i64, [c_i64, c_i64, c_p, c_i64, c_i64, c_i64, c_i64, c_dbl, c_p]),
            "oracle_sgetrf": (c_i64, [c_i64, c_i64, c_p, c_i64, c_i64, c_i64, c_i64, c_dbl, c_p]),
            "oracle_dgetrs": (None, [c_i64, c_i64, c_p, c_i64, c_p, c_p, c_i64, c_i64, c_int]),
            "oracle_sgetrs": (None, [c_i64, c_i64, c_p, c_i64, c_p, c_p, c_i64, c_i64, c_int]),
            "oracle_dgesv": (c_i64, [c_i64, c_i64, c_p, c_i64, c_p, c_p, c_i64, c_i64, c_i64, c_i64]),
            "oracle_dpotrf": (c_i64, [c_i64, c_p, c_i64, c_i64]),
            "oracle_spotrf": (c_i64, [c_i64, c_p, c_i64, c_i64]),
            "oracle_dpotrs": (None, [c_i64, c_i64, c_p, c_i64, c_p, c_i64, c_i64]),
            "oracle_dposv": (c_i64, [c_i64, c_i64, c_p, c_i64, c_p, c_i64, c_i64]),
            "oracle_dsgesv": (c_i64, [c_i64, c_i64, c_p, c_i64, c_p, c_p, c_i64, c_p, c_i64,
                                      c_i64, c_i64, c_i64, c_i64, c_dbl, c_int, c_p]),
            "oracle_dsposv": (c_i64, [c_i64, c_i64, c_p, c_i64, c_p, c_i64, c_p, c_i64,
                                      c_i64, c_i64, c_dbl, c_int, c_p]),
            "oracle_dsgesv_gmres": (c_i64, [c_i64, c_p, c_i64, c_p, c_p, c_p, c_i64, c_i64,
                                            c_i64, c_i64, c_dbl, c_int, c_p]),
            "oracle_dsposv_gmres": (c_i64, [c_i64, c_p, c_i64, c_p, c_p, c_i64, c_i64, c_dbl,
                                            c_int, c_p]),
            "oracle_norm_inf": (c_dbl, [c_i64, c_i64, c_p, c_i64]),
            "oracle_norm_inf_sy": (c_dbl, [c_i64, c_p, c_i64]),
            "oracle_residual": (None, [c_i64, c_i64, c_p, c_i64, c_int, c_p, c_i64, c_p, c_i64,
                                       c_p, c_i64]),
            "oracle_rotg": (None, [c_p, c_p, c_p, c_p]),
            "oracle_rbt_weights": (None, [c_i64, c_i64, c_p]),
            "oracle_rbt_half": (c_i64, [c_i64, c_i64, c_i64, c_i64]),
            "oracle_gerbt": (None, [c_i64, c_i64, c_i64, c_p, c_p, c_p, c_i64]),
            "oracle_gerbt_rhs": (None, [c_int, c_i64, c_i64, c_i64, c_i64, c_p, c_p, c_i64]),
            "oracle_dgesv_rbt": (c_i64, [c_i64, c_i64, c_p, c_i64, c_p, c_i64, c_p, c_i64, c_i64,
                                         c_i64, c_i64, c_i64, c_dbl, c_int, c_p]),
            "oracle_set_blas": (c_int, [ctypes.c_char_p]),
            "oracle_set_update_precision": (c_int, [c_int, c_int, c_i64]),
            "oracle_sgemm_update": (None, [c_i64, c_i64, c_i64, c_p, c_i64, c_p, c_i64, c_p, c_i64]),
            "oracle_e4m3_decode": (c_dbl, [c_int]),
            "oracle_e4m3_round": (c_dbl, [c_dbl]),
            "oracle_e4m3_encode": (c_int, [c_dbl]),
            "oracle_e4m3_scale_exp": (c_int, [c_dbl]),
            "oracle_bf16_rne": (ctypes.c_float, [ctypes.c_float]),
            "oracle_gmres_first_cycle": (c_i64, [c_i64, c_p, c_i64, c_int, c_p, c_p, c_i64, c_i64, c_int,
                                                 c_p, c_p, c_p, c_p]),
            "oracle_num_threads": (c_int, []),
            "oracle_set_num_threads": (None, [c_int]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _fmat(a, dtype=np.float64):
    a = np.asarray(a, dtype=dtype)
    return np.asfortranarray(a)


DIST_UNIFORM, DIST_RANDS, DIST_NORMAL = 1, 2, 3


def generate(dist, seed, m, n, ioff=0, joff=0):
    """matgen entry (i, j) = f(seed, i+ioff, j+joff); returns an m x n Fortran array."""
    A = np.empty((m, n), dtype=np.float64, order="F")
    lib().oracle_generate(dist, seed, m, n, ioff, joff, _ptr(A), max(m, 1))
    return A


def rands(seed, m, n):
    return generate(DIST_RANDS, seed, m, n)


def randn(seed, m, n):
    return generate(DIST_NORMAL, seed, m, n)


def philox(s0, s1, seed):
    out = np.zeros(2, dtype=np.uint64)
    lib().oracle_philox(s0, s1, seed, _ptr(out))
    return int(out[0]), int(out[1])


def getrf(A, nb=256, ib=16, nthreads=1, threshold=1.0, dtype=np.float64):
    """Tile LU; returns (LU, ipiv 1-based int64, info)."""
    A = _fmat(A, dtype).copy(order="F")
    m, n = A.shape
    ipiv = np.zeros(min(m, n), dtype=np.int64)
    f = lib().oracle_dgetrf if dtype == np.float64 else lib().oracle_sgetrf
    info = f(m, n, _ptr(A), max(m, 1), nb, ib, nthreads, threshold, _ptr(ipiv))
    return A, ipiv, int(info)


def getrs(LU, ipiv, B, nb=256, trans=False):
    dtype = LU.dtype
    LU = _fmat(LU, dtype)
    B = _fmat(B, dtype).copy(order="F")
    if B.ndim == 1:
        B = B.reshape(-1, 1, order="F")
    n = LU.shape[0]
    f = lib().oracle_dgetrs if dtype == np.float64 else lib().oracle_sgetrs
    f(n, B.shape[1], _ptr(LU), max(n, 1), _ptr(np.ascontiguousarray(ipiv, dtype=np.int64)),
      _ptr(B), max(n, 1), nb, int(bool(trans)))
    return B


def gesv(A, B, nb=256, ib=16, nthreads=1):
    A = _fmat(A).copy(order="F")
    B = _fmat(B).copy(order="F")
    if B.ndim == 1:
        B = B.reshape(-1, 1, order="F")
    n = A.shape[0]
    ipiv = np.zeros(n, dtype=np.int64)
    info = lib().oracle_dgesv(n, B.shape[1], _ptr(A), max(n, 1), _ptr(ipiv), _ptr(B),
                              max(n, 1), nb, ib, nthreads)
    return A, ipiv, B, int(info)


def potrf(A, nb=256, dtype=np.float64):
    """Lower tile Cholesky; returns (L in the lower triangle, info)."""
    A = _fmat(A, dtype).copy(order="F")
    n = A.shape[0]
    f = lib().oracle_dpotrf if dtype == np.float64 else lib().oracle_spotrf
    info = f(n, _ptr(A), max(n, 1), nb)
    return A, int(info)


def potrs(L, B, nb=256):
    L = _fmat(L)
    B = _fmat(B).copy(order="F")
    if B.ndim == 1:
        B = B.reshape(-1, 1, order="F")
    n = L.shape[0]
    lib().oracle_dpotrs(n, B.shape[1], _ptr(L), max(n, 1), _ptr(B), max(n, 1), nb)
    return B


def posv(A, B, nb=256):
    A = _fmat(A).copy(order="F")
    B = _fmat(B).copy(order="F")
    if B.ndim == 1:
        B = B.reshape(-1, 1, order="F")
    n = A.shape[0]
    info = lib().oracle_dposv(n, B.shape[1], _ptr(A), max(n, 1), _ptr(B), max(n, 1), nb)
    return A, B, int(info)


def gesv_mixed(A, B, nb=256, ib=16, nthreads=1, itermax=30, tol=0.0, use_fallback=True):
    """Classic IR with a float factor; returns (A_out, ipiv, X, info, iter)."""
    A = _fmat(A).copy(order="F")
    B = _fmat(B)
    if B.ndim == 1:
        B = B.reshape(-1, 1, order="F")
    n, nrhs = B.shape
    X = np.zeros((n, nrhs), order="F")
    ipiv = np.zeros(n, dtype=np.int64)
    it = ctypes.c_int(0)
    info = lib().oracle_dsgesv(n, nrhs, _ptr(A), n, _ptr(ipiv), _ptr(B), n, _ptr(X), n, nb, ib,
                               nthreads, itermax, tol, int(use_fallback), ctypes.byref(it))
    return A, ipiv, X, int(info), it.value


def posv_mixed(A, B, nb=256, itermax=30, tol=0.0, use_fallback=True):
    A = _fmat(A).copy(order="F")
    B = _fmat(B)
    if B.ndim == 1:
        B = B.reshape(-1, 1, order="F")
    n, nrhs = B.shape
    X = np.zeros((n, nrhs), order="F")
    it = ctypes.c_int(0)
    info = lib().oracle_dsposv(n, nrhs, _ptr(A), n, _ptr(B), n, _ptr(X), n, nb, itermax, tol,
                               int(use_fallback), ctypes.byref(it))
    return A, X, int(info), it.value


def gesv_mixed_gmres(A, b, nb=256, ib=16, nthreads=1, itermax=30, tol=0.0, use_fallback=True):
    A = _fmat(A).copy(order="F")
    b = np.ascontiguousarray(np.asarray(b, dtype=np.float64).reshape(-1))
    n = A.shape[0]
    x = np.zeros(n)
    ipiv = np.zeros(n, dtype=np.int64)
    it = ctypes.c_int(0)
    info = lib().oracle_dsgesv_gmres(n, _ptr(A), n, _ptr(ipiv), _ptr(b), _ptr(x), nb, ib,
                                     nthreads, itermax, tol, int(use_fallback), ctypes.byref(it))
    return A, ipiv, x, int(info), it.value


def posv_mixed_gmres(A, b, nb=256, itermax=30, tol=0.0, use_fallback=True):
    A = _fmat(A).copy(order="F")
    b = np.ascontiguousarray(np.asarray(b, dtype=np.float64).reshape(-1))
    n = A.shape[0]
    x = np.zeros(n)
    it = ctypes.c_int(0)
    info = lib().oracle_dsposv_gmres(n, _ptr(A), n, _ptr(b), _ptr(x), nb, itermax, tol,
                                     int(use_fallback), ctypes.byref(it))
    return A, x, int(info), it.value


def rbt_weights(seed, n):
    """Butterfly weights exp(r/20), r = rands(seed) column 0 (internal_rbt_generate.cc:45-51)."""
    u = np.empty(n)
    lib().oracle_rbt_weights(seed, n, _ptr(u))
    return u


def rbt_half(n, nb, depth, k):
    return int(lib().oracle_rbt_half(n, nb, depth, k))


def gerbt(u, A, v, depth, nb):
    """A := U^T A V (src/gerbt.cc:124-217); returns a new Fortran array."""
    A = _fmat(A).copy(order="F")
    n = A.shape[0]
    u = np.ascontiguousarray(u, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    lib().oracle_gerbt(n, nb, depth, _ptr(u), _ptr(v), _ptr(A), max(n, 1))
    return A


def gerbt_rhs(u, B, depth, nb, trans):
    """B := U^T B (trans) or U B (src/gerbt.cc:251-323); returns a new Fortran array."""
    B = _fmat(B).copy(order="F")
    if B.ndim == 1:
        B = B.reshape(-1, 1, order="F")
    n, nrhs = B.shape
    u = np.ascontiguousarray(u, dtype=np.float64)
    lib().oracle_gerbt_rhs(int(bool(trans)), n, nrhs, nb, depth, _ptr(u), _ptr(B), max(n, 1))
    return B


def gesv_rbt(A, B, nb=256, ib=16, depth=2, itermax=30, tol=0.0, use_fallback=True):
    """src/gesv_rbt.cc:94-236; returns (A_out, X, info, iter)."""
    A = _fmat(A).copy(order="F")
    B = _fmat(B)
    if B.ndim == 1:
        B = B.reshape(-1, 1, order="F")
    n, nrhs = B.shape
    X = np.zeros((n, nrhs), order="F")
    it = ctypes.c_int(0)
    info = lib().oracle_dgesv_rbt(n, nrhs, _ptr(A), n, _ptr(B), n, _ptr(X), n, nb, ib, depth, itermax,
                                  tol, int(use_fallback), ctypes.byref(it))
    return A, X, int(info), it.value


def sigma(dist, cond, n):
    """matgen/generate_sigma.hh:61-111: 1 arith, 2 geo, 3 cluster0, 4 cluster1."""
    i = np.arange(n, dtype=np.float64)
    t = i / (n - 1) if n > 1 else np.zeros(n)
    if dist == 1:
        return 1.0 - t * (1.0 - 1.0 / cond)
    if dist == 2:
        return cond ** (-t)
    if dist == 3:
        s = np.full(n, 1.0 / cond)
        s[0] = 1.0
        return s
    if dist == 4:
        s = np.ones(n)
        s[-1] = 1.0 / cond
        return s
    raise ValueError("unknown sigma distribution %r" % dist)


def reflector_product(seed, n):
    """Q = H_1 H_2 ... H_n, H_i = I - 2 v_i v_i^T / (v_i^T v_i), v_i = column i
    of matgen randn(seed) (n x n), applied one reflector at a time."""
    V = randn(seed, n, n)
    Q = np.eye(n)
    for i in range(n - 1, -1, -1):          # Q := H_i Q, last reflector first
        v = V[:, i]
        Q -= np.outer(v, (2.0 / (v @ v)) * (v @ Q))
    return Q


def generate_svd(sigma_dist, cond, seed, n, sym=False):
    """A = U diag(sigma) V^T with U = reflector_product(seed + 1, n) and
    V = reflector_product(seed, n); sym: A = V diag(sigma) V^T."""
    sg = sigma(sigma_dist, cond, n)
    V = reflector_product(seed, n)
    U = V if sym else reflector_product(seed + 1, n)
    return (U * sg) @ V.T


# ---- selector MLP (float64 restatement of slate_amd/csrc/mlp.hip) ----------
MLP_F, MLP_H, MLP_K = 24, 64, 8


def mlp_unpack(P):
    """Views of the parameter array: W1 (F x H), b1, W2 (H x H), b2, W3 (H x K), b3."""
    F, H, K = MLP_F, MLP_H, MLP_K
    o = 0
    out = []
    for shape in ((F, H), (H,), (H, H), (H,), (H, K), (K,)):
        size = int(np.prod(shape))
        out.append(P[o:o + size].reshape(shape))
        o += size
    assert o == P.size
    return out


def mlp_num_params():
    return MLP_F * MLP_H + MLP_H + MLP_H * MLP_H + MLP_H + MLP_H * MLP_K + MLP_K


def mlp_forward(P, X):
    W1, b1, W2, b2, W3, b3 = mlp_unpack(np.asarray(P, dtype=np.float64))
    h1 = np.maximum(X @ W1 + b1, 0)
    h2 = np.maximum(h1 @ W2 + b2, 0)
    z = h2 @ W3 + b3
    z = z - z.max(axis=1, keepdims=True)
    p = np.exp(z)
    return p / p.sum(axis=1, keepdims=True)


def mlp_loss_grad(P, X, y):
    """Mean softmax cross-entropy and its gradient w.r.t. the flat parameters."""
    P = np.asarray(P, dtype=np.float64)
    X = np.asarray(X, dtype=np.float64)
    W1, b1, W2, b2, W3, b3 = mlp_unpack(P)
    B = X.shape[0]
    a1 = X @ W1 + b1
    h1 = np.maximum(a1, 0)
    a2 = h1 @ W2 + b2
    h2 = np.maximum(a2, 0)
    z = h2 @ W3 + b3
    z = z - z.max(axis=1, keepdims=True)
    p = np.exp(z)
    p /= p.sum(axis=1, keepdims=True)
    loss = -np.mean(np.log(np.maximum(p[np.arange(B), y], 1e-300)))
    d3 = p.copy()
    d3[np.arange(B), y] -= 1.0
    d3 /= B
    gW3 = h2.T @ d3
    gb3 = d3.sum(axis=0)
    d2 = (d3 @ W3.T) * (a2 > 0)
    gW2 = h1.T @ d2
    gb2 = d2.sum(axis=0)
    d1 = (d2 @ W2.T) * (a1 > 0)
    gW1 = X.T @ d1
    gb1 = d1.sum(axis=0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g


def adam_step(P, g, m, v, lr, t, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    P = P - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return P, m, v


def norm_inf(A):
    A = _fmat(A)
    return lib().oracle_norm_inf(A.shape[0], A.shape[1], _ptr(A), max(A.shape[0], 1))


def rotg(a, b):
    va, vb, vc, vs = (ctypes.c_double(a), ctypes.c_double(b), ctypes.c_double(0), ctypes.c_double(0))
    lib().oracle_rotg(ctypes.byref(va), ctypes.byref(vb), ctypes.byref(vc), ctypes.byref(vs))
    return va.value, vb.value, vc.value, vs.value


def scaled_residual(A, X, B):
    """Tester check of test/test_gesv.cc:330-377: ||B-AX||_1 / (n ||A||_1 ||X||_1), in units of eps."""
    A = np.asarray(A, dtype=np.float64)
    X = np.asarray(X, dtype=np.float64).reshape(A.shape[0], -1)
    B = np.asarray(B, dtype=np.float64).reshape(A.shape[0], -1)
    n = A.shape[0]
    R = B - A @ X
    r1 = np.abs(R).sum(axis=0).max()
    a1 = np.abs(A).sum(axis=0).max()
    x1 = np.abs(X).sum(axis=0).max()
    return float(r1 / (n * a1 * x1) / np.finfo(np.float64).eps)


def ipiv_from_lapack(ipiv_lapack):
    """LAPACK 1-based int32 ipiv -> int64 array (same convention as the oracle)."""
    return np.asarray(ipiv_lapack, dtype=np.int64) + 0


# ---- low-precision update emulation, FP8 e4m3, tile BLAS ---------------------
UPD_FLOAT, UPD_BF16, UPD_FP8 = 0, 1, 2


class update_precision:
    """Context manager: the float factor's trailing updates use split-BF16
    (kind UPD_BF16, terms 1-3) or per-tile scaled FP8 e4m3 (UPD_FP8, terms
    1-2, tiles of `qrows` rows) operands, as the GPU factor precisions do."""

    def __init__(self, kind, terms, qrows=128):
        self.args = (int(kind), int(terms), int(qrows))

    def __enter__(self):
        if lib().oracle_set_update_precision(*self.args) != 0:
            raise ValueError("bad update precision %r" % (self.args,))
        return self

    def __exit__(self, *exc):
        lib().oracle_set_update_precision(0, 0, 128)


def sgemm_update(A, B, C):
    """C - A B (float32) under the current update precision; returns a new array."""
    A = _fmat(A, np.float32)
    B = _fmat(B, np.float32)
    C = _fmat(C, np.float32).copy(order="F")
    m, k = A.shape
    n = B.shape[1]
    lib().oracle_sgemm_update(m, n, k, _ptr(A), m, _ptr(B), k, _ptr(C), m)
    return C


def e4m3_decode(b):
    return lib().oracle_e4m3_decode(int(b))


def e4m3_round(x):
    return lib().oracle_e4m3_round(float(x))


def e4m3_encode(x):
    return lib().oracle_e4m3_encode(float(x))


def e4m3_scale_exp(amax):
    return lib().oracle_e4m3_scale_exp(float(amax))


def bf16_rne(x):
    return lib().oracle_bf16_rne(float(x))


def openblas_path():
    """scipy's bundled LP64 OpenBLAS (exports scipy_dgemm_ ...), or None."""
    import glob
    try:
        import scipy
    except ImportError:
        return None
    d = os.path.join(os.path.dirname(os.path.dirname(scipy.__file__)), "scipy.libs")
    hits = sorted(glob.glob(os.path.join(d, "libscipy_openblas-*.so")))
    return hits[0] if hits else None


def set_blas(path):
    """Tile gemm/trsm through the OpenBLAS at `path` (None: plain loops)."""
    rc = lib().oracle_set_blas(path.encode() if path else None)
    if rc != 0:
        raise RuntimeError("oracle_set_blas(%r) failed" % (path,))


class tile_blas:
    """Context manager: tile gemm/trsm through scipy's bundled OpenBLAS, as
    SLATE's HostTask tiles call a vendor BLAS (src/internal/internal_gemm.cc).
    A BLAS sgemm accumulates each product sum before adding it to C, which is
    how the GPU's float updates round too; the plain loops update C one
    rank-1 term at a time, which is more accurate on some inputs (a
    rank-one-dominated matrix loses far less to cancellation)."""

    def __enter__(self):
        path = openblas_path()
        if path is None:
            raise RuntimeError("tile_blas: scipy's OpenBLAS not found")
        set_blas(path)
        return self

    def __exit__(self, *exc):
        set_blas(None)


def gmres_first_cycle(A, b, nb=256, itermax=30, sym=False, cgs_passes=2):
    """First restart cycle of the oracle's gesv_mixed_gmres (posv_mixed_gmres
    with sym): returns (V n x (j+1), W n x j, Hbar (j+1) x j, |S| history).
    cgs_passes=1 plants a one-pass classical Gram-Schmidt."""
    A = _fmat(A)
    n = A.shape[0]
    b = np.ascontiguousarray(np.asarray(b, dtype=np.float64).reshape(n))
    X = np.zeros(n)
    R1 = min(30, itermax, min(nb, n) - 1) + 1
    V = np.zeros((n, R1), order="F")
    W = np.zeros((n, R1), order="F")
    H = np.zeros((R1, R1), order="F")
    S = np.zeros(R1)
    j = lib().oracle_gmres_first_cycle(n, _ptr(A), n, int(sym), _ptr(b), _ptr(X), nb, itermax,
                                       int(cgs_passes), _ptr(V), _ptr(W), _ptr(H), _ptr(S))
    return V[:, :j + 1], W[:, :j], H[:j + 1, :j], S[:j]
