"""slate_amd — MI355X-native single-GPU factor-and-solve (getrf/getrs/gesv,
potrf/potrs/posv, gesv_mixed, gesv_mixed_gmres, posv_mixed, posv_mixed_gmres).

Python binding over the C ABI declared in ``include/slate_amd.h`` and
implemented by ``slate_amd/libslate_amd.so`` (hand-written HIP kernels for
gfx950).  Device memory comes from torch tensors; matrices are column-major
(``A.stride(0) == 1``), exactly the layout ``slate::Matrix::fromLAPACK`` and
``fromDevices`` describe.  Pivots are LAPACK-style 1-based int64 arrays.

The entry points follow SLATE's driver signatures (include/slate/slate.hh:
503-745): they factor in place, return ``info`` (0, or the 1-based index of a
zero pivot / non-SPD minor), and for the mixed drivers also ``iter``.  Errors
raise ``SlateError``; there is no CPU fallback — if the library is missing or
no GPU is present, calls fail loudly.
"""
import atexit
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libslate_amd.so")

PREC_FP32, PREC_BF16, PREC_BF16X2, PREC_FP64, PREC_BF16X3, PREC_FP8, PREC_FP8X2 = 0, 1, 2, 3, 4, 5, 6
LOWER, UPPER = 0, 1
NOTRANS, TRANS = 0, 1
DIST_UNIFORM, DIST_RANDS, DIST_RANDN = 1, 2, 3
SIGMA_ARITH, SIGMA_GEO, SIGMA_CLUSTER0, SIGMA_CLUSTER1 = 1, 2, 3, 4
LU_PARTIALPIV, LU_NOPIV, LU_CALU = 0, 1, 2

# Public C symbols (the C-ABI test checks that the library exports each one).
C_SYMBOLS = [
    "slate_amd_init", "slate_amd_finalize", "slate_amd_last_error", "slate_amd_options_init",
    "slate_amd_last_time_ms",
    "slate_amd_dgetrf", "slate_amd_sgetrf", "slate_amd_dgetrs", "slate_amd_sgetrs", "slate_amd_dgetrs_noalloc",
    "slate_amd_dgesv", "slate_amd_dpotrf", "slate_amd_spotrf", "slate_amd_dpotrs",
    "slate_amd_dposv", "slate_amd_dsgesv", "slate_amd_dsgesv_gmres", "slate_amd_dsposv",
    "slate_amd_dsposv_gmres", "slate_amd_dnorm_inf", "slate_amd_generate", "slate_amd_generate_svd",
    "slate_amd_dgecondest", "slate_amd_dpocondest", "slate_amd_mlp_num_params",
    "slate_amd_selector_features", "slate_amd_mlp_forward", "slate_amd_mlp_train_step",
    "slate_amd_dsgesv_auto", "slate_amd_tune_set_nb", "slate_amd_tune_get_nb", "slate_amd_dgesv_batched",
    "slate_amd_dgemm", "slate_amd_sgemm", "slate_amd_sgemm_bf16", "slate_amd_sgemm_fp8", "slate_amd_profile_enable",
    "slate_amd_profile_read", "slate_amd_profile_read_kernel", "slate_amd_dgesv_rbt", "slate_amd_drbt_generate", "slate_amd_dgerbt",
    "slate_amd_dgerbt_rhs", "slate_amd_dgetrf_tiles", "slate_amd_dgetrs_tiles", "slate_amd_dgesv_tiles",
    "slate_amd_dpotrf_tiles", "slate_amd_dpotrs_tiles", "slate_amd_dposv_tiles", "slate_amd_dsgesv_tiles",
    "slate_amd_dsgesv_gmres_tiles", "slate_amd_dsposv_tiles", "slate_amd_dsposv_gmres_tiles",
    "slate_amd_graph_stats", "slate_amd_dgesv_batched_auto", "slate_amd_batched_graph_stats",
    "slate_amd_tune_set_batched", "slate_amd_tune_get_batched",
]


class SlateError(RuntimeError):
    pass


class TileDesc(ctypes.Structure):
    """Mirror of ``slate_amd_tiles`` (SLATE tile storage on one device)."""
    _fields_ = [
        ("m", ctypes.c_int64), ("n", ctypes.c_int64), ("mb", ctypes.c_int64), ("nb", ctypes.c_int64),
        ("tiles", ctypes.POINTER(ctypes.c_void_p)), ("lds", ctypes.POINTER(ctypes.c_int64)),
    ]


class Options(ctypes.Structure):
    """Mirror of ``slate_amd_options`` (the slate::Option values of the path)."""
    _fields_ = [
        ("block_size", ctypes.c_int64),
        ("inner_blocking", ctypes.c_int64),
        ("lookahead", ctypes.c_int64),
        ("pivot_threshold", ctypes.c_double),
        ("max_iterations", ctypes.c_int64),
        ("tolerance", ctypes.c_double),
        ("use_fallback", ctypes.c_int32),
        ("factor_precision", ctypes.c_int32),
        ("method_lu", ctypes.c_int32),
        ("depth", ctypes.c_int32),
        ("use_graph", ctypes.c_int32),
    ]

    def __init__(self, **kw):
        super().__init__()
        lib().slate_amd_options_init(ctypes.byref(self))
        for k, v in kw.items():
            if not hasattr(self, k):
                raise TypeError("unknown option %r" % k)
            setattr(self, k, v)


_LIB = None
_CTX = {}


def lib():
    """Load libslate_amd.so (built by slate_amd/build.py); raise if absent."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise SlateError("libslate_amd.so not built (run `python -m slate_amd.build`)")
        L = ctypes.CDLL(LIB_PATH)
        vp, i64, i32, dbl = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double
        pi64 = ctypes.POINTER(ctypes.c_int64)
        pint = ctypes.POINTER(ctypes.c_int)
        po = ctypes.POINTER(Options)
        ptd = ctypes.POINTER(TileDesc)
        sig = {
            "slate_amd_init": (i32, [ctypes.POINTER(vp), i32]),
            "slate_amd_finalize": (i32, [vp]),
            "slate_amd_last_error": (ctypes.c_char_p, []),
            "slate_amd_options_init": (None, [po]),
            "slate_amd_last_time_ms": (dbl, [vp]),
            "slate_amd_dgetrf": (i64, [vp, i64, vp, i64, pi64, po]),
            "slate_amd_sgetrf": (i64, [vp, i64, vp, i64, pi64, po]),
            "slate_amd_dgetrs": (i64, [vp, i32, i64, i64, vp, i64, pi64, vp, i64, po]),
            "slate_amd_sgetrs": (i64, [vp, i32, i64, i64, vp, i64, pi64, vp, i64, po]),
            "slate_amd_dgetrs_noalloc": (i64, [vp, i64, vp, i64, pi64, vp]),
            "slate_amd_dgesv": (i64, [vp, i64, i64, vp, i64, pi64, vp, i64, po]),
            "slate_amd_dpotrf": (i64, [vp, i32, i64, vp, i64, po]),
            "slate_amd_spotrf": (i64, [vp, i32, i64, vp, i64, po]),
            "slate_amd_dpotrs": (i64, [vp, i32, i64, i64, vp, i64, vp, i64, po]),
            "slate_amd_dposv": (i64, [vp, i32, i64, i64, vp, i64, vp, i64, po]),
            "slate_amd_dsgesv": (i64, [vp, i64, i64, vp, i64, pi64, vp, i64, vp, i64, pint, po]),
            "slate_amd_dsgesv_gmres": (i64, [vp, i64, i64, vp, i64, pi64, vp, i64, vp, i64, pint, po]),
            "slate_amd_dsposv": (i64, [vp, i32, i64, i64, vp, i64, vp, i64, vp, i64, pint, po]),
            "slate_amd_dsposv_gmres": (i64, [vp, i32, i64, i64, vp, i64, vp, i64, vp, i64, pint, po]),
            "slate_amd_dgesv_rbt": (i64, [vp, i64, i64, vp, i64, vp, i64, vp, i64, pint, po]),
            "slate_amd_drbt_generate": (i32, [vp, i64, i64, vp]),
            "slate_amd_dgerbt": (i32, [vp, i64, i32, i64, vp, vp, vp, i64]),
            "slate_amd_dgerbt_rhs": (i32, [vp, i32, i64, i64, i32, i64, vp, vp, i64]),
            "slate_amd_dnorm_inf": (i32, [vp, i32, i64, vp, i64, ctypes.POINTER(dbl)]),
            "slate_amd_generate": (i32, [vp, i32, i64, i64, i64, i64, i64, vp, i64]),
            "slate_amd_generate_svd": (i32, [vp, i32, dbl, i64, i32, i64, vp, i64]),
            "slate_amd_dgecondest": (i64, [vp, i32, i64, vp, i64, pi64, dbl, ctypes.POINTER(dbl)]),
            "slate_amd_dpocondest": (i64, [vp, i32, i64, vp, i64, dbl, ctypes.POINTER(dbl)]),
            "slate_amd_mlp_num_params": (i32, []),
            "slate_amd_tune_set_nb": (i32, [vp, i32, i32, pi64, pi64]),
            "slate_amd_tune_get_nb": (i64, [vp, i32, i64]),
            "slate_amd_dgesv_batched": (i64, [vp, i32, vp, vp, vp, vp, vp, i32, vp, vp, vp, po]),
            "slate_amd_dgesv_batched_auto": (i64, [vp, i32, vp, vp, vp, vp, vp, i32, vp, vp, vp, vp, vp, vp, po]),
            "slate_amd_batched_graph_stats": (i32, [vp, pi64, pi64]),
            "slate_amd_tune_set_batched": (i32, [vp, i32, i32, i32]),
            "slate_amd_tune_get_batched": (i32, [vp, pint, pint, pint, pint]),
            "slate_amd_selector_features": (i64, [vp, i64, vp, i64, vp, vp, po, vp]),
            "slate_amd_mlp_forward": (i32, [vp, vp, vp, i32, vp]),
            "slate_amd_mlp_train_step": (i32, [vp, vp, vp, vp, vp, vp, vp, i32, ctypes.c_float, i32,
                                               ctypes.POINTER(ctypes.c_float)]),
            "slate_amd_dsgesv_auto": (i64, [vp, i64, vp, i64, pi64, vp, vp, vp, i32, pint, pint, vp, po]),
            "slate_amd_dgetrf_tiles": (i64, [vp, ptd, pi64, po]),
            "slate_amd_dgetrs_tiles": (i64, [vp, i32, ptd, pi64, ptd, po]),
            "slate_amd_dgesv_tiles": (i64, [vp, ptd, pi64, ptd, po]),
            "slate_amd_dpotrf_tiles": (i64, [vp, ptd, po]),
            "slate_amd_dpotrs_tiles": (i64, [vp, ptd, ptd, po]),
            "slate_amd_dposv_tiles": (i64, [vp, ptd, ptd, po]),
            "slate_amd_dsgesv_tiles": (i64, [vp, ptd, pi64, ptd, ptd, pint, po]),
            "slate_amd_dsgesv_gmres_tiles": (i64, [vp, ptd, pi64, ptd, ptd, pint, po]),
            "slate_amd_dsposv_tiles": (i64, [vp, ptd, ptd, ptd, pint, po]),
            "slate_amd_dsposv_gmres_tiles": (i64, [vp, ptd, ptd, ptd, pint, po]),
            "slate_amd_profile_enable": (i32, [vp, i32]),
            "slate_amd_graph_stats": (i32, [vp, pi64, pi64, pi64]),
            "slate_amd_profile_read": (i32, [vp, ctypes.POINTER(dbl), ctypes.POINTER(dbl), pi64]),
            "slate_amd_profile_read_kernel": (i32, [vp, i32, ctypes.POINTER(dbl), ctypes.POINTER(dbl), pi64]),
            "slate_amd_dgemm": (i32, [vp, i64, i64, i64, dbl, vp, i64, vp, i64, dbl, vp, i64]),
            "slate_amd_sgemm": (i32, [vp, i64, i64, i64, ctypes.c_float, vp, i64, vp, i64,
                                      ctypes.c_float, vp, i64]),
            "slate_amd_sgemm_bf16": (i32, [vp, i32, i64, i64, i64, ctypes.c_float, vp, i64, vp, i64,
                                           vp, i64]),
            "slate_amd_sgemm_fp8": (i32, [vp, i32, i64, i64, i64, ctypes.c_float, vp, i64, vp, i64,
                                          vp, i64]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def context(device=0):
    """Per-device library context (streams + workspaces), created on first use."""
    if device not in _CTX:
        h = ctypes.c_void_p()
        rc = lib().slate_amd_init(ctypes.byref(h), device)
        if rc != 0:
            raise SlateError("slate_amd_init: " + last_error())
        _CTX[device] = h
        from . import autotune
        autotune.load(device=device)       # measured block sizes, when tuned on this machine
    return _CTX[device]


@atexit.register
def release_contexts():
    """slate_amd_finalize for every context this process created: waits for
    the device, then frees the workspaces, events and streams (including the
    CU-masked update stream) while the HIP runtime is still loaded."""
    while _CTX:
        _, h = _CTX.popitem()
        lib().slate_amd_finalize(h)


def last_error():
    return lib().slate_amd_last_error().decode()


def last_time_ms(device=0):
    return lib().slate_amd_last_time_ms(context(device))


def _check(rc):
    if rc < 0:
        raise SlateError("slate_amd error %d: %s" % (rc, last_error()))
    return int(rc)


def _mat(A, dtypes=("float64",)):
    """Return (ptr, lda, device) for a column-major CUDA matrix tensor."""
    import torch
    if not isinstance(A, torch.Tensor) or not A.is_cuda:
        raise SlateError("expected a CUDA tensor")
    if str(A.dtype).replace("torch.", "") not in dtypes:
        raise SlateError("unsupported dtype %s" % A.dtype)
    if A.dim() == 1:
        if A.shape[0] > 1 and A.stride(0) != 1:
            raise SlateError("vector must be contiguous (stride(0) == 1)")
        return A.data_ptr(), max(1, A.shape[0]), A.device.index or 0
    if A.dim() != 2:
        raise SlateError("expected a matrix or a vector")
    if A.stride(0) != 1 and A.shape[0] > 1:
        raise SlateError("matrix must be column-major (stride(0) == 1)")
    lda = A.stride(1) if A.shape[1] > 1 else max(1, A.shape[0])
    return A.data_ptr(), max(lda, 1), A.device.index or 0


def _ncols(B):
    return 1 if B.dim() == 1 else B.shape[1]


def _rows(A, *mats):
    """A must be square (n x n) and every right-hand side / solution must
    have exactly n rows (a shorter view with a large stride(1) would pass the
    C ABI's ld check and be written past its end)."""
    n = A.shape[0]
    if A.dim() != 2 or A.shape[1] != n:
        raise SlateError("expected a square matrix, got %s" % (tuple(A.shape),))
    for M in mats:
        if M.shape[0] != n:
            raise SlateError("expected %d rows, got %d" % (n, M.shape[0]))


def colmajor(t):
    """Column-major copy of a 2-D tensor (same device, same dtype)."""
    return t.t().contiguous().t()


def empty_colmajor(m, n, dtype, device="cuda", ld=None):
    import torch
    ld = ld or m
    return torch.empty((n, ld), dtype=dtype, device=device).t()[:m]


def _opts(opts):
    return ctypes.byref(opts if opts is not None else Options())


def _ipiv_out(n):
    return np.zeros(max(n, 1), dtype=np.int64)


def _p64(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


#---------------------------------------------------------------- LU
def getrf(A, opts=None):
    """slate::getrf: A = P L U in place; returns (info, ipiv 1-based)."""
    ptr, lda, dev = _mat(A, ("float64", "float32"))
    n = A.shape[0]
    if A.shape[1] != n:
        raise SlateError("getrf: square matrices only")
    ipiv = _ipiv_out(n)
    f = lib().slate_amd_dgetrf if A.dtype.itemsize == 8 else lib().slate_amd_sgetrf
    info = _check(f(context(dev), n, ptr, lda, _p64(ipiv), _opts(opts)))
    return info, ipiv[:n]


def getrs(LU, ipiv, B, trans=False, opts=None):
    """slate::getrs: B := op(A)^{-1} B using the factors of getrf."""
    pa, lda, dev = _mat(LU, ("float64", "float32"))
    pb, ldb, _ = _mat(B, ("float64", "float32"))
    _rows(LU, B)
    n = LU.shape[0]
    ip = np.ascontiguousarray(ipiv, dtype=np.int64)
    f = lib().slate_amd_dgetrs if LU.dtype.itemsize == 8 else lib().slate_amd_sgetrs
    _check(f(context(dev), TRANS if trans else NOTRANS, n, _ncols(B), pa, lda, _p64(ip), pb, ldb,
             _opts(opts)))


def gesv(A, B, opts=None):
    """slate::gesv: factor A and overwrite B with the solution; returns (info, ipiv)."""
    pa, lda, dev = _mat(A)
    pb, ldb, _ = _mat(B)
    _rows(A, B)
    n = A.shape[0]
    ipiv = _ipiv_out(n)
    info = _check(lib().slate_amd_dgesv(context(dev), n, _ncols(B), pa, lda, _p64(ipiv), pb, ldb,
                                        _opts(opts)))
    return info, ipiv[:n]


#---------------------------------------------------------------- Cholesky
def potrf(A, uplo=LOWER, opts=None):
    """slate::potrf on a HermitianMatrix stored in the `uplo` triangle; returns info."""
    pa, lda, dev = _mat(A, ("float64", "float32"))
    f = lib().slate_amd_dpotrf if A.dtype.itemsize == 8 else lib().slate_amd_spotrf
    return _check(f(context(dev), uplo, A.shape[0], pa, lda, _opts(opts)))


def potrs(L, B, uplo=LOWER, opts=None):
    pa, lda, dev = _mat(L)
    pb, ldb, _ = _mat(B)
    _rows(L, B)
    _check(lib().slate_amd_dpotrs(context(dev), uplo, L.shape[0], _ncols(B), pa, lda, pb, ldb,
                                  _opts(opts)))


def posv(A, B, uplo=LOWER, opts=None):
    pa, lda, dev = _mat(A)
    pb, ldb, _ = _mat(B)
    _rows(A, B)
    return _check(lib().slate_amd_dposv(context(dev), uplo, A.shape[0], _ncols(B), pa, lda, pb, ldb,
                                        _opts(opts)))


#---------------------------------------------------------------- mixed precision
def gesv_mixed(A, B, X, opts=None):
    """slate::gesv_mixed: returns (info, iter, ipiv); X receives the solution."""
    pa, lda, dev = _mat(A)
    pb, ldb, _ = _mat(B)
    px, ldx, _ = _mat(X)
    _rows(A, B, X)
    n = A.shape[0]
    ipiv = _ipiv_out(n)
    it = ctypes.c_int(0)
    info = _check(lib().slate_amd_dsgesv(context(dev), n, _ncols(B), pa, lda, _p64(ipiv), pb, ldb,
                                         px, ldx, ctypes.byref(it), _opts(opts)))
    return info, it.value, ipiv[:n]


def gesv_mixed_gmres(A, B, X, opts=None):
    pa, lda, dev = _mat(A)
    pb, ldb, _ = _mat(B)
    px, ldx, _ = _mat(X)
    _rows(A, B, X)
    n = A.shape[0]
    ipiv = _ipiv_out(n)
    it = ctypes.c_int(0)
    info = _check(lib().slate_amd_dsgesv_gmres(context(dev), n, _ncols(B), pa, lda, _p64(ipiv), pb,
                                               ldb, px, ldx, ctypes.byref(it), _opts(opts)))
    return info, it.value, ipiv[:n]


def posv_mixed(A, B, X, uplo=LOWER, opts=None):
    pa, lda, dev = _mat(A)
    pb, ldb, _ = _mat(B)
    px, ldx, _ = _mat(X)
    _rows(A, B, X)
    it = ctypes.c_int(0)
    info = _check(lib().slate_amd_dsposv(context(dev), uplo, A.shape[0], _ncols(B), pa, lda, pb, ldb,
                                         px, ldx, ctypes.byref(it), _opts(opts)))
    return info, it.value


def posv_mixed_gmres(A, B, X, uplo=LOWER, opts=None):
    pa, lda, dev = _mat(A)
    pb, ldb, _ = _mat(B)
    px, ldx, _ = _mat(X)
    _rows(A, B, X)
    it = ctypes.c_int(0)
    info = _check(lib().slate_amd_dsposv_gmres(context(dev), uplo, A.shape[0], _ncols(B), pa, lda,
                                               pb, ldb, px, ldx, ctypes.byref(it), _opts(opts)))
    return info, it.value


#---------------------------------------------------------------- random butterflies
def gesv_rbt(A, B, X, opts=None):
    """slate::gesv_rbt (src/gesv_rbt.cc): returns (info, iter); X receives the solution."""
    pa, lda, dev = _mat(A)
    pb, ldb, _ = _mat(B)
    px, ldx, _ = _mat(X)
    _rows(A, B, X)
    it = ctypes.c_int(0)
    info = _check(lib().slate_amd_dgesv_rbt(context(dev), A.shape[0], _ncols(B), pa, lda, pb, ldb,
                                            px, ldx, ctypes.byref(it), _opts(opts)))
    return info, it.value


def rbt_generate(n, seed, device=None):
    """Butterfly weights exp(r/20), r = matgen rands (seed, i, 0): a float64 CUDA vector."""
    import torch
    u = torch.empty(n, dtype=torch.float64, device=device or "cuda")
    _check(lib().slate_amd_drbt_generate(context(u.device.index or 0), n, seed, u.data_ptr()))
    return u


def gerbt(u, A, v, depth, nb):
    """slate::gerbt(U, A, V): A := U^T A V in place."""
    pa, lda, dev = _mat(A)
    pu, pv = _weights(u, A.shape[0]), _weights(v, A.shape[0])
    _check(lib().slate_amd_dgerbt(context(dev), A.shape[0], depth, nb, pu, pv, pa, lda))


def gerbt_rhs(u, B, depth, nb, trans):
    """slate::gerbt(U, B): B := U^T B (trans) or U B (not trans) in place."""
    pb, ldb, dev = _mat(B)
    _check(lib().slate_amd_dgerbt_rhs(context(dev), TRANS if trans else NOTRANS, B.shape[0], _ncols(B),
                                      depth, nb, _weights(u, B.shape[0]), pb, ldb))


def _weights(u, n):
    if u.dim() != 1 or u.shape[0] < n or u.stride(0) != 1:
        raise SlateError("butterfly weights must be a contiguous vector of length >= n")
    return _mat(u)[0]


#---------------------------------------------------------------- utilities
def gemm(alpha, A, B, beta, C):
    """slate::gemm: C = alpha A B + beta C (column-major CUDA matrices)."""
    pa, lda, dev = _mat(A, ("float64", "float32"))
    pb, ldb, _ = _mat(B, ("float64", "float32"))
    pc, ldc, _ = _mat(C, ("float64", "float32"))
    f = lib().slate_amd_dgemm if C.dtype.itemsize == 8 else lib().slate_amd_sgemm
    _check(f(context(dev), C.shape[0], _ncols(C), A.shape[1], alpha, pa, lda, pb, ldb, beta, pc, ldc))


def gemm_bf16(terms, alpha, A, B, C):
    """C += alpha A B (float32, column-major) on the split-BF16 MFMA kernel
    with `terms` (1, 2 or 3) bf16 terms per operand."""
    pa, lda, dev = _mat(A, ("float32",))
    pb, ldb, _ = _mat(B, ("float32",))
    pc, ldc, _ = _mat(C, ("float32",))
    _check(lib().slate_amd_sgemm_bf16(context(dev), int(terms), C.shape[0], _ncols(C), A.shape[1],
                                      alpha, pa, lda, pb, ldb, pc, ldc))


def gemm_fp8(terms, alpha, A, B, C):
    """C += alpha A B (float32, column-major) on the e4m3 MFMA kernel: A
    quantized per 128-row tile, B per 128-column tile (power-of-two scales),
    `terms` (1 or 2) planes per operand."""
    pa, lda, dev = _mat(A, ("float32",))
    pb, ldb, _ = _mat(B, ("float32",))
    pc, ldc, _ = _mat(C, ("float32",))
    _check(lib().slate_amd_sgemm_fp8(context(dev), int(terms), C.shape[0], _ncols(C), A.shape[1],
                                     alpha, pa, lda, pb, ldb, pc, ldc))


def profile_enable(on=True, device=0):
    """Start (or stop) timing the trailing-update GEMMs; resets the totals."""
    _check(lib().slate_amd_profile_enable(context(device), int(on)))


def graph_stats(device=0):
    """(captured, failed, replays) of the use_graph LU sequences."""
    a, b, r = ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_int64(0)
    _check(lib().slate_amd_graph_stats(context(device), ctypes.byref(a), ctypes.byref(b), ctypes.byref(r)))
    return a.value, b.value, r.value


def profile_read(device=0):
    """(gemm_ms, gemm_flops, launches) accumulated since profile_enable."""
    ms, fl, n = ctypes.c_double(0), ctypes.c_double(0), ctypes.c_int64(0)
    _check(lib().slate_amd_profile_read(context(device), ctypes.byref(ms), ctypes.byref(fl),
                                        ctypes.byref(n)))
    return ms.value, fl.value, n.value


PROF_OTHER, PROF_LT, PROF_SP = 0, 1, 2


def profile_read_kernel(kernel, device=0):
    """(gemm_ms, gemm_flops, launches) of one trailing-update kernel
    (PROF_LT: 256 x 256 tiles over bf16 planes, PROF_SP: 128 x 128 split at
    staging, PROF_OTHER: the rest) accumulated since profile_enable."""
    ms, fl, ln = ctypes.c_double(0), ctypes.c_double(0), ctypes.c_int64(0)
    _check(lib().slate_amd_profile_read_kernel(context(device), int(kernel), ctypes.byref(ms), ctypes.byref(fl),
                                               ctypes.byref(ln)))
    return ms.value, fl.value, ln.value


def norm_inf(A, sym=False):
    pa, lda, dev = _mat(A)
    out = ctypes.c_double(0)
    _check(lib().slate_amd_dnorm_inf(context(dev), int(sym), A.shape[0], pa, lda, ctypes.byref(out)))
    return out.value


# ---- learned selection ------------------------------------------------------
SELECTOR_FEATURES = 24
# (factor precision, refinement) of each selector choice; the probe factor
# (BF16X2) is kept by choices 0 and 1
CHOICES = ("BF16X2 + IR", "BF16X2 + GMRES-IR", "BF16X3 + IR", "BF16X3 + GMRES-IR", "FP32 + IR",
           "FP32 + GMRES-IR", "FP8X2 + GMRES-IR", "FP64")
CHOICE_PREC = (PREC_BF16X2, PREC_BF16X2, PREC_BF16X3, PREC_BF16X3, PREC_FP32, PREC_FP32, PREC_FP8X2, PREC_FP64)
CHOICE_GMRES = (False, True, False, True, False, True, True, False)
SELECTOR_PROBE = PREC_BF16X2
DEFAULT_CHOICE = 4            # SLATE's gesv_mixed: float factor + classic IR


def mlp_num_params():
    return int(lib().slate_amd_mlp_num_params())


def selector_features(A, B, opts=None):
    """The selector's 24 features of the system A x = B (column-major float64
    CUDA A, one right-hand side; factors A in float internally, A and B are
    not modified).  Returns (features, info)."""
    import torch
    pa, lda, dev = _mat(A, ("float64",))
    pb, _, _ = _mat(B, ("float64",))
    _rows(A, B)
    X = torch.empty(A.shape[0], dtype=torch.float64, device=A.device)
    f = np.zeros(SELECTOR_FEATURES, dtype=np.float32)
    info = _check(lib().slate_amd_selector_features(context(dev), A.shape[0], pa, lda, pb, X.data_ptr(), _opts(opts),
                                                    f.ctypes.data_as(ctypes.c_void_p)))
    return f, info


def mlp_forward(params, X):
    """Class probabilities (B x 3) for features X (B x 24), both CUDA float32."""
    import torch
    B = X.shape[0]
    probs = torch.empty(B, len(CHOICES), dtype=torch.float32, device=X.device)
    _check(lib().slate_amd_mlp_forward(context(X.device.index or 0), params.data_ptr(), X.data_ptr(), B,
                                       probs.data_ptr()))
    return probs


def mlp_train_step(params, grad, m, v, X, y, lr, step):
    """One minibatch step on the GPU; returns the mean loss.  lr = 0 only
    computes the gradient into grad."""
    loss = ctypes.c_float(0)
    _check(lib().slate_amd_mlp_train_step(context(X.device.index or 0), params.data_ptr(), grad.data_ptr(),
                                          m.data_ptr(), v.data_ptr(), X.data_ptr(), y.data_ptr(),
                                          X.shape[0], float(lr), int(step), ctypes.byref(loss)))
    return loss.value


def gesv_auto(A, B, X, params=None, opts=None, force=None):
    """gesv_mixed with the learned selector: params (numpy float32 MLP
    parameters) picks one of CHOICES from the probe factor's features; force
    (an index into CHOICES) runs that choice instead; neither: SLATE's default
    (float factor + classic IR, no probe).  Returns
    (info, iter, choice, ipiv, features)."""
    pa, lda, dev = _mat(A, ("float64",))
    pb, _, _ = _mat(B, ("float64",))
    px, _, _ = _mat(X, ("float64",))
    _rows(A, B, X)
    n = A.shape[0]
    ipiv = _ipiv_out(n)
    it, ch = ctypes.c_int(0), ctypes.c_int(-1)
    feat = np.zeros(SELECTOR_FEATURES, dtype=np.float32)
    pp = None
    if params is not None:
        params = np.ascontiguousarray(params, dtype=np.float32)
        if params.size != mlp_num_params():
            raise SlateError("selector parameters: expected %d floats" % mlp_num_params())
        pp = params.ctypes.data_as(ctypes.c_void_p)
    info = _check(lib().slate_amd_dsgesv_auto(context(dev), n, pa, lda, _p64(ipiv), pb, px, pp,
                                              -1 if force is None else int(force), ctypes.byref(it),
                                              ctypes.byref(ch), feat.ctypes.data_as(ctypes.c_void_p),
                                              _opts(opts)))
    return info, it.value, ch.value, ipiv[:n], feat


def gesv_batched(ns, A, B, X, factor_precision=PREC_FP32, mode=None, opts=None):
    """Batched uneven solves: ns (sizes), A/B/X packed float64 CUDA buffers
    (A_i column-major back to back, then b_i / x_i).  factor_precision FP32,
    BF16 or FP64; mode[i] = 1 sends system i to FP64.  Returns (iters, info,
    berr) numpy arrays."""
    ns = np.ascontiguousarray(ns, dtype=np.int32)
    cnt = int(ns.size)
    iters = np.zeros(cnt, dtype=np.int32)
    info = np.zeros(cnt, dtype=np.int32)
    berr = np.zeros(cnt, dtype=np.float64)
    md = None if mode is None else np.ascontiguousarray(mode, dtype=np.int32)
    _batched_check(ns, A, B, X)
    dev = A.device.index or 0
    vp = ctypes.c_void_p
    _check(lib().slate_amd_dgesv_batched(context(dev), cnt, ns.ctypes.data_as(vp), A.data_ptr(), B.data_ptr(),
                                         X.data_ptr(), None if md is None else md.ctypes.data_as(vp),
                                         int(factor_precision), iters.ctypes.data_as(vp),
                                         info.ctypes.data_as(vp), berr.ctypes.data_as(vp), _opts(opts)))
    return iters, info, berr


# choices of the batched selector (slate_amd_dgesv_batched_auto)
BATCHED_CHOICES = ("BF16 + IR", "FP32 + IR", "FP64")
BATCHED_DEFAULT = 1           # SLATE's gesv_mixed form: FP32 factor + classic IR
BATCHED_PARAMS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "selector_batched_params.npy")


def batched_params():
    """The committed batched-selector parameters (None if absent)."""
    if not os.path.exists(BATCHED_PARAMS):
        return None
    return np.load(BATCHED_PARAMS, allow_pickle=False).astype(np.float32)


def gesv_batched_auto(ns, A, B, X, params=None, force=None, opts=None, want_features=False, want_ms=False):
    """Batched uneven solves with the precision chosen per system by the
    selector network (params: its parameters, default the committed ones);
    force (0..2, BATCHED_CHOICES) runs one choice for every system.  Returns
    (iters, info, berr, choice, features or None, ms or None)."""
    ns = np.ascontiguousarray(ns, dtype=np.int32)
    cnt = int(ns.size)
    _batched_check(ns, A, B, X)
    if force is None:
        if params is None:
            params = batched_params()
        if params is None:
            raise SlateError("no batched selector parameters (pass params or force a choice)")
        params = np.ascontiguousarray(params, dtype=np.float32)
        if params.size != mlp_num_params():
            raise SlateError("selector parameters: expected %d floats" % mlp_num_params())
    iters = np.zeros(cnt, dtype=np.int32)
    info = np.zeros(cnt, dtype=np.int32)
    berr = np.zeros(cnt, dtype=np.float64)
    choice = np.zeros(cnt, dtype=np.int32)
    feat = np.zeros((cnt, SELECTOR_FEATURES), dtype=np.float32) if want_features else None
    ms = np.zeros(cnt, dtype=np.float64) if want_ms else None
    vp = ctypes.c_void_p
    dev = A.device.index or 0
    _check(lib().slate_amd_dgesv_batched_auto(
        context(dev), cnt, ns.ctypes.data_as(vp), A.data_ptr(), B.data_ptr(), X.data_ptr(),
        None if force is not None else params.ctypes.data_as(vp), -1 if force is None else int(force),
        iters.ctypes.data_as(vp), info.ctypes.data_as(vp), berr.ctypes.data_as(vp), choice.ctypes.data_as(vp),
        None if feat is None else feat.ctypes.data_as(vp), None if ms is None else ms.ctypes.data_as(vp),
        _opts(opts)))
    return iters, info, berr, choice, feat, ms


def _batched_check(ns, A, B, X):
    import torch
    for t in (A, B, X):
        if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.float64 or not t.is_contiguous():
            raise SlateError("batched: A, B, X must be contiguous float64 CUDA tensors")
    if int(np.sum(ns.astype(np.int64) ** 2)) > A.numel() or int(np.sum(ns)) > min(B.numel(), X.numel()):
        raise SlateError("batched: packed buffers smaller than the systems")


def tune_set_batched(nbf, bt, wgs_per_cu, device=0):
    _check(lib().slate_amd_tune_set_batched(context(device), int(nbf), int(bt), int(wgs_per_cu)))


def tune_get_batched(device=0):
    """(nbf, bt, wgs_per_cu, max_wgs_per_cu) of the batched solver."""
    v = [ctypes.c_int(0) for _ in range(4)]
    _check(lib().slate_amd_tune_get_batched(context(device), *[ctypes.byref(x) for x in v]))
    return tuple(x.value for x in v)


def batched_graph_stats(device=0):
    b, r = ctypes.c_int64(0), ctypes.c_int64(0)
    _check(lib().slate_amd_batched_graph_stats(context(device), ctypes.byref(b), ctypes.byref(r)))
    return b.value, r.value


def gecondest(LU, ipiv, anorm, norm="1"):
    """slate::gecondest: reciprocal condition estimate from getrf's factors."""
    pa, lda, dev = _mat(LU, ("float64",))
    ip = np.ascontiguousarray(ipiv, dtype=np.int64)
    out = ctypes.c_double(0)
    _check(lib().slate_amd_dgecondest(context(dev), ord(norm), LU.shape[0], pa, lda, _p64(ip), float(anorm),
                                      ctypes.byref(out)))
    return out.value


def pocondest(L, anorm, uplo=LOWER):
    """slate::pocondest: reciprocal condition estimate from potrf's factor."""
    pa, lda, dev = _mat(L, ("float64",))
    out = ctypes.c_double(0)
    _check(lib().slate_amd_dpocondest(context(dev), uplo, L.shape[0], pa, lda, float(anorm),
                                      ctypes.byref(out)))
    return out.value


def generate_svd(A, sigma_dist, cond, seed, sym=False):
    """A = U diag(sigma) V^T (sym: V diag(sigma) V^T) with U, V products of n
    Householder reflectors from matgen randn columns; sigma per
    matgen/generate_sigma.hh with condition number cond."""
    pa, lda, dev = _mat(A, ("float64",))
    n = A.shape[0]
    if _ncols(A) != n:
        raise SlateError("generate_svd: square matrices only")
    _check(lib().slate_amd_generate_svd(context(dev), int(sigma_dist), float(cond), int(seed), int(sym), n,
                                        pa, lda))
    return A


def generate(A, dist, seed, ioffset=0, joffset=0):
    """Fill a column-major float64 CUDA matrix with matgen Philox entries."""
    pa, lda, dev = _mat(A)
    m = A.shape[0]
    n = 1 if A.dim() == 1 else A.shape[1]
    _check(lib().slate_amd_generate(context(dev), dist, seed, m, n, ioffset, joffset, pa, lda))
    return A


#---------------------------------------------------------------- SLATE tile storage
class TiledMatrix:
    """A matrix held the way slate::Matrix::insertLocalTiles(Target::Devices)
    holds it on one device: every mb x nb tile its own column-major CUDA
    buffer with its own leading dimension (include/slate/internal/
    MatrixStorage.hh:151).  lower=True stores only tiles (i, j) with i >= j
    (a HermitianMatrix).  pad adds rows of padding to every tile's ld."""

    def __init__(self, m, n, mb, nb, lower=False, pad=0, device="cuda"):
        import torch
        self.m, self.n, self.mb, self.nb, self.lower = m, n, mb, nb, lower
        self.mt, self.nt = -(-m // mb), -(-n // nb)
        self.tiles = {}
        for j in range(self.nt):
            for i in range(self.mt):
                if lower and i < j:
                    continue
                r = min(mb, m - i * mb)
                c = min(nb, n - j * nb)
                self.tiles[(i, j)] = torch.zeros((c, r + pad), dtype=torch.float64, device=device).t()[:r]
        cnt = self.mt * self.nt
        self._ptr = (ctypes.c_void_p * cnt)()
        self._ld = (ctypes.c_int64 * cnt)()
        for (i, j), t in self.tiles.items():
            self._ptr[i + j * self.mt] = t.data_ptr()
            self._ld[i + j * self.mt] = t.stride(1)
        for k in range(cnt):
            if self._ld[k] == 0:
                self._ld[k] = 1
        self.desc = TileDesc(m, n, mb, nb, ctypes.cast(self._ptr, ctypes.POINTER(ctypes.c_void_p)),
                             ctypes.cast(self._ld, ctypes.POINTER(ctypes.c_int64)))

    @classmethod
    def from_dense(cls, M, mb, nb, lower=False, pad=0):
        T = cls(M.shape[0], 1 if M.dim() == 1 else M.shape[1], mb, nb, lower, pad, M.device)
        Md = M if M.dim() == 2 else M.reshape(-1, 1)
        for (i, j), t in T.tiles.items():
            t.copy_(Md[i * mb:i * mb + t.shape[0], j * nb:j * nb + t.shape[1]])
        return T

    def to_dense(self):
        import torch
        dev = next(iter(self.tiles.values())).device
        M = empty_colmajor(self.m, self.n, torch.float64, device=dev)
        M.zero_()
        for (i, j), t in self.tiles.items():
            M[i * self.mb:i * self.mb + t.shape[0], j * self.nb:j * self.nb + t.shape[1]] = t
        return M

    def ref(self):
        return ctypes.byref(self.desc)

    def device(self):
        return next(iter(self.tiles.values())).device.index or 0


def getrf_tiles(A, opts=None):
    ipiv = _ipiv_out(A.n)
    info = _check(lib().slate_amd_dgetrf_tiles(context(A.device()), A.ref(), _p64(ipiv), _opts(opts)))
    return info, ipiv[:A.n]


def getrs_tiles(A, ipiv, B, trans=False, opts=None):
    ip = np.ascontiguousarray(ipiv, dtype=np.int64)
    _check(lib().slate_amd_dgetrs_tiles(context(A.device()), TRANS if trans else NOTRANS, A.ref(), _p64(ip),
                                        B.ref(), _opts(opts)))


def gesv_tiles(A, B, opts=None):
    ipiv = _ipiv_out(A.n)
    info = _check(lib().slate_amd_dgesv_tiles(context(A.device()), A.ref(), _p64(ipiv), B.ref(), _opts(opts)))
    return info, ipiv[:A.n]


def potrf_tiles(A, opts=None):
    return _check(lib().slate_amd_dpotrf_tiles(context(A.device()), A.ref(), _opts(opts)))


def potrs_tiles(A, B, opts=None):
    _check(lib().slate_amd_dpotrs_tiles(context(A.device()), A.ref(), B.ref(), _opts(opts)))


def posv_tiles(A, B, opts=None):
    return _check(lib().slate_amd_dposv_tiles(context(A.device()), A.ref(), B.ref(), _opts(opts)))


def gesv_mixed_tiles(A, B, X, gmres=False, opts=None):
    ipiv = _ipiv_out(A.n)
    it = ctypes.c_int(0)
    f = lib().slate_amd_dsgesv_gmres_tiles if gmres else lib().slate_amd_dsgesv_tiles
    info = _check(f(context(A.device()), A.ref(), _p64(ipiv), B.ref(), X.ref(), ctypes.byref(it), _opts(opts)))
    return info, it.value, ipiv[:A.n]


def posv_mixed_tiles(A, B, X, gmres=False, opts=None):
    it = ctypes.c_int(0)
    f = lib().slate_amd_dsposv_gmres_tiles if gmres else lib().slate_amd_dsposv_tiles
    info = _check(f(context(A.device()), A.ref(), B.ref(), X.ref(), ctypes.byref(it), _opts(opts)))
    return info, it.value
