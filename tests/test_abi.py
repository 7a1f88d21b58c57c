"""CPU checks of the drop-in boundary: the C-ABI library loads, exports every
symbol that include/slate_amd.h declares, and reports misuse as negative
codes with a message instead of throwing (no GPU calls are made here)."""
import ctypes
import os
import re

import slate_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols(header="slate_amd.h", pattern=r"\b(slate_amd_\w+)\s*\("):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(pattern, text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(slate_amd.LIB_PATH)
    decl = declared_symbols()
    assert len(decl) >= 20
    missing = [s for s in decl if not hasattr(lib, s)]
    assert missing == []
    assert sorted(slate_amd.C_SYMBOLS) == decl


def test_library_exports_every_lapack_symbol():
    lib = ctypes.CDLL(slate_amd.LIB_PATH)
    decl = declared_symbols("slate_amd_lapack.h", r"\b(slate_\w+_)\s*\(")
    assert len(decl) >= 11
    missing = [s for s in decl if not hasattr(lib, s)]
    assert missing == []


def test_lapack_argument_errors_need_no_gpu():
    """LAPACK-style argument checks run before any device call (info = -i)."""
    lib = ctypes.CDLL(slate_amd.LIB_PATH)
    i = ctypes.c_int
    info = i(0)
    n, nrhs, lda, ldb = i(4), i(1), i(2), i(4)
    lib.slate_dgesv_(ctypes.byref(n), ctypes.byref(nrhs), None, ctypes.byref(lda), None, None,
                     ctypes.byref(ldb), ctypes.byref(info))
    assert info.value == -4
    lda = i(4)
    ldb = i(1)
    lib.slate_dgesv_(ctypes.byref(n), ctypes.byref(nrhs), None, ctypes.byref(lda), None, None,
                     ctypes.byref(ldb), ctypes.byref(info))
    assert info.value == -7
    uplo = ctypes.c_char(b"X")
    lib.slate_dpotrf_(ctypes.byref(uplo), ctypes.byref(n), None, ctypes.byref(lda), ctypes.byref(info))
    assert info.value == -1
    trans = ctypes.c_char(b"Q")
    lib.slate_dgetrs_(ctypes.byref(trans), ctypes.byref(n), ctypes.byref(nrhs), None, ctypes.byref(lda), None,
                      None, ctypes.byref(lda), ctypes.byref(info))
    assert info.value == -1
    neg = i(-1)
    it = i(0)
    lib.slate_dsgesv_(ctypes.byref(neg), ctypes.byref(nrhs), None, ctypes.byref(lda), None, None,
                      ctypes.byref(lda), None, ctypes.byref(lda), None, None, ctypes.byref(it), ctypes.byref(info))
    assert info.value == -1


def test_options_defaults_mirror_slate():
    o = slate_amd.Options()
    assert o.lookahead == 1            # src/getrf.cc:39
    assert o.pivot_threshold == 1.0    # src/getrf.cc:38
    assert o.max_iterations == 30      # src/gesv_mixed.cc:125
    assert o.use_fallback == 1         # src/gesv_mixed.cc:127
    assert o.tolerance == 0.0          # 0 -> eps*sqrt(n), src/gesv_mixed.cc:126
    assert o.factor_precision == slate_amd.PREC_FP32


def test_null_context_is_an_error_code_not_a_crash():
    lib = slate_amd.lib()
    ipiv = (ctypes.c_int64 * 4)()
    rc = lib.slate_amd_dgetrf(None, 4, None, 4, ipiv, None)
    assert rc < 0
    assert "context" in slate_amd.last_error()
    it = ctypes.c_int(0)
    rc = lib.slate_amd_dsgesv(None, 4, 1, None, 4, ipiv, None, 4, None, 4, ctypes.byref(it), None)
    assert rc < 0


def test_python_layer_rejects_host_tensors():
    import numpy as np
    import pytest
    import torch
    with pytest.raises(slate_amd.SlateError):
        slate_amd.getrf(torch.from_numpy(np.eye(3)))
