"""CPU checks of the drop-in boundary: the C-ABI library loads, exports every
symbol that include/slate_amd.h declares, and reports misuse as negative
codes with a message instead of throwing (no GPU calls are made here)."""
import ctypes
import os
import re

import slate_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "slate_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(slate_amd_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(slate_amd.LIB_PATH)
    decl = declared_symbols()
    assert len(decl) >= 20
    missing = [s for s in decl if not hasattr(lib, s)]
    assert missing == []
    assert sorted(slate_amd.C_SYMBOLS) == decl


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
