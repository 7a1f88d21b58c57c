"""slate_amd_dgetrs_noalloc: the single-right-hand-side getrs whose workspace
(device pivots, interchange plans, inverted diagonal blocks, look-back
exchange words) is the library's static device scratch instead of allocated
buffers.  Called through the C ABI, it must give the same solution, bit for
bit, as slate_amd_dgetrs on the same factors, on two calls in a row (the
second smaller, on different factors, so the scratch holds the first call's
contents) and reject sizes beyond the scratch limit before any work."""
import ctypes

import numpy as np
import pytest
import torch

import slate_amd as S
from helpers import dev_rands, say

pytestmark = pytest.mark.gpu


def _noalloc(LU, ipiv, b):
    f = S.lib().slate_amd_dgetrs_noalloc
    f.restype = ctypes.c_int64
    f.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64,
                  ctypes.POINTER(ctypes.c_int64), ctypes.c_void_p]
    ip = np.ascontiguousarray(ipiv, dtype=np.int64)
    n = LU.shape[0]
    return f(S.context(0), n, LU.data_ptr(), LU.stride(1), ip.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
             b.data_ptr())


@pytest.mark.parametrize("sizes", [(2000, 700), (4096, 1500)])
def test_noalloc_getrs_equals_getrs(gpu, sizes):
    for rnd, n in enumerate(sizes):
        say("dgetrs_noalloc n=%d (call %d)" % (n, rnd))
        A = dev_rands(40 + rnd, n, n)
        info, ipiv = S.getrf(A)
        assert info == 0
        b = dev_rands(50 + rnd, n, 1)
        x_ref = b.clone()
        S.getrs(A, ipiv, x_ref)
        x = b[:, 0].clone()
        rc = _noalloc(A, ipiv, x)
        assert rc == 0, S.lib().slate_amd_last_error()
        assert torch.equal(x, x_ref[:, 0])


def test_noalloc_getrs_rejects_large_n(gpu):
    n = 16385
    A = torch.zeros((1, 1), dtype=torch.float64, device="cuda")     # never read: rejected first
    x = torch.zeros(1, dtype=torch.float64, device="cuda")
    f = S.lib().slate_amd_dgetrs_noalloc
    f.restype = ctypes.c_int64
    f.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64,
                  ctypes.POINTER(ctypes.c_int64), ctypes.c_void_p]
    ip = np.arange(1, n + 1, dtype=np.int64)
    rc = f(S.context(0), n, A.data_ptr(), n, ip.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), x.data_ptr())
    assert rc == -3
