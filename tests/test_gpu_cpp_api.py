"""The slate_amd:: C++ interface on the GPU (tests/cpp/test_slate_api.cc):
gesv/getrs/getrf/trans getrs/posv and the four mixed drivers through
Matrix::fromDevices / fromLAPACK, with SLATE's tester bound (scaled residual
<= 3), and the n = 2000 gesv pivots equal to LAPACK dgetrf's on the same
matgen matrix (committed fixture)."""
import os
import subprocess

import numpy as np
import pytest

from helpers import say

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "test_slate_api")


@pytest.mark.gpu
def test_cpp_interface(gpu, tmp_path):
    assert os.path.exists(BIN), "tests/cpp/test_slate_api is not built (run __graft_entry__.build())"
    say("cpp interface: running %s" % BIN)
    r = subprocess.run([BIN, str(tmp_path)], capture_output=True, text=True, timeout=300)
    say(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    ipiv = np.loadtxt(tmp_path / "ipiv_gesv.txt", dtype=np.int64)
    ref = np.load(os.path.join(HERE, "golden", "ipiv_lapack_n2000.npy")).astype(np.int64)
    assert ipiv.shape == ref.shape
    assert np.array_equal(ipiv, ref), "first mismatch at %d" % int(np.argmax(ipiv != ref))
