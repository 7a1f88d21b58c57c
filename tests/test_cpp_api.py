"""Host-only checks of the C++ interface (include/slate_amd.hh): pivots
conversion (lapack_api/lapack_gesv.cc:90-100), Options translation and
argument errors.  Builds tests/cpp/test_host_api with its Makefile."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CPP = os.path.join(HERE, "cpp")


def test_cpp_host_api():
    subprocess.check_call(["make", "-s", "-C", CPP, "test_host_api"])
    r = subprocess.run([os.path.join(CPP, "test_host_api")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout
