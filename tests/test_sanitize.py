"""Host-only AddressSanitizer + UndefinedBehaviorSanitizer run over the CPU
oracle (tests/cpp/oracle_sanitize.cc): every solver family on small ragged
inputs, in its own process; any sanitizer report fails the test."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CPP = os.path.join(HERE, "cpp")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_oracle_under_asan_ubsan():
    subprocess.check_call(["make", "-s", "-C", CPP, "oracle_sanitize"])
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="2")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([os.path.join(CPP, "oracle_sanitize")], capture_output=True, text=True, env=env,
                       timeout=600)
    out = r.stdout + r.stderr
    assert "ERROR: AddressSanitizer" not in out and "runtime error:" not in out, out[-4000:]
    assert r.returncode == 0, out[-4000:]
    assert "0 failures" in r.stdout
