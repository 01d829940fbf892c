"""ThreadSanitizer and AddressSanitizer + UBSan runs of the host code (CPU only).

tests/native/Makefile builds libhasty_hip with its host pass instrumented and drives it
from 8 threads (tests/native/host_threads.c: options, LLR shift and workspace queries,
error paths and thread-local messages), and runs the oracle's C restatement under ASan +
UBSan (tests/native/oracle_stress.c). Building the instrumented library takes a few
minutes, so these tests run only with HASTY_SANITIZE=1 in the environment.
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))

pytestmark = pytest.mark.skipif(os.environ.get("HASTY_SANITIZE") != "1",
                                reason="sanitizer builds take minutes; set HASTY_SANITIZE=1 to run them")


@pytest.mark.parametrize("kind", ["tsan", "asan"])
def test_host_code_under_sanitizer(kind):
    r = subprocess.run(["make", "-C", os.path.join(HERE, "native"), "-j8", kind], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=1800)
    assert r.returncode == 0, r.stdout[-4000:]
    assert "WARNING: ThreadSanitizer" not in r.stdout and "ERROR: AddressSanitizer" not in r.stdout
    assert "runtime error" not in r.stdout  # UBSan
    assert "host_threads: 8 threads x 2000 rounds, 0 failures" in r.stdout
    if kind == "asan":
        assert "oracle_stress: 0 failures" in r.stdout
