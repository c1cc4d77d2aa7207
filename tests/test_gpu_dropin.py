"""The C++ drop-in headers (include/ecl_wavelet/) used the way the
reference's gtest files use its classes; runs tests/cpp/test_dropin."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_dropin")


def test_cpp_dropin(gpu):
    assert os.path.exists(BIN), "built by __graft_entry__.build()"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=240)
    print(r.stdout[-4000:])
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
    assert "ALL OK" in r.stdout
