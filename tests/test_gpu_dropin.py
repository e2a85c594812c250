"""GPU: a C++ caller written against the reference's calaman::expm template
interface (include/calaman/expm.hpp) runs unchanged on the library
(tests/cpp/expm_dropin.cpp, built by __graft_entry__.build())."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(__file__), "cpp", "expm_dropin")


def test_cpp_dropin_caller(dev):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} not built: run __graft_entry__.build()")
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASSED" in r.stdout
