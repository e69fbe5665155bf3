"""Runs the reference-style C++ pipeline (tests/cpp/test_dropin.cpp) through the
drop-in classes on the GPU; the binary checks every stage against the oracle."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_dropin")


def test_dropin_pipeline(tmp_path):
    assert os.path.exists(BIN), "tests/cpp/test_dropin is built by __graft_entry__.build()"
    r = subprocess.run([BIN, str(tmp_path)], capture_output=True, text=True, timeout=300)
    print(r.stdout[-3000:], r.stderr[-3000:])
    assert r.returncode == 0, r.stderr[-2000:]
    assert "OK: 0 failure(s)" in r.stdout
