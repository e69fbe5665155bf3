"""Host-only parts of the drop-in header against the reference's own code
compiled in place (oracle/_ref): builds tests/cpp/dropin_host_check.cpp with
g++ and runs it.  No GPU."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libscref.so")
LIB_DIR = os.path.join(ROOT, "second-candidate_amd", "lib")


@pytest.mark.skipif(not os.path.exists(REF_LIB), reason="oracle/_ref not built")
def test_dropin_distortion_matches_reference(tmp_path):
    exe = str(tmp_path / "dropin_host_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Wno-unused-parameter",
                    os.path.join(HERE, "cpp", "dropin_host_check.cpp"), "-o", exe,
                    REF_LIB, "-L" + LIB_DIR, "-lsc_amd",
                    "-Wl,-rpath," + os.path.dirname(REF_LIB), "-Wl,-rpath," + LIB_DIR],
                   check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "OK: 0 failure(s)" in r.stdout, r.stdout + r.stderr
