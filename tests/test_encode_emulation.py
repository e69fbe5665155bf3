"""CPU emulation of the MFMA encode kernel (SC_PREC_BF16) under sanitizers.

The device code of second-candidate_amd/csrc/sc_encode_mfma.hip is compiled
for the host against tests/emu/shim.h, whose wave-collective MFMA follows the
v_mfma_f32_16x16x16bf16_1k and v_mfma_f32_16x16x32_bf16 operand layouts (FP32
sums; any summation order is inside the kernel's error bound).  Each case runs
the codebook preparation kernels, the tiled coarse assignment (256-thread
workgroups of 128 rows) where the library uses it, and one emulated wave per
16 vectors, then compares every cell id and code with the oracle's encoder bit
for bit.  AddressSanitizer checks every
access against exact-size buffers (LDS included); ThreadSanitizer checks the
wave's LDS traffic.
"""
import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "emu"))
import make_emu  # noqa: E402

CXX = os.environ.get("EMU_CXX", "/opt/rocm/llvm/bin/clang++")
pytestmark = pytest.mark.skipif(not shutil.which(CXX) and not os.path.exists(CXX),
                                reason="no clang++ for the host emulation")

CASES = [  # n dim mc kc mp kp nc noisy ties tiles
    ("sc top-2, K=256, planted ties", "96 128 1 256 8 256 2 1 1 1"),
    ("ivf top-1, K=256", "48 128 1 256 8 256 1 1 1 1"),
    ("imi two halves, K=64, per-wave kernel", "48 128 2 64 8 256 1 1 1 0"),
    ("imi two halves, K=64, tiled", "150 128 2 64 8 256 1 1 1 1"),
    ("two halves top-2, K=300 (partial chunk), tiled, planted ties", "140 128 2 300 8 256 2 1 1 1"),
    ("32-dim halves top-2, K=128, tiled", "130 64 2 128 4 256 2 1 0 1"),
    ("sc K=1000: several groups, partial tile", "40 64 1 1000 4 256 2 1 1 1"),
    ("D=960, integer data", "20 960 2 64 15 256 1 0 0 1"),
]


@pytest.fixture(scope="module", params=["thread", "address"])
def emu(request, tmp_path_factory):
    out = str(tmp_path_factory.mktemp("emu_enc") / ("emu_encode_" + request.param))
    make_emu.build(out, request.param, CXX, encode=True)
    return out


@pytest.mark.parametrize("name,args", CASES, ids=[c[0] for c in CASES])
def test_mfma_encode_emulation(emu, name, args):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", TSAN_OPTIONS="report_signal_unsafe=0")
    r = subprocess.run([emu] + args.split(), capture_output=True, text=True, timeout=900, env=env)
    log = r.stdout + r.stderr
    assert "WARNING: ThreadSanitizer" not in log, log[-4000:]
    assert "ERROR: AddressSanitizer" not in log, log[-4000:]
    assert r.returncode == 0, log[-4000:]
    assert "emu_encode: 0 cell ids and 0 codes differ" in log, log[-4000:]
