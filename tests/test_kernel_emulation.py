"""CPU emulation of the fused search kernel under sanitizers.

The kernel's device code (second-candidate_amd/csrc/sc_search.hip) is compiled
for the host against tests/emu/shim.h: one workgroup runs as 256 std::threads,
__syncthreads is a barrier, wave intrinsics exchange values per wave.  Each
case builds an index with the oracle, runs the emulated kernel and compares
every query's ids, distances and candidate count with the oracle's search.
ThreadSanitizer reports any LDS access two waves make without a barrier
between them; AddressSanitizer reports accesses outside the exact-size
buffers (the dynamic LDS region included).
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

CASES = [  # kind n kc R w T nq threads trained precomputed_lut mixed [margin ncomp]
    ("sc plain, trained codebooks", "2 20000 32 100 300 1500 4 256 1 0 0"),
    ("sc mixed, precomputed LUTs", "2 20000 64 100 300 1500 4 256 0 1 1"),
    ("ivf plain, precomputed LUTs", "0 20000 64 100 12 3000 4 256 0 1 0"),
    ("sc mixed, forced exact redo", "2 20000 32 100 300 1500 3 256 1 0 1 1000000"),
    ("sc plain, 512 threads", "2 20000 32 100 300 1500 3 512 0 0 0"),
    ("sc plain, traversal past the rank window", "2 20000 64 100 3000 6000 3 256 1 0 0"),
    ("drop-in data: 16-component mixture, 40 queries", "2 20000 32 100 300 1500 40 256 1 0 0 1 16"),
    ("sc plain, pop budget 5: long traversals handed on", "2 20000 32 100 300 1500 4 256 1 0 0 1 64 5"),
]


@pytest.fixture(scope="module", params=["thread", "address"])
def emu(request, tmp_path_factory):
    out = str(tmp_path_factory.mktemp("emu") / ("emu_" + request.param))
    make_emu.build(out, request.param, CXX)
    return request.param, out


@pytest.mark.parametrize("name,args", CASES, ids=[c[0] for c in CASES])
def test_fused_kernel_emulation(emu, name, args):
    san, binary = emu
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", TSAN_OPTIONS="report_signal_unsafe=0")
    r = subprocess.run([binary] + args.split(), capture_output=True, text=True, timeout=900,
                       env=env)
    log = r.stdout + r.stderr
    assert "WARNING: ThreadSanitizer" not in log, log[-4000:]
    assert "ERROR: AddressSanitizer" not in log, log[-4000:]
    assert r.returncode == 0, log[-4000:]
    nq = int(args.split()[6])
    assert ("emu: 0 of %d queries mismatch" % nq) in log, log[-4000:]


TRAV3_CASES = [  # n kc w T nq
    ("budget T binds", "5000 12 300 1500 16"),
    ("heap capacity 3: inserts dropped", "4000 8 3 100000 16"),
    ("whole 4^3 grid popped", "1500 4 5000 100000 8"),
    ("cell budget binds", "6000 12 40 300 32"),
    ("kc 16", "8000 16 2000 4000 24"),
]


@pytest.fixture(scope="module")
def trav3(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("emu3") / "trav3")
    make_emu.build(out, "address", CXX, trav3=True)
    return out


@pytest.mark.parametrize("name,args", TRAV3_CASES, ids=[c[0] for c in TRAV3_CASES])
def test_traverse3_kernel_emulation(trav3, name, args):
    """The three-candidate traversal kernel, compiled for the host under
    AddressSanitizer, against the oracle's traced search query by query."""
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([trav3] + args.split(), capture_output=True, text=True, timeout=600,
                       env=env)
    log = r.stdout + r.stderr
    assert "ERROR: AddressSanitizer" not in log, log[-4000:]
    assert r.returncode == 0, log[-4000:]
    assert ("OK: 0 of %s queries differ" % args.split()[4]) in log, log[-4000:]


TRAVW_CASES = [  # kind n kc w T nq [hcap]
    ("imi", "1 8000 32 300 1500 8"),
    ("imi heap capacity 5: inserts dropped", "1 8000 32 5 100000 8"),
    ("imi whole 8x8 grid", "1 3000 8 1000 100000 6"),
    ("sc", "2 8000 32 300 1500 8"),
    ("sc cell budget binds", "2 8000 32 4 100000 8"),
    ("imi on-chip heap 3: flagged", "1 8000 32 300 1500 8 3"),
    ("sc on-chip heap 4: flagged", "2 8000 32 300 1500 8 4"),
    ("imi kc 300, within the staged ranks", "1 8000 300 300 1500 4"),
    ("imi kc 300, past the staged ranks: flagged", "1 3000 300 100000 100000 2"),
]


@pytest.fixture(scope="module", params=["thread", "address"])
def travw(request, tmp_path_factory):
    out = str(tmp_path_factory.mktemp("emuw") / ("travw_" + request.param))
    make_emu.build(out, request.param, CXX, travw=True)
    return request.param, out


@pytest.mark.parametrize("name,args", TRAVW_CASES, ids=[c[0] for c in TRAVW_CASES])
def test_traverse_wave_kernel_emulation(travw, name, args):
    """The one-workgroup-per-query traversal kernel (64 host threads per
    workgroup) under ThreadSanitizer and AddressSanitizer, against the
    oracle's traced search; queries whose heap outgrows the on-chip capacity
    must be flagged, never wrong."""
    san, binary = travw
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", TSAN_OPTIONS="report_signal_unsafe=0")
    r = subprocess.run([binary] + args.split(), capture_output=True, text=True, timeout=600,
                       env=env)
    log = r.stdout + r.stderr
    assert "WARNING: ThreadSanitizer" not in log, log[-4000:]
    assert "ERROR: AddressSanitizer" not in log, log[-4000:]
    assert r.returncode == 0, log[-4000:]
    assert ("OK: 0 of %s queries differ" % args.split()[5]) in log, log[-4000:]


# The per-query scan kernel: n kc R w T nq mixed [margin_scale].  More than
# kRound candidates per query exercise the buffer compaction (bounded
# select, register-held move), runs of more than kRunCells cells the run
# staging, T below R the position-ordered output, and a widened margin the
# mixed form's exact-redo flag (as does an R whose kept set outgrows the
# half of the buffer a compaction may keep).
SCAN_CASES = [
    ("plain, compactions", "8000 32 100 1000 3000 4 0"),
    ("mixed, compactions", "8000 32 100 1000 3000 4 1"),
    ("mixed, margin x40", "8000 32 100 1000 3000 4 1 40"),
    ("mixed, margin x2000: redo flags", "8000 32 100 1000 3000 4 1 2000"),
    ("plain, fewer candidates than R", "8000 32 100 300 60 4 0"),
    ("mixed, fewer candidates than R", "8000 32 100 300 60 4 1"),
    ("plain, R 8", "8000 32 8 2000 5000 3 0"),
    ("mixed, R 700", "8000 32 700 2000 5000 3 1"),
    ("mixed, R 1000: redo flags", "8000 32 1000 2000 5000 3 1"),
]


@pytest.fixture(scope="module", params=["thread", "address"])
def scan(request, tmp_path_factory):
    out = str(tmp_path_factory.mktemp("emus") / ("scan_" + request.param))
    make_emu.build(out, request.param, CXX, scan=True)
    return request.param, out


@pytest.mark.parametrize("name,args", SCAN_CASES, ids=[c[0] for c in SCAN_CASES])
def test_scan_kernel_emulation(scan, name, args):
    """The per-query ADC scan and top-R kernel (256 host threads per
    workgroup) under ThreadSanitizer and AddressSanitizer, against the
    oracle's search over the same traced cells: identical ids, distances and
    candidate counts; queries flagged for an exact redo are not compared
    (and with the bound's own margin none is)."""
    san, binary = scan
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", TSAN_OPTIONS="report_signal_unsafe=0")
    r = subprocess.run([binary] + args.split(), capture_output=True, text=True, timeout=600,
                       env=env)
    log = r.stdout + r.stderr
    assert "WARNING: ThreadSanitizer" not in log, log[-4000:]
    assert "ERROR: AddressSanitizer" not in log, log[-4000:]
    assert r.returncode == 0, log[-4000:]
    nq = args.split()[5]
    assert ("OK: 0 of %s queries differ" % nq) in log, log[-4000:]
    if len(args.split()) < 8 and "redo flags" not in name:
        assert "(0 flagged" in log, log[-4000:]
    if "redo flags" in name:
        assert ("(%s flagged" % nq) in log, log[-4000:]
