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


def _build(tmp_path):
    exe = str(tmp_path / "dropin_host_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Wno-unused-parameter",
                    os.path.join(HERE, "cpp", "dropin_host_check.cpp"), "-o", exe,
                    REF_LIB, "-L" + LIB_DIR, "-lsc_amd",
                    "-Wl,-rpath," + os.path.dirname(REF_LIB), "-Wl,-rpath," + LIB_DIR],
                   check=True)
    return exe


@pytest.mark.skipif(not os.path.exists(REF_LIB), reason="oracle/_ref not built")
def test_dropin_distortion_matches_reference(tmp_path):
    exe = _build(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "OK: 0 failure(s)" in r.stdout, r.stdout + r.stderr


@pytest.mark.skipif(not os.path.exists(REF_LIB), reason="oracle/_ref not built")
def test_dropin_refuses_malformed_inverted_file(tmp_path):
    """PQQuery::load_encoded_data (src/query.cpp:82-165) on truncated or
    inconsistent `_ivf.edat_` files: every read is bounded by the file size
    and every cell by N, so the loader stops with a message instead of
    reading or writing past its buffers (before any device work)."""
    import numpy as np
    import sco
    exe = _build(tmp_path)
    kc, mp, dim = 4, 2, 8
    cq, pq = str(tmp_path / "c.ctr_"), str(tmp_path / "p.ctr_")
    sco.write_codebook(cq, np.zeros((kc, dim), np.float32), kc, 1, dim)
    sco.write_codebook(pq, np.zeros((mp * 16, dim // mp), np.float32), 16, mp, dim)

    def edat(cells, n):
        out = [np.int32(sum(1 for c in cells if c)).tobytes(), np.int32(n).tobytes()]
        for c in cells:
            out.append(np.int32(len(c)).tobytes())
            out.append(np.asarray(c, np.int32).tobytes())
            out.append(bytes(len(c) * mp))
        return b"".join(out)

    good = edat([[0, 1], [], [2], [3, 4]], 5)
    cases = {
        "truncated (cell body)": good[:-3],
        "truncated (cell length)": good[:8 + 4 + 2 * 4 + 2 * mp + 2],
        "cell lengths exceed N": edat([[0, 1], [], [2], [3, 4]], 4),
        "do not sum to N": edat([[0, 1], [], [2], [3, 4]], 6),
        "too short for its header": good[:6],
    }
    for why, blob in cases.items():
        f = tmp_path / "x_ivf.edat_"
        f.write_bytes(blob)
        r = subprocess.run([exe, "edat", cq, pq, str(f)], capture_output=True, text=True,
                           timeout=60)
        assert r.returncode == 1 and why in r.stderr, (why, r.returncode, r.stderr)
