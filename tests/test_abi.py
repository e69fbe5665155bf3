"""CPU checks of the C ABI: the library loads and exports every entry point
include/sc_hip.h declares.  No compute call is made here (no GPU)."""
import ctypes as C
import re

import sc_amd


def test_library_exports_every_declared_symbol():
    names = sc_amd.header_symbols()
    assert len(names) >= 20
    lib = C.CDLL(sc_amd.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []
    assert set(names) <= set(sc_amd._SIGS)          # the binding covers all of them


def test_version_and_error_strings():
    v = sc_amd.lib().sc_version().decode()
    assert "gfx950" in v
    assert isinstance(sc_amd.lib().sc_last_error(), bytes)


def test_header_is_plain_c():
    text = open(sc_amd.HEADER).read()
    assert 'extern "C"' in text
    for bad in ("torch", "std::", "hipStream_t", "template"):
        assert bad not in re.sub(r"/\*.*?\*/", "", text, flags=re.S)


def test_status_codes_match_header():
    text = open(sc_amd.HEADER).read()
    assert re.search(r"#define SC_OK 0\b", text)
    assert re.search(r"#define SC_KIND_SC 2\b", text)
    assert sc_amd.KIND_SC == 2 and sc_amd.FORM_MIXED == 1 and sc_amd.PREC_BF16 == 1


def test_index_load_refuses_malformed_inverted_file():
    """sc_index_load validates L and pid on the host before any device work:
    L must be a non-decreasing inclusive prefix ending at n and every pid in
    [0, n); otherwise SC_ERR_ARG with the reason (no GPU needed to reach it)."""
    import numpy as np
    L = sc_amd.lib()
    kc, mp, kp, dim = 4, 2, 4, 8
    cq = np.zeros((kc, dim), np.float32)
    pq = np.zeros((mp, kp, dim // mp), np.float32)
    out = C.c_void_p()

    def load(Lc, pid):
        Lc = np.ascontiguousarray(Lc, np.int32)
        pid = np.ascontiguousarray(pid, np.int32)
        codes = np.zeros((pid.size, mp), np.uint8)
        rc = L.sc_index_load(None, Lc.ctypes.data, Lc.size, pid.ctypes.data, codes.ctypes.data,
                             pid.size, dim, cq.ctypes.data, 1, kc, pq.ctypes.data, mp, kp, 1,
                             C.byref(out))
        return rc, L.sc_last_error().decode()

    rc, msg = load([2, 1, 3, 5], [0, 1, 2, 3, 4])
    assert rc == sc_amd.SC_ERR_ARG and "non-decreasing" in msg
    rc, msg = load([2, 2, 3, 4], [0, 1, 2, 3, 4])
    assert rc == sc_amd.SC_ERR_ARG and "L[last]" in msg
    rc, msg = load([2, 2, 3, 5], [0, 1, 7, 3, 4])
    assert rc == sc_amd.SC_ERR_ARG and "pid[2]=7" in msg
    rc, msg = load([2, 2, 3, 5], [0, 1, -1, 3, 4])
    assert rc == sc_amd.SC_ERR_ARG and "pid[2]=-1" in msg
    rc, msg = load([2, 2, 3, 5], [0, 1, 2, 3, 4])       # well formed: stops at the context
    assert rc == sc_amd.SC_ERR_ARG and "null context" in msg
