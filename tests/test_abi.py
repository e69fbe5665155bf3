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
