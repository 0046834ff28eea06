"""CPU-side checks of the C ABI: the library loads and exports every entry point of
include/wf.h; the oracle-independent host logic behaves.  No GPU compute is called."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "wf.h")
LIB = os.path.join(ROOT, "mathpub_amd", "libwf.so")


def declared_functions():
    src = open(HDR).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(wf_[a-z_]+)\s*\(", src)))


def test_header_declares_the_abi():
    names = declared_functions()
    for want in ("wf_open", "wf_close", "wf_scan", "wf_scan_batch", "wf_quotients", "wf_primes",
                 "wf_serve_submit", "wf_serve_poll", "wf_strerror"):
        assert want in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libwf.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert missing == []


@pytest.mark.skipif(not os.path.exists(LIB), reason="libwf.so not built")
def test_strerror_codes():
    from mathpub_amd import wf
    L = wf.load()
    assert L.wf_strerror(0) == b"ok"
    assert b"argument" in L.wf_strerror(-1)
    assert b"not supported" in L.wf_strerror(-4)


@pytest.mark.skipif(not os.path.exists(LIB), reason="libwf.so not built")
def test_argument_errors_without_gpu_work():
    from mathpub_amd import wf
    L = wf.load()
    # NULL output pointer and bad device are rejected before any GPU work
    assert L.wf_open(0, None, None) == -1
    h = ctypes.c_void_p()
    assert L.wf_open(-5, None, ctypes.byref(h)) == -2
    assert L.wf_scan(None, 0, 10, 0, None, None, 0, 0, None, 0) == -1
