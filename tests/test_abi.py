"""The C-ABI boundary without a GPU: libhydist.so loads, exports every
function include/hydist.h declares, and the Python binding declares exactly
that set.  Only host-side functions are called (version, error strings)."""
import ctypes
import os
import re

import pytest

from hydist_amd import _lib, cnn

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = {"hydist.h": _lib.EXPORTS, "hydist_cnn.h": cnn.EXPORTS}


def declared(header=None):
    names = set()
    for h in ([header] if header else HEADERS):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"^[A-Za-z_][\w \*]*?\b(hd_\w+)\s*\(", text, flags=re.M))
    return names


def test_headers_and_bindings_agree():
    assert "hd_create" in declared("hydist.h") and "hd_ensemble_run" in declared("hydist.h")
    for h, exports in HEADERS.items():
        assert declared(h) == set(exports), h


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libhydist.so not built (run __graft_entry__.build())")
    so = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(declared()) if not hasattr(so, n)]
    assert not missing, missing


def test_host_only_entry_points():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libhydist.so not built")
    L = _lib.load()
    assert b"gfx950" in L.hd_version()
    assert L.hd_error_string(-5) == b"a line has more wet runs than segment_limit"
    assert L.hd_error_string(0) == b"ok"
