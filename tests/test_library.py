"""CPU checks of the C ABI: the library loads and exports every declared symbol."""

import ctypes
import os
import re

from tenax_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            txt = open(os.path.join(inc, fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(tnx_[a-z0-9_]+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    lib = _lib.load(require_gpu=False)
    names = declared_symbols()
    assert names, "no declarations found"
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_table_matches_headers():
    assert set(_lib.SIGNATURES) == declared_symbols()


def test_version_and_error_text():
    lib = _lib.load(require_gpu=False)
    assert lib.tnx_version() >= 1000
    buf = ctypes.create_string_buffer(64)
    assert lib.tnx_last_error(buf, 64) == 0
