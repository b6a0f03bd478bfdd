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


def test_copy2d_rejects_rectangles_outside_their_tensors():
    """The batched copy's host-side checks run before any device work."""
    import pytest
    import torch

    a = torch.zeros(100, dtype=torch.float64)
    b = torch.zeros(100, dtype=torch.float64)
    with pytest.raises(ValueError, match="outside a tensor"):
        _lib.copy2d([(a, 90, 10, b, 0, 10, 2, 10, 1.0)], a.device)  # source rows 90..109
    with pytest.raises(ValueError, match="outside a tensor"):
        _lib.copy2d([(a, 0, 10, b, 0, 5, 2, 10, 1.0)], a.device)  # pitch below the row length
    with pytest.raises(ValueError, match="contiguous float64"):
        _lib.copy2d([(a.float(), 0, 10, b, 0, 10, 1, 10, 1.0)], a.device)
