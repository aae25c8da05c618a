"""The C-ABI library loads and exports every function include/golay_gpu.h declares
(no GPU calls).  The product fails loudly without the library."""
import ctypes
import os
import re

import pytest

import golaypair_amd as g


def _declared():
    src = open(g.HEADER_PATH).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(gp_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_binding_surface():
    assert _declared() == sorted(g.SIGNATURES)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(g.LIB_PATH):
        pytest.skip("library not built")
    lib = ctypes.CDLL(g.LIB_PATH)
    for name in _declared():
        assert hasattr(lib, name), name
    g.load_library()


def test_missing_library_raises(tmp_path):
    saved = g._lib
    try:
        g._lib = None
        with pytest.raises(g.GolayError):
            g.load_library(str(tmp_path / "nope.so"))
    finally:
        g._lib = saved
