"""The C-ABI library loads and exports every function include/gst.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

from gpu4gst import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gst_[a-z_]+|gg_[a-z_]+)\s*\(", text)))


def test_libgst_exports_header():
    names = _declared("gst.h")
    assert "gst_solve" in names and "gst_backward" in names
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(_lib.EXPORTS) == names


def test_libgstgen_exports_header():
    names = _declared("gst_gen.h")
    lib = ctypes.CDLL(os.path.join(os.path.dirname(_lib.LIB_PATH), "libgstgen.so"))
    for n in names:
        assert hasattr(lib, n), n


def test_version_symbol():
    assert _lib.lib().gst_version() >= 1
