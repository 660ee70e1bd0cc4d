"""The C ABI: libgle.so loads and exports every symbol include/gle.h declares."""
import os
import re

from garli_amd import engine as E

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "gle.h")


def declared():
    txt = open(HDR).read()
    return sorted(set(re.findall(r"\b(gle_[a-z0-9_]+)\s*\(", txt)))


def test_header_symbols_exported():
    lib = E.lib()
    names = declared()
    assert len(names) > 40
    for n in names:
        assert hasattr(lib, n), n


def test_bindings_cover_header():
    assert sorted(E.SIGNATURES) == declared()
