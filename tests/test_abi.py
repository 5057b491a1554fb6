"""The C ABI: the built library exports every function include/hermes_amd.h declares, and the
ctypes binding covers them (no GPU calls)."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hermes_amd.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hermes_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = _declared()
    for n in ("hermes_project", "hermes_pair_reduce", "hermes_multipole_gram_f32", "hermes_mc_binned",
              "hermes_capture_begin", "hermes_graph_launch"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from pyhermes_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libhermes_amd.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_binding_covers_every_declaration():
    from pyhermes_amd import _lib
    bound = set(_lib.declared_symbols())
    missing = [n for n in _declared() if n not in bound]
    assert not missing, missing
