"""The C-ABI library exists, loads and exports every entry point declared in
include/bmagwa_hip.h (no compute calls: runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bmagwa_hip.h")
LIB = os.path.join(ROOT, "bmagwa_amd", "libbmagwa_hip.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bmg_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "bmg_ctx_create" in names and "bmg_batch_run" in names and "bmg_model_add_term" in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="engine not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="engine not built")
def test_binding_matches_header():
    import bmagwa_amd._lib as L
    assert set(L.SIGNATURES) == set(declared())
