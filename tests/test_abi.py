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


def test_header_compiles_as_c_and_cpp(tmp_path):
    """include/bmagwa_hip.h is a plain C header: a C11 and a C++17 file that
    include it and take the address of every declared entry point compile
    (host compilers only)."""
    import subprocess
    names = declared()
    body = "#include \"bmagwa_hip.h\"\nvoid* table[] = {%s};\nint main(void) { return table[0] == 0; }\n" \
        % ", ".join("(void*)&" + n for n in names)
    for ext, cc, std in (("c", "gcc", "-std=c11"), ("cpp", "g++", "-std=c++17")):
        src = tmp_path / ("t." + ext)
        src.write_text(body)
        subprocess.check_call([cc, std, "-Wall", "-Werror", "-c", "-I", os.path.join(ROOT, "include"), str(src),
                               "-o", str(tmp_path / ("t_" + ext + ".o"))])


@pytest.mark.skipif(not os.path.exists(LIB), reason="engine not built")
def test_cpp_model_caller_links(tmp_path):
    """The C++ HipModel caller of tests/cpp compiles against the header and
    links against the library (it runs in the GPU tests)."""
    import subprocess
    src = os.path.join(ROOT, "tests", "cpp", "model_caller.cpp")
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                           src, "-o", str(tmp_path / "caller"), "-L", os.path.join(ROOT, "bmagwa_amd"),
                           "-lbmagwa_hip", "-Wl,-rpath-link,/opt/rocm/lib"])
