"""C ABI checks that need no GPU: the library loads and exports every symbol
declared in include/fdtd_hip.h."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "fdtd_hip.h")).read()
    return sorted(set(re.findall(r"\b(fdtd_[a-z_]+)\s*\(", txt)))


def test_header_matches_binding_list():
    from fdtd_amd.grid import ABI_SYMBOLS
    assert sorted(ABI_SYMBOLS) == declared_symbols()


def test_library_exports_every_declared_symbol():
    from fdtd_amd.grid import LIB_PATH, load_library
    if not os.path.exists(LIB_PATH):
        pytest.fail("libfdtd_hip.so not built (run __graft_entry__.build())")
    lib = load_library()
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_last_error_without_gpu_is_a_string():
    from fdtd_amd.grid import load_library
    assert isinstance(load_library().fdtd_last_error(), bytes)
