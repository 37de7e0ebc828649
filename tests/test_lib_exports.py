"""CPU checks of the C-ABI library: it loads and exports every symbol that
include/psy_gpu.h declares; no compute calls are made here."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    with open(os.path.join(ROOT, "include", "psy_gpu.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:int|int32_t|void|const char\*|const psy_\w+\*)\s+(psy_\w+)\s*\(", text, re.M)))


def test_header_symbols_listed_in_abi_module():
    from psychopath_cpp_amd import abi
    assert header_symbols() == sorted(abi.EXPORTED_SYMBOLS)


def test_library_exports_every_header_symbol():
    from psychopath_cpp_amd import gpu
    if not os.path.exists(gpu.LIB_PATH):
        pytest.fail("libpsy_gpu.so not built: run `python -m psychopath_cpp_amd.build`")
    lib = C.CDLL(gpu.LIB_PATH)
    missing = [s for s in header_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    lib.psy_abi_version.restype = C.c_int
    assert lib.psy_abi_version() == 3


def test_oracle_is_not_linked_by_product():
    from psychopath_cpp_amd import gpu
    with open(gpu.LIB_PATH, "rb") as f:
        blob = f.read()
    assert b"liboracle" not in blob and b"ora_trace" not in blob
