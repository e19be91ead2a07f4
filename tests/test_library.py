"""The C-ABI library loads and exports every entry point of include/simc_gpu.h
(no compute calls: runs without a GPU)."""
import ctypes
import os
import re

from simc_gfortran_amd import api

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(REPO, "include", "simc_gpu.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|int64_t|const char \*)\s*\*?\s*(simc_\w+)\s*\(",
                                 src, re.M)))


def test_header_lists_every_export():
    assert declared_symbols() == sorted(api.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(api.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_version_string():
    assert api.library().simc_version().decode().startswith("simc_gpu")
