"""The C-ABI library loads without a GPU and exports every function that
include/imc_hip.h declares (no compute calls here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not fn.endswith(".h"):
            continue
        txt = open(os.path.join(ROOT, "include", fn)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*([a-z][a-z0-9_]*)\s*\(", txt, flags=re.M)
    return names


def test_header_parses():
    names = declared_functions()
    assert "imc2d_transport" in names and "imc_last_error" in names
    assert len(names) == len(set(names))


def test_library_exports_all_declared_symbols():
    from radtranbook_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libimc_hip.so not built")
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    # every declared symbol is bound by the Python loader and vice versa
    assert set(declared_functions()) == set(_lib.SIGNATURES)


def test_loader_and_device_count_without_gpu():
    from radtranbook_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libimc_hip.so not built")
    L = _lib.load()
    assert L.imc_device_count() >= 0
    assert isinstance(L.imc_last_error(), (bytes, type(None)))
