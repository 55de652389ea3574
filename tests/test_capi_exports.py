"""The C-ABI library loads and exports every function include/helios_ccsd.h declares (no GPU)."""
import ctypes
import os
import re

import pytest

from helios_amd import capi


def declared(root):
    text = open(os.path.join(root, "include", "helios_ccsd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hcc_[a-z0-9_]+)\s*\(", text)))


def test_header_lists_the_interface(root):
    names = declared(root)
    for must in ("hcc_create", "hcc_set_molecule", "hcc_iterate", "hcc_energy", "hcc_solve",
                 "hcc_batch_load", "hcc_batch_solve", "hcc_destroy", "hcc_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol(root):
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("library not built")
    lib = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in declared(root) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(declared(root)) <= set(capi.SIGNATURES)


def test_argument_errors_without_device():
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("library not built")
    lib = capi.load()
    h = ctypes.c_void_p()
    assert lib.hcc_create(ctypes.byref(h), 0, 7, 20) == -1       # bad precision
    assert lib.hcc_create(ctypes.byref(h), 0, 0, 1) == -1        # DIIS depth below 2
    assert lib.hcc_create(ctypes.byref(h), 0, 0, 99) == -1       # DIIS depth above the cap
    assert lib.hcc_create(None, 0, 0, 20) == -1
    assert lib.hcc_last_error(None) == b"null context"
