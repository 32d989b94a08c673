"""The C ABI library loads and exports every entry point its header declares."""
import ctypes
import os
import re

from conftest import REPO

import pawsim_amd


def declared_symbols():
    text = open(os.path.join(REPO, "include", "pawsim_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pg_[a-z_]+)\s*\(", text)))


def test_header_declares_binding_symbols():
    assert sorted(pawsim_amd.ABI_SYMBOLS) == declared_symbols()


def test_library_exports_declared_symbols():
    lib = ctypes.CDLL(pawsim_amd.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_loads():
    lib = pawsim_amd.load_library()
    assert lib.pg_last_error(None) == b"null handle"


def test_driver_built():
    assert os.access(pawsim_amd.DRIVER_PATH, os.X_OK)
