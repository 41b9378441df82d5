"""The C-ABI library loads and exports every entry point of include/wtb_hip.h
(no device calls: runs without a GPU)."""
import ctypes
import os
import re

from conftest import ROOT
from wtb_amd import hip


def header_functions():
    src = open(os.path.join(ROOT, "include", "wtb_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(wtb_\w+)\s*\(", src, flags=re.M)))


def test_header_declares_the_boundary():
    fns = header_functions()
    assert set(hip.EXPORTS) == set(fns)


def test_library_exports_every_symbol():
    lib = ctypes.CDLL(hip.LIB_PATH)
    for name in header_functions():
        assert hasattr(lib, name), name


def test_binding_loads():
    lib = hip.load()
    assert lib.wtb_last_error(None) == b"null context"
