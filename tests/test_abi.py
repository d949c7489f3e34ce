"""CPU tests of the C-ABI library: it loads, exports every entry point that
include/*.h declares, and reports errors through ugx_last_error (no device
computation is run here)."""
import ctypes as C
import glob
import os
import re

import numpy as np

import ugcore_amd as U

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(ugx_\w+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(U.LIB_PATH)
    names = declared_entry_points()
    assert len(names) > 50
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_bindings_cover_the_abi():
    assert set(declared_entry_points()) == set(U._SIGNATURES)


def test_host_numbering_error_path():
    out = np.zeros(8, np.int32)
    rc = U.lib().ugx_hierarchy_numbering(0, 3, 7, out.ctypes.data, None)
    assert rc < 0
    assert b"out of range" in U.lib().ugx_last_error()
