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


def test_q2_argument_validation_before_any_device_work():
    """short coefficient arrays and lattice sizes whose pattern would not fit
    int32 are refused before a device buffer is made (no GPU needed)"""
    import pytest
    X = np.zeros((3, 8, 3))
    with pytest.raises(ValueError, match="one coefficient per element"):
        U.q2_hex_element_matrices(X, np.ones(2))
    with pytest.raises(ValueError, match="one coefficient per element"):
        U.q2_hex_element_matrices(X, 1.0)
    with pytest.raises(ValueError, match="shape"):
        U.q2_hex_element_matrices(np.zeros((3, 4, 3)))
    with pytest.raises(ValueError, match="one value per element"):
        U.q2_assemble_lattice(3, np.zeros(26))
    for n in (0, U.Q2_LATTICE_MAX_N + 1, 600):
        with pytest.raises(ValueError, match="int32"):
            U.q2_assemble_lattice(n)
    # the C entry point itself refuses n = 162 ((8n+1)^3 > 2^31 - 1) up front
    assert (8 * 161 + 1) ** 3 < 2 ** 31 <= (8 * 162 + 1) ** 3
    assert U.lib().ugx_q2_assemble_lattice(162, None, 1.0, None, None) < 0
    assert b"1..161" in U.lib().ugx_last_error()
    assert U.lib().ugx_q2_hex_element_matrices(-1, None, None, 0.0, None, None) < 0
