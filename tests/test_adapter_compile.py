"""The bridge adapter (include/ug_gpu/ug_bridge_adapter.hpp) compiles against
the reference's own interface headers (ILinearOperator, ILinearIterator,
IPreconditioner, MatrixOperator, AlgebraType, SetDirichletRow and the matrix
product templates), and the reference's generic code instantiates with its
classes (tests/adapter/adapter_check.cpp). Skipped where the reference
checkout is absent; nothing here is linked into libugcore_amd.so."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
UGBASE = "/root/reference/ugbase"


@pytest.mark.skipif(not os.path.isdir(UGBASE), reason="reference checkout absent")
@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ absent")
def test_adapter_compiles_against_reference_interfaces():
    cmd = ["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-DNDEBUG", f"-I{UGBASE}",
           "-I" + os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "adapter", "adapter_check.cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
