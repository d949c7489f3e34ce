"""The C client (tests/c_client/ugx_client.c) uses only include/ugcore_amd.h
and libugcore_amd.so, as a binding on the reference side would: it must link
against the in-tree library (CPU) and solve a Q1 problem to 1e-9 (GPU)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLIENT = os.path.join(ROOT, "tests", "c_client", "_build", "ugx_client")


def test_client_is_built_and_links_the_library():
    assert os.path.exists(CLIENT), "build with __graft_entry__.build() (make -C tests/c_client)"
    out = subprocess.run(["ldd", CLIENT], capture_output=True, text=True, check=True).stdout
    line = [l for l in out.splitlines() if "libugcore_amd.so" in l]
    assert line and "not found" not in line[0]
    assert os.path.realpath(line[0].split("=>")[1].split()[0]) == \
        os.path.realpath(os.path.join(ROOT, "ugcore_amd", "libugcore_amd.so"))


@pytest.mark.gpu
@pytest.mark.parametrize("refs", [4, 6])
def test_client_solves(refs):
    print(f"case C client refs={refs}", flush=True)
    r = subprocess.run([CLIENT, str(refs)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    m = re.search(r"converged=(\d) steps=(\d+) lattice_path=(\d) relres=(\S+)", r.stdout)
    assert m and m.group(1) == "1" and m.group(3) == "1"
    assert float(m.group(4)) <= 1e-9
    assert 1 <= int(m.group(2)) <= 30
