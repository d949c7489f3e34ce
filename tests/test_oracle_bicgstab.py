"""BiCGStab restarts in the oracle (BiCGStab::set_restart and
set_min_orthogonality, bicgstab.h:157-215, 386-390). The reference solver
needs boost and cannot be built here; pins: a restart every k steps leaves the
first k defects bit-identical to the run without restarts (nothing differs
before step k), every variant solves the system (scipy direct solve), and
min_orthogonality 1 restarts at every step (|(r0,r)| <= ||r0|| ||r|| by
Cauchy-Schwarz) and still converges."""
import numpy as np
import pytest
import scipy.sparse.linalg as spl

import oracle_lib as O

VEL = (20.0, 10.0, 5.0)


def tet_problem(refs=2):
    p = O.Problem(2, refs)
    p.set_theta(O.make_theta(3 * 2 ** refs, 5, 3))
    p.set_tet_params(vel=VEL, jitter_seed=11, jitter_amp=0.2 / np.sqrt(3.0))
    p.assemble_all()
    return p


@pytest.mark.parametrize("k", [3, 5])
def test_fixed_restart_keeps_the_first_steps(k):
    p = tet_problem()
    r0 = p.solve("bicgstab", base_lev=0, omega=0.6, rel_red=1e-10)
    p.set_bicgstab_options(num_restarts=k)
    r1 = p.solve("bicgstab", base_lev=0, omega=0.6, rel_red=1e-10)
    assert r0["ok"] and r1["ok"]
    # BiCGStab records two defects per step (s and r)
    n = 2 * k + 1
    assert np.array_equal(r0["defects"][:n], r1["defects"][:n])
    assert not np.array_equal(r0["defects"][: n + 2], r1["defects"][: n + 2])
    x_direct = spl.spsolve(p.A(2).to_scipy().tocsc(), p.rhs(2))
    for r in (r0, r1):
        assert np.linalg.norm(r["x"] - x_direct) <= 1e-8 * np.linalg.norm(x_direct)


@pytest.mark.parametrize("m", [0.3, 1.0])
def test_min_orthogonality_restarts(m):
    p = tet_problem()
    r0 = p.solve("bicgstab", base_lev=0, omega=0.6, rel_red=1e-10)
    p.set_bicgstab_options(min_ortho=m)
    r1 = p.solve("bicgstab", base_lev=0, omega=0.6, rel_red=1e-10)
    assert r1["ok"]
    x_direct = spl.spsolve(p.A(2).to_scipy().tocsc(), p.rhs(2))
    assert np.linalg.norm(r1["x"] - x_direct) <= 1e-8 * np.linalg.norm(x_direct)
    # the first step is the same (restart at step 0 either way)
    assert np.array_equal(r0["defects"][:3], r1["defects"][:3])
    if m == 1.0:
        assert not np.array_equal(r0["defects"], r1["defects"])
