"""The script-style front end assembles exactly what the direct device API
assembles (CSR pattern, values and right-hand side bit for bit), and a
GMG-preconditioned solve on it converges."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("domain,refs", [("unit_cube_hex", 3), ("unit_square_tri", 5), ("unit_cube_tet", 2)])
def test_script_matches_direct_api(domain, refs):
    say(f"case script {domain} refs={refs}")
    import ugcore_amd as U
    from ugcore_amd import script as S
    dom = S.Domain(domain, refs)
    kind = S._KINDS[domain]
    ne = dom.num_elements()
    theta = np.random.default_rng(refs).normal(size=ne) * 0.5
    approx = S.ApproximationSpace(dom)
    approx.add_fct("c", "Lagrange", 1)
    e = S.ConvectionDiffusion("c", "Inner", "fe" if kind == 0 else "fv1")
    e.set_diffusion(S.ExpField(theta))
    if kind == 2:
        e.set_velocity((20.0, 10.0, 5.0))
        e.set_upwind(S.FullUpwind())
        dom.set_jitter(11, 0.2 / np.sqrt(3.0))
    e.set_source(1.0)
    bnd = S.DirichletBoundary()
    bnd.add(0.0, "c", "Boundary")
    sd = S.DomainDiscretization(approx)
    sd.add(e)
    sd.add(bnd)
    A, b = sd.assemble_linear()

    d = U.DomainDiscretization(kind, refs)
    d.set_theta(theta)
    if kind == 2:
        d.set_velocity((20.0, 10.0, 5.0))
        d.set_jitter(11, 0.2 / np.sqrt(3.0))
    d.set_source(np.ones(d.num_dofs()) if kind == 1 else 1.0)
    A2 = U.SparseMatrix()
    b2 = U.Vector(d.num_dofs())
    d.assemble_linear(A2, b2)
    for x, y in zip(A.download_csr(), A2.download_csr()):
        assert np.array_equal(x, y)
    assert np.array_equal(b.numpy(), b2.numpy())

    g = U.GeometricMultiGrid(sd.product, base_level=0, precision="f64", omega=0.6 if kind == 2 else 0.66)
    solver = (U.BiCGStab if kind == 2 else U.CG)(g, max_steps=200, reduction=1e-10)
    solver.init(A)
    x = U.Vector(sd.product.num_dofs())
    assert solver.apply(x, b)
    r = U.Vector(values=b.numpy())
    A.apply_sub(r, x)
    assert np.linalg.norm(r.numpy()) <= 1e-9 * np.linalg.norm(b.numpy())


@pytest.mark.parametrize("domain,refs", [("unit_cube_hex", 4), ("unit_square_tri", 6), ("unit_cube_tet", 2)])
def test_script_against_the_oracle(domain, refs):
    """the script path (not the direct API) against the oracle: pattern
    bit-exact, entries within 1e-12 of the row maximum, right-hand side within
    1e-12, and the script-configured GMG solve within 1e-9 of the oracle's
    solve with Krylov steps within +-1"""
    say(f"case script vs oracle {domain} refs={refs}")
    import ugcore_amd as U
    from ugcore_amd import script as S
    import oracle_lib as O
    dom = S.Domain(domain, refs)
    kind = S._KINDS[domain]
    n_fine = (3 if kind == 2 else 1) << refs
    theta = O.make_theta(n_fine, 5, 3) if kind != 1 else \
        np.random.default_rng(refs).normal(size=dom.num_elements()) * 0.5
    approx = S.ApproximationSpace(dom)
    approx.add_fct("c", "Lagrange", 1)
    e = S.ConvectionDiffusion("c", "Inner", "fe" if kind == 0 else "fv1")
    e.set_diffusion(S.ExpField(theta))
    vel = (20.0, 10.0, 5.0)
    jit = (11, 0.2 / np.sqrt(3.0))
    if kind == 2:
        e.set_velocity(vel)
        e.set_upwind(S.FullUpwind())
        dom.set_jitter(*jit)
    e.set_source(1.0)
    bnd = S.DirichletBoundary()
    bnd.add(0.0, "c", "Boundary")
    sd = S.DomainDiscretization(approx)
    sd.add(e)
    sd.add(bnd)
    A, b = sd.assemble_linear()

    p = O.Problem(kind, refs)
    p.set_theta(theta)
    if kind == 2:
        p.set_tet_params(vel=vel, jitter_seed=jit[0], jitter_amp=jit[1])
    if kind == 1:
        p.set_source_top(np.ones(p.num_dofs(refs)))    # the FV1 source is given per DoF
    p.assemble_all(1.0)
    rp, col, val = A.download_csr()
    R = p.A(refs)
    assert np.array_equal(rp.astype(np.int64), R.rowptr) and np.array_equal(col, R.col)
    worst = 0.0
    for r in range(len(rp) - 1):
        a, c = val[rp[r]:rp[r + 1]], R.val[rp[r]:rp[r + 1]]
        worst = max(worst, np.abs(a - c).max() / max(np.abs(c).max(), 1e-300))
    say(f"  entries: worst row-scaled difference {worst:.2e}")
    assert worst <= 1e-12
    bo = p.rhs(refs)
    assert np.abs(b.numpy() - bo).max() <= 1e-12 * np.abs(bo).max()

    omega = 0.6 if kind == 2 else 0.66
    solver_name = "bicgstab" if kind == 2 else "cg"
    ref = p.solve(solver=solver_name, base_lev=0, omega=omega, rel_red=1e-10)
    assert ref["ok"]
    g = U.GeometricMultiGrid(sd.product, base_level=0, precision="f64", omega=omega)
    solver = (U.BiCGStab if kind == 2 else U.CG)(g, max_steps=200, reduction=1e-10)
    solver.init(A)
    x = U.Vector(sd.product.num_dofs())
    assert solver.apply(x, b)
    err = np.linalg.norm(x.numpy() - ref["x"]) / np.linalg.norm(ref["x"])
    say(f"  {solver_name}: {solver.steps()} steps (oracle {ref['steps']}), solution rel L2 difference {err:.2e}")
    assert abs(solver.steps() - ref["steps"]) <= 1
    assert err <= 1e-9
