"""GPU Galerkin coarse operators (AssembledMultiGridCycle::set_rap,
mg_solver.h:149; init_rap_operator, mg_solver_impl.hpp:829-1013) against the
oracle's R A P levels, which are pinned bit for bit to the reference's
AddMultiplyOf. Tolerances: level entries 1e-12 of the row scale, GMG
application 1e-12, solution 1e-9 (relative L2), Krylov steps +-1."""
import sys

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


def U():
    import ugcore_amd
    return ugcore_amd


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def problems(kind, refs, seed=7):
    """Oracle with Galerkin levels and the GPU discretization of the same
    problem (hexahedra: seeded theta; triangles: the seeded Poisson source)."""
    u = U()
    p = O.Problem(kind, refs)
    n = 2 ** refs
    theta = f = None
    if kind == 0:
        theta = O.make_theta(n, seed, 3)
        p.set_theta(theta)
    else:
        from ugcore_amd.synthetic import poisson_source_lattice
        fl = poisson_source_lattice(n, seed)
        lat = p.dof_lattice(refs)
        f = fl[lat[:, 1] * (n + 1) + lat[:, 0]]
        p.set_source_top(f)
    p.assemble_all()
    p.set_rap()
    d = u.DomainDiscretization(kind, refs)
    if theta is not None:
        d.set_theta(theta)
    if f is not None:
        d.set_source(f)
    A = u.SparseMatrix()
    b = u.Vector(d.num_dofs())
    d.assemble_linear(A, b)
    return p, d, A, b


@pytest.mark.parametrize("kind,refs", [(0, 3), (1, 5)])
def test_galerkin_level_operators(kind, refs):
    """Every coarse level operator equals the oracle's R A P entry by entry
    (interior couplings to Dirichlet columns eliminated, as in the lattice
    storage); entries outside the pattern are 0."""
    say(f"case rap levels kind={kind} refs={refs}")
    u = U()
    p, d, A, b = problems(kind, refs)
    g = u.GeometricMultiGrid(d, base_level=0, precision="f64", rap=True)
    g.init(A)
    dim = 3 if kind == 0 else 2
    for lev in range(refs):
        W = 2 ** lev + 1
        dia = g.level_dia(lev)
        lat = p.dof_lattice(lev) // (2 ** (refs - lev))
        li = (lat[:, 2] * W + lat[:, 1]) * W + lat[:, 0] if dim == 3 else lat[:, 1] * W + lat[:, 0]
        M = p.A(lev)
        dm = p.dirichlet(lev).astype(bool)
        seen = np.zeros_like(dia, dtype=bool)
        for r in range(M.n_rows):
            cols = M.col[M.rowptr[r]:M.rowptr[r + 1]]
            vals = M.val[M.rowptr[r]:M.rowptr[r + 1]]
            scale = np.abs(vals).max()
            for c, v in zip(cols, vals):
                if not dm[r] and dm[c]:
                    v = 0.0
                dl = lat[c] - lat[r]
                k = (dl[2] + 1) * 9 + (dl[1] + 1) * 3 + (dl[0] + 1) if dim == 3 else (dl[1] + 1) * 3 + (dl[0] + 1)
                assert abs(dia[k, li[r]] - v) <= 1e-12 * scale, (lev, r, c)
                seen[k, li[r]] = True
        assert np.all(dia[~seen] == 0.0)


@pytest.mark.parametrize("kind,refs", [(0, 4), (1, 6)])
def test_galerkin_gmg_application(kind, refs):
    say(f"case rap gmg apply kind={kind} refs={refs}")
    u = U()
    p, d, A, b = problems(kind, refs)
    g = u.GeometricMultiGrid(d, base_level=0, nu1=2, nu2=2, omega=0.66, precision="f64", rap=True)
    g.init(A)
    rng = np.random.default_rng(5)
    dv = rng.normal(size=p.num_dofs(refs))
    dv[p.dirichlet(refs) == 1] = 0.0
    c_orc = p.gmg_apply(dv, base_lev=0)
    c = u.Vector(len(dv))
    g.apply(c, u.Vector(values=dv))
    assert np.linalg.norm(c.numpy() - c_orc) <= 1e-12 * np.linalg.norm(c_orc)


@pytest.mark.parametrize("kind,refs,precision", [(0, 4, "f64"), (0, 5, "f64"), (1, 7, "f64"), (0, 5, "f32")])
def test_galerkin_gmg_cg_solve(kind, refs, precision):
    say(f"case rap solve kind={kind} refs={refs} {precision}")
    u = U()
    p, d, A, b = problems(kind, refs)
    r_orc = p.solve(solver="cg", base_lev=0, rel_red=1e-10, max_steps=200)
    assert r_orc["ok"]
    g = u.GeometricMultiGrid(d, base_level=0, precision=precision, rap=True)
    S = u.CG(g, max_steps=200, reduction=1e-10)
    S.init(A)
    x = u.Vector(d.num_dofs())
    assert S.apply(x, b)
    if precision == "f64":
        assert abs(S.steps() - r_orc["steps"]) <= 1
    else:
        assert S.steps() <= r_orc["steps"] + 2
    xo = r_orc["x"]
    assert np.linalg.norm(x.numpy() - xo) <= 1e-9 * np.linalg.norm(xo)


def test_galerkin_rejected_on_tetrahedra():
    u = U()
    d = u.DomainDiscretization(2, 1)
    A = u.SparseMatrix()
    b = u.Vector(d.num_dofs())
    d.assemble_linear(A, b)
    g = u.GeometricMultiGrid(d, base_level=0, precision="f64", rap=True)
    with pytest.raises(u.UGError):
        g.init(A)
