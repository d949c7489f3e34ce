"""GPU parity tests of the tetrahedral path (unit cube as 3x3x3 hexahedra x 6
tetrahedra, regular refinement; FV1 convection-diffusion with full upwind,
velocity 20*(1, 0.5, 0.25), k = exp(theta) per fine cell, jittered top level;
BiCGStab + GMG V(2,2) with damped Jacobi 0.6; adjoint through the transposed
operator) against the CPU oracle on the same seeded inputs.

Tolerances as for the hexahedral path: pattern and DoF indexing bit-exact,
entries <= 1e-12 of the row maximum, solutions <= 1e-9 (relative L2),
iteration counts within +-1.
"""
import sys

import numpy as np
import pytest
import scipy.sparse.linalg as spl

import oracle_lib as O

pytestmark = pytest.mark.gpu

VEL = (20.0, 10.0, 5.0)
# jitter: displacement length <= 0.2h (+-0.2h/sqrt(3) per component)
JSEED, JAMP = 11, 0.2 / np.sqrt(3.0)
OMEGA = 0.6


def U():
    import ugcore_amd
    return ugcore_amd


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


_cache = {}


def oracle_problem(refs, seed=7):
    if refs not in _cache:
        p = O.Problem(2, refs)
        theta = O.make_theta(p.n_fine, seed, 3)
        p.set_theta(theta)
        p.set_tet_params(vel=VEL, jitter_seed=JSEED, jitter_amp=JAMP)
        p.assemble_all(1.0)
        _cache[refs] = (p, theta)
    return _cache[refs]


def gpu_problem(refs, theta):
    u = U()
    d = u.DomainDiscretization(2, refs)
    d.set_theta(theta)
    d.set_velocity(VEL)
    d.set_jitter(JSEED, JAMP)
    d.set_source(1.0)
    A = u.SparseMatrix()
    b = u.Vector(d.num_dofs())
    d.assemble_linear(A, b)
    return d, A, b


def rows_close(rp, val_a, val_b, tol=1e-12):
    worst = 0.0
    for r in range(len(rp) - 1):
        a = val_a[rp[r]:rp[r + 1]]
        bb = val_b[rp[r]:rp[r + 1]]
        if bb.size:
            worst = max(worst, np.abs(a - bb).max() / max(np.abs(bb).max(), 1e-300))
    return worst <= tol, worst


@pytest.mark.parametrize("refs", [1, 2, 3])
def test_surface_matrix_and_rhs(refs):
    say(f"case tet surface refs={refs}")
    p, theta = oracle_problem(refs)
    d, A, b = gpu_problem(refs, theta)
    rp, col, val = A.download_csr()
    ref = p.A(refs)
    assert np.array_equal(rp.astype(np.int64), ref.rowptr)
    assert np.array_equal(col, ref.col)
    ok, worst = rows_close(ref.rowptr, val, ref.val)
    assert ok, worst
    bo = p.rhs(refs)
    assert np.abs(b.numpy() - bo).max() <= 1e-12 * np.abs(bo).max()


def test_level_operators():
    """every level operator (full 27-diagonal lattice storage, Dirichlet rows
    identity) equals the oracle's level matrix entry by entry"""
    refs = 2
    say(f"case tet levels refs={refs}")
    u = U()
    p, theta = oracle_problem(refs)
    d, A, b = gpu_problem(refs, theta)
    g = u.GeometricMultiGrid(d, base_level=0, omega=OMEGA, precision="f64")
    g.init(A)
    for lev in range(refs + 1):
        W = d.lattice_width(lev)
        dia = g.level_dia(lev)
        assert dia.shape[0] == 27
        lat = p.dof_lattice(lev) // (2 ** (refs - lev))
        li = (lat[:, 2] * W + lat[:, 1]) * W + lat[:, 0]
        M = p.A(lev)
        seen = np.zeros_like(dia, dtype=bool)
        for r in range(M.n_rows):
            cols = M.col[M.rowptr[r]:M.rowptr[r + 1]]
            vals = M.val[M.rowptr[r]:M.rowptr[r + 1]]
            scale = np.abs(vals).max()
            for c, v in zip(cols, vals):
                dl = lat[c] - lat[r]
                k = (dl[2] + 1) * 9 + (dl[1] + 1) * 3 + (dl[0] + 1)
                assert abs(dia[k, li[r]] - v) <= 1e-12 * scale
                seen[k, li[r]] = True
        assert np.all(dia[~seen] == 0.0)


@pytest.mark.parametrize("refs", [2, 3])
def test_gmg_application(refs):
    say(f"case tet gmg apply refs={refs}")
    u = U()
    p, theta = oracle_problem(refs)
    d, A, b = gpu_problem(refs, theta)
    g = u.GeometricMultiGrid(d, base_level=0, nu1=2, nu2=2, omega=OMEGA, precision="f64")
    g.init(A)
    rng = np.random.default_rng(5)
    dv = rng.normal(size=p.num_dofs(refs))
    dv[p.dirichlet(refs) == 1] = 0.0
    c_orc = p.gmg_apply(dv, base_lev=0, omega=OMEGA)
    c = u.Vector(len(dv))
    g.apply(c, u.Vector(values=dv))
    assert np.linalg.norm(c.numpy() - c_orc) <= 1e-12 * np.linalg.norm(c_orc)


@pytest.mark.parametrize("refs,precision", [(2, "f64"), (3, "f64"), (2, "f32"), (3, "f32")])
def test_bicgstab_gmg_solve(refs, precision):
    say(f"case tet solve refs={refs} precision={precision}")
    u = U()
    p, theta = oracle_problem(refs)
    r_orc = p.solve(solver="bicgstab", base_lev=0, omega=OMEGA, rel_red=1e-10, max_steps=300)
    assert r_orc["ok"]
    d, A, b = gpu_problem(refs, theta)
    g = u.GeometricMultiGrid(d, base_level=0, omega=OMEGA, precision=precision)
    S = u.BiCGStab(g, max_steps=300, reduction=1e-10)
    S.init(A)
    x = u.Vector(d.num_dofs())
    assert S.apply(x, b)
    xo = r_orc["x"]
    assert np.linalg.norm(x.numpy() - xo) <= 1e-9 * np.linalg.norm(xo)
    dg, do = S.defects(), r_orc["defects"]
    say(f"  steps gpu {S.steps()} oracle {r_orc['steps']}; defects gpu {len(dg)} oracle {len(do)}")
    if precision == "f64":
        # the same iteration: the defect histories agree to round-off
        m = min(len(dg), len(do))
        rel = np.abs(dg[:m] - do[:m]) / do[:m]
        assert rel[: m // 2].max() <= 1e-9, rel.max()
        assert abs(S.steps() - r_orc["steps"]) <= 1, (S.steps(), r_orc["steps"])
    else:
        # FP32 level operators: the same solution, comparable steps
        assert S.steps() <= r_orc["steps"] + 3, (S.steps(), r_orc["steps"])


def test_adjoint_operator_and_transposed_solve():
    refs = 2
    say(f"case tet adjoint refs={refs}")
    u = U()
    p, theta = oracle_problem(refs)
    d, A, b = gpu_problem(refs, theta)
    AT = d.adjoint_operator(A)
    rp, col, val = AT.download_csr()
    ref = p.A(refs)
    assert np.array_equal(rp.astype(np.int64), ref.rowptr) and np.array_equal(col, ref.col)
    import scipy.sparse as sp
    dm = p.dirichlet(refs).astype(bool)
    inner = np.where(~dm)[0]
    def elim_T(rp_, col_, val_):
        n = len(rp_) - 1
        rows = np.repeat(np.arange(n), np.diff(rp_))
        keep = ~((~dm[rows]) & dm[col_])            # Dirichlet columns of interior rows
        E = sp.csr_matrix((val_ * keep, (rows, col_)), shape=(n, n))
        return E.T.toarray()

    dense_got = np.zeros((len(dm), len(dm)))
    for r in range(len(dm)):
        dense_got[r, col[rp[r]:rp[r + 1]]] = val[rp[r]:rp[r + 1]]
    # exactly the transpose of the device's own (eliminated) operator ...
    arp, acol, aval = A.download_csr()
    assert np.abs(dense_got - elim_T(arp, acol, aval)).max() == 0.0
    # ... and of the oracle's to round-off
    ETo = elim_T(ref.rowptr, ref.col, ref.val)
    assert np.abs(dense_got - ETo).max() <= 1e-12 * np.abs(ETo).max()
    # transposed GMG + BiCGStab against the oracle's adjoint solve
    rng = np.random.default_rng(9)
    rhs = rng.normal(size=len(dm))
    lam_o, steps_o = p.solve_adjoint(rhs, omega=OMEGA, rel_red=1e-10)
    gT = u.GeometricMultiGrid(d, base_level=0, omega=OMEGA, precision="f64", transposed=True)
    ST = u.BiCGStab(gT, max_steps=300, reduction=1e-10)
    ST.init(AT)
    r = u.Vector(values=rhs)
    d.adjust_defect(r)
    lam = u.Vector(len(dm))
    assert ST.apply(lam, r)
    assert abs(ST.steps() - steps_o) <= 1, (ST.steps(), steps_o)
    assert np.linalg.norm(lam.numpy() - lam_o) <= 1e-9 * np.linalg.norm(lam_o)
    # the interior part solves A_II^T lambda_I = rhs_I
    AII = ref.to_scipy()[inner][:, inner]
    li = spl.spsolve(AII.T.tocsc(), rhs[inner])
    assert np.linalg.norm(lam.numpy()[inner] - li) <= 1e-8 * np.linalg.norm(li)


@pytest.mark.parametrize("refs", [1, 2])
def test_adjoint_gradient_kernel(refs):
    say(f"case tet gradient refs={refs}")
    u = U()
    p, theta = oracle_problem(refs)
    d, A, b = gpu_problem(refs, theta)
    rng = np.random.default_rng(3)
    n = p.num_dofs(refs)
    uu = rng.normal(size=n)
    ll = rng.normal(size=n)
    g_o = p.adjoint_gradient(uu, ll)
    g = d.adjoint_gradient(u.Vector(values=uu), u.Vector(values=ll))
    assert g.shape == g_o.shape
    assert np.abs(g - g_o).max() <= 1e-12 * np.abs(g_o).max()


def test_gradient_through_transposed_operator_matches_oracle():
    """dJ/dtheta for J = 1/2 |u(obs) - d|^2: forward BiCGStab+GMG, adjoint with
    the transposed operator and transposed GMG, element kernel; against the
    oracle's two-solve gradient"""
    refs = 2
    say(f"case tet two-solve gradient refs={refs}")
    u = U()
    p, theta = oracle_problem(refs)
    n = p.num_dofs(refs)
    dm = p.dirichlet(refs).astype(bool)
    rng = np.random.default_rng(21)
    obs = rng.choice(np.where(~dm)[0], 50, replace=False)
    dobs = rng.normal(size=50) * 1e-3
    r_o = p.solve(solver="bicgstab", base_lev=0, omega=OMEGA, rel_red=1e-12, max_steps=300)
    u_o = r_o["x"]
    dJ = np.zeros(n)
    dJ[obs] = u_o[obs] - dobs
    lam_o, _ = p.solve_adjoint(dJ, omega=OMEGA, rel_red=1e-12)
    g_o = p.adjoint_gradient(u_o, lam_o)

    d, A, b = gpu_problem(refs, theta)
    g = u.GeometricMultiGrid(d, base_level=0, omega=OMEGA, precision="f64")
    S = u.BiCGStab(g, max_steps=300, reduction=1e-12)
    S.init(A)
    x = u.Vector(n)
    assert S.apply(x, b)
    xv = x.numpy()
    dJg = np.zeros(n)
    dJg[obs] = xv[obs] - dobs
    AT = d.adjoint_operator(A)
    gT = u.GeometricMultiGrid(d, base_level=0, omega=OMEGA, precision="f64", transposed=True)
    ST = u.BiCGStab(gT, max_steps=300, reduction=1e-12)
    ST.init(AT)
    r = u.Vector(values=dJg)
    d.adjust_defect(r)
    lam = u.Vector(n)
    assert ST.apply(lam, r)
    grad = d.adjoint_gradient(x, lam)
    assert np.linalg.norm(grad - g_o) <= 1e-6 * np.linalg.norm(g_o)


def test_full_size_96_properties():
    """the 96^3 x 6 tetrahedra configuration (912,673 DoFs, 13,465,441 nnz):
    FP32- and FP64-preconditioned BiCGStab both reach 1e-10 and agree"""
    refs = 5
    say(f"case tet 96^3 refs={refs}")
    u = U()
    from ugcore_amd.synthetic import make_theta
    d = u.DomainDiscretization(2, refs)
    d.set_theta(make_theta(d.n_fine, 7, 3))
    d.set_velocity(VEL)
    d.set_jitter(JSEED, JAMP)
    d.set_source(1.0)
    A = u.SparseMatrix()
    b = u.Vector(d.num_dofs())
    d.assemble_linear(A, b)
    assert d.num_dofs() == 97 ** 3
    assert A.nnz() == 13465441
    sols = {}
    for prec in ("f32", "f64"):
        g = u.GeometricMultiGrid(d, base_level=0, omega=OMEGA, precision=prec)
        S = u.BiCGStab(g, max_steps=500, reduction=1e-10)
        S.init(A)
        x = u.Vector(d.num_dofs())
        assert S.apply(x, b), prec
        res = u.Vector(d.num_dofs())
        A.axpy(res, 1.0, b, -1.0, x)
        assert np.linalg.norm(res.numpy()) <= 1.01e-10 * np.linalg.norm(b.numpy())
        sols[prec] = x.numpy()
        say(f"  {prec}: {S.steps()} steps")
    assert np.linalg.norm(sols["f32"] - sols["f64"]) <= 1e-8 * np.linalg.norm(sols["f64"])


@pytest.mark.parametrize("num_restarts,min_ortho", [(3, 0.0), (0, 0.3), (4, 0.2)])
def test_bicgstab_restarts_match_oracle(num_restarts, min_ortho):
    """BiCGStab::set_restart / set_min_orthogonality (bicgstab.h:157-215):
    the same restarted iteration on the GPU and in the oracle (FP64 levels)."""
    say(f"case tet bicgstab restarts={num_restarts} min_ortho={min_ortho}")
    u = U()
    refs = 2
    p, theta = oracle_problem(refs)
    p.set_bicgstab_options(num_restarts=num_restarts, min_ortho=min_ortho)
    try:
        r_orc = p.solve(solver="bicgstab", base_lev=0, omega=OMEGA, rel_red=1e-10, max_steps=300)
    finally:
        p.set_bicgstab_options()
    assert r_orc["ok"]
    d, A, b = gpu_problem(refs, theta)
    g = u.GeometricMultiGrid(d, base_level=0, omega=OMEGA, precision="f64")
    S = u.BiCGStab(g, max_steps=300, reduction=1e-10)
    S.set_restart(num_restarts)
    S.set_min_orthogonality(min_ortho)
    S.init(A)
    x = u.Vector(d.num_dofs())
    assert S.apply(x, b)
    assert abs(S.steps() - r_orc["steps"]) <= 1, (S.steps(), r_orc["steps"])
    xo = r_orc["x"]
    assert np.linalg.norm(x.numpy() - xo) <= 1e-9 * np.linalg.norm(xo)
    dg, do = S.defects(), r_orc["defects"]
    m = min(len(dg), len(do))
    assert (np.abs(dg[:m] - do[:m]) / do[:m])[: m // 2].max() <= 1e-9
