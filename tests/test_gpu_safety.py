"""GPU tests of the paths that must not give a silent wrong answer, and of
failure handling:

* a matrix edited through its raw CSR pointers is no longer the
  discretization's assembly: CG solves the edited system on the surface CSR
  (never the matrix-free / lattice-order operator derived from the assembly);
* the F-cycle (mg_solver_impl.hpp:2115-2118) and gamma-cycles against the
  oracle, and invalid cycle types, precisions and smoothing counts refused;
* failure cases of the reference's solvers stop with false or an error:
  a NaN defect (StdConvCheck, convergence_check_impl.h:164), a zero diagonal
  (Jacobi set-up, jacobi.h:198-206), a singular base level (LU set-up,
  lu.h:236-269), CG breakdown lambda = 0 (cg.h:172-184).
"""
import ctypes as C
import sys

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

import oracle_lib as O

pytestmark = pytest.mark.gpu


def U():
    import ugcore_amd
    return ugcore_amd


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def q1_problem(refs, seed=7):
    p = O.Problem(0, refs)
    theta = O.make_theta(2 ** refs, seed, 3)
    p.set_theta(theta)
    p.assemble_all()
    u = U()
    d = u.DomainDiscretization(0, refs)
    d.set_theta(theta)
    A = u.SparseMatrix()
    b = u.Vector(d.num_dofs())
    d.assemble_linear(A, b)
    return p, d, A, b


def edit_diagonal(A, row, factor):
    """scale A(row, row) through the raw device pointers"""
    u = U()
    rp, col, val = A.download_csr()
    # Python ints: rp is int32 and a device address does not fit in it
    k = int(rp[row]) + int(np.nonzero(col[rp[row]:rp[row + 1]] == row)[0][0])
    new = np.array([val[k] * factor])
    _, _, vptr = A.device_csr()
    assert u.lib().ugx_memcpy_htod(C.c_void_p(vptr + 8 * k), new.ctypes.data_as(C.c_void_p), 8) == 0
    val[k] = new[0]
    return sp.csr_matrix((val, col, rp), shape=(len(rp) - 1, len(rp) - 1))


@pytest.mark.parametrize("order", ["init_then_edit", "edit_then_init"])
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_edited_matrix_is_solved_as_edited(order, precision):
    say(f"case edited matrix order={order} precision={precision}")
    u = U()
    refs = 4
    p, d, A, b = q1_problem(refs)
    g = u.GeometricMultiGrid(d, base_level=0, precision=precision)
    S = u.CG(g, max_steps=200, reduction=1e-12)
    ni = d.num_inner()
    row = ni // 2                        # an interior DoF
    if order == "init_then_edit":
        S.init(A)
        x0 = u.Vector(d.num_dofs())
        assert S.apply(x0, b) and S.used_lattice_path()
        As = edit_diagonal(A, row, 1.5)
        A.finalize()
    else:
        As = edit_diagonal(A, row, 1.5)
        A.finalize()
        S.init(A)
    x = u.Vector(d.num_dofs())
    assert S.apply(x, b)
    assert not S.used_lattice_path()
    xs = spl.spsolve(As.tocsc(), b.numpy())
    err = np.linalg.norm(x.numpy() - xs) / np.linalg.norm(xs)
    say(f"  steps {S.steps()}, rel err vs direct solve of the edited matrix {err:.2e}")
    assert err <= 1e-9
    # the unedited system has a different solution, so the check can fail
    x_orig = spl.spsolve(p.A(refs).to_scipy().tocsc(), b.numpy())
    assert np.linalg.norm(x_orig - xs) / np.linalg.norm(xs) > 1e-6


def test_raw_pointer_edit_without_finalize_uses_the_csr():
    """between ugx_matrix_device_csr and ugx_matrix_finalize SpMV reads the
    edited CSR itself"""
    say("case raw edit without finalize")
    u = U()
    p, d, A, b = q1_problem(3)
    As = edit_diagonal(A, d.num_inner() // 3, 2.0)
    xv = np.random.default_rng(3).normal(size=d.num_dofs())
    y = u.Vector(d.num_dofs())
    A.apply(y, u.Vector(values=xv))
    ref = As @ xv
    assert np.linalg.norm(y.numpy() - ref) <= 1e-13 * np.linalg.norm(ref)


@pytest.mark.parametrize("kind,refs,cycle,nu1,nu2", [
    (0, 4, -1, 2, 2), (0, 4, 3, 2, 2), (0, 4, -1, 1, 0), (0, 3, -1, 0, 1), (1, 6, -1, 2, 2), (1, 6, 3, 1, 1)])
def test_f_and_gamma_cycle_application(kind, refs, cycle, nu1, nu2):
    """c = B d for the F-cycle (one F-cycle then one V-cycle per coarser
    level) and a gamma = 3 cycle against the oracle's
    AssembledMultiGridCycle::apply"""
    say(f"case gmg cycle={cycle} kind={kind} refs={refs} nu=({nu1},{nu2})")
    u = U()
    p = O.Problem(kind, refs)
    n = 2 ** refs
    d = u.DomainDiscretization(kind, refs)
    if kind == 0:
        theta = O.make_theta(n, 7, 3)
        p.set_theta(theta)
        d.set_theta(theta)
    p.assemble_all()
    A = u.SparseMatrix()
    b = u.Vector(d.num_dofs())
    d.assemble_linear(A, b)
    rng = np.random.default_rng(11)
    dv = rng.normal(size=p.num_dofs(refs))
    dv[p.dirichlet(refs) == 1] = 0.0
    c_orc = p.gmg_apply(dv, base_lev=0, nu1=nu1, nu2=nu2, cycle=cycle)
    c_v = p.gmg_apply(dv, base_lev=0, nu1=nu1, nu2=nu2, cycle=1)
    assert np.linalg.norm(c_orc - c_v) > 1e-6 * np.linalg.norm(c_v)
    for precision, tol in (("f64", 1e-12), ("f32", 1e-5)):
        g = u.GeometricMultiGrid(d, base_level=0, nu1=nu1, nu2=nu2, cycle=cycle, precision=precision)
        g.init(A)
        c = u.Vector(len(dv))
        g.apply(c, u.Vector(values=dv))
        err = np.linalg.norm(c.numpy() - c_orc) / np.linalg.norm(c_orc)
        assert err <= tol, (precision, err)


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("cycle", ["F", "W"])
def test_f_and_w_cycle_cg_on_the_lattice_path(cycle, precision):
    """GMG-CG with an F- or W-cycle on the lattice path (graph replay of the
    cycle) against the oracle's solve: steps +-1, solution 1e-9"""
    say(f"case cg cycle={cycle} precision={precision}")
    u = U()
    refs = 5
    p, d, A, b = q1_problem(refs)
    cyc = {"F": -1, "W": 2}[cycle]
    r = p.solve(base_lev=0, cycle=cyc, rel_red=1e-10)
    assert r["ok"]
    g = u.GeometricMultiGrid(d, base_level=0, cycle=cycle, precision=precision)
    S = u.CG(g, max_steps=200, reduction=1e-10)
    S.init(A)
    x = u.Vector(d.num_dofs())
    assert S.apply(x, b)
    assert S.used_lattice_path()
    say(f"  steps {S.steps()} (oracle {r['steps']})")
    assert abs(S.steps() - r["steps"]) <= (1 if precision == "f64" else 2)
    assert np.linalg.norm(x.numpy() - r["x"]) <= 1e-9 * np.linalg.norm(r["x"])


def test_invalid_gmg_arguments_refused():
    say("case invalid gmg arguments")
    u = U()
    d = u.DomainDiscretization(0, 2)
    L = u.lib()
    for cycle in (0, -2, -7):
        assert not L.ugx_gmg_create(d.h, 0, 2, 2, cycle, 0.66, 0)
        assert b"cycle" in L.ugx_last_error()
        with pytest.raises(u.UGError):
            u.GeometricMultiGrid(d, cycle=cycle)
    for prec in (2, -1, 7):
        assert not L.ugx_gmg_create(d.h, 0, 2, 2, 1, 0.66, prec)
        assert b"precision" in L.ugx_last_error()
    with pytest.raises(ValueError):
        u.GeometricMultiGrid(d, precision="bf16")
    with pytest.raises(ValueError):
        u.GeometricMultiGrid(d, cycle="X")
    for nu1, nu2 in ((-1, 2), (2, -1)):
        assert not L.ugx_gmg_create(d.h, 0, nu1, nu2, 1, 0.66, 0)
    assert not L.ugx_gmg_create(d.h, -1, 2, 2, 1, 0.66, 0)
    A = u.SparseMatrix()
    b = u.Vector(d.num_dofs())
    d.assemble_linear(A, b)
    g = u.GeometricMultiGrid(d, base_level=3)          # above the top level (2)
    with pytest.raises(u.UGError):
        g.init(A)


def small_csr(dense):
    M = sp.csr_matrix(np.asarray(dense, np.float64))
    u = U()
    A = u.SparseMatrix()
    A.upload_csr(M.indptr, M.indices, M.data)
    return A


def test_nan_defect_stops_the_solver():
    """StdConvCheck ends the iteration on a non-finite defect and reports
    failure (convergence_check_impl.h:164)"""
    say("case NaN defect")
    u = U()
    n = 12
    T = 2 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    A = small_csr(T)
    bv = np.ones(n)
    bv[5] = np.nan
    for Solver in (u.CG, u.BiCGStab):
        S = Solver(u.Jacobi(0.66), max_steps=50, reduction=1e-10)
        S.init(A)
        x = u.Vector(n)
        assert not S.apply(x, u.Vector(values=bv))
        assert S.steps() <= 1
    # the oracle stops the same way
    ok, steps, _ = O.csr_solve(sp.csr_matrix(T), bv, max_steps=50)
    assert not ok and steps == 0


def test_zero_diagonal_fails_jacobi_setup():
    """Jacobi::preprocess fails on a zero diagonal entry (jacobi.h:198-206)"""
    say("case zero diagonal")
    u = U()
    T = np.array([[2.0, -1.0, 0.0], [-1.0, 0.0, -1.0], [0.0, -1.0, 2.0]])
    A = small_csr(T)
    S = u.CG(u.Jacobi(0.66), max_steps=20)
    with pytest.raises(u.UGError, match="zero diagonal"):
        S.init(A)
    # the raw kernel entry point reports it as -2
    rp, col, val = A.device_csr()
    dinv = u.DeviceArray(np.zeros(3))
    assert u.lib().ugx_jacobi_diag_inv_f64(3, rp, col, val, 0.66, dinv.ptr, None) == -2


def test_singular_base_level_fails_multigrid_setup():
    """a base level without a usable pivot (here k = 0 everywhere) makes the
    multigrid set-up fail with an error (LU::init_lu, lu.h:236-269; or the
    smoother's zero diagonals above it), not a garbage correction"""
    say("case singular base level")
    u = U()
    d = u.DomainDiscretization(0, 2)
    d.set_theta(np.full(4 ** 3, -np.inf))
    A = u.SparseMatrix()
    b = u.Vector(d.num_dofs())
    d.assemble_linear(A, b)
    for base in (0, 1):
        g = u.GeometricMultiGrid(d, base_level=base)
        with pytest.raises(u.UGError):
            g.init(A)


def test_cg_breakdown_returns_false():
    """lambda = (A p, p) = 0 ends CG with false (cg.h:172-184)"""
    say("case CG breakdown")
    u = U()
    T = np.array([[1.0, -1.0], [-1.0, 1.0]])
    A = small_csr(T)
    S = u.CG(None, max_steps=20)
    S.init(A)
    x = u.Vector(2)
    assert not S.apply(x, u.Vector(values=np.ones(2)))
    ok, steps, _ = O.csr_solve(sp.csr_matrix(T), np.ones(2), max_steps=20)
    assert not ok
