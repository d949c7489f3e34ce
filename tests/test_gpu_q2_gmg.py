"""GPU tests of the Q2 hierarchy against the oracle: the LEVEL numbering of
every level, the level operators (element loop with the coefficient at the
Gauss points), StdTransfer's prolongations and restrictions with their
Dirichlet adjustments, the multigrid over these matrices (V, W and F cycles,
one application and the GMG-preconditioned CG), and the generic Q2 element
loop against the lattice Q2 assembly."""
import sys

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def U():
    import ugcore_amd
    return ugcore_amd


def rows_close(rp, a, b):
    worst = 0.0
    for r in range(len(rp) - 1):
        x, y = a[rp[r]:rp[r + 1]], b[rp[r]:rp[r + 1]]
        if y.size:
            worst = max(worst, np.abs(x - y).max() / max(np.abs(y).max(), 1e-300))
    return worst


def oracle(refs, theta):
    p = O.Problem(0, refs)
    p.enable_q2()
    if theta is not None:
        p.set_theta(theta)
    p.assemble_all()
    return p


@pytest.mark.parametrize("refs", [1, 2, 3])
def test_numbering_matches_oracle(refs):
    say(f"case q2 numbering refs={refs}")
    u = U()
    q = u.Q2Discretization(refs)
    p = O.Problem(0, refs)
    p.enable_q2()
    for lev in range(refs + 1):
        pos, dr = p.q2_dofs(lev)
        X = q.dof_positions(lev)
        assert np.array_equal(np.rint(X * 2 * 2 ** refs).astype(np.int64), pos), lev
        assert np.array_equal(q.dirichlet(lev), dr), lev


@pytest.mark.parametrize("refs,seed", [(2, 3), (3, 4)])
def test_operators_and_transfers_match_oracle(refs, seed):
    say(f"case q2 hierarchy refs={refs}")
    u = U()
    theta = np.random.default_rng(seed).normal(size=8 ** refs) * 0.4
    p = oracle(refs, theta)
    q = u.Q2Discretization(refs)
    q.set_theta(theta)
    q.set_source(1.0)
    A, P, R, b = q.hierarchy()
    for lev in range(refs + 1):
        rp, col, val = A[lev].download_csr()
        o = p.A(lev)
        assert np.array_equal(rp, o.rowptr) and np.array_equal(col, o.col), lev
        assert rows_close(rp, val, o.val) <= 1e-12, lev
        if lev > 0:
            for mine, ref in ((P[lev], p.P(lev)), (R[lev], p.R(lev))):
                rp2, col2, val2 = mine.download_csr()
                assert np.array_equal(rp2, ref.rowptr) and np.array_equal(col2, ref.col), lev
                assert np.array_equal(val2, ref.val), lev
    bo = p.rhs(refs)
    assert np.abs(b.numpy() - bo).max() <= 1e-12 * np.abs(bo).max()


@pytest.mark.parametrize("refs,base,cycle", [(2, 0, 1), (3, 0, 1), (3, 1, 2), (3, 0, -1)])
def test_multigrid_matches_oracle(refs, base, cycle):
    say(f"case q2 multigrid refs={refs} base={base} cycle={cycle}")
    u = U()
    theta = np.random.default_rng(refs + 10).normal(size=8 ** refs) * 0.4
    p = oracle(refs, theta)
    q = u.Q2Discretization(refs)
    q.set_theta(theta)
    q.set_source(1.0)
    A, P, R, b = q.hierarchy()
    N = q.num_dofs(refs)
    dr = q.dirichlet(refs)
    d = np.random.default_rng(1).normal(size=N)
    d[dr == 1] = 0.0
    G = u.MatrixMultiGrid(A, P, R, base_level=base, nu1=2, nu2=2, cycle=cycle, omega=0.66)
    G.init(A[refs])
    c = u.Vector(N)
    G.apply(c, u.Vector(values=d))
    co = p.gmg_apply(d, base_lev=base, nu1=2, nu2=2, cycle=cycle, omega=0.66)
    assert np.abs(c.numpy() - co).max() <= 1e-11 * np.abs(co).max()
    # the preconditioned CG
    ro = p.solve(base_lev=base, cycle=cycle, rel_red=1e-10)
    S = u.CG(G, max_steps=100, reduction=1e-10)
    S.init(A[refs])
    x = u.Vector(N)
    assert S.apply(x, b)
    assert S.steps() == ro["steps"]
    assert np.abs(x.numpy() - ro["x"]).max() <= 1e-9 * np.abs(ro["x"]).max()


def test_generic_q2_element_loop_equals_the_lattice_assembly():
    """ElementDiscretization('hexahedron_q2') fed the lattice connectivity
    (nodes x fastest) gives the lattice Q2 assembly; its defect is A u - b"""
    say("case q2 generic vs lattice")
    u = U()
    n = 4
    W = 2 * n + 1
    theta = np.random.default_rng(7).normal(size=n ** 3) * 0.3
    mi = O.q2_hex_multi_index()
    E = []
    for ez in range(n):
        for ey in range(n):
            for ex in range(n):
                E.append([((2 * ez + m[2]) * W + (2 * ey + m[1])) * W + (2 * ex + m[0]) for m in mi])
    E = np.array(E, np.int32)
    g = np.arange(W) / (2.0 * n)
    X = np.stack(np.meshgrid(g, g, g, indexing="ij")[::-1], -1).reshape(-1, 3)
    bnd = np.flatnonzero(np.any((X == 0) | (X == 1), axis=1)).astype(np.int32)
    e = u.ElementDiscretization("hexahedron_q2", E, X)
    e.set_dirichlet(bnd)
    e.set_diffusion(np.exp(theta))
    e.set_source(1.0)
    A, b = u.SparseMatrix(), u.Vector(W ** 3)
    e.assemble_linear(A, b)
    L, bl = u.q2_assemble_lattice(n, theta, 1.0)
    rp, col, val = A.download_csr()
    rl, cl, vl = L.download_csr()
    assert np.array_equal(rp, rl) and np.array_equal(col, cl)
    assert rows_close(rp, val, vl) <= 1e-12
    assert np.abs(b.numpy() - bl.numpy()).max() <= 1e-12 * np.abs(bl.numpy()).max()
    # the coefficient at the Gauss points (here constant per element) gives the same
    e.set_diffusion_at_points(np.repeat(np.exp(theta), 27))
    A2, b2 = u.SparseMatrix(), u.Vector(W ** 3)
    e.assemble_linear(A2, b2)
    assert rows_close(rp, A2.download_csr()[2], val) <= 1e-15
    # defect
    x = np.random.default_rng(2).normal(size=W ** 3)
    d = u.Vector(W ** 3)
    e.assemble_defect(d, u.Vector(values=x))
    import scipy.sparse as sp
    S = sp.csr_matrix((val, col, rp), shape=(W ** 3, W ** 3))
    ref = S @ x - b.numpy()
    ref[bnd] = 0.0
    assert np.abs(d.numpy() - ref).max() <= 1e-12 * np.abs(ref).max()


def test_pybind_q2_multigrid_solve_equals_the_ctypes_path():
    say("case q2 ug4py")
    u = U()
    from ugcore_amd import ug4py as P
    refs = 2
    theta = np.random.default_rng(9).normal(size=8 ** refs) * 0.3
    P.init()
    q = P.Q2Discretization(refs)
    q.set_theta(theta)
    q.set_source(1.0)
    A, Pm, Rm = [], [None], [None]
    b = None
    for lev in range(refs + 1):
        a, bv = P.SparseMatrix(), P.Vector(q.num_dofs(lev))
        q.assemble_linear(lev, a, bv)
        A.append(a)
        b = bv
        if lev:
            p, r = P.SparseMatrix(), P.SparseMatrix()
            q.prolongation(lev, p)
            q.restriction(lev, p, r)
            Pm.append(p)
            Rm.append(r)
    G = P.MatrixMultiGrid(A, Pm, Rm, base_level=0)
    S = P.CG(G, 100, 1e-10, 1e-50)
    S.init(A[refs])
    x = P.Vector(q.num_dofs(refs))
    assert S.apply(x, b)
    # the ctypes path
    q2 = u.Q2Discretization(refs)
    q2.set_theta(theta)
    q2.set_source(1.0)
    A2, P2, R2, b2 = q2.hierarchy()
    S2 = u.CG(u.MatrixMultiGrid(A2, P2, R2), max_steps=100, reduction=1e-10)
    S2.init(A2[refs])
    x2 = u.Vector(q2.num_dofs(refs))
    assert S2.apply(x2, b2)
    assert S.steps() == S2.steps()
    assert np.array_equal(x.numpy(), x2.numpy())
