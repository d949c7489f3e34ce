"""Pins of the oracle's tetrahedral path (unit cube as 3x3x3 hexahedra x 6
tetrahedra, regular refinement, FV1 convection-diffusion, transposed adjoint).

The reference's ConvectionDiffusion plugin is external and lib_grid cannot be
built here, so the oracle is pinned by what the mathematics fixes:
* hierarchy: every level is a lattice mesh (6 tetrahedra per lattice cell),
  vertex/edge counts follow from the refinement rule, nnz = V + 2E;
* FV1 diffusion on simplices equals the P1 finite-element stiffness matrix
  (barycentric dual), also on the jittered top level;
* full-upwind convection: zero row sums on interior rows (closed dual cells,
  constant velocity), zero column sums away from the boundary (conservative),
  off-diagonals <= 0 and diagonals >= 0 (upwinding);
* source: scv volumes of a tetrahedron sum to its volume;
* transfer: P rows are the vertex copy (1) or the two endpoints of the parent
  edge (1/2 each), P*1 = 1 on interior rows;
* solve and adjoint: BiCGStab+GMG against a direct solve, the transposed
  adjoint against scipy, the gradient against central differences.
"""
import itertools

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

import oracle_lib as O

VEL = (20.0, 10.0, 5.0)
# jitter: displacement length <= 0.2h, i.e. +-0.2h/sqrt(3) per component (the
# refined tetrahedra have minimum altitude h/sqrt(3); +-0.2h per component
# inverts some of them)
JITTER = 0.2 / np.sqrt(3.0)


def make(refs, vel=VEL, jitter=JITTER, theta_seed=5, assemble=True):
    p = O.Problem(2, refs)
    if theta_seed is not None:
        p.set_theta(O.make_theta(p.n_fine, theta_seed, 3))
    p.set_tet_params(vel=vel, jitter_seed=11, jitter_amp=jitter)
    if assemble:
        p.assemble_all(1.0)
    return p


def p1_stiffness(pos, tets, n):
    """P1 FE stiffness (k = 1) assembled from positions and tetrahedra"""
    rows, cols, vals = [], [], []
    for t in tets:
        X = pos[t]
        J = (X[1:] - X[0]).T
        Ji = np.linalg.inv(J)
        G = np.vstack([-Ji.sum(axis=0), Ji])          # gradients of phi_0..3 (rows)
        V = abs(np.linalg.det(J)) / 6.0
        K = V * G @ G.T
        for a, b in itertools.product(range(4), range(4)):
            rows.append(t[a]); cols.append(t[b]); vals.append(K[a, b])
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def test_hierarchy_lattice_counts():
    p = O.Problem(2, 3)
    for lev in range(p.num_levels):
        W = 3 * 2 ** lev + 1
        assert p.num_dofs(lev) == W ** 3
        lat = p.dof_lattice(lev)
        step = 2 ** (3 - lev)
        assert np.all(lat % step == 0)
        assert len({tuple(r) for r in lat}) == W ** 3
        E = p.elements(lev)
        assert E.shape == (6 * (W - 1) ** 3, 4)
        # every tetrahedron inside one lattice cell, positively oriented
        L = lat[E] // step
        assert np.all(L.max(axis=1) - L.min(axis=1) <= 1)
        d = np.linalg.det((L[:, 1:] - L[:, :1]).astype(float))
        assert np.all(d > 0)
        # Dirichlet = boundary vertices, numbered last (subset-major)
        dm = p.dirichlet(lev).astype(bool)
        bnd = np.any((lat == 0) | (lat == 3 * 2 ** 3), axis=1)
        assert np.array_equal(dm, bnd)
        assert not dm[: (~dm).sum()].any()


def test_diffusion_equals_p1_stiffness_on_jittered_mesh():
    p = make(2, vel=(0, 0, 0), jitter=JITTER, theta_seed=None)
    top = p.num_levels - 1
    A = p.A(top).to_scipy()
    pos = p.dof_positions(top)
    n = p.num_dofs(top)
    K = p1_stiffness(pos, p.elements(top), n)
    dm = p.dirichlet(top).astype(bool)
    inner = np.where(~dm)[0]
    D = (A - K)[inner]
    assert abs(D).max() <= 1e-13 * abs(K).max()
    # pattern = vertex pairs of the tetrahedra (explicit zeros kept), nnz = V + 2E
    E = p.elements(top)
    pairs = {(int(a), int(b)) for t in E for a in t for b in t}
    assert A.nnz == len(pairs)
    # jitter moved the interior vertices, not the boundary ones
    h = 1.0 / p.n_fine
    off = pos - p.dof_lattice(top) * h
    assert np.abs(off[dm]).max() <= 1e-15
    assert np.abs(off[~dm]).max() <= JITTER * h + 1e-15 and np.abs(off[~dm]).max() > 0.9 * JITTER * h
    assert np.linalg.norm(off[~dm], axis=1).max() <= 0.2 * h + 1e-15


def test_diffusion_coefficient_scales_element_matrices():
    # theta constant c => A_diff = exp(c) * A(k=1)
    p1 = make(1, vel=(0, 0, 0), theta_seed=None)
    p2 = O.Problem(2, 1)
    p2.set_theta(np.full(p2.n_fine ** 3, 0.7))
    p2.set_tet_params(vel=(0, 0, 0), jitter_seed=11, jitter_amp=JITTER)
    p2.assemble_all(1.0)
    top = p1.num_levels - 1
    dm = p1.dirichlet(top).astype(bool)
    A1 = p1.A(top).to_scipy()[~dm]
    A2 = p2.A(top).to_scipy()[~dm]
    assert abs(A2 - np.exp(0.7) * A1).max() <= 1e-13 * abs(A2).max()


def test_convection_upwind_structure():
    pv = make(2, vel=VEL)
    p0 = make(2, vel=(0, 0, 0))
    top = pv.num_levels - 1
    C = (pv.A(top).to_scipy() - p0.A(top).to_scipy()).tocsr()
    dm = pv.dirichlet(top).astype(bool)
    inner = np.where(~dm)[0]
    Ci = C[inner]
    # row sums zero on interior rows
    assert np.abs(np.asarray(Ci.sum(axis=1))).max() <= 1e-13 * abs(Ci).max()
    # off-diagonals <= 0, diagonals >= 0
    for i in inner[:400]:
        row = C.getrow(i).toarray().ravel()
        d = row[i]
        row[i] = 0
        assert d >= -1e-15 and row.max() <= 1e-15
    # column sums zero for columns whose neighbours are all interior
    lat = pv.dof_lattice(top)
    deep = np.all((lat >= 2) & (lat <= pv.n_fine - 2), axis=1)
    cs = np.asarray(C.sum(axis=0)).ravel()
    assert np.abs(cs[deep]).max() <= 1e-13 * abs(C).max()
    assert abs(C).max() > 0


def test_jitter_keeps_every_tetrahedron_valid():
    # displacement <= 0.2h: no inversion (minimum altitude h/sqrt(3)); the
    # literal +-0.2h per component would invert tetrahedra from 24^3 on
    p = O.Problem(2, 3)
    top = p.num_levels - 1
    E = p.elements(top)
    for amp, expect_inverted in ((JITTER, False), (0.2, True)):
        p.set_tet_params(vel=VEL, jitter_seed=11, jitter_amp=amp)
        pos = p.dof_positions(top)
        V = np.linalg.det(pos[E[:, 1:]] - pos[E[:, :1]]) / 6.0
        assert bool((V <= 0).any()) == expect_inverted


def test_scv_volumes_sum_to_domain():
    # b = f * sum of scv volumes; with f = 1 and no Dirichlet rows removed the
    # interior sum plus the boundary scv volumes is the unit cube: check via a
    # solve-free identity on the assembled rhs of two problems with f = 1 / f = 2
    p = make(2)
    top = p.num_levels - 1
    b = p.rhs(top)
    p2 = make(2, assemble=False)
    p2.assemble_all(2.0)
    assert np.allclose(p2.rhs(top), 2 * b, rtol=0, atol=1e-15)
    # scv volume of each interior vertex equals a quarter of its tetrahedra (barycentric dual)
    pos = p.dof_positions(top)
    E = p.elements(top)
    V = np.abs(np.linalg.det(pos[E[:, 1:]] - pos[E[:, :1]])) / 6.0
    lumped = np.zeros(p.num_dofs(top))
    np.add.at(lumped, E.ravel(), np.repeat(V / 4.0, 4))
    dm = p.dirichlet(top).astype(bool)
    assert np.abs(b[~dm] - lumped[~dm]).max() <= 1e-12 * lumped.max()
    assert abs(V.sum() - 1.0) < 1e-12


def test_prolongation_edges():
    p = make(2)
    for lev in (1, 2):
        P = p.P(lev).to_scipy()
        dm = p.dirichlet(lev).astype(bool)
        vals = P.data[P.data != 0]
        assert set(np.unique(vals)) <= {0.5, 1.0}
        rs = np.asarray(P.sum(axis=1)).ravel()
        assert np.allclose(rs[~dm], 1.0)
        nz = np.diff(P.indptr)
        assert set(np.unique(nz[~dm])) <= {1, 2}
        # an edge child lies at the midpoint of its two parents
        latf = p.dof_lattice(lev)
        latc = p.dof_lattice(lev - 1)
        for r in np.where(~dm)[0][:300]:
            cols = P.indices[P.indptr[r]:P.indptr[r + 1]]
            assert np.array_equal(2 * latf[r], latc[cols].sum(axis=0) * (2 // len(cols)))


@pytest.mark.parametrize("refs", [1, 2, 3])
def test_bicgstab_gmg_solves(refs):
    p = make(refs)
    r = p.solve(solver="bicgstab", base_lev=0, omega=0.6, rel_red=1e-10)
    top = p.num_levels - 1
    A = p.A(top).to_scipy().tocsc()
    x = spl.spsolve(A, p.rhs(top))
    assert r["ok"] and r["steps"] <= 30
    assert np.linalg.norm(r["x"] - x) <= 1e-8 * np.linalg.norm(x)


def test_adjoint_transposed_solve():
    p = make(2)
    top = p.num_levels - 1
    rng = np.random.default_rng(3)
    g = rng.standard_normal(p.num_dofs(top))
    lam, steps = p.solve_adjoint(g, rel_red=1e-13)
    dm = p.dirichlet(top).astype(bool)
    inner = np.where(~dm)[0]
    AII = p.A(top).to_scipy()[inner][:, inner]
    li = spl.spsolve(AII.T.tocsc(), g[inner])
    assert np.linalg.norm(lam[inner] - li) <= 1e-10 * np.linalg.norm(li)
    assert np.all(lam[dm] == 0)


def test_adjoint_gradient_central_differences():
    refs = 1
    p = make(refs)
    n = p.n_fine
    theta = O.make_theta(n, 5, 3)
    top = p.num_levels - 1
    N = p.num_dofs(top)
    rng = np.random.default_rng(7)
    obs = rng.choice(np.where(p.dirichlet(top) == 0)[0], 20, replace=False)
    d = rng.standard_normal(20) * 0.01

    def forward(th):
        q = O.Problem(2, refs)
        q.set_theta(th)
        q.set_tet_params(vel=VEL, jitter_seed=11, jitter_amp=JITTER)
        q.assemble_all(1.0)
        A = q.A(top).to_scipy().tocsc()
        u = spl.spsolve(A, q.rhs(top))
        return q, u

    def J(th):
        _, u = forward(th)
        return 0.5 * np.sum((u[obs] - d) ** 2)

    q, u = forward(theta)
    dJdu = np.zeros(N)
    dJdu[obs] = u[obs] - d
    lam, _ = q.solve_adjoint(dJdu, rel_red=1e-14)
    grad = q.adjoint_gradient(u, lam)
    for k in range(4):
        v = rng.standard_normal(n ** 3)
        h = 1e-6
        fd = (J(theta + h * v) - J(theta - h * v)) / (2 * h)
        assert abs(fd - grad @ v) <= 1e-6 * abs(fd) + 1e-12
