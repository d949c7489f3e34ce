"""GPU tests of the element-loop assembly from caller-supplied connectivity
(ElementDiscretization, ugx_elemdisc_*) and of DomainDiscretization's
assemble_jacobian / assemble_defect.

* The three lattice domains fed through the generic path with the oracle's own
  top-level connectivity (element -> DoF table in the reference's local vertex
  order) and DoF coordinates: pattern bit-exact, entries within 1e-12 of the
  row maximum, right-hand side within 1e-12 against the oracle's
  assemble_linear (Q1 hexahedra, FV1 triangles with a per-DoF source, FV1
  tetrahedra with full upwind and jittered vertices).
* assemble_defect against d = A u - b with the oracle's A and b and the
  Dirichlet entries 0 (the element sum of A_e u_e - f_e in exact arithmetic),
  within 1e-12 of each row's sum |A||u| + |b|; through the generic path and
  through DomainDiscretization.assemble_defect.
* assemble_jacobian = the assembled matrix, bit for bit.
* An unstructured Delaunay tetrahedral mesh: the FV1 diffusion matrix of P1
  elements equals the P1 Galerkin stiffness matrix vol * grad(phi_i).grad(phi_j)
  (barycentric dual control volumes), assembled here with numpy.
"""
import sys

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

VEL = (20.0, 10.0, 5.0)
JSEED, JAMP = 11, 0.2 / np.sqrt(3.0)
TYPES = {0: "hexahedron_q1", 1: "triangle_fv1", 2: "tetrahedron_fv1"}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def U():
    import ugcore_amd
    return ugcore_amd


def oracle(kind, refs, seed=5):
    p = O.Problem(kind, refs)
    theta = O.make_theta(p.n_fine, seed, 3) if kind != 1 else \
        np.random.default_rng(seed).normal(size=p.n_fine ** 2) * 0.5
    p.set_theta(theta)
    f = None
    if kind == 1:
        f = np.random.default_rng(seed + 1).normal(size=p.num_dofs(refs)) + 2.0
        p.set_source_top(f)
    if kind == 2:
        p.set_tet_params(vel=VEL, jitter_seed=JSEED, jitter_amp=JAMP)
    p.assemble(refs, 1.0)
    return p, theta, f


def element_data(p, kind, refs, theta):
    """element -> DoF table, DoF coordinates and k = exp(theta[cell]) per element"""
    E = p.elements(refs)
    lat = p.dof_lattice(refs)
    n = p.n_fine
    dim = 2 if kind == 1 else 3
    cell = lat[E].min(axis=1)                         # lower corner of the element's lattice cell
    if dim == 2:
        ci = cell[:, 1] * n + cell[:, 0]
    else:
        ci = (cell[:, 2] * n + cell[:, 1]) * n + cell[:, 0]
    k = np.exp(theta[ci])
    X = p.dof_positions(refs)[:, :dim]
    return E, X, k


def generic(p, kind, refs, theta, f):
    u = U()
    E, X, k = element_data(p, kind, refs, theta)
    ed = u.ElementDiscretization(TYPES[kind], E, X)
    ed.set_diffusion(k)
    ed.set_source(f if kind == 1 else 1.0)
    if kind == 2:
        ed.set_velocity(VEL)
    ed.set_dirichlet(np.nonzero(p.dirichlet(refs))[0])
    return ed


def row_scaled(rp, a, b):
    worst = 0.0
    for r in range(len(rp) - 1):
        x, y = a[rp[r]:rp[r + 1]], b[rp[r]:rp[r + 1]]
        if y.size:
            worst = max(worst, np.abs(x - y).max() / max(np.abs(y).max(), 1e-300))
    return worst


CASES = [(0, 3), (0, 4), (1, 5), (2, 1), (2, 2)]


@pytest.mark.parametrize("kind,refs", CASES)
def test_generic_assembly_matches_oracle(kind, refs):
    say(f"case generic assembly kind={kind} refs={refs}")
    u = U()
    p, theta, f = oracle(kind, refs)
    ed = generic(p, kind, refs, theta, f)
    A = u.SparseMatrix()
    b = u.Vector(p.num_dofs(refs))
    ed.assemble_linear(A, b)
    R = p.A(refs)
    rp, col, val = A.download_csr()
    assert ed.nnz() == R.nnz
    assert np.array_equal(rp.astype(np.int64), R.rowptr) and np.array_equal(col, R.col)
    err = row_scaled(rp, val, R.val)
    bo = p.rhs(refs)
    say(f"  {ed.num_colours()} colours, nnz {ed.nnz()}, entries {err:.2e}, "
        f"rhs {np.abs(b.numpy() - bo).max() / np.abs(bo).max():.2e}")
    assert err <= 1e-12
    assert np.abs(b.numpy() - bo).max() <= 1e-12 * np.abs(bo).max()
    # the Jacobian of the linear discretisation is the matrix itself
    J = u.SparseMatrix()
    ed.assemble_jacobian(J, u.Vector(values=np.ones(p.num_dofs(refs))))
    for x, y in zip(J.download_csr(), (rp, col, val)):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("kind,refs", CASES)
def test_defect_matches_oracle(kind, refs):
    say(f"case defect kind={kind} refs={refs}")
    u = U()
    p, theta, f = oracle(kind, refs)
    N = p.num_dofs(refs)
    S = p.A(refs).to_scipy()
    bo = p.rhs(refs)
    dm = p.dirichlet(refs) == 1
    uv = np.random.default_rng(7).normal(size=N)
    expect = S @ uv - bo
    expect[dm] = 0.0
    scale = np.abs(S) @ np.abs(uv) + np.abs(bo)
    # through the generic path
    ed = generic(p, kind, refs, theta, f)
    d = u.Vector(N)
    ed.assemble_defect(d, u.Vector(values=uv))
    err1 = np.max(np.abs(d.numpy() - expect) / scale)
    # through the lattice discretisation
    dd = u.DomainDiscretization(kind, refs)
    dd.set_theta(theta)
    if kind == 1:
        dd.set_source(f)
    else:
        dd.set_source(1.0)
    if kind == 2:
        dd.set_velocity(VEL)
        dd.set_jitter(JSEED, JAMP)
    d2 = u.Vector(N)
    dd.assemble_defect(d2, u.Vector(values=uv))
    err2 = np.max(np.abs(d2.numpy() - expect) / scale)
    say(f"  defect vs A u - b: generic {err1:.2e}, DomainDiscretization {err2:.2e}")
    assert err1 <= 1e-12 and err2 <= 1e-12
    assert np.all(d.numpy()[dm] == 0.0) and np.all(d2.numpy()[dm] == 0.0)
    # assemble_jacobian of the lattice discretisation = its assemble_linear matrix
    A = u.SparseMatrix()
    b = u.Vector(N)
    dd.assemble_linear(A, b)
    J = u.SparseMatrix()
    dd.assemble_jacobian(J, u.Vector(values=uv))
    for x, y in zip(J.download_csr(), A.download_csr()):
        assert np.array_equal(x, y)


def test_unstructured_tetrahedra_equal_p1_stiffness():
    """FV1 diffusion on P1 tetrahedra = the P1 Galerkin stiffness matrix"""
    say("case unstructured Delaunay tetrahedra")
    from scipy.spatial import Delaunay
    u = U()
    rng = np.random.default_rng(3)
    corners = np.array([[x, y, z] for x in (0, 1) for y in (0, 1) for z in (0, 1)], float)
    X = np.vstack([corners, rng.uniform(0.05, 0.95, size=(400, 3))])
    T = Delaunay(X).simplices.astype(np.int32)
    # positive orientation, as a mesh generator hands them over
    J = np.stack([X[T[:, i]] - X[T[:, 0]] for i in (1, 2, 3)], axis=2)
    neg = np.linalg.det(J) < 0
    T[neg, 1], T[neg, 2] = T[neg, 2].copy(), T[neg, 1].copy()
    k = np.exp(rng.normal(scale=0.5, size=len(T)))
    ed = u.ElementDiscretization("tetrahedron_fv1", T, X)
    ed.set_diffusion(k)
    ed.set_source(0.0)
    A = u.SparseMatrix()
    b = u.Vector(len(X))
    ed.assemble_linear(A, b)
    rp, col, val = A.download_csr()
    import scipy.sparse as sp
    G = sp.csr_matrix((val, col, rp), shape=(len(X), len(X))).toarray()
    K = np.zeros((len(X), len(X)))
    for e, t in enumerate(T):
        P = X[t]
        M = np.hstack([np.ones((4, 1)), P])
        grads = np.linalg.inv(M)[1:, :].T           # grad phi_i, rows i
        vol = abs(np.linalg.det(M)) / 6.0
        K[np.ix_(t, t)] += k[e] * vol * grads @ grads.T
    say(f"  {len(T)} tetrahedra, {ed.num_colours()} colours: max |FV1 - P1 FE| / max|K| = "
        f"{np.abs(G - K).max() / np.abs(K).max():.2e}")
    assert np.abs(G - K).max() <= 1e-12 * np.abs(K).max()
    assert not np.any(b.numpy())
