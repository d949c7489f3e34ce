"""Oracle Q2 hierarchy: LEVEL numbering, level operators, StdTransfer
prolongation / restriction with the Dirichlet adjustments, and the
GMG-preconditioned CG over them, pinned by a hand trace, closed forms and
the mathematics (quadratic reproduction, consistency, symmetry, a direct
solve)."""
import json
import os

import numpy as np
import pytest
import scipy.sparse.linalg as sla

import oracle_lib as O

TRACE = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "numbering_trace.json")))


def q2_problem(refs, theta_seed=None):
    p = O.Problem(0, refs)
    p.enable_q2()
    if theta_seed is not None:
        rng = np.random.default_rng(theta_seed)
        p.set_theta(rng.normal(size=8 ** refs) * 0.4)
    p.assemble_all()
    return p


def test_q2_numbering_level0_is_the_traced_one():
    p = O.Problem(0, 2)
    p.enable_q2()
    pos, dr = p.q2_dofs(0)
    t = TRACE["q2_hexahedron_level0"]
    # doubled finest units (finest scale 4 at refs 2) -> units of 1/2
    assert np.array_equal(pos // 4, np.array(t["dof"], np.int64))
    assert np.array_equal(dr, np.array(t["dirichlet"], np.uint8))


@pytest.mark.parametrize("refs", [1, 2, 3])
def test_q2_numbering_structure(refs):
    """every node of the doubled lattice once; the vertex DoFs come first and
    in the P1 order; within each object type Inner before Boundary"""
    p = O.Problem(0, refs)
    p.enable_q2()
    p1_problem = O.Problem(0, refs)
    for lev in range(refs + 1):
        pos, dr = p.q2_dofs(lev)
        step = 2 ** (refs - lev)   # doubled finest units per doubled level unit
        assert (pos % step == 0).all()
        q = pos // step
        W = 2 ** (lev + 1) + 1
        idx = q[:, 0] + W * q[:, 1] + W * W * q[:, 2]
        assert np.array_equal(np.sort(idx), np.arange(W ** 3))
        odd = (q % 2).sum(axis=1)
        assert (np.diff(odd) >= 0).all()                     # vertices, edges, faces, volumes
        for t in range(4):
            d = dr[odd == t].astype(int)
            assert (np.diff(d) >= 0).all()                   # Inner then Boundary
        nv = (odd == 0).sum()
        p1 = p1_problem.dof_lattice(lev) * 2 // step
        assert np.array_equal(q[:nv], p1)
        bnd = ((q == 0) | (q == W - 1)).any(axis=1)
        assert np.array_equal(dr.astype(bool), bnd)


def _q2_values(P, pf, pc, fine_row, coarse_pos):
    """P entry (fine_row, coarse DoF at position coarse_pos)"""
    j = int(np.flatnonzero((pc == np.array(coarse_pos)).all(axis=1))[0])
    return P[fine_row, j]


def test_q2_prolongation_closed_forms():
    p = q2_problem(2)
    P = p.P(2).to_scipy().tocsr()
    R = p.R(2).to_scipy().tocsr()
    pf, df = p.q2_dofs(2)
    pc, dc = p.q2_dofs(1)
    # an interior fine node at the quarter of a coarse element along x only
    # (y, z on coarse vertex planes): the 1D weights 3/8, 3/4, -1/8 on the
    # coarse edge's three nodes, explicit zeros nowhere else
    r = int(np.flatnonzero((pf == [1 * 1, 4, 4]).all(axis=1))[0])   # fine doubled units = 1/8
    row = P.getrow(r)
    assert row.nnz == 3
    got = sorted(zip(pc[row.indices][:, 0], row.data))
    assert got == [(0, 0.375), (2, 0.75), (4, -0.125)]
    # the coarse element centre (all axes at t = 1/2): 27 stored entries, one 1
    r = int(np.flatnonzero((pf == [2, 2, 2]).all(axis=1))[0])
    row = P.getrow(r)
    assert row.nnz == 27 and sorted(row.data)[-1] == 1.0 and np.count_nonzero(row.data) == 1
    # quadratic reproduction on the non-Dirichlet rows
    q = lambda X: X[:, 0] ** 2 + 0.3 * X[:, 1] * X[:, 2] - X[:, 2] + 2.0
    e = np.abs(P @ q(pc.astype(float)) - q(pf.astype(float)))[df == 0]
    assert e.max() < 1e-12
    # Dirichlet rows: the pattern kept, a single 1 at the parent's own DoF
    for r in np.flatnonzero(df)[:50]:
        row = P.getrow(r)
        assert np.count_nonzero(row.data) == 1 and row.data.max() == 1.0
    # R: P^T except the Dirichlet rows, which hold one 1
    PT = P.T.tocsr()
    for c in range(R.shape[0]):
        a, b = R.getrow(c), PT.getrow(c)
        assert np.array_equal(a.indices, b.indices)
        if dc[c]:
            assert np.count_nonzero(a.data) == 1 and a.data.max() == 1.0
            one = a.indices[a.data == 1.0][0]
            # the last Dirichlet child: the largest fine Dirichlet DoF whose P
            # row holds its 1 at this coarse DoF
            kids = [f for f in np.flatnonzero(df) if P[f, c] == 1.0 and np.count_nonzero(P.getrow(f).data) == 1]
            assert one == max(kids)
        else:
            assert np.array_equal(a.data, b.data)


@pytest.mark.parametrize("refs", [2, 3])
def test_q2_operator_consistency(refs):
    """constant k: symmetric on the interior, rows annihilate linear
    functions, the load integrates the source"""
    p = q2_problem(refs)
    top = refs
    A = p.A(top).to_scipy().tocsr()
    pos, dr = p.q2_dofs(top)
    X = pos / (2.0 * 2 ** refs)
    inner = dr == 0
    Ai = A[inner][:, inner]
    assert abs(Ai - Ai.T).max() < 1e-14 * abs(Ai).max()
    lin = 1.0 + X[:, 0] - 2 * X[:, 1] + 0.5 * X[:, 2]
    # rows of DoFs whose couplings are all inner: A lin = 0
    full = np.array([not dr[A.indices[A.indptr[r]:A.indptr[r + 1]]].any() for r in range(A.shape[0])])
    assert np.abs((A @ lin)[full & inner]).max() < 1e-12
    # f = 1: the load of a DoF is the integral of its shape function (the
    # Dirichlet entries are zeroed); on the interior they add up to the
    # integral of the sum of the interior shapes, and the full sum over an
    # unconstrained assembly would be 1
    b = p.rhs(top)
    assert np.abs(b[dr == 1]).max() == 0.0
    assert 0.0 < b.sum() < 1.0


@pytest.mark.parametrize("refs,seed", [(2, 1), (3, 2)])
def test_q2_gmg_cg_converges_to_the_direct_solution(refs, seed):
    p = q2_problem(refs, seed)
    r = p.solve(base_lev=0, rel_red=1e-10)
    assert r["ok"] and r["steps"] <= 15
    A = p.A(refs).to_scipy().tocsc()
    xd = sla.spsolve(A, p.rhs(refs))
    assert np.abs(r["x"] - xd).max() <= 1e-8 * np.abs(xd).max()
