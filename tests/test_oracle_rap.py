"""The oracle's Galerkin coarse operators (AssembledMultiGridCycle::set_rap,
mg_solver.h:149; init_rap_operator, mg_solver_impl.hpp:829-1013) pinned
against the reference's own AddMultiplyOf compiled from its headers
(sparsematrix_util.h:167-240, bit-exact), against scipy's sparse triple
product, and against the invariants the Dirichlet-adjusted transfer fixes."""
import numpy as np
import pytest

import oracle_lib as O

ref = O.ref_lib()
needs_ref = pytest.mark.skipif(ref is None, reason="oracle/_ref not built (reference checkout absent)")

CASES = [(0, 2), (0, 3), (1, 3), (1, 4)]


def problem(kind, refs, rap):
    p = O.Problem(kind, refs)
    n = 2 ** refs
    if kind == 0:
        p.set_theta(O.make_theta(n, 5, 3))
    p.assemble_all()
    if rap:
        p.set_rap()
    return p


def to_ref(M):
    h = ref.ref_sm_create(M.n_rows, M.n_cols)
    for r in range(M.n_rows):
        for k in range(M.rowptr[r], M.rowptr[r + 1]):
            ref.ref_sm_set(h, r, int(M.col[k]), float(M.val[k]))
    return h


def ref_csr(h):
    n = ref.ref_sm_rows(h)
    nnz = ref.ref_sm_nnz(h)
    rp = np.zeros(n + 1, np.int64)
    col = np.zeros(nnz, np.int32)
    val = np.zeros(nnz, np.float64)
    ref.ref_sm_csr(h, rp, col, val)
    return rp, col, val


@needs_ref
@pytest.mark.parametrize("kind,refs", CASES)
def test_galerkin_product_matches_reference_addmultiplyof(kind, refs):
    """R_l A_l P_l of the oracle equals AddMultiplyOf on the same three
    matrices, pattern and values bit for bit, on every level."""
    p0 = problem(kind, refs, rap=False)
    p1 = problem(kind, refs, rap=True)
    A = p0.A(refs)
    for lev in range(refs, 0, -1):
        R, P = p0.R(lev), p0.P(lev)
        hR, hA, hP = to_ref(R), to_ref(A), to_ref(P)
        hM = ref.ref_sm_multiply3(hR, hA, hP)
        rp, col, val = ref_csr(hM)
        mine = p1.A(lev - 1)
        assert np.array_equal(rp, mine.rowptr)
        assert np.array_equal(col, mine.col)
        assert np.array_equal(val, mine.val)
        for h in (hR, hA, hP, hM):
            ref.ref_sm_destroy(h)
        A = mine


@pytest.mark.parametrize("kind,refs", CASES)
def test_galerkin_product_matches_scipy(kind, refs):
    p0 = problem(kind, refs, rap=False)
    p1 = problem(kind, refs, rap=True)
    A = p0.A(refs).to_scipy()
    for lev in range(refs, 0, -1):
        Ac = (p0.R(lev).to_scipy() @ A @ p0.P(lev).to_scipy()).tocsr()
        mine = p1.A(lev - 1).to_scipy()
        assert abs(Ac - mine).max() <= 1e-14 * abs(Ac).max()
        A = mine


@pytest.mark.parametrize("kind,refs", CASES)
def test_galerkin_levels_invariants(kind, refs):
    """Coarse Dirichlet rows are identity rows; the interior block of R A P is
    symmetric for the symmetric Q1 and FV1-triangle operators; the row sums of
    the interior block of a pure-diffusion operator vanish away from the
    boundary (constants are in the kernel and P reproduces them)."""
    p1 = problem(kind, refs, rap=True)
    for lev in range(refs - 1, -1, -1):
        M = p1.A(lev).to_scipy().tolil()
        dm = p1.dirichlet(lev).astype(bool)
        n = M.shape[0]
        for i in np.where(dm)[0]:
            row = M.getrow(i).toarray().ravel()
            assert row[i] == 1.0
            assert np.count_nonzero(row) == 1
        inner = np.where(~dm)[0]
        if inner.size == 0:
            continue
        B = M.tocsr()[inner][:, inner].toarray()
        assert np.abs(B - B.T).max() <= 1e-13 * np.abs(B).max()
        if lev >= 2:
            full = M.tocsr()
            lat = p1.dof_lattice(lev) // (2 ** (refs - lev))
            W = 2 ** lev
            away = [i for i in inner if np.all(lat[i, : (3 if kind == 0 else 2)] >= 2)
                    and np.all(lat[i, : (3 if kind == 0 else 2)] <= W - 2)]
            rs = np.asarray(full[away].sum(axis=1)).ravel()
            assert np.abs(rs).max() <= 1e-12 * np.abs(full[away].toarray()).max()


@pytest.mark.parametrize("kind,refs", [(0, 3), (1, 5)])
def test_galerkin_gmg_cg_converges(kind, refs):
    """CG with the Galerkin GMG reaches 1e-10 in no more steps than with the
    assembled levels plus two, and both solve the same top-level system."""
    p0 = problem(kind, refs, rap=False)
    p1 = problem(kind, refs, rap=True)
    r0 = p0.solve("cg", rel_red=1e-10)
    r1 = p1.solve("cg", rel_red=1e-10)
    assert r0["ok"] and r1["ok"]
    assert r1["steps"] <= r0["steps"] + 2
    assert np.linalg.norm(r1["x"] - r0["x"]) <= 1e-8 * np.linalg.norm(r0["x"])
