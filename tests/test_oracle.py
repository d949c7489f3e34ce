"""CPU tests: pin the oracle against the reference's golden vectors, against
the reference's own SparseMatrix (compiled into oracle/_ref where the checkout
exists), and against closed forms; cross-check the product's host-side DoF
numbering with the oracle's independent restatement."""
import json
import os

import numpy as np
import pytest

import oracle_lib as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sparsematrix_golden.json")


def golden():
    with open(GOLDEN) as f:
        return json.load(f)


# ---------------------------------------------------------------------------
# CSR semantics vs the reference's golden outputs
# ---------------------------------------------------------------------------

def sm_transpose_test0_entries(N, M):
    """insertion sequence of tests/sm_transpose.cc test0 (values set, not added)"""
    e = []
    for i in range(1, min(N, M)):
        e += [(i, i, 2.0), (i, i - 1, 1.0), (i - 1, i, 1.0)]
    if N > 6 and M > 6:
        e += [(5, 5, 0.0), (5, 0, 0.0), (1, 6, 1.0)]
    return e


def sm_transpose_test3_entries(N):
    return [(i, j, float(i + j)) for i in range(N) for j in range(N)]


@pytest.mark.parametrize("case", range(6))
def test_sm_transpose_golden(case):
    g = golden()["sm_transpose"][case]
    rows, cols = g["input_shape"]
    if g["name"].startswith("test3"):
        ent = sm_transpose_test3_entries(rows)
        keep_expected, drop_expected = g["nnz_second"], g["nnz_first"]
    else:
        ent = sm_transpose_test0_entries(rows, cols)
        keep_expected, drop_expected = g["nnz_first"], g["nnz_second"]
    keep = O.builder_csr(ent, rows, cols, mode=1, add=False)
    drop = O.builder_csr(ent, rows, cols, mode=2, add=False)
    assert keep.nnz == keep_expected
    assert drop.nnz == drop_expected
    # both transposes agree on the non-zero entries (is_equal in sm_transpose.cc)
    a = keep.to_scipy().toarray() if rows * cols <= 1e6 else None
    if a is not None:
        assert np.array_equal(a, drop.to_scipy().toarray())


def test_boost_test0_golden():
    g = golden()["boost_test0"]
    N = 7
    ent = [(0, N - 1, -1.0), (1, N - 1, 0.0), (0, 0, 2.0)]
    for i in range(1, N):
        ent += [(i, i, 2.0), (i, i - 1, 1.0), (i - 1, i, 1.0)]
    A = O.builder_csr(ent, N, N, mode=0, add=False)
    B = O.builder_csr(ent, N, N, mode=1, add=False)
    # explicit zero A(1,6) is stored but not adjacent in the graph view
    assert 6 in A.col[A.rowptr[1]:A.rowptr[2]]
    for r in range(N):
        cols = A.col[A.rowptr[r]:A.rowptr[r + 1]]
        vals = A.val[A.rowptr[r]:A.rowptr[r + 1]]
        nz = [int(c) for c, v in zip(cols, vals) if v != 0.0]
        assert nz == g["A_nonzero_adjacency"][str(r)]
        assert len(nz) == g["A_out_degree"][str(r)]
    cols = B.col[B.rowptr[1]:B.rowptr[2]]
    vals = B.val[B.rowptr[1]:B.rowptr[2]]
    assert [int(c) for c in cols] == g["B_row1_adjacent"]
    assert [[1, int(c), float(v)] for c, v in zip(cols, vals)] == g["B_row1_edges"]


# ---------------------------------------------------------------------------
# oracle vs the reference's own SparseMatrix/Vector (oracle/_ref)
# ---------------------------------------------------------------------------

ref = O.ref_lib()
needs_ref = pytest.mark.skipif(ref is None, reason="oracle/_ref not built (reference checkout absent)")


def ref_csr(h):
    n = ref.ref_sm_rows(h)
    nnz = ref.ref_sm_nnz(h)
    rp = np.zeros(n + 1, np.int64)
    col = np.zeros(nnz, np.int32)
    val = np.zeros(nnz, np.float64)
    ref.ref_sm_csr(h, rp, col, val)
    return rp, col, val


@needs_ref
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_insertion_semantics_match_reference(seed):
    rng = np.random.default_rng(seed)
    rows, cols = 40, 35
    ent = []
    for _ in range(600):
        r, c = int(rng.integers(rows)), int(rng.integers(cols))
        v = float(rng.choice([0.0, rng.normal()]))
        ent.append((r, c, v))
    h = ref.ref_sm_create(rows, cols)
    for r, c, v in ent:
        ref.ref_sm_add(h, r, c, v)
    for mode in (0, 1, 2):
        hd = ref.ref_sm_derive(h, mode)
        rp, col, val = ref_csr(hd)
        mine = O.builder_csr(ent, rows, cols, mode=mode, add=True)
        assert np.array_equal(rp, mine.rowptr)
        assert np.array_equal(col, mine.col)
        assert np.array_equal(val, mine.val)
        ref.ref_sm_destroy(hd)
    # SpMV variants: axpy, apply_ignore_zero_rows, axpy_transposed
    A = O.builder_csr(ent, rows, cols, mode=0, add=True)
    for op in (0, 1, 2):
        nd, nw = (cols, rows) if op == 2 else (rows, cols)
        w = rng.normal(size=nw)
        v1 = rng.normal(size=nd)
        for alpha, beta in ((0.0, 1.0), (1.0, -1.0), (0.5, 2.0)):
            d_ref = rng.normal(size=nd)
            d_orc = d_ref.copy()
            ref.ref_sm_op(h, op, alpha, v1.ctypes.data, beta, w, d_ref)
            O.lib().orc_csr_op(op, rows, cols, A.rowptr, A.col, A.val, alpha, v1.ctypes.data, beta, w, d_orc)
            assert np.allclose(d_orc, d_ref, rtol=1e-14, atol=1e-14)
    x = rng.normal(size=100)
    y = rng.normal(size=100)
    assert abs(ref.ref_vec_dot(100, x, y) - float(np.dot(x, y))) < 1e-12
    ref.ref_sm_destroy(h)


@needs_ref
def test_oracle_level_matrix_rebuilt_by_reference_sparsematrix():
    """Scatter the oracle's assembled Q1 operator entry by entry into the
    reference SparseMatrix in shuffled order: the reference's defragmented CSR
    must equal the oracle's (sorted rows, explicit zeros kept)."""
    p = O.Problem(0, 3)
    p.set_theta(O.make_theta(8, 3, 3))
    p.assemble_all()
    A = p.A(3)
    rng = np.random.default_rng(0)
    rows = np.repeat(np.arange(A.n_rows), np.diff(A.rowptr))
    order = rng.permutation(A.nnz)
    h = ref.ref_sm_create(A.n_rows, A.n_rows)
    for k in order:
        ref.ref_sm_add(h, int(rows[k]), int(A.col[k]), float(A.val[k]))
    rp, col, val = ref_csr(h)
    assert np.array_equal(rp, A.rowptr)
    assert np.array_equal(col, A.col)
    assert np.array_equal(val, A.val)
    ref.ref_sm_destroy(h)


# ---------------------------------------------------------------------------
# closed forms and invariants of the discretisation
# ---------------------------------------------------------------------------

def lattice_index(lat, W, dim=3):
    if dim == 3:
        return (lat[:, 2] * W + lat[:, 1]) * W + lat[:, 0]
    return lat[:, 1] * W + lat[:, 0]


@pytest.mark.parametrize("refs", [2, 3])
def test_q1_closed_form_stencil(refs):
    p = O.Problem(0, refs)
    p.assemble_all()
    top = refs
    n = 2 ** refs
    h = 1.0 / n
    A = p.A(top).to_scipy().tocsr()
    lat = p.dof_lattice(top)
    dm = p.dirichlet(top)
    W = n + 1
    assert A.nnz == (3 * W - 2) ** 3
    expect = {0: 8.0 / 3.0, 1: 0.0, 2: -1.0 / 6.0, 3: -1.0 / 12.0}
    b = p.rhs(top)
    for r in range(A.shape[0]):
        row = A.getrow(r)
        if dm[r]:
            assert np.array_equal(row.toarray().ravel()[row.indices], (row.indices == r).astype(float))
            assert b[r] == 0.0
            continue
        assert abs(row.sum()) < 1e-14
        for c, v in zip(row.indices, row.data):
            d = int(np.abs(lat[c] - lat[r]).sum() // (2 ** (refs - top)))
            assert abs(v - h * expect[d]) < 1e-14
        assert abs(b[r] - h ** 3) < 1e-15


def test_fv1_p1_closed_form_stencil():
    """FV1 on P1 triangles equals the P1 stiffness: 5-point stencil
    (4,-1,-1,-1,-1) with explicit zeros on the (1,1)-diagonal couplings."""
    refs = 4
    p = O.Problem(1, refs)
    n = 2 ** refs
    f = np.zeros(p.num_dofs(refs))
    p.set_source_top(f)
    p.assemble_all()
    A = p.A(refs).to_scipy().tocsr()
    lat = p.dof_lattice(refs)
    dm = p.dirichlet(refs)
    V = (n + 1) ** 2
    E = 3 * n * n + 2 * n   # axis edges 2n(n+1) + diagonals n^2
    assert A.nnz == V + 2 * E
    for r in np.where(dm == 0)[0]:
        row = A.getrow(r)
        for c, v in zip(row.indices, row.data):
            d = lat[c] - lat[r]
            if (d == 0).all():
                assert abs(v - 4.0) < 1e-13
            elif abs(d).sum() == 1:
                assert abs(v + 1.0) < 1e-13
            else:
                assert d[0] == d[1] and abs(v) < 1e-13


def test_prolongation_restriction_invariants():
    refs = 3
    p = O.Problem(0, refs)
    p.assemble_all()
    for lev in range(1, refs + 1):
        P = p.P(lev).to_scipy().tocsr()
        R = p.R(lev).to_scipy().tocsr()
        dmf = p.dirichlet(lev)
        dmc = p.dirichlet(lev - 1)
        assert set(np.unique(P.data)) <= {0.0, 0.125, 0.25, 0.5, 1.0}
        rs = np.asarray(P.sum(axis=1)).ravel()
        assert np.allclose(rs[dmf == 0], 1.0)
        # R = P^T except coarse Dirichlet rows, which are injections
        PT = P.T.tocsr()
        assert np.array_equal(PT.indptr, R.indptr) and np.array_equal(PT.indices, R.indices)
        Rd = R.toarray()
        PTd = PT.toarray()
        assert np.array_equal(Rd[dmc == 0], PTd[dmc == 0])
        assert np.allclose(Rd[dmc == 1].sum(axis=1), 1.0)


def test_base_lu_solves_exactly():
    """GMG with base level == top level is the dense LU: CG converges in one
    step and matches a direct solve (textbook LU, numpy)."""
    p = O.Problem(0, 2)
    p.set_theta(O.make_theta(4, 11, 3))
    p.assemble_all()
    r = p.solve(base_lev=2)
    assert r["ok"] and r["steps"] == 1
    A = p.A(2).to_scipy().toarray()
    x = np.linalg.solve(A, p.rhs(2))
    assert np.allclose(r["x"], x, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("kind,refs", [(0, 3), (0, 4), (1, 5)])
def test_gmg_cg_matches_direct_solve(kind, refs):
    import scipy.sparse.linalg as spla
    p = O.Problem(kind, refs)
    n = 2 ** refs
    if kind == 0:
        p.set_theta(O.make_theta(n, 5, 3))
    else:
        p.set_source_top(np.ones(p.num_dofs(refs)))
    p.assemble_all()
    r = p.solve(base_lev=0)
    assert r["ok"]
    assert r["steps"] <= 12
    A = p.A(refs).to_scipy().tocsc()
    x = spla.spsolve(A, p.rhs(refs))
    assert np.linalg.norm(r["x"] - x) / np.linalg.norm(x) < 1e-8
    d = r["defects"]
    assert d[-1] / d[0] < 1e-10


def test_gmg_mesh_independent_convergence():
    steps = []
    for refs in (3, 4, 5):
        p = O.Problem(0, refs)
        p.assemble_all()
        steps.append(p.solve(base_lev=0)["steps"])
    assert max(steps) - min(steps) <= 2


def test_bicgstab_gmg():
    p = O.Problem(0, 4)
    p.set_theta(O.make_theta(16, 2, 3))
    p.assemble_all()
    r = p.solve(solver="bicgstab", base_lev=0)
    rc = p.solve(solver="cg", base_lev=0)
    assert r["ok"] and rc["ok"]
    assert np.linalg.norm(r["x"] - rc["x"]) / np.linalg.norm(rc["x"]) < 1e-8


def test_fv1_manufactured_solution_second_order():
    errs = []
    for refs in (4, 5, 6):
        p = O.Problem(1, refs)
        lat = p.dof_lattice(refs).astype(float) / 2 ** refs
        x, y = lat[:, 0], lat[:, 1]
        p.set_source_top(2 * np.pi ** 2 * np.sin(np.pi * x) * np.sin(np.pi * y))
        p.assemble_all()
        r = p.solve(base_lev=0)
        u = np.sin(np.pi * x) * np.sin(np.pi * y)
        errs.append(np.abs(r["x"] - u).max())
    rates = [np.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert all(1.8 < q < 2.2 for q in rates), rates


# ---------------------------------------------------------------------------
# product host-side numbering vs the oracle (no device needed)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("kind,refs", [(0, 1), (0, 2), (0, 3), (0, 4), (1, 3), (1, 5), (1, 6)])
def test_product_numbering_matches_oracle(kind, refs):
    import ugcore_amd as U
    p = O.Problem(kind, refs)
    dim = 3 if kind == 0 else 2
    for lev in range(refs + 1):
        W = 2 ** lev + 1
        lat = p.dof_lattice(lev) // (2 ** (refs - lev))
        expect = np.full(W ** dim, -1, np.int64)
        expect[lattice_index(lat, W, dim)] = np.arange(len(lat))
        got, n_inner = U.hierarchy_numbering(kind, refs, lev)
        assert np.array_equal(got, expect)
        assert n_inner == int((p.dirichlet(lev) == 0).sum())


def test_theta_generators_agree():
    from ugcore_amd.synthetic import make_theta
    for n in (4, 8, 16):
        assert np.abs(make_theta(n, 9, 3) - O.make_theta(n, 9, 3)).max() < 1e-13


def _tet_cells_from_oracle(p, lev):
    """oracle tetrahedra of a level as per-cell lists of packed corner codes"""
    refs = p.num_refs
    step = 2 ** (refs - lev)
    n = 3 * 2 ** lev
    lat = p.dof_lattice(lev) // step
    E = p.elements(lev)
    cells = [[] for _ in range(n ** 3)]
    for t in E:
        q = lat[t]
        mn = q.min(axis=0)
        c = q - mn
        code = 0
        for j in range(4):
            code |= int(c[j, 0] | (c[j, 1] << 1) | (c[j, 2] << 2)) << (3 * j)
        cells[(mn[2] * n + mn[1]) * n + mn[0]].append(code)
    return cells


@pytest.mark.parametrize("refs", [1, 2, 3])
def test_product_tet_hierarchy_matches_oracle(refs):
    import ugcore_amd as U
    p = O.Problem(2, refs)
    p.assemble_all(1.0)
    for lev in range(refs + 1):
        W = 3 * 2 ** lev + 1
        lat = p.dof_lattice(lev) // (2 ** (refs - lev))
        expect = np.full(W ** 3, -1, np.int64)
        expect[lattice_index(lat, W, 3)] = np.arange(len(lat))
        got, n_inner = U.hierarchy_numbering(2, refs, lev)
        assert np.array_equal(got, expect)
        assert n_inner == int((p.dirichlet(lev) == 0).sum())
        cells, par = U.hierarchy_tets(refs, lev)
        ocells = _tet_cells_from_oracle(p, lev)
        assert all(len(c) == 6 for c in ocells)
        assert np.array_equal(cells, np.array(ocells, np.uint16))
        if lev > 0:
            # parent codes: lattice points with all-even coordinates are copies
            lw = np.arange(W ** 3)
            x, y, z = lw % W, (lw // W) % W, lw // (W * W)
            copy = (x % 2 == 0) & (y % 2 == 0) & (z % 2 == 0)
            assert np.all(par[copy] == 0) and np.all(par[~copy] > 0) and par.max() <= 13
            # interior edge children: the code is the direction of the two
            # coarse parents in the oracle's P
            P = p.P(lev).to_scipy().tocsr()
            latc = p.dof_lattice(lev - 1) // (2 ** (refs - lev + 1))
            dm = p.dirichlet(lev).astype(bool)
            for r in np.where(~dm)[0]:
                cols = P.indices[P.indptr[r]:P.indptr[r + 1]]
                li = lattice_index(lat[r:r + 1], W, 3)[0]
                if len(cols) == 1:
                    assert par[li] == 0
                    continue
                d = latc[cols[1]] - latc[cols[0]]
                c = (d[0] + 1) + 3 * (d[1] + 1) + 9 * (d[2] + 1)
                c = c if c > 13 else 26 - c
                assert par[li] == c - 13
