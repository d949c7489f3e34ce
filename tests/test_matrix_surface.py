"""The SparseMatrix surface of GPUSparseMatrix that needs no device: the
host-side pattern build (resize_and_clear, operator()(r, c) as insert-or-get,
const access, row access, resize_and_keep_values, clear_retain_structure,
scale and SetDirichletRow on a host build) against the oracle's SparseMatrix
restatement (oracle_lib.builder_csr), which test_oracle.py pins to the
reference's own SparseMatrix compiled into oracle/_ref."""
import numpy as np
import pytest

import oracle_lib as O
import ugcore_amd as U


def random_entries(seed, rows, cols, count):
    rng = np.random.default_rng(seed)
    ent = []
    for _ in range(count):
        r, c = int(rng.integers(rows)), int(rng.integers(cols))
        v = float(rng.choice([0.0, rng.normal()]))
        ent.append((r, c, v))
    return ent


def host_rows(A):
    rows = []
    for r in range(A.num_rows()):
        rows.append(A.row(r))
    return rows


def rows_of_csr(M):
    return [(M.col[M.rowptr[r]:M.rowptr[r + 1]], M.val[M.rowptr[r]:M.rowptr[r + 1]]) for r in range(M.n_rows)]


@pytest.mark.parametrize("seed,add", [(0, True), (1, True), (2, False)])
def test_host_build_matches_insertion_semantics(seed, add):
    rows, cols = 40, 35
    ent = random_entries(seed, rows, cols, 600)
    A = U.SparseMatrix()
    A.resize_and_clear(rows, cols)
    assert A.building() and A.num_rows() == rows and A.num_cols() == cols
    for r, c, v in ent:
        if add:
            A.add(r, c, v)
        else:
            A[r, c] = v
    ref = O.builder_csr(ent, rows, cols, mode=0, add=add)
    assert A.nnz() == ref.nnz
    for (c1, v1), (c2, v2) in zip(host_rows(A), rows_of_csr(ref)):
        assert np.array_equal(c1, c2)          # sorted, explicit zeros kept
        assert np.array_equal(v1, v2)          # same additions in the same order
    # const access: stored value or 0
    for r, c, _ in ent[:50]:
        assert A.has_connection(r, c)
    M = ref.to_scipy().toarray()
    for r in range(rows):
        for c in range(0, cols, 7):
            assert A[r, c] == M[r, c]


def test_resize_keep_values_and_clear_on_a_host_build():
    ent = random_entries(5, 20, 30, 200)
    A = U.SparseMatrix()
    A.resize_and_clear(20, 30)
    for r, c, v in ent:
        A.add(r, c, v)
    before = host_rows(A)
    A.resize_and_keep_values(25, 18)           # drop columns >= 18, add 5 empty rows
    after = host_rows(A)
    assert len(after) == 25 and all(len(c) == 0 for c, _ in after[20:])
    for (c0, v0), (c1, v1) in zip(before, after[:20]):
        keep = c0 < 18
        assert np.array_equal(c0[keep], c1) and np.array_equal(v0[keep], v1)
    assert A.nnz() == sum(len(c) for c, _ in after)
    A.scale(-2.0)
    for (c0, v0), (c1, v1) in zip(after, host_rows(A)):
        assert np.array_equal(c0, c1) and np.array_equal(-2.0 * v0, v1)
    A.clear_retain_structure()
    for (c0, _), (c1, v1) in zip(after, host_rows(A)):
        assert np.array_equal(c0, c1) and not np.any(v1)


def test_dirichlet_rows_on_a_host_build():
    """SetDirichletRow: the row's entries 0 and A(i,i) = 1, created if absent"""
    A = U.SparseMatrix()
    A.resize_and_clear(4, 4)
    A[0, 1] = 3.0
    A[1, 0] = -1.0
    A[1, 1] = 2.0
    A[1, 3] = 5.0
    A.set_dirichlet_rows([0, 1])
    c, v = A.row(0)
    assert list(c) == [0, 1] and list(v) == [1.0, 0.0]
    c, v = A.row(1)
    assert list(c) == [0, 1, 3] and list(v) == [0.0, 1.0, 0.0]


def test_host_build_errors():
    A = U.SparseMatrix()
    A.resize_and_clear(3, 3)
    with pytest.raises(U.UGError, match="out of range"):
        A[3, 0] = 1.0
    with pytest.raises(U.UGError, match="out of range"):
        A.add(0, -1, 1.0)
    with pytest.raises(U.UGError, match="built on the host"):
        A.download_csr()
    with pytest.raises(U.UGError, match="built on the host"):
        A.device_csr()
    with pytest.raises(U.UGError):
        A.resize_and_clear(-1, 2)


def test_element_discretization_argument_checks():
    """connectivity errors are refused before any device work"""
    X = np.zeros((4, 3))
    with pytest.raises(ValueError, match="elem_type"):
        U.ElementDiscretization("prism", np.zeros((1, 6), np.int32), np.zeros((6, 3)))
    with pytest.raises(ValueError, match="shape"):
        U.ElementDiscretization("tetrahedron_fv1", np.zeros((1, 3), np.int32), X)
    with pytest.raises(ValueError, match="shape"):
        U.ElementDiscretization("triangle_fv1", np.zeros((1, 3), np.int32), X)
    with pytest.raises(U.UGError, match="out of range"):
        U.ElementDiscretization("tetrahedron_fv1", np.array([[0, 1, 2, 4]], np.int32), X)
    with pytest.raises(U.UGError, match="out of range"):
        U.ElementDiscretization("tetrahedron_fv1", np.array([[0, 1, -1, 3]], np.int32), X)
