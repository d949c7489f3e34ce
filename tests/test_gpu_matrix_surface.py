"""GPU tests of the SparseMatrix surface of GPUSparseMatrix against the oracle's
SparseMatrix restatement (oracle_lib.builder_csr / orc_csr_op, pinned to the
reference's SparseMatrix in oracle/_ref by test_oracle.py):

* a host build moved to the device by finalize(): CSR bit-exact, SpMV equal;
  operator() refused afterwards;
* set_as_transpose_of (explicit zeros kept) and set_as_transpose_of2 (zeros
  dropped) on the device: pattern and values bit-exact, also scaled, empty
  rows and columns, an empty matrix and the assembled 32^3 Q1 operator;
* axpy_transposed: every alpha/beta case of the reference (alpha 0, 1, other;
  dest aliasing v1), zero entries skipped (a NaN in w behind a zero entry
  stays out), and the 64^3 Q1 operator against scipy; 1e-14 relative;
* SetDirichletRow on the device, set_as_copy_of, scale, clear_retain_structure,
  resize_and_keep_values from a device CSR.
"""
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


def random_entries(seed, rows, cols, count):
    rng = np.random.default_rng(seed)
    ent = []
    for _ in range(count):
        r, c = int(rng.integers(rows)), int(rng.integers(cols))
        v = float(rng.choice([0.0, rng.normal()]))
        ent.append((r, c, v))
    return ent


def built(ent, rows, cols):
    u = U()
    A = u.SparseMatrix()
    A.resize_and_clear(rows, cols)
    for r, c, v in ent:
        A.add(r, c, v)
    A.finalize()
    return A


def same_csr(A, M):
    rp, col, val = A.download_csr()
    assert np.array_equal(rp.astype(np.int64), M.rowptr)
    assert np.array_equal(col, M.col)
    assert np.array_equal(val, M.val)


@pytest.mark.parametrize("seed", [0, 1])
def test_finalized_host_build(seed):
    say(f"case host build -> device seed={seed}")
    u = U()
    rows, cols = 60, 45
    ent = random_entries(seed, rows, cols, 900)
    A = built(ent, rows, cols)
    assert not A.building()
    ref = O.builder_csr(ent, rows, cols, mode=0, add=True)
    same_csr(A, ref)
    rng = np.random.default_rng(seed)
    w = rng.normal(size=cols)
    v1 = rng.normal(size=rows)
    for alpha, beta in ((0.0, 1.0), (1.0, -1.0), (0.5, 2.0)):
        d = u.Vector(values=v1)
        A.axpy(d, alpha, u.Vector(values=v1), beta, u.Vector(values=w))
        d_orc = rng.normal(size=rows)
        O.lib().orc_csr_op(0, rows, cols, ref.rowptr, ref.col, ref.val, alpha, v1.ctypes.data, beta, w, d_orc)
        assert np.allclose(d.numpy(), d_orc, rtol=1e-14, atol=1e-14)
    with pytest.raises(u.UGError, match="finalised on the device"):
        A[0, 0] = 1.0
    # const access and rows of the device CSR
    M = ref.to_scipy().toarray()
    for r in range(0, rows, 5):
        c, v = A.row(r)
        assert np.array_equal(c, ref.col[ref.rowptr[r]:ref.rowptr[r + 1]])
        assert np.array_equal(v, ref.val[ref.rowptr[r]:ref.rowptr[r + 1]])
        for cc in range(0, cols, 4):
            assert A[r, cc] == M[r, cc]


@pytest.mark.parametrize("seed,rows,cols,count", [(0, 40, 35, 600), (1, 300, 200, 4000), (2, 7, 90, 30)])
def test_transposes_match_oracle(seed, rows, cols, count):
    say(f"case transposes seed={seed} {rows}x{cols}")
    u = U()
    ent = random_entries(seed, rows, cols, count)
    A = built(ent, rows, cols)
    for mode, fn in ((1, "set_as_transpose_of"), (2, "set_as_transpose_of2")):
        ref = O.builder_csr(ent, rows, cols, mode=mode, add=True)
        for scale in (1.0, 2.5):
            T = u.SparseMatrix()
            getattr(T, fn)(A, scale)
            assert T.num_rows() == cols and T.num_cols() == rows
            rp, col, val = T.download_csr()
            assert np.array_equal(rp.astype(np.int64), ref.rowptr), fn
            assert np.array_equal(col, ref.col), fn
            assert np.array_equal(val, scale * ref.val), fn


def test_transpose_of_empty_matrix_and_empty_rows():
    say("case transposes of empty matrices")
    u = U()
    A = built([], 5, 3)
    assert A.nnz() == 0
    T = u.SparseMatrix()
    T.set_as_transpose_of(A)
    rp, col, _ = T.download_csr()
    assert T.num_rows() == 3 and np.array_equal(rp, np.zeros(4, np.int32)) and col.size == 0
    A = built([(1, 2, 0.0), (3, 2, 1.5)], 5, 4)      # only zeros and one value in one column
    T2 = u.SparseMatrix()
    T2.set_as_transpose_of2(A)
    rp, col, val = T2.download_csr()
    assert list(rp) == [0, 0, 0, 1, 1] and list(col) == [3] and list(val) == [1.5]


def test_transpose_of_assembled_operator():
    say("case transpose of the 32^3 Q1 operator")
    u = U()
    refs = 5
    theta = O.make_theta(32, 3, 3)
    p = O.Problem(0, refs)
    p.set_theta(theta)
    p.assemble(refs)
    d = u.DomainDiscretization(0, refs)
    d.set_theta(theta)
    A = u.SparseMatrix()
    b = u.Vector(d.num_dofs())
    d.assemble_linear(A, b)
    R = p.A(refs)
    Rs = R.to_scipy()
    for keep in (True, False):
        T = u.SparseMatrix()
        (T.set_as_transpose_of if keep else T.set_as_transpose_of2)(A)
        rp, col, val = T.download_csr()
        St = Rs.T.tocsr()
        if not keep:
            St.eliminate_zeros()
        St.sort_indices()
        assert np.array_equal(rp, St.indptr) and np.array_equal(col, St.indices)
        assert np.abs(val - St.data).max() <= 1e-12 * np.abs(St.data).max()


@pytest.mark.parametrize("seed", [0, 3])
def test_axpy_transposed_matches_oracle(seed):
    say(f"case axpy_transposed seed={seed}")
    u = U()
    rows, cols = 70, 55
    ent = random_entries(seed, rows, cols, 1100)
    A = built(ent, rows, cols)
    ref = O.builder_csr(ent, rows, cols, mode=0, add=True)
    rng = np.random.default_rng(seed + 10)
    w = rng.normal(size=rows)
    v1 = rng.normal(size=cols)
    for alpha, beta in ((0.0, 1.0), (1.0, -1.0), (0.5, 2.0)):
        for alias in (False, True):
            start = rng.normal(size=cols)
            d = u.Vector(values=v1 if alias else start)
            A.axpy_transposed(d, alpha, d if alias else u.Vector(values=v1), beta, u.Vector(values=w))
            d_orc = (v1 if alias else start).copy()
            O.lib().orc_csr_op(2, rows, cols, ref.rowptr, ref.col, ref.val, alpha,
                               d_orc.ctypes.data if alias else v1.ctypes.data, beta, w, d_orc)
            assert np.allclose(d.numpy(), d_orc, rtol=1e-14, atol=1e-14), (alpha, beta, alias)
    # a NaN in w behind zero entries only stays out of the result (the
    # reference skips zero entries)
    A2 = built([(0, 1, 0.0), (0, 3, 0.0), (1, 1, 2.0)], 2, 4)
    d = u.Vector(4)
    A2.apply_transposed(d, u.Vector(values=np.array([np.nan, 3.0])))
    assert np.array_equal(d.numpy(), [0.0, 6.0, 0.0, 0.0])


def test_axpy_transposed_on_assembled_operator_and_after_edits():
    """the 64^3 Q1 operator (274,625 rows): A^T w against scipy; after scale()
    the cached transpose follows the new values"""
    say("case axpy_transposed 64^3")
    u = U()
    refs = 6
    theta = O.make_theta(64, 5, 3)
    d = u.DomainDiscretization(0, refs)
    d.set_theta(theta)
    A = u.SparseMatrix()
    b = u.Vector(d.num_dofs())
    d.assemble_linear(A, b)
    rp, col, val = A.download_csr()
    import scipy.sparse as sp
    S = sp.csr_matrix((val, col, rp), shape=(len(rp) - 1,) * 2)
    w = np.random.default_rng(1).normal(size=S.shape[0])
    y = u.Vector(S.shape[0])
    A.apply_transposed(y, u.Vector(values=w))
    ref = S.T @ w
    assert np.linalg.norm(y.numpy() - ref) <= 1e-13 * np.linalg.norm(ref)
    A.scale(3.0)
    A.apply_transposed(y, u.Vector(values=w))
    assert np.linalg.norm(y.numpy() - 3.0 * ref) <= 1e-13 * np.linalg.norm(3.0 * ref)


def test_device_dirichlet_rows_copy_scale_clear_and_keep_values():
    say("case device dirichlet rows, copy, scale, clear, keep values")
    u = U()
    rows, cols = 30, 30
    ent = random_entries(7, rows, cols, 300) + [(r, r, 4.0) for r in range(rows)]
    A = built(ent, rows, cols)
    ref = O.builder_csr(ent, rows, cols, mode=0, add=True)
    M = ref.to_scipy().toarray()
    dr = [0, 5, 17, 29]
    A.set_dirichlet_rows(dr)
    for r in dr:
        M[r, :] = 0.0
        M[r, r] = 1.0
    rp, col, val = A.download_csr()
    assert np.array_equal(rp.astype(np.int64), ref.rowptr) and np.array_equal(col, ref.col)   # pattern kept
    Md = np.zeros((rows, cols))
    for r in range(rows):
        Md[r, col[rp[r]:rp[r + 1]]] = val[rp[r]:rp[r + 1]]
    assert np.array_equal(Md, M)
    # a row without its diagonal cannot become a Dirichlet row on the device
    B = built([(0, 1, 1.0), (1, 1, 1.0)], 2, 2)
    with pytest.raises(u.UGError, match="diagonal"):
        B.set_dirichlet_rows([0])
    # copy, scale, clear
    C2 = u.SparseMatrix()
    C2.set_as_copy_of(A, -0.5)
    rp2, col2, val2 = C2.download_csr()
    assert np.array_equal(rp2, rp) and np.array_equal(col2, col) and np.array_equal(val2, -0.5 * val)
    C2.scale(-2.0)
    assert np.array_equal(C2.download_csr()[2], val)
    C2.clear_retain_structure()
    rp3, col3, val3 = C2.download_csr()
    assert np.array_equal(col3, col) and not np.any(val3)
    # resize_and_keep_values moves the device CSR into a host build
    A.resize_and_keep_values(rows, 20)
    assert A.building()
    A.add(0, 19, 1.0)
    A.finalize()
    rp4, col4, val4 = A.download_csr()
    for r in range(rows):
        c_old = col[rp[r]:rp[r + 1]]
        keep = c_old < 20
        c_new = col4[rp4[r]:rp4[r + 1]]
        expect = c_old[keep] if r else np.unique(np.append(c_old[keep], 19))
        assert np.array_equal(c_new, expect)
