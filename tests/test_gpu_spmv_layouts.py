"""GPU tests of the SpMV layouts of GPUSparseMatrix on the operators the
lattice discretisations assemble: the sliced ELL with a column dictionary per
slice (64 rows of a 4x4x4 / 8x8 block, one-byte column positions) gives the
sliced-ELL result (bit for bit where the slices have the same widths), the CSR
result within 1e-14 relative, and the oracle's SpMV within 1e-14; alpha/beta
variants, apply_sub, and a rebuild after raw CSR edits."""
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


@pytest.mark.parametrize("kind,refs", [(0, 3), (0, 5), (1, 6), (2, 2)])
def test_dictionary_layout_equals_sliced_ell(kind, refs):
    say(f"case spmv layouts kind={kind} refs={refs}")
    u = U()
    d = u.DomainDiscretization(kind, refs)
    rng = np.random.default_rng(refs)
    if kind != 1:
        d.set_theta(rng.normal(size=d.n_fine ** d.dim) * 0.5)
    if kind == 2:
        d.set_velocity((20.0, 10.0, 5.0))
    A = u.SparseMatrix()
    b = u.Vector(d.num_dofs())
    d.assemble_linear(A, b)
    assert A.spmv_layout() == "dict"
    N = d.num_dofs()
    rp, col, val = A.download_csr()
    x = rng.normal(size=N)
    v = rng.normal(size=N)
    ref = np.zeros(N)
    O.lib().orc_csr_op(0, N, N, rp.astype(np.int64), col, val, 0.0, None, 1.0, x, ref)
    res = {}
    for layout in ("dict", "sell", "csr"):
        A.set_spmv_layout(layout)
        assert A.spmv_layout() == layout
        for alpha, beta in ((0.0, 1.0), (0.5, -2.0)):
            y = u.Vector(values=v)
            A.axpy(y, alpha, u.Vector(values=v), beta, u.Vector(values=x))
            res[(layout, alpha)] = y.numpy()
        y = u.Vector(values=v)
        A.apply_sub(y, u.Vector(values=x))
        res[(layout, "sub")] = y.numpy()
    # both layouts sum a row's entries in two sums grouped by the slice width;
    # the hexahedral and triangular operators have slices of equal width in
    # both placements (bit for bit), the tetrahedral ones do not (rounding)
    for key in ((0.0,), (0.5,), ("sub",)):
        if kind != 2:
            assert np.array_equal(res[("dict",) + key], res[("sell",) + key]), key
        else:
            assert np.allclose(res[("dict",) + key], res[("sell",) + key], rtol=1e-15,
                               atol=1e-15 * np.abs(ref).max()), key
        assert np.allclose(res[("dict",) + key], res[("csr",) + key], rtol=1e-14, atol=1e-14 * np.abs(ref).max())
    assert np.abs(res[("dict", 0.0)] - ref).max() <= 1e-14 * np.abs(ref).max() * 10
    # raw edits drop the layout; finalize rebuilds it from the edited CSR
    A.set_spmv_layout("dict")
    _, _, vptr = A.device_csr()
    assert A.spmv_layout() == "csr"
    A.finalize()
    assert A.spmv_layout() == "dict"


def _banded_csr(rng, n, half_band, max_len):
    """random CSR whose row r has up to max_len columns in [r - half_band, r + half_band]
    (some rows empty), columns sorted"""
    rp, cols = [0], []
    for r in range(n):
        lo, hi = max(0, r - half_band), min(n, r + half_band + 1)
        k = int(rng.integers(0, min(max_len, hi - lo) + 1)) if r % 17 else 0
        cols.append(np.sort(rng.choice(np.arange(lo, hi), size=k, replace=False)))
        rp.append(rp[-1] + k)
    col = np.concatenate(cols).astype(np.int32)
    return np.array(rp, np.int32), col, rng.normal(size=col.size)


def _slices_with_holes(rng, n, rows_per_slice):
    """slice order: consecutive groups of rows_per_slice rows, shuffled inside
    their 64 slots (the rest -1), plus one empty slice"""
    order = []
    for g in range(0, n, rows_per_slice):
        slots = np.full(64, -1, np.int32)
        rows = np.arange(g, min(n, g + rows_per_slice), dtype=np.int32)
        slots[rng.choice(64, size=rows.size, replace=False)] = rows
        order.append(slots)
    order.append(np.full(64, -1, np.int32))
    return np.concatenate(order)


def _spmv_all(u, A, x, v):
    out = {}
    for alpha, beta in ((0.0, 1.0), (0.5, -2.0)):
        y = u.Vector(values=v)
        A.axpy(y, alpha, u.Vector(values=v), beta, u.Vector(values=x))
        out[alpha] = y.numpy()
    y = u.Vector(values=v)
    A.apply_sub(y, u.Vector(values=x))
    out["sub"] = y.numpy()
    return out


@pytest.mark.parametrize("seed,rows_per_slice", [(0, 64), (1, 50), (2, 7)])
def test_dictionary_layout_of_a_generic_matrix_with_given_slices(seed, rows_per_slice):
    say(f"case generic dictionary seed={seed} rows_per_slice={rows_per_slice}")
    u = U()
    rng = np.random.default_rng(seed)
    n = 2500
    rp, col, val = _banded_csr(rng, n, 90, 40)   # <= 64 + 180 distinct columns per slice
    A = u.SparseMatrix()
    A.upload_csr(rp, col, val)
    A.set_slice_order(_slices_with_holes(rng, n, rows_per_slice))
    assert A.spmv_layout() == "dict"
    x, v = rng.normal(size=n), rng.normal(size=n)
    import scipy.sparse as sp
    S = sp.csr_matrix((val, col, rp), shape=(n, n))
    ax = S @ x
    scale = np.abs(S).max() * np.abs(x).max() * 40
    got = _spmv_all(u, A, x, v)
    nonempty = np.diff(rp) > 0
    assert np.abs(got[0.0] - ax).max() <= 1e-14 * scale
    assert np.abs(got[0.5] - (0.5 * v - 2.0 * ax)).max() <= 1e-14 * scale
    assert np.abs(got["sub"] - (v - ax)).max() <= 1e-14 * scale
    A.set_spmv_layout("csr")
    csr = _spmv_all(u, A, x, v)
    A.set_spmv_layout("dict")
    assert A.spmv_layout() == "dict"
    for k in csr:
        assert np.abs(got[k] - csr[k]).max() <= 1e-14 * scale, k
    # values changed in place keep the layout (values rewritten, positions kept)
    A.scale(2.0)
    assert A.spmv_layout() == "dict"
    assert np.array_equal(_spmv_all(u, A, x, v)[0.0], 2.0 * got[0.0])
    rows = np.flatnonzero(nonempty & (np.arange(n) % 5 == 0))
    diag_rows = [r for r in rows if np.any(col[rp[r]:rp[r + 1]] == r)]
    A.set_dirichlet_rows(np.array(diag_rows, np.int32))
    assert A.spmv_layout() == "dict"
    S2 = (2.0 * S).tolil()
    for r in diag_rows:
        S2.rows[r] = [c for c in S2.rows[r]]
        for j, c in enumerate(S2.rows[r]):
            S2.data[r][j] = 1.0 if c == r else 0.0
    ref2 = S2.tocsr() @ x
    assert np.abs(_spmv_all(u, A, x, v)[0.0] - ref2).max() <= 2e-14 * scale


def test_slices_beyond_the_dictionary_keep_the_sliced_ell():
    say("case dictionary overflow")
    u = U()
    rng = np.random.default_rng(5)
    n = 4000
    rp, col, val = _banded_csr(rng, n, 2000, 30)   # far more than 255 distinct columns per slice
    A = u.SparseMatrix()
    A.upload_csr(rp, col, val)
    A.set_slice_order(_slices_with_holes(rng, n, 64))
    assert A.spmv_layout() == "sell"
    x = rng.normal(size=n)
    import scipy.sparse as sp
    ref = sp.csr_matrix((val, col, rp), shape=(n, n)) @ x
    y = u.Vector(n)
    A.apply(y, u.Vector(values=x))
    assert np.abs(y.numpy() - ref).max() <= 1e-13 * np.abs(ref).max()
    # a row longer than 64 entries in a slice also keeps the sliced ELL
    rp2, col2, val2 = _banded_csr(rng, 300, 100, 80)
    B = u.SparseMatrix()
    B.upload_csr(rp2, col2, val2)
    B.set_slice_order(_slices_with_holes(rng, 300, 64))
    assert B.spmv_layout() == "sell"
    with pytest.raises(u.UGError):
        B.set_slice_order(np.arange(64, dtype=np.int32))   # rows missing


@pytest.mark.parametrize("kind,refs", [(0, 4), (1, 5), (2, 2)])
def test_reassembly_keeps_the_pattern(kind, refs):
    """a second assemble_linear into the same matrix reuses its pattern and
    rewrites only the layout's values; the result equals a fresh assembly"""
    say(f"case reassembly kind={kind} refs={refs}")
    u = U()
    rng = np.random.default_rng(10 + kind)
    d = u.DomainDiscretization(kind, refs)
    if kind == 2:
        d.set_velocity((5.0, -3.0, 2.0))
    N = d.num_dofs()
    ne = d.n_fine ** d.dim
    A, b = u.SparseMatrix(), u.Vector(N)
    d.set_theta(rng.normal(size=ne) * 0.4)
    d.assemble_linear(A, b)
    theta2 = rng.normal(size=ne) * 0.4
    d.set_theta(theta2)
    d.assemble_linear(A, b)
    d2 = u.DomainDiscretization(kind, refs)
    if kind == 2:
        d2.set_velocity((5.0, -3.0, 2.0))
    d2.set_theta(theta2)
    A2, b2 = u.SparseMatrix(), u.Vector(N)
    d2.assemble_linear(A2, b2)
    for p, q in zip(A.download_csr(), A2.download_csr()):
        assert np.array_equal(p, q)
    assert np.array_equal(b.numpy(), b2.numpy())
    assert A.spmv_layout() == A2.spmv_layout() == "dict"
    x = rng.normal(size=N)
    y, y2 = u.Vector(N), u.Vector(N)
    A.apply(y, u.Vector(values=x))
    A2.apply(y2, u.Vector(values=x))
    assert np.array_equal(y.numpy(), y2.numpy())
    # after an edit that drops the provenance, assembly rebuilds everything
    A.scale(3.0)
    d.assemble_linear(A, b)
    A.apply(y, u.Vector(values=x))
    assert np.array_equal(y.numpy(), y2.numpy())


def _hex_lattice(n):
    """connectivity and coordinates of n^3 unit-cube hexahedra (reference
    hexahedron vertex order), lattice-numbered DoFs"""
    W = n + 1
    g = np.arange(W) / n
    X = np.stack(np.meshgrid(g, g, g, indexing="ij")[::-1], -1).reshape(-1, 3)
    E = []
    for k in range(n):
        for j in range(n):
            for i in range(n):
                v = lambda a, b, c: ((k + c) * W + (j + b)) * W + (i + a)
                E.append([v(0, 0, 0), v(1, 0, 0), v(1, 1, 0), v(0, 1, 0),
                          v(0, 0, 1), v(1, 0, 1), v(1, 1, 1), v(0, 1, 1)])
    return np.array(E, np.int32), X


def test_element_reassembly_keeps_the_pattern():
    say("case element reassembly")
    u = U()
    rng = np.random.default_rng(3)
    E, X = _hex_lattice(6)
    N = X.shape[0]
    bnd = np.flatnonzero(np.any((X == 0) | (X == 1), axis=1)).astype(np.int32)
    e = u.ElementDiscretization("hexahedron_q1", E, X)
    e.set_dirichlet(bnd)
    e.set_source(1.0)
    A, b = u.SparseMatrix(), u.Vector(N)
    e.assemble_linear(A, b)
    k2 = np.exp(rng.normal(size=E.shape[0]) * 0.3)
    e.set_diffusion(k2)
    e.assemble_linear(A, b)
    e2 = u.ElementDiscretization("hexahedron_q1", E, X)
    e2.set_dirichlet(bnd)
    e2.set_source(1.0)
    e2.set_diffusion(k2)
    B, c = u.SparseMatrix(), u.Vector(N)
    e2.assemble_linear(B, c)
    for p, q in zip(A.download_csr(), B.download_csr()):
        assert np.array_equal(p, q)
    assert np.array_equal(b.numpy(), c.numpy())
    x = rng.normal(size=N)
    y, y2 = u.Vector(N), u.Vector(N)
    A.apply(y, u.Vector(values=x))
    B.apply(y2, u.Vector(values=x))
    assert np.array_equal(y.numpy(), y2.numpy())
    # the Jacobian into the same matrix (same pattern) equals the operator
    J = u.SparseMatrix()
    e.assemble_jacobian(J, u.Vector(values=x))
    e.assemble_jacobian(J, u.Vector(values=x))
    assert np.array_equal(J.download_csr()[2], A.download_csr()[2])
