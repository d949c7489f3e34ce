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
