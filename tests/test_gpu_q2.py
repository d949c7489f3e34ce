"""Q2 hexahedron element matrices on the device (ugx_q2_hex_element_matrices,
one element per wavefront with the FP64 matrix cores) against the oracle's
restatement: jittered trilinear elements, boxes, a batch size that is not a
multiple of anything, per-element and unit coefficients, the empty batch.
Tolerance: 1e-12 relative to the largest entry of each element matrix (the
contraction sums the 81 point-component products in another order)."""
import sys

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

CORNERS = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)


def U():
    import ugcore_amd
    return ugcore_amd


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _check(Kg, fg, Ko, fo):
    for e in range(Ko.shape[0]):
        s = np.max(np.abs(Ko[e]))
        assert np.max(np.abs(Kg[e] - Ko[e])) <= 1e-12 * s, e
        sf = np.max(np.abs(fo[e]))
        if sf > 0:
            assert np.max(np.abs(fg[e] - fo[e])) <= 1e-12 * sf, e


@pytest.mark.parametrize("n,seed", [(1, 0), (37, 1), (1001, 2)])
def test_jittered_elements_match_oracle(n, seed):
    say(f"case q2 elements n={n} seed={seed}")
    rng = np.random.default_rng(seed)
    X = (np.repeat(CORNERS[None], n, axis=0) + rng.uniform(-0.15, 0.15, (n, 8, 3))) * rng.uniform(0.05, 3.0, (n, 1, 1))
    X += rng.normal(size=(n, 1, 3))
    k = np.exp(rng.normal(scale=0.7, size=n))
    Kg, fg = U().q2_hex_element_matrices(X, k, 1.3)
    Ko, fo = O.q2_hex_elements(X, k, 1.3)
    _check(Kg, fg, Ko, fo)
    assert np.max(np.abs(Kg - np.transpose(Kg, (0, 2, 1)))) <= 1e-13 * np.max(np.abs(Kg))


def test_boxes_unit_coefficient_and_no_source():
    say("case q2 boxes")
    h = np.array([[0.5, 0.5, 0.5], [0.25, 1.0, 2.0], [1e-3, 1e-3, 1e-3]])
    X = CORNERS[None] * h[:, None, :]
    Kg, fg = U().q2_hex_element_matrices(X)
    Ko, fo = O.q2_hex_elements(X)
    _check(Kg, fg, Ko, fo)
    assert np.all(fg == 0.0)


def test_null_load_vector_output():
    """fe = NULL (documented in the C ABI): the element matrices alone, equal
    to those of a call that also writes the load vectors"""
    say("case q2 fe=NULL")
    u = U()
    rng = np.random.default_rng(5)
    n = 70
    X = np.repeat(CORNERS[None], n, axis=0) + rng.uniform(-0.1, 0.1, (n, 8, 3))
    k = np.exp(rng.normal(size=n))
    K_ref, _ = u.q2_hex_element_matrices(X, k, 2.0)
    dx, dk = u.DeviceArray(np.ascontiguousarray(X)), u.DeviceArray(k)
    ke = u.DeviceArray(np.zeros((n, 27, 27)))
    assert u.lib().ugx_q2_hex_element_matrices(n, dx.ptr, dk.ptr, 2.0, ke.ptr, None) == 0
    u.synchronize()
    assert np.array_equal(ke.numpy(), K_ref)


def test_empty_batch():
    say("case q2 empty batch")
    Kg, fg = U().q2_hex_element_matrices(np.zeros((0, 8, 3)))
    assert Kg.shape == (0, 27, 27) and fg.shape == (0, 27)


def _oracle_lattice_system(n, theta, f):
    """the Q2 system of the unit cube of n^3 elements assembled from the
    oracle's element matrices, nodes of the (2n+1)^3 lattice x fastest, the
    pattern of all element couplings, Dirichlet rows = unit rows"""
    import scipy.sparse as sp
    W = 2 * n + 1
    e = np.arange(n ** 3)
    ex, ey, ez = e % n, (e // n) % n, e // (n * n)
    h = 1.0 / n
    X = (np.stack([ex, ey, ez], axis=1)[:, None, :] + CORNERS[None]) * h
    Ke, fe = O.q2_hex_elements(X, np.exp(theta), f)
    mi = O.q2_hex_multi_index()
    nodes = ((2 * ez[:, None] + mi[None, :, 2]) * W + (2 * ey[:, None] + mi[None, :, 1])) * W + (2 * ex[:, None] + mi[None, :, 0])
    rows = np.repeat(nodes, 27, axis=1).ravel()
    cols = np.tile(nodes, (1, 27)).ravel()
    N = W ** 3
    A = sp.csr_matrix((Ke.reshape(len(e), -1).ravel(), (rows, cols)), shape=(N, N))
    A.sum_duplicates()
    A.sort_indices()
    b = np.zeros(N)
    np.add.at(b, nodes.ravel(), fe.ravel())
    x, y, z = np.arange(N) % W, (np.arange(N) // W) % W, np.arange(N) // (W * W)
    bnd = (x == 0) | (y == 0) | (z == 0) | (x == W - 1) | (y == W - 1) | (z == W - 1)
    A = A.tolil()
    for r in np.nonzero(bnd)[0]:
        for c in list(A.rows[r]):   # a copy: storing 0 drops the entry from the row list
            A[r, c] = 1.0 if c == r else 0.0
    b[bnd] = 0.0
    A = A.tocsr()
    A.sort_indices()
    # the pattern keeps every element coupling, also the zeros just written
    P = sp.csr_matrix((np.ones(len(rows)), (rows, cols)), shape=(N, N))
    P.sum_duplicates()
    P.sort_indices()
    dense = A.toarray()
    r_idx = np.repeat(np.arange(N), np.diff(P.indptr))
    return P.indptr.astype(np.int64), P.indices.astype(np.int64), dense[r_idx, P.indices], b, A


# n = 7: several workgroups per colour launch, and odd n leaves colour slots
# without elements
@pytest.mark.parametrize("n", [1, 2, 3, 4, 7])
def test_lattice_assembly_matches_oracle_elements(n):
    say(f"case q2 lattice assembly n={n}")
    theta = np.random.default_rng(n).normal(scale=0.5, size=n ** 3)
    A, b = U().q2_assemble_lattice(n, theta, 1.0)
    rp, col, val = A.download_csr()
    rp_o, col_o, val_o, b_o, _ = _oracle_lattice_system(n, theta, 1.0)
    assert np.array_equal(rp, rp_o) and np.array_equal(col, col_o)
    s = np.max(np.abs(val_o))
    assert np.max(np.abs(val - val_o)) <= 1e-12 * s
    assert np.max(np.abs(b.numpy() - b_o)) <= 1e-12 * np.max(np.abs(b_o))


def test_lattice_q2_solve_matches_direct_solve():
    """CG with Jacobi on the device-assembled Q2 system against scipy's direct
    solve of the oracle-assembled one (the Q2 path has no multigrid transfer)"""
    import scipy.sparse.linalg as sla
    n = 6
    say(f"case q2 lattice solve n={n}")
    theta = np.random.default_rng(9).normal(scale=0.5, size=n ** 3)
    A, b = U().q2_assemble_lattice(n, theta, 1.0)
    _, _, _, b_o, A_o = _oracle_lattice_system(n, theta, 1.0)
    x_o = sla.spsolve(A_o.tocsc(), b_o)
    u = U()
    S = u.CG(u.Jacobi(0.66), max_steps=3000, reduction=1e-13)
    S.init(A)
    x = u.Vector(len(b_o))
    x.set(0.0)
    assert S.apply(x, b)
    xs = x.numpy()
    assert np.linalg.norm(xs - x_o) <= 1e-8 * np.linalg.norm(x_o)
