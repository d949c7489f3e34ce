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


def test_empty_batch():
    say("case q2 empty batch")
    Kg, fg = U().q2_hex_element_matrices(np.zeros((0, 8, 3)))
    assert Kg.shape == (0, 27, 27) and fg.shape == (0, 27)
