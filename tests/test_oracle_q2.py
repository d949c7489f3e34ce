"""The oracle's Q2 (triquadratic Lagrange) hexahedron element matrices, pinned
by closed forms: the reference's local shape order (lagrange.cpp:236-253 on the
reference hexahedron of reference_element.cpp), the tensor-product 1D mass and
stiffness matrices on boxes, the exact metric-tensor formula on sheared
parallelepipeds, and invariants of the trilinear case (symmetry, zero row
sums, one-dimensional kernel, rigid-motion invariance, linear scaling)."""
import numpy as np
import numpy.polynomial.polynomial as npp

import oracle_lib as O

CORNERS = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)


def test_multi_index_order():
    mi = O.q2_hex_multi_index()
    assert np.array_equal(mi[:8], 2 * CORNERS.astype(np.int32))
    edges = [(0, 1), (1, 2), (2, 3), (3, 0), (0, 4), (1, 5), (2, 6), (3, 7), (4, 5), (5, 6), (6, 7), (7, 4)]
    for e, (a, b) in enumerate(edges):
        assert np.array_equal(mi[8 + e], (CORNERS[a] + CORNERS[b]).astype(np.int32))
    faces = [(1, 1, 0), (1, 0, 1), (2, 1, 1), (1, 2, 1), (0, 1, 1), (1, 1, 2)]
    assert np.array_equal(mi[20:26], np.array(faces, np.int32))
    assert np.array_equal(mi[26], [1, 1, 1])
    assert len({tuple(r) for r in mi}) == 27


def _poly1d():
    nodes = [0.0, 0.5, 1.0]
    L = []
    for i in range(3):
        c = np.array([1.0])
        for m in range(3):
            if m != i:
                c = npp.polymul(c, np.array([-nodes[m], 1.0]) / (nodes[i] - nodes[m]))
        L.append(c)
    return L


def _int01(c):
    ci = npp.polyint(c)
    return npp.polyval(1.0, ci) - npp.polyval(0.0, ci)


def _mats1d():
    L = _poly1d()
    D = [npp.polyder(c) for c in L]
    M = np.array([[_int01(npp.polymul(L[a], L[b])) for b in range(3)] for a in range(3)])
    S = np.array([[_int01(npp.polymul(D[a], D[b])) for b in range(3)] for a in range(3)])
    Cm = np.array([[_int01(npp.polymul(D[a], L[b])) for b in range(3)] for a in range(3)])   # int phi_a' phi_b
    m = np.array([_int01(c) for c in L])
    return M, S, Cm, m


def _affine_exact(A, k, f):
    """exact Q2 matrices of the parallelepiped x = A xi (xi in the unit cube)"""
    mi = O.q2_hex_multi_index()
    M, S, Cm, m = _mats1d()
    G = np.linalg.inv(A) @ np.linalg.inv(A).T      # grad_x phi_a . grad_x phi_b = dxi_a^T G dxi_b
    det = abs(np.linalg.det(A))
    K = np.zeros((27, 27))
    for a in range(27):
        for b in range(27):
            s = 0.0
            for c in range(3):
                for d in range(3):
                    t = 1.0
                    for e in range(3):
                        i, j = mi[a][e], mi[b][e]
                        if e == c and e == d:
                            t *= S[i, j]
                        elif e == c:
                            t *= Cm[i, j]
                        elif e == d:
                            t *= Cm[j, i]
                        else:
                            t *= M[i, j]
                    s += G[c, d] * t
            K[a, b] = k * det * s
    fe = np.array([f * det * m[mi[a][0]] * m[mi[a][1]] * m[mi[a][2]] for a in range(27)])
    return K, fe


def test_box_matches_tensor_products():
    M, S, _, m = _mats1d()
    # the 1D Q2 matrices themselves: h = 1
    assert np.allclose(S * 3, [[7, -8, 1], [-8, 16, -8], [1, -8, 7]], atol=1e-14)
    assert np.allclose(M * 30, [[4, 2, -1], [2, 16, 2], [-1, 2, 4]], atol=1e-14)
    a, b, c, k, f = 0.3, 0.7, 1.9, 2.5, 1.75
    X = (CORNERS * [a, b, c] + [1.0, -2.0, 0.5])[None]
    Ke, fe = O.q2_hex_elements(X, [k], f)
    mi = O.q2_hex_multi_index()
    K = np.zeros((27, 27))
    h = (a, b, c)
    for p in range(27):
        for q in range(27):
            t = 0.0
            for d in range(3):
                prod = 1.0
                for e in range(3):
                    i, j = mi[p][e], mi[q][e]
                    prod *= S[i, j] / h[e] if e == d else M[i, j] * h[e]
                t += prod
            K[p, q] = k * t
    assert np.max(np.abs(Ke[0] - K)) <= 1e-13 * np.max(np.abs(K))
    F = np.array([f * m[mi[p][0]] * m[mi[p][1]] * m[mi[p][2]] * a * b * c for p in range(27)])
    assert np.max(np.abs(fe[0] - F)) <= 1e-14 * np.max(np.abs(F))


def test_sheared_parallelepiped_matches_metric_formula():
    rng = np.random.default_rng(11)
    for _ in range(3):
        A = np.eye(3) * 0.5 + rng.uniform(-0.15, 0.15, (3, 3))
        k, f = rng.uniform(0.5, 2.0), rng.uniform(-1, 1)
        X = (CORNERS @ A.T + rng.normal(size=3))[None]
        Ke, fe = O.q2_hex_elements(X, [k], f)
        K, F = _affine_exact(A, k, f)
        assert np.max(np.abs(Ke[0] - K)) <= 1e-12 * np.max(np.abs(K))
        assert np.max(np.abs(fe[0] - F)) <= 1e-12 * np.max(np.abs(F))


def _jittered(rng, n, amp=0.12):
    X = np.repeat(CORNERS[None], n, axis=0) + rng.uniform(-amp, amp, (n, 8, 3))
    return X * rng.uniform(0.2, 2.0, (n, 1, 1))


def test_trilinear_invariants():
    rng = np.random.default_rng(3)
    X = _jittered(rng, 6)
    k = rng.uniform(0.5, 3.0, 6)
    Ke, fe = O.q2_hex_elements(X, k, 1.0)
    for e in range(6):
        K = Ke[e]
        s = np.max(np.abs(K))
        assert np.max(np.abs(K - K.T)) <= 1e-14 * s
        assert np.max(np.abs(K.sum(axis=1))) <= 1e-12 * s
        ev = np.linalg.eigvalsh(K)
        assert abs(ev[0]) <= 1e-12 * s and ev[1] > 1e-6 * s      # only the constants are in the kernel
        # rigid motion: the same matrix
        Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        K2, _ = O.q2_hex_elements((X[e] @ Q.T + rng.normal(size=3))[None], [k[e]], 1.0)
        assert np.max(np.abs(K2[0] - K)) <= 1e-12 * s
        # scaling by t: the 3D stiffness scales with t
        K3, f3 = O.q2_hex_elements((2.5 * X[e])[None], [k[e]], 1.0)
        assert np.max(np.abs(K3[0] - 2.5 * K)) <= 1e-12 * 2.5 * s
        assert np.max(np.abs(f3[0] - 2.5 ** 3 * fe[e])) <= 1e-12 * 2.5 ** 3 * np.max(np.abs(fe[e]))


def test_planted_mistakes_are_caught():
    """a transposed inverse Jacobian or a swapped pair of local shapes would not
    pass the metric-formula pin"""
    A = np.array([[0.6, 0.2, 0.0], [0.0, 0.5, 0.1], [0.15, 0.0, 0.7]])
    X = (CORNERS @ A.T)[None]
    Ke, _ = O.q2_hex_elements(X, [1.0], 0.0)
    K, _ = _affine_exact(A, 1.0, 0.0)
    Kt, _ = _affine_exact(A.T, 1.0, 0.0)     # what an A^-T / A^-1 mix-up computes
    assert np.max(np.abs(Ke[0] - K)) <= 1e-12 * np.max(np.abs(K))
    assert np.max(np.abs(Kt - K)) > 1e-3 * np.max(np.abs(K))
    perm = np.arange(27)
    perm[[9, 10]] = perm[[10, 9]]
    assert np.max(np.abs(K[np.ix_(perm, perm)] - K)) > 1e-3 * np.max(np.abs(K))
