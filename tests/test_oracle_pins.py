"""Pins of the oracle's multigrid cycle and Krylov iterate sequences by what
the mathematics fixes (CPU only; the oracle is the checker of the HIP path).

GMG (mg_solver_impl.hpp:175-275, 1686-2136):
* V-cycle: one application equals (I - E_l) A_l^{-1} d with the textbook
  error propagation E_l = (I - S_l A_l)^nu2 (I - P_l B_{l-1} R_l A_l)
  (I - S_l A_l)^nu1, S_l = omega diag(A_l)^{-1}, B_base = A_base^{-1},
  built from dense P, R, A (those are pinned by closed forms elsewhere), on
  hierarchies of 2 to 4 levels and for a base level above 0.
* W-, gamma- and F-cycles: the coarse level is visited more than once. The
  reference skips the defect update after the last post-smoothing step on
  every level below the finest full-refinement level (mg_solver_impl.hpp:
  1954), so a revisit smooths a defect that lacks that last correction. The
  dense composition below carries exactly that term; a test also checks that
  it differs from the textbook W-cycle (so a restatement that "fixed" it
  would fail). With two levels the base level is solved once per visit and
  every cycle type equals the V-cycle.
CG (cg.h:103-242): the k-th iterate minimises the A-norm of the error over
  x0 + K_k(BA, B r0) (B = identity or the dense multigrid operator).
BiCGStab (bicgstab.h:112-385): unpreconditioned residuals equal
  psi_k(A) phi_k(A) r0 with the BiCG polynomial phi_k fixed by
  (r0, A^i phi_k(A) r0) = 0 and psi_k = prod(1 - omega_j t) from the
  recorded steps; with the multigrid as right preconditioner the half-step
  residual is orthogonal to r0 and the full-step residual to t = A q; a
  restart equals a fresh solve from the iterate it restarts at.
"""
import numpy as np
import pytest

import oracle_lib as O

OMEGA = 0.66


def problem(kind, refs, seed=5):
    p = O.Problem(kind, refs)
    if kind == 0:
        p.set_theta(O.make_theta(p.n_fine, seed, 3))
    elif kind == 1:
        p.set_source_top(np.ones(p.num_dofs(refs)))
    else:
        p.set_theta(O.make_theta(p.n_fine, seed, 3))
        p.set_tet_params(vel=(20.0, 10.0, 5.0), jitter_seed=11, jitter_amp=0.2 / np.sqrt(3.0))
    p.assemble_all()
    return p


def dense(M):
    return M.to_scipy().toarray()


class DenseHierarchy:
    """Dense level operators of an oracle problem"""

    def __init__(self, p, base, nu1, nu2, omega=OMEGA):
        self.top = p.num_levels - 1
        self.base = base
        self.nu1, self.nu2 = nu1, nu2
        self.A = {l: dense(p.A(l)) for l in range(base, self.top + 1)}
        self.P = {l: dense(p.P(l)) for l in range(base + 1, self.top + 1)}
        self.R = {l: dense(p.R(l)) for l in range(base + 1, self.top + 1)}
        self.S = {l: omega * np.diag(1.0 / np.diag(self.A[l])) for l in range(base + 1, self.top + 1)}

    def textbook(self, l, gamma=1):
        """B_l of the textbook cycle with gamma coarse iterations"""
        A = self.A[l]
        if l == self.base:
            return np.linalg.inv(A)
        n = A.shape[0]
        I = np.eye(n)
        Bc = np.linalg.inv(self.A[l - 1]) if l - 1 == self.base else self.textbook(l - 1, gamma)
        if l - 1 != self.base and gamma > 1:
            Ac = self.A[l - 1]
            Ec = np.eye(Ac.shape[0]) - Bc @ Ac
            Bc = (np.eye(Ac.shape[0]) - np.linalg.matrix_power(Ec, gamma)) @ np.linalg.inv(Ac)
        Es = I - self.S[l] @ A
        E = np.linalg.matrix_power(Es, self.nu2) @ (I - self.P[l] @ Bc @ self.R[l] @ A) @ \
            np.linalg.matrix_power(Es, self.nu1)
        return (I - E) @ np.linalg.inv(A)

    def visit(self, l, typ):
        """one visit of level l as linear maps of its defect d: the correction
        c = C d and the defect left on the level, D d (the last correction is
        not subtracted below the top level)"""
        A = self.A[l]
        n = A.shape[0]
        S = self.S[l]
        C = np.zeros((n, n))
        D = np.eye(n)
        for _ in range(self.nu1):
            T = S @ D
            D = D - A @ T
            C = C + T
        Rd = self.R[l] @ D
        if l - 1 == self.base:
            Cc = np.linalg.inv(self.A[l - 1])
        elif typ == -1:
            C1, D1 = self.visit(l - 1, -1)
            C2, _ = self.visit(l - 1, 1)
            Cc = C1 + C2 @ D1
        else:
            nc = self.A[l - 1].shape[0]
            Cc = np.zeros((nc, nc))
            Dc = np.eye(nc)
            for _ in range(typ):
                Ci, Di = self.visit(l - 1, typ)
                Cc = Cc + Ci @ Dc
                Dc = Di @ Dc
        T = self.P[l] @ Cc @ Rd
        C = C + T
        for _ in range(self.nu2):
            D = D - A @ T
            T = S @ D
            C = C + T
        if l == self.top:
            D = D - A @ T
        return C, D


def oracle_B(p, cols, base, nu1, nu2, cycle):
    return np.stack([p.gmg_apply(c, base_lev=base, nu1=nu1, nu2=nu2, cycle=cycle, omega=OMEGA) for c in cols.T],
                    axis=1)


def probe_vectors(n, k=4, seed=3):
    rng = np.random.default_rng(seed)
    V = rng.normal(size=(n, k))
    return np.hstack([V, np.eye(n)[:, [0, n // 2, n - 1]]])


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("kind,refs,base,nu1,nu2", [
    (0, 2, 0, 2, 2), (0, 3, 0, 2, 2), (0, 3, 1, 2, 2), (0, 3, 0, 1, 0), (0, 3, 0, 0, 1), (0, 3, 0, 3, 1),
    (1, 3, 0, 2, 2), (1, 4, 0, 1, 2), (2, 1, 0, 2, 2)])
def test_v_cycle_is_the_textbook_cycle(kind, refs, base, nu1, nu2):
    p = problem(kind, refs)
    H = DenseHierarchy(p, base, nu1, nu2)
    V = probe_vectors(H.A[H.top].shape[0])
    B = oracle_B(p, V, base, nu1, nu2, 1)
    Bt = H.textbook(H.top) @ V
    assert rel(B, Bt) <= 1e-12
    # the dense composition agrees as well (no revisits in a V-cycle)
    C, _ = H.visit(H.top, 1)
    assert rel(C @ V, Bt) <= 1e-12


@pytest.mark.parametrize("kind,refs", [(0, 1), (1, 1), (2, 1)])
def test_two_levels_every_cycle_type_is_the_v_cycle(kind, refs):
    """with the base level directly below the top it is solved once
    (mg_solver_impl.hpp:2111-2113), whatever the cycle type"""
    p = problem(kind, refs)
    H = DenseHierarchy(p, 0, 2, 2)
    V = probe_vectors(H.A[H.top].shape[0])
    Bt = H.textbook(H.top) @ V
    for cycle in (1, 2, 3, -1):
        assert rel(oracle_B(p, V, 0, 2, 2, cycle), Bt) <= 1e-12, cycle


@pytest.mark.parametrize("kind,refs,base,cycle,nu1,nu2", [
    (0, 2, 0, 2, 2, 2), (0, 3, 0, 2, 2, 2), (0, 3, 0, -1, 2, 2), (0, 3, 0, 3, 2, 2), (0, 3, 0, -1, 1, 0),
    (0, 3, 0, 2, 0, 1), (1, 4, 0, -1, 2, 2), (1, 4, 0, 2, 1, 1), (1, 4, 1, -1, 2, 2)])
def test_w_gamma_f_cycles(kind, refs, base, cycle, nu1, nu2):
    p = problem(kind, refs)
    H = DenseHierarchy(p, base, nu1, nu2)
    V = probe_vectors(H.A[H.top].shape[0])
    C, _ = H.visit(H.top, cycle)
    assert rel(oracle_B(p, V, base, nu1, nu2, cycle), C @ V) <= 1e-12


def test_w_cycle_revisit_keeps_the_reference_defect():
    """the revisit of a level starts from the defect without its last
    correction (mg_solver_impl.hpp:1954); the textbook W-cycle differs"""
    p = problem(0, 3)
    H = DenseHierarchy(p, 0, 2, 2)
    V = probe_vectors(H.A[H.top].shape[0])
    B = oracle_B(p, V, 0, 2, 2, 2)
    assert rel(B, H.textbook(H.top, gamma=2) @ V) > 1e-6
    # the F-cycle is neither the V- nor the W-cycle
    F = oracle_B(p, V, 0, 2, 2, -1)
    assert rel(F, B) > 1e-6 and rel(F, oracle_B(p, V, 0, 2, 2, 1)) > 1e-6


@pytest.mark.parametrize("cycle", [0, -2, -5])
def test_invalid_cycle_refused(cycle):
    p = problem(0, 2)
    with pytest.raises(RuntimeError):
        p.gmg_apply(np.ones(p.num_dofs(2)), cycle=cycle)


def krylov_basis(M, v, k):
    """orthonormal basis of span{v, Mv, ..., M^(k-1) v} (Arnoldi, twice
    orthogonalised)"""
    Q = np.zeros((len(v), k))
    Q[:, 0] = v / np.linalg.norm(v)
    for j in range(1, k):
        w = M @ Q[:, j - 1]
        for _ in range(2):
            w = w - Q[:, :j] @ (Q[:, :j].T @ w)
        Q[:, j] = w / np.linalg.norm(w)
    return Q


@pytest.mark.parametrize("precond", ["none", "gmg"])
def test_cg_iterates_minimise_the_energy_error(precond):
    """x_k = argmin over x0 + K_k(BA, B r0) of ||x - x*||_A (interior block;
    x and b vanish on the Dirichlet DoFs)"""
    refs = 3
    p = problem(0, refs)
    top = refs
    A = dense(p.A(top))
    b = p.rhs(top)
    inner = p.dirichlet(top) == 0
    assert np.all(b[~inner] == 0.0)
    AII = A[np.ix_(inner, inner)]
    n = AII.shape[0]
    if precond == "gmg":
        H = DenseHierarchy(p, 0, 2, 2)
        Bfull = H.textbook(top)
        assert np.abs(Bfull[np.ix_(~inner, inner)]).max() <= 1e-14
        B = Bfull[np.ix_(inner, inner)]
        kmax = 5
    else:
        B = np.eye(n)
        kmax = 10
    xs, res = p.iterates(solver="cg", gmg=(precond == "gmg"), base_lev=0, rel_red=1e-10, max_steps=kmax)
    xstar = np.linalg.solve(AII, b[inner])
    r0 = b[inner]
    e0 = np.sqrt(xstar @ AII @ xstar)
    M = B @ AII
    for k in range(1, min(kmax, len(xs) - 1) + 1):
        Q = krylov_basis(M, B @ r0, k)
        y = np.linalg.solve(Q.T @ AII @ Q, Q.T @ r0)
        xk = Q @ y
        diff = xs[k][inner] - xk
        assert np.sqrt(diff @ AII @ diff) <= 1e-8 * e0, k
        assert np.all(xs[k][~inner] == 0.0)


def bicg_problem():
    p = problem(2, 1)
    top = 1
    return p, dense(p.A(top)), p.rhs(top), top


def test_bicgstab_residual_polynomials():
    """r_k = psi_k(A) phi_k(A) r0 and s_k = psi_k(A) phi_{k+1}(A) r0
    (unpreconditioned; bicgstab.h:231-330 as van der Vorst's method)"""
    p, A, b, top = bicg_problem()
    nrm = np.linalg.norm(A, 2)
    As = A / nrm                       # polynomials in the scaled matrix
    xs, res = p.iterates(solver="bicgstab", gmg=False, rel_red=1e-30, max_steps=8)
    r0 = b.copy()
    # half-step and full-step residuals from the recorded iterates
    s = [b - A @ xs[2 * j + 1] for j in range((len(xs) - 1) // 2)]
    r = [b - A @ xs[2 * j] for j in range((len(xs) + 1) // 2)]
    # omega_j from x_{j+1} - x_{j+1/2} = omega_j s_j
    om = []
    for j in range(len(s)):
        dx = xs[2 * j + 2] - xs[2 * j + 1]
        om.append((dx @ s[j]) / (s[j] @ s[j]) * nrm)   # omega for the scaled matrix
    K = np.column_stack([np.linalg.matrix_power(As, i) @ r0 for i in range(12)])

    def phi(k):
        if k == 0:
            return r0.copy()
        H = np.array([[r0 @ K[:, i + j] for j in range(1, k + 1)] for i in range(k)])
        c = np.linalg.solve(H, -np.array([r0 @ K[:, i] for i in range(k)]))
        return r0 + K[:, 1:k + 1] @ c

    def psi(k, v):
        for j in range(k):
            v = v - om[j] * (As @ v)
        return v

    for k in range(1, len(r)):
        assert rel(r[k], psi(k, phi(k))) <= 1e-7, k
    for k in range(len(s)):
        assert rel(s[k], psi(k, phi(k + 1))) <= 1e-7, k


def test_bicgstab_gmg_orthogonality():
    """right-preconditioned BiCGStab: (r0, s_k) = 0 by the choice of alpha and
    (r_{k+1}, A q) = 0 by the choice of omega, with q = x_{k+1} - x_{k+1/2}
    up to the factor omega"""
    p, A, b, top = bicg_problem()
    xs, res = p.iterates(solver="bicgstab", gmg=True, base_lev=0, omega=OMEGA, rel_red=1e-10, max_steps=20)
    assert res["ok"]
    r0 = b.copy()
    nb = np.linalg.norm(b)
    # residuals are recomputed as b - A x; near convergence that carries a
    # rounding floor of order 1e-16 ||b||, hence the second term
    for j in range((len(xs) - 1) // 2):
        s = b - A @ xs[2 * j + 1]
        assert abs(r0 @ s) <= (1e-10 * np.linalg.norm(s) + 1e-14 * nb) * np.linalg.norm(r0), j
        if 2 * j + 2 < len(xs):
            r = b - A @ xs[2 * j + 2]
            t = A @ (xs[2 * j + 2] - xs[2 * j + 1])
            if np.linalg.norm(t) > 0:
                assert abs(r @ t) <= (1e-10 * np.linalg.norm(r) + 1e-14 * nb) * np.linalg.norm(t), j


def test_bicgstab_restart_equals_fresh_solve():
    """a restart (set_restart, bicgstab.h:157-170) continues as a fresh solve
    from the iterate it restarts at. The step counter counts half steps, so
    num_restarts = 6 restarts after 3 full steps."""
    import scipy.sparse as sp
    p, A, b, top = bicg_problem()
    kr = 6
    p.set_bicgstab_options(num_restarts=kr)
    try:
        xs, res = p.iterates(solver="bicgstab", gmg=False, rel_red=1e-30, max_steps=kr + 4)
    finally:
        p.set_bicgstab_options(0, 0.0)
    xs0, _ = p.iterates(solver="bicgstab", gmg=False, rel_red=1e-30, max_steps=kr + 4)
    assert np.array_equal(xs0[:kr + 1], xs[:kr + 1])
    for m in range(1, 5):
        _, _, xm = O.csr_solve(sp.csr_matrix(A), b, x0=xs[kr], solver="bicgstab", rel_red=1e-30, max_steps=m)
        assert rel(xm, xs[kr + m]) <= 1e-9, m
    assert rel(xs0[kr + 2], xs[kr + 2]) > 1e-8
