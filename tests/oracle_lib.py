"""ctypes access to the CPU oracle (oracle/_build/liboracle.so) and to the
reference-built SparseMatrix wrapper (oracle/_ref/libref_sm.so).

Test infrastructure only: the oracle is the checker, never the product path.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref_sm.so")

_lib = None
_ref = None

P = C.c_void_p
I64 = C.c_int64
F64P = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
I64P = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
I32P = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
U8P = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(ORACLE_SO)
        sig = {
            "orc_last_error": (C.c_char_p, []),
            "orc_problem_create": (P, [C.c_int, C.c_int]),
            "orc_problem_destroy": (None, [P]),
            "orc_num_levels": (C.c_int, [P]),
            "orc_num_dofs": (I64, [P, C.c_int]),
            "orc_num_elements": (I64, [P, C.c_int]),
            "orc_dof_lattice": (None, [P, C.c_int, I64P]),
            "orc_dirichlet_mask": (None, [P, C.c_int, U8P]),
            "orc_elements": (None, [P, C.c_int, I32P]),
            "orc_make_theta": (None, [C.c_int, C.c_uint64, C.c_int, F64P]),
            "orc_gauss": (C.c_double, [C.c_uint64, C.c_uint64]),
            "orc_set_theta": (None, [P, F64P]),
            "orc_set_source_top": (None, [P, F64P]),
            "orc_assemble": (C.c_int, [P, C.c_int, C.c_double]),
            "orc_matrix_nnz": (I64, [P, C.c_int, C.c_int]),
            "orc_matrix_rows": (I64, [P, C.c_int, C.c_int]),
            "orc_matrix_get": (None, [P, C.c_int, C.c_int, I64P, I32P, F64P]),
            "orc_rhs_get": (None, [P, C.c_int, F64P]),
            "orc_solve": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.c_double, C.c_double, C.c_double, C.c_int]),
            "orc_solution_get": (None, [P, F64P]),
            "orc_gmg_apply": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, F64P, F64P]),
            "orc_defects_get": (C.c_int, [P, C.c_void_p, C.c_int]),
            "orc_solve_rhs": (C.c_int, [P, F64P, F64P, C.c_int, C.c_double, C.c_int]),
            "orc_adjoint_gradient": (None, [P, F64P, F64P, F64P]),
            "orc_set_tet_params": (None, [P, C.c_double, C.c_double, C.c_double, C.c_uint64, C.c_double]),
            "orc_dof_positions": (None, [P, C.c_int, F64P]),
            "orc_solve_adjoint": (C.c_int, [P, C.c_int, F64P, F64P, C.c_int, C.c_int, C.c_int, C.c_double,
                                            C.c_double, C.c_int]),
            "orc_timings": (None, [P, F64P]),
            "orc_set_rap": (C.c_int, [P]),
            "orc_set_bicgstab_options": (None, [P, C.c_int, C.c_double]),
            "orc_builder_create": (P, [I64, I64]),
            "orc_builder_set": (None, [P, I64, I64, C.c_double]),
            "orc_builder_add": (None, [P, I64, I64, C.c_double]),
            "orc_builder_destroy": (None, [P]),
            "orc_builder_csr": (I64, [P, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
            "orc_csr_op": (None, [C.c_int, I64, I64, I64P, I32P, F64P, C.c_double, C.c_void_p,
                                  C.c_double, F64P, F64P]),
            "orc_q2_hex_elements": (None, [I64, P, P, C.c_double, P, P, P]),
            "orc_set_record_x": (None, [P, C.c_int]),
            "orc_iterates_count": (C.c_int, [P]),
            "orc_iterate_get": (C.c_int, [P, C.c_int, F64P]),
            "orc_csr_solve": (C.c_int, [C.c_int, I64, I64P, I32P, F64P, F64P, F64P, C.c_double, C.c_double,
                                        C.c_int]),
            "orc_q2_enable": (C.c_int, [P]),
            "orc_q2_num_dofs": (I64, [P, C.c_int]),
            "orc_q2_dof_positions": (None, [P, C.c_int, I64P, U8P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ref_lib():
    """The reference's own SparseMatrix compiled into oracle/_ref (None if absent)."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_SO):
            return None
        L = C.CDLL(REF_SO)
        sig = {
            "ref_sm_create": (P, [I64, I64]),
            "ref_sm_destroy": (None, [P]),
            "ref_sm_set": (None, [P, I64, I64, C.c_double]),
            "ref_sm_add": (None, [P, I64, I64, C.c_double]),
            "ref_sm_derive": (P, [P, C.c_int]),
            "ref_sm_rows": (I64, [P]),
            "ref_sm_nnz": (I64, [P]),
            "ref_sm_csr": (None, [P, I64P, I32P, F64P]),
            "ref_sm_op": (None, [P, C.c_int, C.c_double, C.c_void_p, C.c_double, F64P, F64P]),
            "ref_sm_multiply3": (P, [P, P, P]),
            "ref_vec_serialize": (I64, [I64, F64P, C.c_char_p, I64]),
            "ref_vec_dot": (C.c_double, [I64, F64P, F64P]),
            "ref_vec_norm": (C.c_double, [I64, F64P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _ref = L
    return _ref


class CSR:
    def __init__(self, rowptr, col, val, n_cols):
        self.rowptr = rowptr
        self.col = col
        self.val = val
        self.n_rows = len(rowptr) - 1
        self.n_cols = n_cols

    @property
    def nnz(self):
        return len(self.col)

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.val, self.col, self.rowptr), shape=(self.n_rows, self.n_cols))


class Problem:
    """Oracle problem: mesh hierarchy + level operators + solver (kind 0: unit
    cube from one hexahedron, Q1 FE; kind 1: unit square from two triangles, FV1;
    kind 2: unit cube from 3x3x3 hexahedra x 6 tetrahedra, FV1 convection-diffusion)."""

    def __init__(self, kind, num_refs):
        L = lib()
        self.h = L.orc_problem_create(kind, num_refs)
        if not self.h:
            raise RuntimeError(L.orc_last_error().decode())
        self.kind = kind
        self.num_refs = num_refs
        self.dim = 2 if kind == 1 else 3
        # fine cells (elements for kind 0/1) per axis
        self.n_fine = (3 if kind == 2 else 1) << num_refs

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_problem_destroy(self.h)
            self.h = None

    @property
    def num_levels(self):
        return lib().orc_num_levels(self.h)

    def num_dofs(self, lev):
        if getattr(self, "q2", False):
            return lib().orc_q2_num_dofs(self.h, lev)
        return lib().orc_num_dofs(self.h, lev)

    def dof_lattice(self, lev):
        n = self.num_dofs(lev)
        out = np.zeros(3 * n, np.int64)
        lib().orc_dof_lattice(self.h, lev, out)
        return out.reshape(n, 3)

    def dirichlet(self, lev):
        out = np.zeros(self.num_dofs(lev), np.uint8)
        lib().orc_dirichlet_mask(self.h, lev, out)
        return out

    def elements(self, lev):
        ne = lib().orc_num_elements(self.h, lev)
        nv = {0: 8, 1: 3, 2: 4}[self.kind]
        out = np.zeros(ne * nv, np.int32)
        lib().orc_elements(self.h, lev, out)
        return out.reshape(ne, nv)

    def set_theta(self, theta):
        lib().orc_set_theta(self.h, np.ascontiguousarray(theta, np.float64))

    def set_source_top(self, f):
        lib().orc_set_source_top(self.h, np.ascontiguousarray(f, np.float64))

    def assemble(self, lev, f_const=1.0):
        if lib().orc_assemble(self.h, lev, f_const) != 0:
            raise RuntimeError(lib().orc_last_error().decode())

    def assemble_all(self, f_const=1.0):
        for lev in range(self.num_levels):
            self.assemble(lev, f_const)

    def matrix(self, which, lev):
        L = lib()
        nnz = L.orc_matrix_nnz(self.h, which, lev)
        rows = L.orc_matrix_rows(self.h, which, lev)
        rp = np.zeros(rows + 1, np.int64)
        col = np.zeros(nnz, np.int32)
        val = np.zeros(nnz, np.float64)
        L.orc_matrix_get(self.h, which, lev, rp, col, val)
        ncols = self.num_dofs(lev - 1) if which == 1 else (self.num_dofs(lev) if which == 2 else rows)
        if which == 2:
            ncols = self.num_dofs(lev)
        return CSR(rp, col, val, ncols)

    def A(self, lev):
        return self.matrix(0, lev)

    def P(self, lev):
        return self.matrix(1, lev)

    def R(self, lev):
        return self.matrix(2, lev)

    def rhs(self, lev):
        out = np.zeros(self.num_dofs(lev), np.float64)
        lib().orc_rhs_get(self.h, lev, out)
        return out

    def solve(self, solver="cg", gmg=True, base_lev=0, nu1=2, nu2=2, cycle=1, omega=0.66,
              rel_red=1e-10, min_def=1e-50, max_steps=200):
        s = 0 if solver == "cg" else 1
        r = lib().orc_solve(self.h, s, 1 if gmg else 0, base_lev, nu1, nu2, cycle, omega,
                            rel_red, min_def, max_steps)
        if r == -100000:
            raise RuntimeError(lib().orc_last_error().decode())
        top = self.num_levels - 1
        x = np.zeros(self.num_dofs(top), np.float64)
        lib().orc_solution_get(self.h, x)
        n = lib().orc_defects_get(self.h, None, 0)
        d = np.zeros(n, np.float64)
        lib().orc_defects_get(self.h, d.ctypes.data, n)
        return {"ok": r >= 0, "steps": r if r >= 0 else -1 - r, "x": x, "defects": d}

    def iterates(self, solver="cg", gmg=True, base_lev=0, nu1=2, nu2=2, cycle=1, omega=0.66, rel_red=1e-10,
                 max_steps=200):
        """solve as solve() and return every recorded iterate (CG: x_0, x_1,
        ...; BiCGStab: x_0, then x after each half and each full step) with
        the solve's result"""
        lib().orc_set_record_x(self.h, 1)
        try:
            res = self.solve(solver, gmg, base_lev, nu1, nu2, cycle, omega, rel_red, 1e-50, max_steps)
            n = lib().orc_iterates_count(self.h)
            top = self.num_levels - 1
            xs = np.zeros((n, self.num_dofs(top)), np.float64)
            for k in range(n):
                lib().orc_iterate_get(self.h, k, xs[k])
        finally:
            lib().orc_set_record_x(self.h, 0)
        return xs, res

    def gmg_apply(self, d, base_lev=0, nu1=2, nu2=2, cycle=1, omega=0.66):
        d = np.ascontiguousarray(d, np.float64)
        c = np.zeros_like(d)
        if lib().orc_gmg_apply(self.h, base_lev, nu1, nu2, cycle, omega, d, c) != 0:
            raise RuntimeError(lib().orc_last_error().decode())
        return c

    def solve_rhs(self, rhs, base_lev=0, rel_red=1e-12, zero_dirichlet=True, return_steps=False):
        top = self.num_levels - 1
        x = np.zeros(self.num_dofs(top), np.float64)
        r = lib().orc_solve_rhs(self.h, np.ascontiguousarray(rhs, np.float64), x, base_lev, rel_red,
                                1 if zero_dirichlet else 0)
        if r < 0:
            raise RuntimeError("oracle solve failed")
        return (x, r) if return_steps else x

    def adjoint_gradient(self, u, lam):
        g = np.zeros(self.n_fine ** 3, np.float64)
        lib().orc_adjoint_gradient(self.h, np.ascontiguousarray(u, np.float64),
                                   np.ascontiguousarray(lam, np.float64), g)
        return g

    def set_bicgstab_options(self, num_restarts=0, min_ortho=0.0):
        lib().orc_set_bicgstab_options(self.h, int(num_restarts), float(min_ortho))

    def set_rap(self):
        """replace the assembled coarse matrices by Galerkin products R A P
        (after assemble_all)"""
        if lib().orc_set_rap(self.h) != 0:
            raise RuntimeError(lib().orc_last_error().decode())

    def set_tet_params(self, vel=(0.0, 0.0, 0.0), jitter_seed=0, jitter_amp=0.0):
        lib().orc_set_tet_params(self.h, float(vel[0]), float(vel[1]), float(vel[2]), int(jitter_seed),
                                 float(jitter_amp))

    def enable_q2(self):
        """switch to Q2 Lagrange (kind 0): later assemble calls form the Q2
        level operators, prolongations and restrictions in the Q2 numbering"""
        if lib().orc_q2_enable(self.h) != 0:
            raise RuntimeError(lib().orc_last_error().decode())
        self.q2 = True

    def q2_num_dofs(self, lev):
        return lib().orc_q2_num_dofs(self.h, lev)

    def q2_dofs(self, lev):
        """(positions in doubled finest-lattice units [N, 3], Dirichlet flags [N])"""
        n = self.q2_num_dofs(lev)
        pos = np.zeros(3 * n, np.int64)
        dr = np.zeros(n, np.uint8)
        lib().orc_q2_dof_positions(self.h, lev, pos, dr)
        return pos.reshape(n, 3), dr

    def dof_positions(self, lev):
        n = self.num_dofs(lev)
        out = np.zeros(3 * n, np.float64)
        lib().orc_dof_positions(self.h, lev, out)
        return out.reshape(n, 3)

    def solve_adjoint(self, rhs, solver="bicgstab", base_lev=0, nu1=2, nu2=2, omega=0.6, rel_red=1e-12,
                      max_steps=200):
        top = self.num_levels - 1
        lam = np.zeros(self.num_dofs(top), np.float64)
        r = lib().orc_solve_adjoint(self.h, 0 if solver == "cg" else 1, np.ascontiguousarray(rhs, np.float64),
                                    lam, base_lev, nu1, nu2, omega, rel_red, max_steps)
        if r < 0:
            raise RuntimeError("oracle adjoint solve failed")
        return lam, r

    def timings(self):
        out = np.zeros(3, np.float64)
        lib().orc_timings(self.h, out)
        return {"assemble": out[0], "setup": out[1], "solve": out[2]}


def csr_solve(A, b, x0=None, solver="cg", rel_red=1e-10, min_def=1e-50, max_steps=200):
    """CG or BiCGStab without a preconditioner on a scipy CSR matrix (oracle
    restatement); returns (converged, steps, x)"""
    n = A.shape[0]
    rp = np.ascontiguousarray(A.indptr, np.int64)
    col = np.ascontiguousarray(A.indices, np.int32)
    val = np.ascontiguousarray(A.data, np.float64)
    x = np.zeros(n) if x0 is None else np.array(x0, np.float64)
    r = lib().orc_csr_solve(0 if solver == "cg" else 1, n, rp, col, val, np.ascontiguousarray(b, np.float64), x,
                            rel_red, min_def, max_steps)
    if r == -100000:
        raise RuntimeError(lib().orc_last_error().decode())
    return (r >= 0, r if r >= 0 else -1 - r, x)


def make_theta(n, seed, passes=3):
    out = np.zeros(n ** 3, np.float64)
    lib().orc_make_theta(n, seed, passes, out)
    return out


def builder_csr(entries, rows, cols, mode=0, add=True):
    """Insert (r, c, v) in order with SparseMatrix operator() semantics; return CSR."""
    L = lib()
    h = L.orc_builder_create(rows, cols)
    f = L.orc_builder_add if add else L.orc_builder_set
    for r, c, v in entries:
        f(h, r, c, v)
    nnz = L.orc_builder_csr(h, mode, None, None, None)
    nr = cols if mode else rows
    rp = np.zeros(nr + 1, np.int64)
    col = np.zeros(nnz, np.int32)
    val = np.zeros(nnz, np.float64)
    L.orc_builder_csr(h, mode, rp.ctypes.data, col.ctypes.data, val.ctypes.data)
    L.orc_builder_destroy(h)
    return CSR(rp, col, val, rows if mode else cols)


def q2_hex_elements(X, k=None, f_const=0.0):
    """Oracle Q2 hexahedron element matrices: X [n, 8, 3] corners in the
    reference order, k [n] (None: 1); returns (Ke [n, 27, 27], fe [n, 27])"""
    X = np.ascontiguousarray(X, dtype=np.float64)
    n = X.shape[0]
    kk = None if k is None else np.ascontiguousarray(k, dtype=np.float64)
    Ke = np.zeros((n, 27, 27))
    fe = np.zeros((n, 27))
    lib().orc_q2_hex_elements(n, X.ctypes.data, None if kk is None else kk.ctypes.data, f_const,
                              Ke.ctypes.data, fe.ctypes.data, None)
    return Ke, fe


def q2_hex_multi_index():
    """the 27 local multi-indices of the Q2 hexahedron, in the reference's shape order"""
    mi = np.zeros((27, 3), np.int32)
    lib().orc_q2_hex_elements(0, None, None, 0.0, None, None, mi.ctypes.data)
    return mi
