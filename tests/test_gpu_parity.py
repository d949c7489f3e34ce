"""GPU parity tests: the HIP path (through the C ABI) against the CPU oracle on
the same seeded inputs.

Tolerances (from the path's contract):
  * CSR pattern and DoF indexing: bit-exact
  * matrix entries: |a_gpu - a_orc| <= 1e-12 * max|a| of the row (couplings
    that are zero in exact arithmetic, e.g. Q1 face neighbours, carry round-off
    of order 1e-17 and are compared against that row-scaled floor)
  * solution: relative L2 error <= 1e-9
  * Krylov iteration counts: within +-1
"""
import ctypes as C
import sys

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


def U():
    import ugcore_amd
    return ugcore_amd


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


_cache = {}


def oracle_problem(kind, refs, seed=7):
    key = (kind, refs, seed)
    if key not in _cache:
        p = O.Problem(kind, refs)
        n = 2 ** refs
        if kind == 0:
            theta = O.make_theta(n, seed, 3)
            p.set_theta(theta)
            f = None
        else:
            theta = None
            from ugcore_amd.synthetic import poisson_source_lattice
            fl = poisson_source_lattice(n, seed)
            lat = p.dof_lattice(refs)
            f = fl[lat[:, 1] * (n + 1) + lat[:, 0]]
            p.set_source_top(f)
        p.assemble_all()
        _cache[key] = (p, theta, f)
    return _cache[key]


def gpu_problem(kind, refs, theta, f):
    u = U()
    d = u.DomainDiscretization(kind, refs)
    if theta is not None:
        d.set_theta(theta)
    if f is not None:
        d.set_source(f)
    A = u.SparseMatrix()
    b = u.Vector(d.num_dofs())
    d.assemble_linear(A, b)
    return d, A, b


def rows_close(rp, val_a, val_b, tol=1e-12):
    ok = True
    worst = 0.0
    for r in range(len(rp) - 1):
        a = val_a[rp[r]:rp[r + 1]]
        b = val_b[rp[r]:rp[r + 1]]
        scale = max(np.abs(b).max(), 1e-300) if b.size else 1.0
        err = np.abs(a - b).max() / scale if b.size else 0.0
        worst = max(worst, err)
        ok = ok and err <= tol
    return ok, worst


CASES = [(0, 3), (0, 4), (1, 5), (1, 7)]


@pytest.mark.parametrize("kind,refs", CASES)
def test_surface_matrix_and_rhs(kind, refs):
    say(f"case surface kind={kind} refs={refs}")
    p, theta, f = oracle_problem(kind, refs)
    d, A, b = gpu_problem(kind, refs, theta, f)
    rp, col, val = A.download_csr()
    ref = p.A(refs)
    assert np.array_equal(rp.astype(np.int64), ref.rowptr)
    assert np.array_equal(col, ref.col)
    ok, worst = rows_close(ref.rowptr, val, ref.val)
    assert ok, worst
    bo = p.rhs(refs)
    bg = b.numpy()
    assert np.abs(bg - bo).max() <= 1e-12 * max(np.abs(bo).max(), 1e-300)


@pytest.mark.parametrize("kind,refs", [(0, 3), (1, 5)])
def test_level_operators(kind, refs):
    """Every GMG level operator (stored internally in lattice order as the
    symmetric, Dirichlet-eliminated diagonals) equals the oracle's level matrix
    (reference DoF order) entry by entry, with the couplings of interior rows
    to Dirichlet columns eliminated (level vectors vanish on Dirichlet DoFs,
    so the products are the same); couplings outside the pattern are 0."""
    say(f"case levels kind={kind} refs={refs}")
    u = U()
    p, theta, f = oracle_problem(kind, refs)
    d, A, b = gpu_problem(kind, refs, theta, f)
    g = u.GeometricMultiGrid(d, base_level=0, precision="f64")
    g.init(A)
    dim = 3 if kind == 0 else 2
    for lev in range(refs + 1):
        W = 2 ** lev + 1
        dia = g.level_dia(lev)
        nd = dia.shape[0]
        lat = p.dof_lattice(lev) // (2 ** (refs - lev))
        li = (lat[:, 2] * W + lat[:, 1]) * W + lat[:, 0] if dim == 3 else lat[:, 1] * W + lat[:, 0]
        M = p.A(lev)
        dm = p.dirichlet(lev)
        seen = np.zeros_like(dia, dtype=bool)
        for r in range(M.n_rows):
            cols = M.col[M.rowptr[r]:M.rowptr[r + 1]]
            vals = M.val[M.rowptr[r]:M.rowptr[r + 1]]
            scale = np.abs(vals).max()
            for c, v in zip(cols, vals):
                if not dm[r] and dm[c]:
                    v = 0.0
                dl = lat[c] - lat[r]
                if dim == 3:
                    k = (dl[2] + 1) * 9 + (dl[1] + 1) * 3 + (dl[0] + 1)
                else:
                    k = (dl[1] + 1) * 3 + (dl[0] + 1)
                assert abs(dia[k, li[r]] - v) <= 1e-12 * scale
                seen[k, li[r]] = True
        assert np.all(dia[~seen] == 0.0)


@pytest.mark.parametrize("kind,refs", [(0, 4), (1, 6)])
def test_spmv_kernel_entry_points(kind, refs):
    """ugx_spmv_csr_f64 and the vector entry points on raw device pointers vs
    the oracle's SparseMatrix::axpy / Vector restatement."""
    say(f"case spmv kind={kind} refs={refs}")
    u = U()
    p, theta, f = oracle_problem(kind, refs)
    d, A, b = gpu_problem(kind, refs, theta, f)
    ref = p.A(refs)
    n = ref.n_rows
    rng = np.random.default_rng(1)
    rp, col, val = A.device_csr()
    L = u.lib()
    for alpha, beta in ((0.0, 1.0), (1.0, -1.0), (0.3, 2.5)):
        w = rng.normal(size=n)
        v1 = rng.normal(size=n)
        dw, dv, dd = u.DeviceArray(w), u.DeviceArray(v1), u.DeviceArray(np.zeros(n))
        assert L.ugx_spmv_csr_f64(n, rp, col, val, alpha, dv.ptr, beta, dw.ptr, dd.ptr, None) == 0
        expect = rng.normal(size=n)
        O.lib().orc_csr_op(0, n, n, ref.rowptr, ref.col, ref.val, alpha, v1.ctypes.data, beta, w, expect)
        got = dd.numpy()
        assert np.abs(got - expect).max() <= 1e-13 * np.abs(expect).max()
    x = rng.normal(size=n)
    y = rng.normal(size=n)
    dx, dy = u.DeviceArray(x), u.DeviceArray(y)
    out = C.c_double()
    assert L.ugx_dot_f64(n, dx.ptr, dy.ptr, C.byref(out), None) == 0
    assert abs(out.value - float(np.dot(x, y))) <= 1e-12 * np.abs(x).sum() * np.abs(y).max()
    assert L.ugx_norm_f64(n, dx.ptr, C.byref(out), None) == 0
    assert abs(out.value - float(np.linalg.norm(x))) <= 1e-13 * np.linalg.norm(x)
    dz = u.DeviceArray(np.zeros(n))
    assert L.ugx_vec_scale_add3_f64(n, dz.ptr, 2.0, dx.ptr, -0.5, dy.ptr, 0.25, dx.ptr, None) == 0
    assert np.allclose(dz.numpy(), 2.25 * x - 0.5 * y, rtol=1e-15, atol=1e-15)
    assert L.ugx_vec_scale_add2_f64(n, dz.ptr, 1.0, dz.ptr, -2.25, dx.ptr, None) == 0
    assert np.allclose(dz.numpy(), -0.5 * y, rtol=1e-14, atol=1e-14)
    # Jacobi entry points: dinv = 1/(a_ii/omega); c = dinv .* d
    ddinv = u.DeviceArray(np.zeros(n))
    assert L.ugx_jacobi_diag_inv_f64(n, rp, col, val, 0.66, ddinv.ptr, None) == 0
    diag = p.A(refs).to_scipy().diagonal()
    assert np.allclose(ddinv.numpy(), 1.0 / (diag * (1.0 / 0.66)), rtol=1e-15)
    assert L.ugx_jacobi_apply_f64(n, dz.ptr, ddinv.ptr, dx.ptr, None) == 0
    assert np.allclose(dz.numpy(), ddinv.numpy() * x, rtol=1e-15)


def test_spmv_ignore_zero_rows_and_empty_rows():
    say("case spmv empty rows")
    u = U()
    L = u.lib()
    rng = np.random.default_rng(3)
    rows, cols = 300, 200
    ent = [(int(rng.integers(rows)), int(rng.integers(cols)), float(rng.normal())) for _ in range(900)]
    ent = [e for e in ent if e[0] % 7 != 0]          # rows 0, 7, 14, ... stay empty
    M = O.builder_csr(ent, rows, cols, mode=0, add=True)
    rp = u.DeviceArray(M.rowptr.astype(np.int32))
    cl = u.DeviceArray(M.col)
    vl = u.DeviceArray(M.val)
    w = rng.normal(size=cols)
    dw = u.DeviceArray(w)
    init = rng.normal(size=rows)
    dd = u.DeviceArray(init)
    assert L.ugx_spmv_csr_ignore_zero_rows_f64(rows, rp.ptr, cl.ptr, vl.ptr, 1.5, dw.ptr, dd.ptr, None) == 0
    expect = init.copy()
    O.lib().orc_csr_op(1, rows, cols, M.rowptr, M.col, M.val, 0.0, None, 1.5, w, expect)
    assert np.allclose(dd.numpy(), expect, rtol=1e-14, atol=1e-14)
    # axpy with alpha = 0 writes 0 on empty rows
    dd2 = u.DeviceArray(init)
    assert L.ugx_spmv_csr_f64(rows, rp.ptr, cl.ptr, vl.ptr, 0.0, dd2.ptr, 1.0, dw.ptr, dd2.ptr, None) == 0
    e2 = init.copy()
    O.lib().orc_csr_op(0, rows, cols, M.rowptr, M.col, M.val, 0.0, None, 1.0, w, e2)
    assert np.allclose(dd2.numpy(), e2, rtol=1e-14, atol=1e-14)
    # FP32 variant
    vl32 = u.DeviceArray(M.val.astype(np.float32))
    dw32 = u.DeviceArray(w.astype(np.float32))
    dd32 = u.DeviceArray(np.zeros(rows, np.float32))
    assert L.ugx_spmv_csr_f32(rows, rp.ptr, cl.ptr, vl32.ptr, 0.0, dd32.ptr, 1.0, dw32.ptr, dd32.ptr, None) == 0
    assert np.allclose(dd32.numpy(), e2, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("kind,refs", [(0, 3), (0, 4), (1, 6)])
def test_gmg_application(kind, refs):
    """One V(2,2) application c = B d equals the oracle's
    AssembledMultiGridCycle::apply for a random defect (zero on Dirichlet DoFs)."""
    say(f"case gmg apply kind={kind} refs={refs}")
    u = U()
    p, theta, f = oracle_problem(kind, refs)
    d, A, b = gpu_problem(kind, refs, theta, f)
    g = u.GeometricMultiGrid(d, base_level=0, nu1=2, nu2=2, omega=0.66, precision="f64")
    g.init(A)
    rng = np.random.default_rng(5)
    dv = rng.normal(size=p.num_dofs(refs))
    dv[p.dirichlet(refs) == 1] = 0.0
    c_orc = p.gmg_apply(dv, base_lev=0)
    c = u.Vector(len(dv))
    g.apply(c, u.Vector(values=dv))
    cg = c.numpy()
    assert np.linalg.norm(cg - c_orc) <= 1e-12 * np.linalg.norm(c_orc)


@pytest.mark.parametrize("kind,refs,nu1,nu2", [(0, 4, 2, 2), (1, 6, 2, 2), (0, 4, 1, 0), (0, 3, 0, 1)])
def test_gmg_w_cycle_and_smoothing_counts(kind, refs, nu1, nu2):
    """W-cycle (cycle 2) application and V-cycles with other smoothing counts
    against the oracle's AssembledMultiGridCycle::apply (mg_solver_impl.hpp:
    2090-2136: the coarse level is visited twice, so its correction must be
    complete after the first visit)"""
    say(f"case gmg W-cycle kind={kind} refs={refs} nu1={nu1} nu2={nu2}")
    u = U()
    p, theta, f = oracle_problem(kind, refs)
    d, A, b = gpu_problem(kind, refs, theta, f)
    rng = np.random.default_rng(9)
    dv = rng.normal(size=p.num_dofs(refs))
    dv[p.dirichlet(refs) == 1] = 0.0
    for cycle in (2, 1):
        for precision, tol in (("f64", 1e-12), ("f32", 1e-5)):
            g = u.GeometricMultiGrid(d, base_level=0, nu1=nu1, nu2=nu2, cycle=cycle, omega=0.66,
                                     precision=precision)
            g.init(A)
            c_orc = p.gmg_apply(dv, base_lev=0, nu1=nu1, nu2=nu2, cycle=cycle)
            c = u.Vector(len(dv))
            g.apply(c, u.Vector(values=dv))
            err = np.linalg.norm(c.numpy() - c_orc) / np.linalg.norm(c_orc)
            assert err <= tol, (cycle, precision, err)


@pytest.mark.parametrize("kind,refs,solver", [(0, 3, "cg"), (0, 4, "cg"), (0, 5, "cg"), (1, 7, "cg"),
                                              (0, 4, "bicgstab"), (1, 6, "bicgstab")])
def test_gmg_krylov_solve(kind, refs, solver):
    say(f"case solve kind={kind} refs={refs} solver={solver}")
    u = U()
    p, theta, f = oracle_problem(kind, refs)
    r_orc = p.solve(solver=solver, base_lev=0, rel_red=1e-10, max_steps=200)
    assert r_orc["ok"]
    d, A, b = gpu_problem(kind, refs, theta, f)
    g = u.GeometricMultiGrid(d, base_level=0, precision="f64")
    S = (u.CG if solver == "cg" else u.BiCGStab)(g, max_steps=200, reduction=1e-10)
    S.init(A)
    x = u.Vector(d.num_dofs())
    assert S.apply(x, b)
    if solver == "cg":
        assert S.used_lattice_path()
    assert abs(S.steps() - r_orc["steps"]) <= 1
    xo = r_orc["x"]
    assert np.linalg.norm(x.numpy() - xo) <= 1e-9 * np.linalg.norm(xo)


@pytest.mark.parametrize("kind,refs", [(0, 4), (0, 5), (1, 7)])
def test_fp32_preconditioner(kind, refs):
    """FP32 level operators inside the FP64 outer CG: converges to 1e-10 and
    agrees with the FP64-preconditioned oracle solve to 1e-9 (relative L2)."""
    say(f"case fp32 kind={kind} refs={refs}")
    u = U()
    p, theta, f = oracle_problem(kind, refs)
    r_orc = p.solve(base_lev=0, rel_red=1e-10)
    d, A, b = gpu_problem(kind, refs, theta, f)
    g = u.GeometricMultiGrid(d, base_level=0, precision="f32")
    S = u.CG(g, max_steps=200, reduction=1e-10)
    S.init(A)
    x = u.Vector(d.num_dofs())
    assert S.apply(x, b)
    assert S.steps() <= r_orc["steps"] + 2
    xo = r_orc["x"]
    assert np.linalg.norm(x.numpy() - xo) <= 1e-9 * np.linalg.norm(xo)


def test_jacobi_preconditioned_cg():
    say("case jacobi cg")
    u = U()
    p, theta, f = oracle_problem(0, 3)
    d, A, b = gpu_problem(0, 3, theta, f)
    S = u.CG(u.Jacobi(0.66), max_steps=500, reduction=1e-10)
    S.init(A)
    x = u.Vector(d.num_dofs())
    assert S.apply(x, b)
    import scipy.sparse.linalg as spla
    xs = spla.spsolve(p.A(3).to_scipy().tocsc(), p.rhs(3))
    assert np.linalg.norm(x.numpy() - xs) <= 1e-8 * np.linalg.norm(xs)


@pytest.mark.parametrize("refs", [4, 7])
@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_fused_smoothing_equals_separate_passes(precision, refs):
    """the lattice CG's cycle with its last post-smoothing step fused with the
    final correction update, conversion and (z, r) against the separate
    passes, for V-, W- and F-cycles: the same iteration (the product is only
    summed in another order), the same solution to rounding"""
    say(f"case fused final step precision={precision} refs={refs}")
    u = U()
    from ugcore_amd.synthetic import make_theta
    n = 2 ** refs
    theta = make_theta(n, 5, 3)
    d, A, b = gpu_problem(0, refs, theta, None)
    for cycle in ("V", "W", "F"):
        out = {}
        for fused in (True, False):
            g = u.GeometricMultiGrid(d, base_level=1, cycle=cycle, precision=precision, fused_smoothing=fused)
            S = u.CG(g, max_steps=100, reduction=1e-10)
            S.init(A)
            x = u.Vector(d.num_dofs())
            assert S.apply(x, b)
            assert S.used_lattice_path()
            out[fused] = (S.steps(), x.numpy(), np.asarray(S.defects()))
        assert out[True][0] == out[False][0], cycle
        assert np.linalg.norm(out[True][1] - out[False][1]) <= 1e-12 * np.linalg.norm(out[False][1]), cycle
        # defect norms: the same to rounding, measured against the first one
        # (late norms are small differences of large sums)
        d1, d0 = out[True][2], out[False][2]
        assert np.max(np.abs(d1 - d0)) <= 1e-12 * d0[0], cycle


def test_full_size_128_properties():
    """128^3 Q1 (2,146,689 DoFs): pattern size, Dirichlet identity rows,
    symmetric interior couplings, and a solve whose recomputed residual meets
    the 1e-10 reduction for both preconditioner precisions."""
    say("case 128^3 properties")
    u = U()
    from ugcore_amd.synthetic import make_theta
    refs = 7
    theta = make_theta(128, 7, 3)
    d, A, b = gpu_problem(0, refs, theta, None)
    n = d.num_dofs()
    assert n == 129 ** 3
    assert A.nnz() == 385 ** 3
    steps = {}
    for prec in ("f64", "f32"):
        g = u.GeometricMultiGrid(d, base_level=0, precision=prec)
        S = u.CG(g, max_steps=100, reduction=1e-10)
        S.init(A)
        x = u.Vector(n)
        assert S.apply(x, b)
        r = u.Vector(values=b.numpy())
        A.apply_sub(r, x)
        rn = np.linalg.norm(r.numpy())
        assert rn / np.linalg.norm(b.numpy()) < 1e-10
        steps[prec] = S.steps()
    assert abs(steps["f64"] - steps["f32"]) <= 2


def test_lexicographic_ordering_is_a_renumbering():
    """The OrderLex-style numbering yields the same operator up to the DoF
    permutation: identical solutions mapped through the lattice."""
    say("case lex ordering")
    u = U()
    refs = 4
    p, theta, f = oracle_problem(0, refs)
    sols = {}
    for ordering in ("reference", "lex"):
        d = u.DomainDiscretization(0, refs, ordering=ordering)
        d.set_theta(theta)
        A = u.SparseMatrix()
        b = u.Vector(d.num_dofs())
        d.assemble_linear(A, b)
        g = u.GeometricMultiGrid(d, base_level=0)
        S = u.CG(g, reduction=1e-12)
        S.init(A)
        x = u.Vector(d.num_dofs())
        assert S.apply(x, b)
        dol = d.dof_of_lattice(refs)
        sols[ordering] = x.numpy()[dol]        # lattice order
    assert np.linalg.norm(sols["lex"] - sols["reference"]) <= 1e-10 * np.linalg.norm(sols["reference"])


def test_sliced_ell_and_csr_layouts_agree():
    """GPUSparseMatrix SpMV through the sliced-ELL layout and through the CSR
    kernel against the oracle's SparseMatrix::axpy, on a matrix with empty and
    ragged rows (rows and columns not multiples of 64)."""
    say("case sliced ell")
    u = U()
    rng = np.random.default_rng(11)
    rows, cols = 1000, 777
    ent = []
    for r in range(rows):
        if r % 9 == 0:
            continue
        for _ in range(int(rng.integers(1, 40))):
            ent.append((r, int(rng.integers(cols)), float(rng.normal())))
    M = O.builder_csr(ent, rows, cols, mode=0, add=True)
    A = u.SparseMatrix()
    A.upload_csr(M.rowptr.astype(np.int32), M.col, M.val, n_cols=cols)
    w = rng.normal(size=cols)
    v1 = rng.normal(size=rows)
    order = rng.permutation(rows).astype(np.int32)
    got = {}
    for layout in ("sell", "sell-ordered", "csr"):
        A.set_spmv_layout("csr" if layout == "csr" else "sell")
        A.set_row_order(order if layout == "sell-ordered" else None)
        assert A.has_row_order() == (layout == "sell-ordered")
        for alpha, beta in ((0.0, 1.0), (1.0, -1.0), (0.7, 1.3)):
            d0 = np.random.default_rng(int(alpha * 10)).normal(size=rows)
            dest = u.Vector(values=d0)
            vv = u.Vector(values=v1)
            ww = u.Vector(values=w)
            A.axpy(dest, alpha, vv, beta, ww)
            expect = np.zeros(rows)
            O.lib().orc_csr_op(0, rows, cols, M.rowptr, M.col, M.val, alpha, v1.ctypes.data, beta, w, expect)
            assert np.allclose(dest.numpy(), expect, rtol=1e-13, atol=1e-13), layout
            got[(layout, alpha)] = dest.numpy()
    # the slice order changes which lanes hold a row, not its sum
    for alpha in (0.0, 1.0, 0.7):
        assert np.array_equal(got[("sell", alpha)], got[("sell-ordered", alpha)])
    bad = order.copy()
    bad[3] = bad[4]
    with pytest.raises(u.UGError):
        A.set_row_order(bad)
    with pytest.raises(u.UGError):
        A.set_row_order(order[:-1])


@pytest.mark.parametrize("kind,refs", [(0, 4), (1, 6)])
def test_cg_generic_and_lattice_paths_agree(kind, refs):
    """CG on the reference-ordered vectors (generic path) and on the
    lattice-ordered copies (fast path) give the same iterates."""
    say(f"case cg paths kind={kind} refs={refs}")
    u = U()
    p, theta, f = oracle_problem(kind, refs)
    d, A, b = gpu_problem(kind, refs, theta, f)
    res = {}
    for fast in (True, False):
        g = u.GeometricMultiGrid(d, base_level=0, precision="f64")
        S = u.CG(g, max_steps=200, reduction=1e-10)
        S.set_lattice_path(fast)
        S.init(A)
        x = u.Vector(d.num_dofs())
        assert S.apply(x, b)
        assert S.used_lattice_path() == fast
        res[fast] = (S.steps(), x.numpy(), S.defects())
    assert res[True][0] == res[False][0]
    assert np.linalg.norm(res[True][1] - res[False][1]) <= 1e-12 * np.linalg.norm(res[False][1])
    assert np.allclose(res[True][2], res[False][2], rtol=1e-10)


@pytest.mark.parametrize("kind,refs", [(0, 4), (1, 6)])
def test_defect_with_dirichlet_entries(kind, refs):
    """A right-hand side with non-zero Dirichlet entries (solution not
    adjusted): the GMG switches to the full level storage and CG to the
    generic path; results still equal the oracle."""
    say(f"case dirichlet defect kind={kind} refs={refs}")
    u = U()
    p, theta, f = oracle_problem(kind, refs)
    d, A, b = gpu_problem(kind, refs, theta, f)
    rng = np.random.default_rng(21)
    rhs = rng.normal(size=p.num_dofs(refs))
    # one application of the preconditioner
    g = u.GeometricMultiGrid(d, base_level=0, precision="f64")
    g.init(A)
    c = u.Vector(len(rhs))
    g.apply(c, u.Vector(values=rhs))
    c_orc = p.gmg_apply(rhs, base_lev=0)
    assert np.linalg.norm(c.numpy() - c_orc) <= 1e-12 * np.linalg.norm(c_orc)
    # a full solve
    xo, steps = p.solve_rhs(rhs, base_lev=0, rel_red=1e-10, zero_dirichlet=False, return_steps=True)
    S = u.CG(g, max_steps=200, reduction=1e-10)
    S.init(A)
    x = u.Vector(len(rhs))
    assert S.apply(x, u.Vector(values=rhs))
    assert not S.used_lattice_path()
    assert abs(S.steps() - steps) <= 1
    assert np.linalg.norm(x.numpy() - xo) <= 1e-9 * np.linalg.norm(xo)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_matrix_free_top_level_matches_stored(precision):
    """the matrix-free top-level operator (per-element coefficients) gives the
    same cycle and the same CG solve as the stored level matrix"""
    refs = 4
    say(f"case matrix-free vs stored precision={precision}")
    u = U()
    p, theta, f = oracle_problem(0, refs)
    d, A, b = gpu_problem(0, refs, theta, f)
    rng = np.random.default_rng(11)
    dv = rng.normal(size=d.num_dofs())
    dv[p.dirichlet(refs) == 1] = 0.0
    out = {}
    for mf in (True, False):
        g = u.GeometricMultiGrid(d, base_level=0, precision=precision, matrix_free=mf)
        g.init(A)
        c = u.Vector(len(dv))
        g.apply(c, u.Vector(values=dv))
        S = u.CG(g, max_steps=200, reduction=1e-10)
        S.init(A)
        x = u.Vector(d.num_dofs())
        assert S.apply(x, b)
        assert S.used_lattice_path()
        out[mf] = (c.numpy(), S.steps(), x.numpy())
    tol = 1e-13 if precision == "f64" else 1e-5
    assert np.linalg.norm(out[True][0] - out[False][0]) <= tol * np.linalg.norm(out[False][0])
    assert out[True][1] == out[False][1]
    assert np.linalg.norm(out[True][2] - out[False][2]) <= 1e-11 * np.linalg.norm(out[False][2])
