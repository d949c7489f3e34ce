"""The lattice-order GMG cycle replayed from captured HIP graphs against the
same cycle launched kernel by kernel: identical kernels in identical order, so
solutions, defect histories and step counts must agree bit for bit. Each solve
replays the captured cycle once per CG step with the eager CG kernels and a
host synchronisation in between; the graph and eager solves alternate."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def U():
    import ugcore_amd
    return ugcore_amd


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def problem(kind, refs, seed=3):
    u = U()
    d = u.DomainDiscretization(kind, refs)
    n = 2 ** refs
    if kind == 0:
        from ugcore_amd.synthetic import make_theta
        d.set_theta(make_theta(n, seed, 3))
    else:
        from ugcore_amd.synthetic import poisson_source_lattice
        dol = d.dof_of_lattice(refs)
        f = np.empty(d.num_dofs())
        f[dol] = poisson_source_lattice(n, seed)
        d.set_source(f)
    A = u.SparseMatrix()
    b = u.Vector(d.num_dofs())
    d.assemble_linear(A, b)
    return d, A, b


@pytest.mark.parametrize("kind,refs,precision", [(0, 5, "f32"), (0, 5, "f64"), (1, 7, "f32"), (0, 4, "f64")])
def test_graph_cycle_matches_eager(kind, refs, precision):
    say(f"case graphs kind={kind} refs={refs} {precision}")
    u = U()
    d, A, b = problem(kind, refs)
    solvers = {}
    for graphs in (True, False):
        g = u.GeometricMultiGrid(d, base_level=0, precision=precision, graphs=graphs)
        S = u.CG(g, max_steps=200, reduction=1e-10)
        S.init(A)
        solvers[graphs] = (g, S)
    results = {True: [], False: []}
    for rep in range(3):
        for graphs in (True, False):
            g, S = solvers[graphs]
            x = u.Vector(d.num_dofs())
            assert S.apply(x, b)
            assert S.used_lattice_path()
            u.synchronize()
            results[graphs].append((x.numpy(), S.steps(), S.defects()))
    ref_x, ref_steps, ref_def = results[False][0]
    for graphs in (True, False):
        for x, steps, defs in results[graphs]:
            assert steps == ref_steps
            assert np.array_equal(x, ref_x)
            assert np.array_equal(defs, ref_def)


@pytest.mark.parametrize("refs,precision,graphs", [(5, "f32", True), (5, "f64", True), (5, "f32", False),
                                                   (4, "f64", False)])
def test_lattice_cycle_matches_surface_cycle(refs, precision, graphs):
    """the cycle's lattice entry (its first step reads the FP64 defect
    directly, or an FP32 copy of it, and its last update is fused with the
    conversion) against the surface entry (separate conversion, zeroing and
    final update) on the same defect: identical corrections"""
    say(f"case lattice vs surface cycle refs={refs} {precision} graphs={graphs}")
    u = U()
    L = u.lib()
    d, A, b = problem(0, refs)
    g = u.GeometricMultiGrid(d, base_level=0, precision=precision, graphs=graphs)
    g.init(A)
    N = d.num_dofs()
    dol = d.dof_of_lattice(refs)
    rng = np.random.default_rng(11)
    r = rng.normal(size=N)
    r[d.num_inner(refs):] = 0.0          # Dirichlet DoFs come last
    c_surf = u.Vector(N)
    g.apply(c_surf, u.Vector(values=r))
    H = L.ugx_gmg_lattice_halo(g.h)
    d_lat = u.DeviceArray(np.concatenate([np.zeros(H), r[dol], np.zeros(H)]))
    # the FP32 copy of the defect an FP32 cycle may start from
    d_f32 = u.DeviceArray(np.concatenate([np.zeros(H), r[dol], np.zeros(H)]).astype(np.float32))
    assert L.ugx_gmg_lattice_uses_f32_defect(g.h) == (1 if precision == "f32" else 0)
    c_lat = u.DeviceArray(np.zeros(N + 2 * H))
    for rep in range(2):
        for f32 in (None, d_f32.ptr + 4 * H):
            assert L.ugx_gmg_apply_lattice(g.h, c_lat.ptr + 8 * H, d_lat.ptr + 8 * H, f32) == 0
            u.synchronize()
            c = c_lat.numpy()[H:H + N]
            assert np.array_equal(c, c_surf.numpy()[dol])


@pytest.mark.parametrize("refs", [4, 6])
def test_surface_direction_lattice_equals_its_parts(refs):
    """ugx_gmg_surface_direction_lattice (p_new = z + beta p_old, q = A p_new,
    (q, p_new) in one pass) against the direction update formed on the host,
    the lattice surface operator applied to the kernel's p_new (the same
    operator: bit for bit) and the product formed on the host"""
    say(f"case surface direction lattice refs={refs}")
    u = U()
    L = u.lib()
    d, A, b = problem(0, refs)
    g = u.GeometricMultiGrid(d, base_level=0, precision="f32")
    g.init(A)
    N = d.num_dofs()
    H = L.ugx_gmg_lattice_halo(g.h)
    rng = np.random.default_rng(5)
    n = 2 ** refs
    W = n + 1
    lat = np.zeros((W, W, W), bool)
    lat[1:-1, 1:-1, 1:-1] = True          # the vectors vanish on the Dirichlet boundary
    inner = lat.ravel()
    p_old = np.where(inner, rng.normal(size=N), 0.0)
    z = np.where(inner, rng.normal(size=N), 0.0)
    beta = 0.37
    pad = lambda v: np.concatenate([np.zeros(H), v, np.zeros(H)])
    d_pold, d_z = u.DeviceArray(pad(p_old)), u.DeviceArray(pad(z))
    d_pnew, d_q = u.DeviceArray(np.zeros(N + 2 * H)), u.DeviceArray(np.zeros(N))
    d_beta, d_qp, d_part = u.DeviceArray(np.array([beta])), u.DeviceArray(np.zeros(1)), u.DeviceArray(np.zeros(65536))
    assert L.ugx_gmg_surface_direction_lattice(g.h, d_q.ptr, d_pold.ptr + 8 * H, d_z.ptr + 8 * H, d_beta.ptr,
                                               d_pnew.ptr + 8 * H, d_qp.ptr, d_part.ptr) == 0
    u.synchronize()
    p_new = d_pnew.numpy()[H:H + N]
    # the update (the kernel fuses the multiply-add; equal to rounding)
    expect_p = beta * p_old + z
    assert np.allclose(p_new, expect_p, rtol=0, atol=1e-15 * np.abs(expect_p).max())
    d_y = u.DeviceArray(np.zeros(N))
    d_pn = u.DeviceArray(pad(p_new))
    assert L.ugx_gmg_surface_apply_lattice(g.h, d_y.ptr, d_pn.ptr + 8 * H) == 0
    u.synchronize()
    q = d_q.numpy()
    assert np.array_equal(q, d_y.numpy())
    qp = d_qp.numpy()[0]
    assert abs(qp - float(np.dot(q, p_new))) <= 1e-12 * abs(float(np.dot(np.abs(q), np.abs(p_new))))
