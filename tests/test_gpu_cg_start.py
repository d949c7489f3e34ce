"""How the lattice CG starts (cg.h:103-242 with the initial defect
r = b - A x): from b alone when x is zero, with the Dirichlet-eliminated
lattice operator when x and b vanish on the Dirichlet DoFs, and through the
surface operator otherwise. Each start is checked against the solver on the
surface path (CSR operator, the same GMG), and apply_return_defect returns
the final defect b - A x."""
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


def problem(refs, seed=3):
    u = U()
    from ugcore_amd.synthetic import make_theta
    d = u.DomainDiscretization(0, refs)
    d.set_theta(make_theta(2 ** refs, seed, 3))
    A = u.SparseMatrix()
    b = u.Vector(d.num_dofs())
    d.assemble_linear(A, b)
    return d, A, b


def solve(u, d, A, b, x0, precision, lattice, return_defect=False):
    g = u.GeometricMultiGrid(d, base_level=0, precision=precision)
    S = u.CG(g, max_steps=200, reduction=1e-10)
    S.set_lattice_path(lattice)
    S.init(A)
    x = u.Vector(values=x0)
    bb = u.Vector(values=b.numpy())
    ok = S.apply_return_defect(x, bb) if return_defect else S.apply(x, bb)
    assert ok
    return x.numpy(), S.steps(), bool(S.used_lattice_path()), bb.numpy()


@pytest.mark.parametrize("refs,precision", [(5, "f32"), (5, "f64"), (4, "f64")])
@pytest.mark.parametrize("start", ["zero", "inner", "dirichlet"])
def test_cg_starts_match_surface_path(refs, precision, start):
    say(f"case cg start refs={refs} {precision} start={start}")
    u = U()
    d, A, b = problem(refs)
    N = d.num_dofs()
    ni = d.num_inner(refs)
    x0 = np.zeros(N)
    if start != "zero":
        x0 = np.random.default_rng(4).normal(size=N) * 1e-3
        if start == "inner":
            x0[ni:] = 0.0          # Dirichlet DoFs come last
    xs, ss, used_s, _ = solve(u, d, A, b, x0, precision, lattice=False)
    assert not used_s
    for return_defect in (False, True):
        xl, sl, used_l, r = solve(u, d, A, b, x0, precision, lattice=True, return_defect=return_defect)
        # a Dirichlet value in x leaves a defect on the Dirichlet DoFs: surface path
        assert used_l == (start != "dirichlet")
        assert abs(sl - ss) <= 1
        assert np.linalg.norm(xl - xs) <= 1e-9 * np.linalg.norm(xs)
        if return_defect:
            y = u.Vector(N)
            A.apply(y, u.Vector(values=xl))
            r_true = b.numpy() - y.numpy()
            assert np.linalg.norm(r - r_true) <= 1e-12 * np.linalg.norm(b.numpy())
