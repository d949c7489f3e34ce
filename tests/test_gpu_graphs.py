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
