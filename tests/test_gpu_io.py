"""DoF positions and the GPU-object file conveniences (SaveToFile /
ReadFromFile, SaveMatrixForConnectionViewer) against the oracle."""
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


JAMP = 0.2 / np.sqrt(3.0)


@pytest.mark.parametrize("kind,refs", [(0, 3), (1, 4), (2, 2)])
def test_dof_positions_match_oracle(kind, refs):
    say(f"case dof positions kind={kind} refs={refs}")
    u = U()
    p = O.Problem(kind, refs)
    d = u.DomainDiscretization(kind, refs)
    if kind == 2:
        p.set_tet_params(vel=(20.0, 10.0, 5.0), jitter_seed=11, jitter_amp=JAMP)
        d.set_jitter(11, JAMP)
    dim = 2 if kind == 1 else 3
    for lev in range(refs + 1):
        got = d.dof_positions(lev)
        want = p.dof_positions(lev)[:, :dim]
        assert got.shape == want.shape
        assert np.array_equal(got, want), (lev, np.abs(got - want).max())


def test_save_and_read_gpu_vector(tmp_path):
    say("case SaveToFile/ReadFromFile")
    u = U()
    v = np.random.default_rng(2).normal(size=4097)
    x = u.Vector(values=v)
    path = tmp_path / "x.ug4vector"
    u.SaveToFile(x, path)
    assert np.array_equal(u.read_vector_file(path), v)
    y = u.Vector(3)
    u.ReadFromFile(y, path)
    assert y.size() == v.size
    assert np.array_equal(y.numpy(), v)


def test_save_matrix_for_connection_viewer(tmp_path):
    say("case SaveMatrixForConnectionViewer")
    u = U()
    refs = 2
    p = O.Problem(0, refs)
    theta = O.make_theta(2 ** refs, 7, 3)
    p.set_theta(theta)
    p.assemble_all()
    d = u.DomainDiscretization(0, refs)
    d.set_theta(theta)
    A = u.SparseMatrix()
    b = u.Vector(d.num_dofs())
    d.assemble_linear(A, b)
    path = tmp_path / "A.mat"
    u.SaveMatrixForConnectionViewer(d, A, path)
    lines = path.read_text().splitlines()
    n = p.num_dofs(refs)
    assert lines[:3] == ["1", "3", str(n)]
    pos = p.dof_positions(refs)
    for i in range(n):
        assert lines[3 + i] == " ".join(format(float(c), ".6g") for c in pos[i])
    M = p.A(refs)
    body = lines[4 + n:]
    assert len(body) == len(M.col)
    k = 0
    for r in range(n):
        for q in range(M.rowptr[r], M.rowptr[r + 1]):
            i, j, v = body[k].split()
            assert int(i) == r and int(j) == M.col[q]
            # 6 significant digits of the GPU value; the oracle's entry agrees to 1e-12
            ref_v = M.val[q]
            scale = np.abs(M.val[M.rowptr[r]:M.rowptr[r + 1]]).max()
            assert abs(float(v) - ref_v) <= 5e-6 * abs(ref_v) + 1e-12 * scale
            k += 1
