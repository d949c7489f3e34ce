"""Device refinement of the lattice hierarchies (mesh_kernels.hip) against the
oracle's hierarchy and the host hierarchy: DoF numbering per level (the
reference's LEVEL numbering, GlobalMultiGridRefiner creation order and
DoFDistribution subset order), the Inner counts, the tetrahedra per lattice
cell and the parent-edge codes -- all bit-exact -- and assembled operators
that are identical whichever set-up built the hierarchy."""
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


def lattice_index(lat, W, dim):
    if dim == 3:
        return (lat[:, 2] * W + lat[:, 1]) * W + lat[:, 0]
    return lat[:, 1] * W + lat[:, 0]


@pytest.mark.parametrize("kind,refs", [(0, 0), (0, 1), (0, 3), (0, 4), (1, 0), (1, 3), (1, 6),
                                       (2, 0), (2, 1), (2, 2), (2, 3)])
def test_device_numbering_matches_oracle(kind, refs):
    say(f"case device hierarchy kind={kind} refs={refs}")
    u = U()
    p = O.Problem(kind, refs)
    if kind == 2:
        p.assemble_all(1.0)
    d = u.DomainDiscretization(kind, refs, setup="device")
    dim = 2 if kind == 1 else 3
    base = 3 if kind == 2 else 1
    for lev in range(refs + 1):
        W = base * 2 ** lev + 1
        lat = p.dof_lattice(lev) // (2 ** (refs - lev))
        expect = np.full(W ** dim, -1, np.int64)
        expect[lattice_index(lat, W, dim)] = np.arange(len(lat))
        assert np.array_equal(d.dof_of_lattice(lev), expect), lev
        assert d.num_inner(lev) == int((p.dirichlet(lev) == 0).sum())
        assert d.num_dofs(lev) == W ** dim


def _tet_cells_from_oracle(p, lev):
    refs = p.num_refs
    n = 3 * 2 ** lev
    lat = p.dof_lattice(lev) // (2 ** (refs - lev))
    cells = [[] for _ in range(n ** 3)]
    for t in p.elements(lev):
        q = lat[t]
        mn = q.min(axis=0)
        c = q - mn
        code = 0
        for j in range(4):
            code |= int(c[j, 0] | (c[j, 1] << 1) | (c[j, 2] << 2)) << (3 * j)
        cells[(mn[2] * n + mn[1]) * n + mn[0]].append(code)
    return np.array(cells, np.uint16)


@pytest.mark.parametrize("refs", [1, 2, 3])
def test_device_tetrahedra_match_oracle_and_host(refs):
    say(f"case device tetrahedra refs={refs}")
    u = U()
    p = O.Problem(2, refs)
    p.assemble_all(1.0)
    d = u.DomainDiscretization(2, refs, setup="device")
    for lev in range(refs + 1):
        cells, par = d.level_tets(lev)
        assert np.array_equal(cells, _tet_cells_from_oracle(p, lev)), lev
        hcells, hpar = u.hierarchy_tets(refs, lev)
        assert np.array_equal(cells, hcells.reshape(cells.shape))
        if lev > 0:
            assert np.array_equal(par, hpar)


@pytest.mark.parametrize("kind,refs", [(0, 6), (1, 9), (2, 4)])
@pytest.mark.parametrize("ordering", ["reference", "lex"])
def test_device_setup_matches_host_setup(kind, refs, ordering):
    """larger hierarchies: the device refinement reproduces the host one on
    every level, for both DoF orderings"""
    say(f"case device vs host set-up kind={kind} refs={refs} ordering={ordering}")
    u = U()
    dd = u.DomainDiscretization(kind, refs, ordering=ordering, setup="device")
    dh = u.DomainDiscretization(kind, refs, ordering=ordering, setup="host")
    assert dd.num_levels() == dh.num_levels() == refs + 1
    for lev in range(refs + 1):
        assert np.array_equal(dd.dof_of_lattice(lev), dh.dof_of_lattice(lev)), lev
        assert dd.num_inner(lev) == dh.num_inner(lev)
        if kind == 2:
            cd, pd = dd.level_tets(lev)
            ch, ph = dh.level_tets(lev)
            assert np.array_equal(cd, ch)
            if lev > 0:
                assert np.array_equal(pd, ph)


@pytest.mark.parametrize("kind,refs", [(0, 4), (1, 6), (2, 2)])
def test_assembly_independent_of_setup(kind, refs):
    say(f"case assembly after device set-up kind={kind} refs={refs}")
    u = U()
    out = []
    for setup in ("device", "host"):
        d = u.DomainDiscretization(kind, refs, setup=setup)
        d.set_theta(np.random.default_rng(5).normal(scale=0.5, size=d.n_fine ** d.dim))
        if kind == 2:
            d.set_velocity((3.0, 2.0, 1.0))
        A = u.SparseMatrix()
        b = u.Vector(d.num_dofs())
        d.assemble_linear(A, b)
        out.append((A.download_csr(), b.numpy()))
    (rp0, c0, v0), b0 = out[0]
    (rp1, c1, v1), b1 = out[1]
    assert np.array_equal(rp0, rp1) and np.array_equal(c0, c1)
    assert np.array_equal(v0, v1) and np.array_equal(b0, b1)


def test_unknown_setup_is_rejected():
    u = U()
    with pytest.raises(KeyError):
        u.DomainDiscretization(0, 2, setup="other")
    h = u.lib().ugx_disc_create_setup(0, 2, 0, 7)
    assert not h
    assert "set-up" in u.lib().ugx_last_error().decode()
