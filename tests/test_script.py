"""The script-style front end (ugcore_amd.script): configurations are
validated into device plans on the host; unsupported combinations are
rejected with a message naming them."""
import math

import numpy as np
import pytest

import ugcore_amd as U
from ugcore_amd import script as S


def build(domain="unit_cube_hex", refs=2, disc_type=None, diffusion=1.0, velocity=None, upwind=None, source=1.0,
          dirichlet=((0.0, "c", "Boundary"),), subsets="Inner"):
    dom = S.Domain(domain, refs)
    approx = S.ApproximationSpace(dom)
    approx.add_fct("c", "Lagrange", 1)
    e = S.ConvectionDiffusion("c", subsets, disc_type or ("fe" if domain == "unit_cube_hex" else "fv1"))
    e.set_diffusion(diffusion)
    if velocity is not None:
        e.set_velocity(velocity)
    if upwind is not None:
        e.set_upwind(upwind)
    e.set_source(source)
    d = S.DomainDiscretization(approx)
    d.add(e)
    if dirichlet is not None:
        b = S.DirichletBoundary()
        for entry in dirichlet:
            b.add(*entry)
        d.add(b)
    return d


def test_valid_plans():
    p = build(diffusion=2.5).plan()
    assert p["kind"] == 0 and p["num_refs"] == 2 and p["velocity"] is None
    assert p["theta"].size == 64 and np.all(p["theta"] == math.log(2.5))
    theta = np.random.default_rng(0).normal(size=16 ** 2)
    p = build("unit_square_tri", 4, diffusion=S.ExpField(theta), source=np.ones(17 * 17)).plan()
    assert p["kind"] == 1 and np.array_equal(p["theta"], theta) and p["source"].size == 289
    p = build("unit_cube_tet", 1, velocity=(20, 10, 5), upwind=S.FullUpwind(), source=0.5).plan()
    assert p["kind"] == 2 and p["velocity"] == (20.0, 10.0, 5.0) and p["theta"].size == 6 ** 3
    p = build("unit_cube_tet", 1, velocity=(0, 0, 0)).plan()
    assert p["velocity"] is None


@pytest.mark.parametrize("kwargs,fragment", [
    (dict(disc_type="fv1"), "discretised with 'fe'"),
    (dict(velocity=(1, 0, 0), upwind=S.FullUpwind()), "pure diffusion"),
    (dict(domain="unit_cube_tet", refs=1, velocity=(1, 0, 0)), "FullUpwind"),
    (dict(dirichlet=((1.0, "c", "Boundary"),)), "only value 0"),
    (dict(dirichlet=((0.0, "c", "Top"),)), "only value 0"),
    (dict(dirichlet=None), "DirichletBoundary"),
    (dict(subsets="Inner,Boundary"), "subset 'Inner'"),
    (dict(diffusion=S.ExpField(np.zeros(7))), "64 values expected"),
    (dict(diffusion=-1.0), "positive"),
    (dict(source=np.ones(125)), "constant source"),
])
def test_rejected_configurations(kwargs, fragment):
    with pytest.raises(U.UGError, match=fragment):
        build(**kwargs).plan()


def test_rejected_spaces_and_domains():
    with pytest.raises(U.UGError, match="unknown domain"):
        S.Domain("unit_torus", 2)
    approx = S.ApproximationSpace(S.Domain("unit_cube_hex", 1))
    with pytest.raises(U.UGError, match="Lagrange order 1"):
        approx.add_fct("c", "Lagrange", 2)
    with pytest.raises(U.UGError, match="vertex jitter"):
        S.Domain("unit_cube_hex", 1).set_jitter(1, 0.1)
