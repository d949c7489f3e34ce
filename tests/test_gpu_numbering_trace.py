"""Device-built LEVEL numbering against the hand-traced fixtures of
tests/golden/numbering_trace.json (the reference rules it follows are cited
there and in test_numbering_trace.py)."""
import numpy as np
import pytest

from test_numbering_trace import CASES, TRACE, traced_dof_of_lattice

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,kind,lev", CASES)
@pytest.mark.parametrize("extra", [0, 3])
def test_device_numbering_is_the_traced_one(name, kind, lev, extra):
    import ugcore_amd as U
    d = U.DomainDiscretization(kind, lev + extra)
    assert np.array_equal(d.dof_of_lattice(lev), traced_dof_of_lattice(name))
    W = TRACE[name]["W"]
    pos = d.dof_positions(lev)
    assert np.abs(pos * (W - 1) - np.array(TRACE[name]["dof"], np.float64)).max() < 1e-12
