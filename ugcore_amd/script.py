"""Script-style front end: the names and call sequence of a ugcore script built
on the ConvectionDiffusion element discretisation (external plugin; usage as
in lib_disc/spatial_disc/user_data/common_user_data/lognormal_random_field.h:
33-37), mapped onto the device discretisations of this package.

    dom = Domain("unit_cube_hex", num_refs=6)
    approx = ApproximationSpace(dom)
    approx.add_fct("c", "Lagrange", 1)
    elem = ConvectionDiffusion("c", "Inner", "fe")
    elem.set_diffusion(ExpField(theta))          # k = exp(theta) per element
    elem.set_source(1.0)
    bnd = DirichletBoundary()
    bnd.add(0.0, "c", "Boundary")
    disc = DomainDiscretization(approx)
    disc.add(elem)
    disc.add(bnd)
    A, b = disc.assemble_linear()

Supported (everything else raises UGError naming what is not supported):
  * unit_cube_hex: "fe" (Q1), diffusion constant or ExpField per fine
    element, constant source, no velocity;
  * unit_square_tri: "fv1", diffusion constant or ExpField per fine element,
    constant or per-DoF source, no velocity;
  * unit_cube_tet (3x3x3 hexahedra x 6 tetrahedra): "fv1", diffusion constant
    or ExpField per fine lattice cell, constant velocity with FullUpwind
    (required when the velocity is not zero), constant source; optional
    vertex jitter (Domain.set_jitter);
  * one function "c" of Lagrange order 1 on subset "Inner", homogeneous
    Dirichlet value 0 on subset "Boundary".
"""
import math

import numpy as np

_KINDS = {"unit_cube_hex": 0, "unit_square_tri": 1, "unit_cube_tet": 2}
_DISC_TYPES = {0: "fe", 1: "fv1", 2: "fv1"}


def _error(msg):
    import ugcore_amd as U
    return U.UGError(msg)


class Domain:
    """one of the lattice domains, refined num_refs times"""

    def __init__(self, name, num_refs):
        if name not in _KINDS:
            raise _error(f"Domain: unknown domain '{name}' (supported: {', '.join(_KINDS)})")
        if num_refs < 0:
            raise _error("Domain: num_refs must be >= 0")
        self.name = name
        self.kind = _KINDS[name]
        self.num_refs = int(num_refs)
        self.dim = 2 if self.kind == 1 else 3
        self.jitter = None

    def set_jitter(self, seed, amplitude):
        """displace the interior vertices of the top level (tetrahedra only)"""
        if self.kind != 2:
            raise _error("Domain.set_jitter: only the tetrahedral domain has vertex jitter")
        self.jitter = (int(seed), float(amplitude))

    def num_elements(self):
        n = (3 if self.kind == 2 else 1) << self.num_refs
        return n ** self.dim


class ApproximationSpace:
    def __init__(self, domain):
        self.domain = domain
        self.fcts = []

    def add_fct(self, name, family="Lagrange", order=1):
        if family != "Lagrange" or order != 1:
            raise _error("ApproximationSpace.add_fct: only Lagrange order 1 is supported")
        if self.fcts:
            raise _error("ApproximationSpace.add_fct: one function only")
        self.fcts.append(name)


class FullUpwind:
    """ConvectionShapesFullUpwind (disc_util/conv_shape.h:278-330)"""


class NoUpwind:
    """central convection shapes (not supported with a velocity here)"""


class ExpField:
    """k = exp(theta) with one theta per fine element (hexahedra, triangles)
    or fine lattice cell (tetrahedra), lattice order"""

    def __init__(self, theta):
        self.theta = np.ascontiguousarray(theta, dtype=np.float64)


class ConvectionDiffusion:
    def __init__(self, fct, subsets, disc_type):
        self.fct = fct
        self.subsets = subsets
        self.disc_type = disc_type
        self.diffusion = 1.0
        self.velocity = None
        self.source = 0.0
        self.upwind = None

    def set_diffusion(self, d):
        self.diffusion = d

    def set_velocity(self, v):
        self.velocity = tuple(float(x) for x in v)

    def set_source(self, f):
        self.source = f

    def set_upwind(self, u):
        self.upwind = u


class DirichletBoundary:
    def __init__(self):
        self.entries = []

    def add(self, value, fct, subsets):
        self.entries.append((value, fct, subsets))


class DomainDiscretization:
    """collects the element discretisation and the Dirichlet boundary;
    plan() validates them on the host, assemble_linear() builds the device
    discretisation and assembles A, b"""

    def __init__(self, approx):
        self.approx = approx
        self.elem = None
        self.dirichlet = None
        self.product = None

    def add(self, disc):
        if isinstance(disc, ConvectionDiffusion):
            if self.elem is not None:
                raise _error("DomainDiscretization.add: one ConvectionDiffusion only")
            self.elem = disc
        elif isinstance(disc, DirichletBoundary):
            if self.dirichlet is not None:
                raise _error("DomainDiscretization.add: one DirichletBoundary only")
            self.dirichlet = disc
        else:
            raise _error(f"DomainDiscretization.add: unsupported discretisation {type(disc).__name__}")

    def plan(self):
        """the device configuration: dict(kind, num_refs, theta, velocity,
        source, jitter); raises UGError for anything this path does not do"""
        dom = self.approx.domain
        kind = dom.kind
        if not self.approx.fcts:
            raise _error("DomainDiscretization: the approximation space has no function")
        fct = self.approx.fcts[0]
        e = self.elem
        if e is None:
            raise _error("DomainDiscretization: no ConvectionDiffusion added")
        if e.fct != fct:
            raise _error(f"ConvectionDiffusion: unknown function '{e.fct}'")
        if e.subsets != "Inner":
            raise _error("ConvectionDiffusion: the element discretisation must act on subset 'Inner'")
        if e.disc_type != _DISC_TYPES[kind]:
            raise _error(f"ConvectionDiffusion: '{dom.name}' is discretised with '{_DISC_TYPES[kind]}', "
                         f"not '{e.disc_type}'")
        ne = dom.num_elements()
        if isinstance(e.diffusion, ExpField):
            theta = e.diffusion.theta.ravel()
            if theta.size != ne:
                raise _error(f"ConvectionDiffusion.set_diffusion: {ne} values expected, got {theta.size}")
        elif isinstance(e.diffusion, (int, float)):
            if not e.diffusion > 0:
                raise _error("ConvectionDiffusion.set_diffusion: the diffusion must be positive")
            theta = np.full(ne, math.log(float(e.diffusion)))
        else:
            raise _error("ConvectionDiffusion.set_diffusion: a positive number or an ExpField")
        vel = e.velocity
        if vel is not None and any(v != 0.0 for v in vel):
            if kind != 2:
                raise _error(f"ConvectionDiffusion.set_velocity: '{dom.name}' is a pure diffusion path")
            if len(vel) != 3:
                raise _error("ConvectionDiffusion.set_velocity: three components expected")
            if not isinstance(e.upwind, FullUpwind):
                raise _error("ConvectionDiffusion: a velocity needs set_upwind(FullUpwind())")
        else:
            vel = None
        src = e.source
        if isinstance(src, (int, float)):
            src = float(src)
        else:
            src = np.ascontiguousarray(src, dtype=np.float64)
            if kind != 1:
                raise _error(f"ConvectionDiffusion.set_source: '{dom.name}' takes a constant source")
        d = self.dirichlet
        if d is None or not d.entries:
            raise _error("DomainDiscretization: a DirichletBoundary on 'Boundary' is required")
        for value, f, subsets in d.entries:
            if f != fct or subsets != "Boundary" or float(value) != 0.0:
                raise _error("DirichletBoundary.add: only value 0 for the function on subset 'Boundary'")
        return {"kind": kind, "num_refs": dom.num_refs, "theta": theta, "velocity": vel, "source": src,
                "jitter": dom.jitter}

    def build(self):
        """the device discretisation (ugcore_amd.DomainDiscretization)"""
        import ugcore_amd as U
        p = self.plan()
        d = U.DomainDiscretization(p["kind"], p["num_refs"])
        d.set_theta(p["theta"])
        if p["velocity"] is not None:
            d.set_velocity(p["velocity"])
        if p["jitter"] is not None:
            d.set_jitter(*p["jitter"])
        src = p["source"]
        if p["kind"] == 1 and not isinstance(src, np.ndarray):
            src = np.full(d.num_dofs(), src)
        d.set_source(src)
        self.product = d
        return d

    def assemble_linear(self, A=None, b=None):
        """AssembledLinearOperator / GridFunction pair: returns (A, b)"""
        import ugcore_amd as U
        d = self.product if self.product is not None else self.build()
        A = A if A is not None else U.SparseMatrix()
        b = b if b is not None else U.Vector(d.num_dofs())
        d.assemble_linear(A, b)
        return A, b
