"""Adjoint gradients of the Q1 diffusion solve behind torch.autograd.

Problem: -div(exp(theta) grad u) = f on the unit cube (one hexahedron refined
`num_refs` times), homogeneous Dirichlet boundary, theta piecewise constant per
fine element. For a scalar loss J(u), autograd supplies dJ/du; the backward
pass solves the adjoint A^T lambda = dJ/du (Dirichlet entries zeroed) with the
same GMG-preconditioned CG, then the element kernel forms
dJ/dtheta_e = -lambda_e^T (dA_e/dtheta_e) u_e with A_e recomputed on the fly.
For pure diffusion A is symmetric on the interior DoFs, so the adjoint system
is solved with A itself. Only u (and theta) are kept for the backward pass;
with keep_operator=False the operator is dropped after the forward solve and
re-assembled in backward.

Tensors are host (CPU) tensors; the library moves them to the device. torch
ships its own HIP runtime while this library links the system one, and in one
process only the runtime that initialises first sees the GPU, so device
tensors are rejected with that explanation.

The reference has no adjoint; this module has no counterpart there.
"""
import numpy as np

import ugcore_amd as U

U.lib()            # load the HIP library (and its runtime) before torch
import torch       # noqa: E402


_DEVICE_TENSORS = ("device tensors are not supported: torch ships its own HIP runtime and this library links "
                   "the system one; in one process only the runtime that initialises first sees the GPU. "
                   "Pass host (CPU) tensors: the library moves them to the device itself.")


def _host_tensor(t):
    if t.is_cuda:
        raise U.UGError(_DEVICE_TENSORS)
    return t


def _to_vector(t, vec):
    """copy a float64 host tensor into a GPUVector"""
    t = _host_tensor(t.detach()).contiguous()
    if t.dtype != torch.float64:
        raise TypeError("float64 tensors expected")
    vec.upload(t.numpy())


def _from_vector(vec, like):
    _host_tensor(like)
    return torch.from_numpy(vec.numpy())


class DiffusionProblem:
    """GPU state for one mesh: discretisation, operator, GMG-CG solver."""

    def __init__(self, num_refs, f=1.0, base_level=0, precision="f64", reduction=1e-12, max_steps=200,
                 keep_operator=True):
        self.num_refs = num_refs
        self.n = 2 ** num_refs
        self.disc = U.DomainDiscretization(0, num_refs)
        self.disc.set_source(f)
        self.N = self.disc.num_dofs()
        self.base_level = base_level
        self.precision = precision
        self.reduction = reduction
        self.max_steps = max_steps
        self.keep_operator = keep_operator
        self._A = self._b = self._solver = None
        self._theta_key = None
        self.last_steps = {}

    def _set_theta(self, theta):
        theta = _host_tensor(theta.detach()).contiguous()
        if theta.dtype != torch.float64 or theta.numel() != self.n ** 3:
            raise ValueError("theta must be float64 with n_fine^3 entries")
        self.disc.set_theta(theta.numpy())

    def _assemble(self, theta):
        self._set_theta(theta)
        self._A = U.SparseMatrix()
        self._b = U.Vector(self.N)
        self.disc.assemble_linear(self._A, self._b)
        g = U.GeometricMultiGrid(self.disc, base_level=self.base_level, precision=self.precision)
        self._solver = U.CG(g, max_steps=self.max_steps, reduction=self.reduction)
        self._solver.init(self._A)

    def _release(self):
        self._A = self._b = self._solver = None

    def forward(self, theta):
        self._assemble(theta)
        u = U.Vector(self.N)
        ok = self._solver.apply(u, self._b)
        self.last_steps["forward"] = self._solver.steps()
        if not ok:
            raise U.UGError("forward solve did not converge")
        if not self.keep_operator:
            self._release()
        return u

    def backward(self, theta, u, dJdu):
        if self._solver is None:
            self._assemble(theta)
        rhs = U.Vector(self.N)
        _to_vector(dJdu, rhs)
        self.disc.adjust_defect(rhs)
        lam = U.Vector(self.N)
        ok = self._solver.apply(lam, rhs)
        self.last_steps["adjoint"] = self._solver.steps()
        if not ok:
            raise U.UGError("adjoint solve did not converge")
        return lam


class _Solve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, problem):
        u = problem.forward(theta)
        ut = _from_vector(u, theta)
        ctx.problem = problem
        ctx.save_for_backward(theta, ut)
        return ut

    @staticmethod
    def backward(ctx, grad_u):
        theta, ut = ctx.saved_tensors
        problem = ctx.problem
        lam = problem.backward(theta, ut, grad_u)
        u = U.Vector(problem.N)
        _to_vector(ut, u)
        g = torch.from_numpy(problem.disc.adjoint_gradient(u, lam))
        return g, None


def solve(theta, problem):
    """u(theta) with gradients through the adjoint (torch.autograd)."""
    return _Solve.apply(theta, problem)


def observation_points(n, count, seed):
    """`count` distinct seeded random interior lattice vertices of an n^3 mesh
    (lattice indices, x fastest)."""
    from ugcore_amd.synthetic import splitmix64
    W = n + 1
    interior = (n - 1) ** 3
    keys = splitmix64(np.uint64(seed) ^ np.arange(interior, dtype=np.uint64))
    pick = np.argsort(keys, kind="stable")[:count]
    i = pick % (n - 1) + 1
    j = (pick // (n - 1)) % (n - 1) + 1
    k = pick // ((n - 1) ** 2) + 1
    return ((k * W + j) * W + i).astype(np.int64)
