"""Adjoint gradients of the Q1 diffusion solve behind torch.autograd, on the
pybind11 bindings (ugcore_amd.ug4py).

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

Tensors may be host or device (GPU) tensors. Device tensors are exchanged with
the library by address: theta is copied device to device, u is written into a
new device tensor, dJ/du is copied into the adjoint right-hand side, and the
gradient kernel writes straight into the gradient tensor. That needs one HIP
runtime for torch and the library, which importing ugcore_amd arranges when
torch is installed (ugcore_amd.hip_runtimes() lists the mapped runtimes);
with two runtimes device tensors are refused with that reason.

The reference has no adjoint; this module has no counterpart there.
"""
import numpy as np

import ugcore_amd as U          # binds torch's HIP runtime first (one runtime)
from ugcore_amd import ug4py as P
import torch                    # noqa: E402


def _shared_runtime():
    return len(U.hip_runtimes()) <= 1


def _device_ok(t):
    if t.is_cuda and not _shared_runtime():
        raise U.UGError("device tensors need one HIP runtime for torch and this library; "
                        f"mapped: {U.hip_runtimes()} (import ugcore_amd before anything loads the system runtime)")
    return t.is_cuda


def _to_vector(t, vec):
    """copy a float64 tensor (host or device) into a ug4py Vector"""
    t = t.detach()
    if t.dtype != torch.float64:
        raise TypeError("float64 tensors expected")
    t = t.contiguous()
    if t.numel() != vec.size():
        raise ValueError("size mismatch")
    if _device_ok(t):
        torch.cuda.current_stream(t.device).synchronize()   # t's values are written
        vec.copy_from(t.data_ptr())                         # ordered on the library stream, synchronous
    else:
        vec.upload(t.numpy())


def _from_vector(vec, like):
    """a new tensor on like's device with the vector's values"""
    if _device_ok(like):
        out = torch.empty(vec.size(), dtype=torch.float64, device=like.device)
        torch.cuda.current_stream(like.device).synchronize()
        vec.copy_to(out.data_ptr())
        return out
    return torch.from_numpy(vec.numpy())


class DiffusionProblem:
    """GPU state for one mesh: discretisation, operator, GMG-CG solver."""

    def __init__(self, num_refs, f=1.0, base_level=0, precision="f64", reduction=1e-12, max_steps=200,
                 keep_operator=True):
        P.init()
        self.num_refs = num_refs
        self.n = 2 ** num_refs
        self.disc = P.DomainDiscretization(0, num_refs)
        self.disc.set_source(float(f))
        self.N = self.disc.num_dofs()
        self.base_level = base_level
        self.precision = precision
        self.reduction = reduction
        self.max_steps = max_steps
        self.keep_operator = keep_operator
        self._A = self._b = self._solver = self._gmg = None
        self.last_steps = {}

    def dof_of_lattice(self):
        """lattice point (x fastest) -> DoF of the top level"""
        return self.disc.dof_of_lattice()

    def _set_theta(self, theta):
        theta = theta.detach()
        if theta.dtype != torch.float64 or theta.numel() != self.n ** 3:
            raise ValueError("theta must be float64 with n_fine^3 entries")
        theta = theta.contiguous()
        if _device_ok(theta):
            torch.cuda.current_stream(theta.device).synchronize()
            self.disc.set_theta_device(theta.data_ptr())
            P.synchronize()                                  # the copy has read theta
        else:
            self.disc.set_theta(theta.numpy())

    def _assemble(self, theta):
        self._set_theta(theta)
        self._A = P.SparseMatrix()
        self._b = P.Vector(self.N)
        self.disc.assemble_linear(self._A, self._b)
        self._gmg = P.GeometricMultiGrid(self.disc, base_level=self.base_level, precision=self.precision)
        self._solver = P.CG(self._gmg, max_steps=self.max_steps, reduction=self.reduction)
        self._solver.init(self._A)

    def _release(self):
        self._A = self._b = self._solver = self._gmg = None

    def forward(self, theta):
        self._assemble(theta)
        u = P.Vector(self.N)
        ok = self._solver.apply(u, self._b)
        self.last_steps["forward"] = self._solver.steps()
        if not ok:
            raise U.UGError("forward solve did not converge")
        if not self.keep_operator:
            self._release()
        return u

    def backward(self, theta, dJdu):
        if self._solver is None:
            self._assemble(theta)
        rhs = P.Vector(self.N)
        _to_vector(dJdu, rhs)
        self.disc.adjust_defect(rhs)
        lam = P.Vector(self.N)
        ok = self._solver.apply(lam, rhs)
        self.last_steps["adjoint"] = self._solver.steps()
        if not ok:
            raise U.UGError("adjoint solve did not converge")
        return lam

    def gradient(self, u_tensor, lam, like):
        """dJ/dtheta from the saved u and the adjoint state, on like's device"""
        u = P.Vector(self.N)
        _to_vector(u_tensor, u)
        ne = self.n ** 3
        if _device_ok(like):
            g = torch.empty(ne, dtype=torch.float64, device=like.device)
            torch.cuda.current_stream(like.device).synchronize()
            self.disc.adjoint_gradient(u, lam, g.data_ptr())
            P.synchronize()
            return g
        gv = P.Vector(ne)
        self.disc.adjoint_gradient(u, lam, gv.device_ptr())
        P.synchronize()
        return torch.from_numpy(gv.numpy())


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
        lam = problem.backward(theta, grad_u.to(ut.device))
        return problem.gradient(ut, lam, theta), None


def solve(theta, problem):
    """u(theta) with gradients through the adjoint (torch.autograd); theta a
    host or device float64 tensor, u on the same device."""
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
