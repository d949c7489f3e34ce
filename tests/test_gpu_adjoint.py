"""GPU tests of the adjoint gradient (torch.autograd over the ug4py bindings):
the element kernel against the oracle, the full gradient against central
differences and against the oracle's two-solve gradient, with host tensors and
with device tensors (which give the same gradient bit for bit).

Tolerances: gradient vs finite differences and vs the two-solve gradient
1e-6 relative; element kernel 1e-12 relative."""
import sys

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


DEVICES = ["cpu", "cuda"]


def setup(refs, seed=7, seed_true=11, n_obs=64, noise=1e-4, DEV="cpu"):
    from ugcore_amd import adjoint as ADJ
    import torch
    n = 2 ** refs
    theta = torch.tensor(O.make_theta(n, seed, 3), dtype=torch.float64, device=DEV)
    prob = ADJ.DiffusionProblem(refs, reduction=1e-14, max_steps=200)
    obs_lat = ADJ.observation_points(n, n_obs, 5)
    dol = prob.dof_of_lattice()
    obs = torch.tensor(dol[obs_lat], dtype=torch.int64, device=DEV)
    # data from a second field theta* plus N(0, noise^2)
    with torch.no_grad():
        u_true = ADJ.solve(torch.tensor(O.make_theta(n, seed_true, 3), dtype=torch.float64, device=DEV), prob)
    rng = np.random.default_rng(3)
    d = u_true[obs] + torch.tensor(noise * rng.normal(size=n_obs), device=DEV)
    return ADJ, torch, prob, theta, obs, d


def loss(ADJ, prob, theta, obs, d):
    u = ADJ.solve(theta, prob)
    return 0.5 * ((u[obs] - d) ** 2).sum()


@pytest.mark.parametrize("refs", [3, 4])
def test_gradient_kernel_matches_oracle(refs):
    say(f"case adjoint kernel refs={refs}")
    import ugcore_amd as U
    n = 2 ** refs
    theta = O.make_theta(n, 9, 3)
    p = O.Problem(0, refs)
    p.set_theta(theta)
    d = U.DomainDiscretization(0, refs)
    d.set_theta(theta)
    rng = np.random.default_rng(2)
    N = d.num_dofs()
    u = rng.normal(size=N)
    lam = rng.normal(size=N)
    dm = p.dirichlet(refs)
    u[dm == 1] = 0.0
    g = d.adjoint_gradient(U.Vector(values=u), U.Vector(values=lam))
    go = p.adjoint_gradient(u, lam)
    assert np.abs(g - go).max() <= 1e-12 * np.abs(go).max()


@pytest.mark.parametrize("DEV", DEVICES)
def test_gradient_vs_central_differences(DEV):
    say(f"case adjoint fd {DEV}")
    refs = 4
    ADJ, torch, prob, theta, obs, d = setup(refs, DEV=DEV)
    th = theta.clone().requires_grad_(True)
    J = loss(ADJ, prob, th, obs, d)
    J.backward()
    g = th.grad.detach()
    assert prob.last_steps["adjoint"] > 0
    rng = np.random.default_rng(16)
    h = 1e-6
    worst = 0.0
    with torch.no_grad():
        for k in range(16):
            v = torch.tensor(rng.normal(size=theta.numel()), device=DEV)
            jp = loss(ADJ, prob, theta + h * v, obs, d).item()
            jm = loss(ADJ, prob, theta - h * v, obs, d).item()
            fd = (jp - jm) / (2 * h)
            ad = float((g * v).sum())
            assert g.device.type == DEV
            worst = max(worst, abs(fd - ad) / max(abs(ad), 1e-300))
    say(f"adjoint vs central differences: worst relative difference {worst:.2e}")
    assert worst <= 1e-6


@pytest.mark.parametrize("DEV", DEVICES)
def test_gradient_vs_oracle_two_solves(DEV):
    say(f"case adjoint two-solve {DEV}")
    refs = 4
    ADJ, torch, prob, theta, obs, d = setup(refs, DEV=DEV)
    th = theta.clone().requires_grad_(True)
    J = loss(ADJ, prob, th, obs, d)
    J.backward()
    g = th.grad.detach().cpu().numpy()
    # oracle: forward solve, adjoint solve with the same rhs rule, element gradient
    p = O.Problem(0, refs)
    p.set_theta(theta.detach().cpu().numpy())
    p.assemble_all()
    u = p.solve(base_lev=0, rel_red=1e-14)["x"]
    obs_np = obs.cpu().numpy()
    rhs = np.zeros_like(u)
    np.add.at(rhs, obs_np, u[obs_np] - d.cpu().numpy())
    lam = p.solve_rhs(rhs, base_lev=0, rel_red=1e-14)
    go = p.adjoint_gradient(u, lam)
    assert np.abs(g - go).max() <= 1e-6 * np.abs(go).max()


@pytest.mark.parametrize("DEV", DEVICES)
def test_gradient_without_stored_operator(DEV):
    """keep_operator=False drops A after the forward solve and re-assembles it
    in backward; the gradient is unchanged."""
    say(f"case adjoint recompute {DEV}")
    refs = 3
    from ugcore_amd import adjoint as ADJ
    import torch
    n = 2 ** refs
    theta = torch.tensor(O.make_theta(n, 7, 3), device=DEV)
    obs = torch.arange(0, 200, 7, device=DEV)
    grads = []
    for keep in (True, False):
        prob = ADJ.DiffusionProblem(refs, reduction=1e-14, keep_operator=keep)
        th = theta.clone().requires_grad_(True)
        u = ADJ.solve(th, prob)
        (0.5 * (u[obs] ** 2).sum()).backward()
        grads.append(th.grad.detach().cpu().numpy())
    assert np.abs(grads[0] - grads[1]).max() <= 1e-12 * np.abs(grads[0]).max()


def test_device_and_host_tensors_give_the_same_gradient():
    """the device-tensor path moves data by device copies only; the gradient
    equals the host-tensor one bit for bit, and torch and the library share one
    HIP runtime"""
    say("case adjoint device == host")
    import ugcore_amd as U
    refs = 4
    grads, us = {}, {}
    for DEV in DEVICES:
        ADJ, torch, prob, theta, obs, d = setup(refs, DEV=DEV)
        th = theta.clone().requires_grad_(True)
        u = ADJ.solve(th, prob)
        (0.5 * ((u[obs] - d) ** 2).sum()).backward()
        grads[DEV] = th.grad.detach().cpu().numpy()
        us[DEV] = u.detach().cpu().numpy()
    assert len(U.hip_runtimes()) == 1
    assert np.array_equal(us["cpu"], us["cuda"])
    assert np.array_equal(grads["cpu"], grads["cuda"])
