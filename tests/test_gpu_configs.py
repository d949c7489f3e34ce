"""GPU tests of the full-size benchmark configurations.

* 128^3 Q1 diffusion (2,146,689 DoFs, 57,066,625 nonzeros), FP64 throughout:
  pattern and DoF numbering bit-exact against the oracle, entries within 1e-12
  of the row's largest entry, GMG V(2,2)-CG solution within 1e-9 (relative L2)
  of the oracle's solve, CG steps within +-1.
* 256^3 Q1 diffusion (16,974,593 DoFs, 454,756,609 nonzeros), mixed
  precision: with the FP32 and the FP64 multigrid preconditioner (base level
  2 = 4^3) both solves reach the 1e-10 reduction measured with the assembled
  FP64 CSR, the two solutions agree to 1e-9, and the step counts are printed.
  The oracle finishes this size in minutes, not seconds, so it is compared at
  128^3 (above) and here only through these size-independent properties.
* 64^3 adjoint gradient with 4096 observations (262,144 parameters): the
  autograd gradient against central differences (h = 1e-6) along 16 seeded
  directions and elementwise against the oracle's two-solve gradient, both to
  1e-6 relative (per direction, relative to the directional derivative).
"""
import sys
import time

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def U():
    import ugcore_amd
    return ugcore_amd


def row_scaled_error(rp, a, b):
    """max over rows of max|a - b| / max|b| within the row"""
    rp = np.asarray(rp, np.int64)
    lens = np.diff(rp)
    nz = lens > 0
    starts = rp[:-1][nz]
    scale = np.maximum.reduceat(np.abs(b), starts)
    err = np.maximum.reduceat(np.abs(a - b), starts)
    return float((err / np.maximum(scale, 1e-300)).max())


def test_q1_128_against_the_oracle():
    say("case 128^3 Q1 FP64 against the oracle")
    u = U()
    refs = 7
    n = 128
    theta = O.make_theta(n, 7, 3)
    t0 = time.perf_counter()
    p = O.Problem(0, refs)
    p.set_theta(theta)
    p.assemble_all()
    ref = p.solve(base_lev=0, rel_red=1e-10)
    say(f"  oracle: {ref['steps']} CG steps, {time.perf_counter() - t0:.1f} s")
    assert ref["ok"]

    d = u.DomainDiscretization(0, refs)
    d.set_theta(theta)
    # DoF numbering: the lattice point of every DoF
    lat = p.dof_lattice(refs)
    W = n + 1
    assert np.array_equal(d.dof_of_lattice(refs)[(lat[:, 2] * W + lat[:, 1]) * W + lat[:, 0]],
                          np.arange(d.num_dofs()))
    A = u.SparseMatrix()
    b = u.Vector(d.num_dofs())
    d.assemble_linear(A, b)
    assert d.num_dofs() == 129 ** 3 and A.nnz() == 385 ** 3
    rp, col, val = A.download_csr()
    R = p.A(refs)
    assert np.array_equal(rp.astype(np.int64), R.rowptr)
    assert np.array_equal(col, R.col)
    err_a = row_scaled_error(rp, val, R.val)
    say(f"  matrix entries: worst row-scaled difference {err_a:.2e}")
    assert err_a <= 1e-12
    bo = p.rhs(refs)
    assert np.abs(b.numpy() - bo).max() <= 1e-12 * np.abs(bo).max()
    del rp, col, val, R

    g = u.GeometricMultiGrid(d, base_level=0, precision="f64")
    S = u.CG(g, max_steps=100, reduction=1e-10)
    S.init(A)
    x = u.Vector(d.num_dofs())
    assert S.apply(x, b)
    err = np.linalg.norm(x.numpy() - ref["x"]) / np.linalg.norm(ref["x"])
    say(f"  GPU: {S.steps()} CG steps (oracle {ref['steps']}), solution rel L2 difference {err:.2e}")
    assert abs(S.steps() - ref["steps"]) <= 1
    assert err <= 1e-9


def test_q1_256_mixed_precision_properties():
    say("case 256^3 Q1 mixed precision")
    u = U()
    from ugcore_amd.synthetic import make_theta
    refs = 8
    theta = make_theta(256, 7, 3)
    d = u.DomainDiscretization(0, refs)
    d.set_theta(theta)
    A = u.SparseMatrix()
    b = u.Vector(d.num_dofs())
    d.assemble_linear(A, b)
    N = d.num_dofs()
    assert N == 257 ** 3 and A.nnz() == 769 ** 3
    bn = np.linalg.norm(b.numpy())
    sols, steps = {}, {}
    for prec in ("f32", "f64"):
        g = u.GeometricMultiGrid(d, base_level=2, precision=prec)
        S = u.CG(g, max_steps=100, reduction=1e-10)
        S.init(A)
        x = u.Vector(N)
        assert S.apply(x, b)
        assert S.used_lattice_path()
        r = u.Vector(values=b.numpy())
        A.apply_sub(r, x)                  # r = b - A x with the assembled CSR
        rel = np.linalg.norm(r.numpy()) / bn
        say(f"  {prec} preconditioner: {S.steps()} CG steps, ||b - A x|| / ||b|| = {rel:.2e}")
        assert rel < 1e-10
        sols[prec], steps[prec] = x.numpy(), S.steps()
    diff = np.linalg.norm(sols["f32"] - sols["f64"]) / np.linalg.norm(sols["f64"])
    say(f"  FP32 vs FP64 preconditioned solution: rel L2 difference {diff:.2e}")
    assert diff <= 1e-9
    assert abs(steps["f32"] - steps["f64"]) <= 2


@pytest.mark.parametrize("DEV", ["cuda", "cpu"])
def test_adjoint_64_4096_observations(DEV):
    say(f"case 64^3 adjoint gradient, 4096 observations, {DEV} tensors")
    from ugcore_amd import adjoint as ADJ
    import torch
    refs = 6
    n = 64
    n_obs = 4096
    theta = torch.tensor(O.make_theta(n, 7, 3), dtype=torch.float64, device=DEV)
    prob = ADJ.DiffusionProblem(refs, reduction=1e-14, max_steps=200)
    obs_lat = ADJ.observation_points(n, n_obs, 5)
    obs = torch.tensor(prob.dof_of_lattice()[obs_lat], dtype=torch.int64, device=DEV)
    assert len(set(obs.tolist())) == n_obs
    with torch.no_grad():
        u_true = ADJ.solve(torch.tensor(O.make_theta(n, 11, 3), dtype=torch.float64, device=DEV), prob)
    rng = np.random.default_rng(3)
    dat = u_true[obs] + torch.tensor(1e-2 * rng.normal(size=n_obs), device=DEV)   # N(0, 1e-4): std 1e-2

    def loss(th):
        uu = ADJ.solve(th, prob)
        return 0.5 * ((uu[obs] - dat) ** 2).sum()

    th = theta.clone().requires_grad_(True)
    J = loss(th)
    J.backward()
    assert th.grad.device == theta.device
    g = th.grad.detach().cpu().numpy().copy()
    say(f"  J = {J.item():.6e}, forward {prob.last_steps['forward']} / adjoint {prob.last_steps['adjoint']} "
        f"CG steps")

    # central differences along 16 seeded directions. J(+) - J(-) is formed as
    # 1/2 sum (u+ - u-)(u+ + u- - 2d), which equals the difference of the two
    # losses but does not cancel: the losses themselves agree to about 1e-11
    # relative, so subtracting them would leave only a few ulps of J, below
    # the 1e-6 asked of directions whose derivative is small
    rng = np.random.default_rng(16)
    h = 1e-6
    worst = 0.0
    with torch.no_grad():
        for k in range(16):
            v = rng.normal(size=theta.numel())
            vt = torch.tensor(v, device=DEV)
            up = ADJ.solve(theta + h * vt, prob)[obs]
            um = ADJ.solve(theta - h * vt, prob)[obs]
            fd = float((0.5 * (up - um) * (up + um - 2.0 * dat)).sum()) / (2 * h)
            ad = float(g @ v)
            worst = max(worst, abs(fd - ad) / abs(ad))
    say(f"  adjoint vs central differences (16 directions): worst relative difference {worst:.2e}")
    assert worst <= 1e-6

    # the oracle's two-solve gradient, elementwise
    p = O.Problem(0, refs)
    p.set_theta(theta.cpu().numpy())
    p.assemble_all()
    uo = p.solve(base_lev=0, rel_red=1e-14)["x"]
    o = obs.cpu().numpy()
    rhs = np.zeros_like(uo)
    np.add.at(rhs, o, uo[o] - dat.cpu().numpy())
    lam = p.solve_rhs(rhs, base_lev=0, rel_red=1e-14)
    go = p.adjoint_gradient(uo, lam)
    err = np.abs(g - go).max() / np.abs(go).max()
    say(f"  adjoint vs oracle two-solve gradient: max elementwise difference / max|g| = {err:.2e}")
    assert err <= 1e-6
