#!/usr/bin/env python3
"""Benchmark: 256^3 Q1 GMG-CG time to a 1e-10 relative residual on one MI355X,
FP32-preconditioned vs FP64-preconditioned, plus CSR SpMV HBM bandwidth.

Workload (synthetic, seeded): -div(k grad u) = 1 on the unit cube, one
hexahedron refined 8 times (256^3 Q1 elements, 16,974,593 DoFs, 454,756,609
nnz), k = exp(theta) with theta i.i.d. N(0,1) per element smoothed by 3 passes
of the 27-element box average, homogeneous Dirichlet boundary. Solver: CG with
a GMG V(2,2) preconditioner, damped Jacobi (omega 0.66), dense FP64 LU on the
4^3 base level (level 2), reduction 1e-10, x0 = 0.

A step is one complete solve from x0 = 0 (the preconditioner set-up and the
assembly are done once, before the timed region, and reported separately).
`value` is the FP32-preconditioned time-to-solution in seconds.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "time-to-1e-10 residual, 256³ Q1 GMG-CG: FP32-precond vs FP64; SpMV HBM GB/s"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(refs_cpu, seed, gpu_iters, refs_gpu):
    """Single-core CPU restatement of the reference CPUAlgebra path (the oracle),
    timed on a bounded sample (refs_cpu) and scaled per DoF-iteration to the
    GPU workload."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    t0 = time.time()
    p = O.Problem(0, refs_cpu)
    t_mesh = time.time() - t0
    n = 2 ** refs_cpu
    p.set_theta(O.make_theta(n, seed, 3))
    p.assemble_all()
    # the solve timed three times (median), as the baseline plan asks
    solves = []
    for _ in range(3):
        r = p.solve(base_lev=2, rel_red=1e-10, max_steps=200)
        solves.append(p.timings()["solve"])
    tm = p.timings()
    t_solve = float(np.median(solves))
    dofs_cpu = (n + 1) ** 3
    dofs_gpu = (2 ** refs_gpu + 1) ** 3
    per_dof_iter = t_solve / (dofs_cpu * max(r["steps"], 1))
    est = per_dof_iter * dofs_gpu * gpu_iters
    return {
        "value": est,
        "unit": "s",
        "cores": 1,
        "kind": "port",
        "sample": (f"oracle (single-core C++ restatement of CPUAlgebra GMG-CG, FP64) on {n}^3 Q1 "
                   f"({dofs_cpu} DoFs): solve {t_solve:.2f} s (median of 3: "
                   + ", ".join(f"{t:.2f}" for t in solves) + f") in {r['steps']} CG steps, assembly "
                   f"{tm['assemble']:.2f} s, mesh {t_mesh:.2f} s; value = solve time per DoF-iteration "
                   f"x {dofs_gpu} DoFs x {gpu_iters} iterations"),
        "measured_solve_s": t_solve,
        "measured_solve_runs_s": solves,
        "measured_assemble_s": tm["assemble"],
        "measured_steps": r["steps"],
        "host_cpu": _cpu_model(),
        "host_cores_total": os.cpu_count(),
    }


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE, scaled by a
    1 GiB device-copy calibration) from the committed rocprofv3 summary of this
    workload, or None if the summary has no row for the kernel."""
    path = os.path.join(ROOT, "profiles", "gmg_cg_256_pmc_traffic.md")
    try:
        with open(path) as f:
            cols = None
            for line in f:
                if not line.startswith("|"):
                    continue
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                if "HBM fetch B/launch" in cells:
                    cols = (cells.index("HBM fetch B/launch"), cells.index("HBM write B/launch"))
                elif cols and cells[0] == f"`{kernel}`":
                    return float(cells[cols[0]].replace(",", "")) + float(cells[cols[1]].replace(",", ""))
    except OSError:
        pass
    return None


def _timed(U, fn, reps, warmup=1):
    """mean wall time of fn() over reps runs, device synchronised around each"""
    out = []
    for it in range(warmup + reps):
        U.synchronize()
        t0 = time.perf_counter()
        r = fn()
        U.synchronize()
        if it >= warmup:
            out.append(time.perf_counter() - t0)
    return float(np.mean(out)), r


def bench_poisson_2d(U, seed, reps):
    """2D FV1 Poisson: 2 triangles refined 7 times (16,641 DoFs), f =
    2 pi^2 sin(pi x) sin(pi y) + N(0, 0.1^2) per vertex, CG + GMG V(2,2) Jacobi
    0.66, LU on level 0, FP64, tol 1e-10"""
    from ugcore_amd.synthetic import poisson_source_lattice
    refs = 7
    n = 2 ** refs
    d = U.DomainDiscretization(1, refs)
    dol = d.dof_of_lattice(refs)
    f = np.empty(d.num_dofs())
    f[dol] = poisson_source_lattice(n, seed)
    d.set_source(f)
    A = U.SparseMatrix()
    b = U.Vector(d.num_dofs())
    d.assemble_linear(A, b)
    g = U.GeometricMultiGrid(d, base_level=0, omega=0.66, precision="f64")
    S = U.CG(g, max_steps=200, reduction=1e-10)
    S.init(A)
    x = U.Vector(d.num_dofs())

    def solve():
        x.set(0.0)
        if not S.apply(x, b):
            raise RuntimeError("2D solve did not converge")
    t, _ = _timed(U, solve, reps)
    return {"workload": "2D FV1 Poisson, 2 triangles x 7 refinements, CG + GMG V(2,2), FP64, tol 1e-10",
            "dofs": d.num_dofs(), "nnz": A.nnz(), "solve_s": t, "cg_steps": S.steps(),
            "assembly_s": d.assemble_seconds()}


def bench_q2_elements(U, seed, reps, n_cells=64):
    """Q2 hexahedron element matrices (one element per wavefront, FP64 matrix
    cores) for the n_cells^3 elements of a jittered lattice, k = exp(theta)
    per element: elements/s, the HBM rate of the element-matrix output, and
    the FP64 rate of the contraction (27 x 27 x 81 useful multiply-adds per
    element; 63 MFMA 16x16x4 steps issued: three of the four 16 x 16 tiles of
    the symmetric matrix, 21 steps each)"""
    rng = np.random.default_rng(seed)
    n = n_cells ** 3
    h = 1.0 / n_cells
    e = np.arange(n)
    base = np.stack([e % n_cells, (e // n_cells) % n_cells, e // (n_cells * n_cells)], axis=1) * h
    corners = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    X = base[:, None, :] + corners[None] * h + rng.uniform(-0.1 * h, 0.1 * h, (n, 8, 3))
    k = np.exp(rng.normal(scale=0.5, size=n))
    dx, dk = U.DeviceArray(X), U.DeviceArray(k)
    ke = U.DeviceArray(np.zeros(n * 729))
    fe = U.DeviceArray(np.zeros(n * 27))
    L = U.lib()
    t = U.time_launches(lambda: L.ugx_q2_hex_element_matrices(n, dx.ptr, dk.ptr, 1.0, ke.ptr, fe.ptr), reps)
    bytes_ = n * 8 * (24 + 1 + 729 + 27)
    res = {"elements": n, "ms": t * 1e3, "elements_per_s": n / t, "GBps": bytes_ / t / 1e9,
           "frac_of_8TBps": bytes_ / t / 1e9 / HBM_PEAK_GBS,
           "useful_fp64_TFLOPs": n * 2 * 27 * 27 * 81 / t / 1e12,
           "mfma_issued_fp64_TFLOPs": n * 63 * 16 * 16 * 4 * 2 / t / 1e12}
    log(f"[bench] Q2 element matrices {n_cells}^3: {t*1e3:.3f} ms, {res['GBps']:.0f} GB/s written+read, "
        f"{res['useful_fp64_TFLOPs']:.2f} TFLOP/s useful FP64")
    del dx, dk, ke, fe
    # the whole Q2 system of the unit cube (pattern on the device, host scan of
    # the row lengths, colour-ordered scatter, Dirichlet rows), wall clock
    theta = rng.normal(scale=0.5, size=n)
    U.q2_assemble_lattice(n_cells, theta, 1.0)
    t0 = time.perf_counter()
    A, b = U.q2_assemble_lattice(n_cells, theta, 1.0)
    t_asm = time.perf_counter() - t0
    res["lattice_assembly_s"] = t_asm
    res["lattice_assembly_dofs"] = (2 * n_cells + 1) ** 3
    res["lattice_assembly_nnz"] = A.nnz()
    log(f"[bench] Q2 lattice assembly {n_cells}^3 elements: {t_asm*1e3:.1f} ms, {A.nnz()} nonzeros")
    return res


def bench_galerkin(U, disc, A, b, base_level, reps):
    """the headline 256^3 Q1 problem with Galerkin coarse operators
    (set_rap: A_{l-1} = R_l A_l P_l from the assembled top level) instead of
    the assembled level matrices: set-up (R A P on every level, FP64) and
    solve, both preconditioner precisions"""
    N = disc.num_dofs()
    out = {"workload": "256^3 Q1 FE as the headline solve, GMG with Galerkin coarse operators (set_rap), "
                       "CG tol 1e-10"}
    for prec in ("f32", "f64"):
        g = U.GeometricMultiGrid(disc, base_level=base_level, nu1=2, nu2=2, cycle=1, omega=0.66, precision=prec,
                                 rap=True)
        S = U.CG(g, max_steps=100, reduction=1e-10)
        U.synchronize()
        t0 = time.perf_counter()
        S.init(A)
        U.synchronize()
        setup = time.perf_counter() - t0
        x = U.Vector(N)

        def solve():
            x.set(0.0)
            if not S.apply(x, b):
                raise RuntimeError("Galerkin solve did not converge")
        t, _ = _timed(U, solve, reps)
        out[f"{prec}_precond_solve_s"] = t
        out[f"{prec}_precond_cg_steps"] = S.steps()
        out[f"{prec}_setup_s"] = setup
        del S, g
    return out


def bench_adjoint(U, refs, seed, reps, precision="f32"):
    """adjoint gradient of J = 1/2 |u(obs) - d|^2 for the Q1 log-diffusion
    field: forward GMG-CG solve, adjoint GMG-CG solve (same operator, Dirichlet
    entries of dJ/du zeroed), element gradient kernel; 4096 observations"""
    from ugcore_amd.adjoint import observation_points
    from ugcore_amd.synthetic import make_theta
    n = 2 ** refs
    d = U.DomainDiscretization(0, refs)
    d.set_theta(make_theta(n, seed, 3))
    A = U.SparseMatrix()
    N = d.num_dofs()
    b = U.Vector(N)
    t_asm, _ = _timed(U, lambda: d.assemble_linear(A, b), 1, warmup=0)
    g = U.GeometricMultiGrid(d, base_level=min(2, refs - 1), omega=0.66, precision=precision)
    S = U.CG(g, max_steps=200, reduction=1e-10)
    S.init(A)
    obs_lat = observation_points(n, 4096, seed)
    obs = d.dof_of_lattice(refs)[obs_lat]
    rng = np.random.default_rng(seed)
    dobs = rng.normal(size=obs.size) * 1e-2
    u = U.Vector(N)
    lam = U.Vector(N)
    rhs = U.Vector(N)
    grad = U.DeviceArray(np.zeros(n ** 3))

    def forward():
        u.set(0.0)
        if not S.apply(u, b):
            raise RuntimeError("forward solve did not converge")
        return S.steps()

    # dJ/du = u - d on the observed DoFs, formed once on the host and kept on
    # the device (as the torch.autograd path with device tensors hands it
    # over); the timed backward is device work only
    t_f, st_f = _timed(U, forward, reps)
    t0 = time.perf_counter()
    uv = u.numpy()
    r = np.zeros(N)
    r[obs] = uv[obs] - dobs
    dJdu = U.Vector(values=r)
    t_host_misfit = time.perf_counter() - t0

    def backward():
        U.lib().ugx_vec_scale_add2_f64(N, rhs.device_ptr(), 1.0, dJdu.device_ptr(), 0.0, dJdu.device_ptr(), None)
        d.adjust_defect(rhs)
        lam.set(0.0)
        if not S.apply(lam, rhs):
            raise RuntimeError("adjoint solve did not converge")
        d.adjoint_gradient(u, lam, grad_ptr=grad.ptr)
        return S.steps()

    t_b, st_b = _timed(U, backward, reps)
    t_k = U.time_launches(lambda: d.adjoint_gradient(u, lam, grad_ptr=grad.ptr), 10)
    return {"workload": f"adjoint gradient, {n}^3 Q1 log-diffusion, 4096 observations, GMG({precision})-CG tol 1e-10",
            "parameters": n ** 3, "forward_solve_s": t_f, "backward_s": t_b, "gradient_kernel_ms": t_k * 1e3,
            "forward_cg_steps": st_f, "adjoint_cg_steps": st_b, "assembly_s": t_asm,
            "host_misfit_download_upload_s": t_host_misfit,
            "gradient_total_s_excl_assembly": t_f + t_b}


def bench_tet(U, refs, seed, reps):
    """96^3 x 6 tetrahedra, FV1 full upwind, v = 20*(1, 0.5, 0.25), interior
    vertices displaced by at most 0.2h (+-0.2h/sqrt(3) per component), BiCGStab + GMG V(2,2) Jacobi 0.6 (FP32 and FP64 levels), tol 1e-10;
    adjoint solve with the transposed operator and transposed GMG"""
    from ugcore_amd.synthetic import make_theta
    d = U.DomainDiscretization(2, refs)
    d.set_theta(make_theta(d.n_fine, seed, 3))
    d.set_velocity((20.0, 10.0, 5.0))
    d.set_jitter(11, 0.2 / np.sqrt(3.0))      # displacement length <= 0.2h
    d.set_source(1.0)
    N = d.num_dofs()
    A = U.SparseMatrix()
    b = U.Vector(N)
    d.assemble_linear(A, b)
    res = {"workload": f"{d.n_fine}^3 x 6 tetrahedra FV1 convection-diffusion (full upwind, Pe 20, jitter |d| <= 0.2h), "
                       "BiCGStab + GMG V(2,2) Jacobi 0.6, tol 1e-10",
           "dofs": N, "nnz": A.nnz(), "assembly_s": d.assemble_seconds()}
    xs = {}
    for prec in ("f32", "f64"):
        g = U.GeometricMultiGrid(d, base_level=0, omega=0.6, precision=prec)
        S = U.BiCGStab(g, max_steps=500, reduction=1e-10)
        S.init(A)
        x = U.Vector(N)

        def solve():
            x.set(0.0)
            if not S.apply(x, b):
                raise RuntimeError("tetrahedral solve did not converge")
        t, _ = _timed(U, solve, reps)
        res[f"{prec}_precond_s"] = t
        res[f"{prec}_bicgstab_steps"] = S.steps()
        xs[prec] = x.numpy()
    res["rel_l2_diff_fp32_vs_fp64_precond_solution"] = float(
        np.linalg.norm(xs["f32"] - xs["f64"]) / np.linalg.norm(xs["f64"]))
    AT = d.adjoint_operator(A)
    gT = U.GeometricMultiGrid(d, base_level=0, omega=0.6, precision="f32", transposed=True)
    ST = U.BiCGStab(gT, max_steps=500, reduction=1e-10)
    ST.init(AT)
    rng = np.random.default_rng(seed)
    r = U.Vector(values=rng.normal(size=N))
    d.adjust_defect(r)
    lam = U.Vector(N)

    def adj():
        lam.set(0.0)
        if not ST.apply(lam, r):
            raise RuntimeError("transposed solve did not converge")
    t_adj, _ = _timed(U, adj, reps)
    grad = U.DeviceArray(np.zeros(d.n_fine ** 3))
    u = U.Vector(values=xs["f64"])
    t_k = U.time_launches(lambda: d.adjoint_gradient(u, lam, grad_ptr=grad.ptr), 10)
    res.update({"adjoint_transposed_f32_precond_s": t_adj, "adjoint_bicgstab_steps": ST.steps(),
                "gradient_kernel_ms": t_k * 1e3})
    return res


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--refs", type=int, default=8)
    ap.add_argument("--base-level", type=int, default=2)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--cpu-refs", type=int, default=7)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--kernel-reps", type=int, default=20)
    ap.add_argument("--skip-extra", action="store_true",
                    help="skip the secondary configurations (2D FV1, adjoint, tetrahedra)")
    ap.add_argument("--pmc-calibration", action="store_true",
                    help="issue one 1 GiB device copy first (counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (requested --gpus {args.gpus}); "
              "multi-GPU sharding is a design note, not an implemented path.", file=sys.stderr)
        sys.exit(2)

    import ugcore_amd as U
    from ugcore_amd.synthetic import make_theta

    U.init()
    L = U.lib()
    if args.pmc_calibration:
        # a plain 1 GiB device copy (runtime copy kernel): reference dispatch
        # for scaling FETCH_SIZE / WRITE_SIZE counters to bytes
        nb = 1 << 30
        src = L.ugx_malloc(nb)
        dst = L.ugx_malloc(nb)
        assert src and dst and L.ugx_memcpy_dtod(dst, src, nb) == 0
        L.ugx_free(src)
        L.ugx_free(dst)
        log("[bench] 1 GiB calibration copy done")
    refs = args.refs
    n = 2 ** refs
    log(f"[bench] building {n}^3 Q1 hierarchy + DoF numbering")
    t0 = time.time()
    disc = U.DomainDiscretization(0, refs)
    t_mesh = time.time() - t0
    # the same hierarchy refined and numbered on the host, then uploaded
    t0 = time.time()
    disc_host = U.DomainDiscretization(0, refs, setup="host")
    t_mesh_host = time.time() - t0
    same_numbering = bool(np.array_equal(disc.dof_of_lattice(refs), disc_host.dof_of_lattice(refs)))
    del disc_host
    log(f"[bench] hierarchy + numbering: device {t_mesh:.3f} s, host {t_mesh_host:.3f} s, "
        f"identical={same_numbering}")
    theta = make_theta(n, args.seed, 3)
    disc.set_theta(theta)
    A = U.SparseMatrix()
    b = U.Vector(disc.num_dofs())
    log(f"[bench] assembling surface operator")
    disc.assemble_linear(A, b)
    t_asm = disc.assemble_seconds()
    # again into the same matrix: its pattern and SpMV layout are kept and
    # only the values (and the layout's copy of them) are rewritten
    disc.assemble_linear(A, b)
    t_reasm = disc.assemble_seconds()
    N = disc.num_dofs()
    nnz = A.nnz()
    log(f"[bench] N={N} nnz={nnz} assembly {t_asm:.3f} s, re-assembly {t_reasm:.4f} s "
        f"(mesh+numbering {t_mesh:.2f} s)")

    # SpMV set-up: CSR kernel lanes-per-row sweep, then the sliced-ELL layout
    # (the GPUSparseMatrix default); the faster one is used by the solves
    xs = U.Vector(values=np.random.default_rng(0).normal(size=N))
    ys = U.Vector(N)
    A.set_spmv_layout("csr")
    sweep = {}
    for lanes in (8, 16, 32):
        A.set_spmv_lanes(lanes)
        sweep[lanes] = U.time_launches(lambda: A.apply(ys, xs), max(3, args.kernel_reps // 4))
    best_lanes = min(sweep, key=sweep.get)
    A.set_spmv_lanes(best_lanes)
    t_csr = U.time_launches(lambda: A.apply(ys, xs), args.kernel_reps)
    A.set_spmv_layout("sell")
    assert A.has_row_order() and A.spmv_layout() == "sell"
    t_sell = U.time_launches(lambda: A.apply(ys, xs), args.kernel_reps)
    # the same slices filled in DoF order instead of the lattice walk
    A.set_row_order(None)
    t_sell_dof_order = U.time_launches(lambda: A.apply(ys, xs), args.kernel_reps)
    A.set_row_order(disc.dof_of_lattice(refs))
    # sliced ELL with a column dictionary per 4x4x4 block of rows
    A.set_spmv_layout("dict")
    assert A.spmv_layout() == "dict"
    t_dict = U.time_launches(lambda: A.apply(ys, xs), args.kernel_reps)
    layout = min((("dict", t_dict), ("sell", t_sell), ("csr", t_csr)), key=lambda kv: kv[1])[0]
    A.set_spmv_layout(layout)
    log(f"[bench] CSR SpMV lanes sweep (ms): " + ", ".join(f"{k}: {v*1e3:.3f}" for k, v in sweep.items())
        + f"; sliced ELL {t_sell*1e3:.3f} ms (slices in DoF order {t_sell_dof_order*1e3:.3f} ms)"
        + f"; dictionary slices {t_dict*1e3:.3f} ms -> {layout}")

    solvers = {}
    setup_s = {}
    for prec in ("f32", "f64"):
        g = U.GeometricMultiGrid(disc, base_level=args.base_level, nu1=2, nu2=2, cycle=1, omega=0.66,
                                 precision=prec)
        S = U.CG(g, max_steps=100, reduction=1e-10)
        t0 = time.time()
        S.init(A)
        U.synchronize()
        setup_s[prec] = time.time() - t0
        solvers[prec] = (g, S)
        log(f"[bench] GMG {prec} set-up {setup_s[prec]:.3f} s")

    x = {p: U.Vector(N) for p in solvers}
    times = {p: [] for p in solvers}
    iters = {}
    for prec, (g, S) in solvers.items():
        for it in range(args.warmup + args.steps):
            x[prec].set(0.0)
            U.synchronize()
            t0 = time.perf_counter()
            ok = S.apply(x[prec], b)
            U.synchronize()
            dt = time.perf_counter() - t0
            if not ok:
                raise RuntimeError(f"{prec} solve did not converge")
            if it >= args.warmup:
                times[prec].append(dt)
            iters[prec] = S.steps()
        log(f"[bench] {prec}-preconditioned: {np.mean(times[prec])*1e3:.2f} ms, {iters[prec]} CG steps")

    x32 = x["f32"].numpy()
    x64 = x["f64"].numpy()
    err = float(np.linalg.norm(x32 - x64) / np.linalg.norm(x64))

    # the same FP32-preconditioned solve from a non-zero initial guess (zero on
    # the Dirichlet DoFs): the initial defect b - A x is formed, on the lattice
    # operator, where the zero start takes b itself
    g32s, S32s = solvers["f32"]
    x0 = np.random.default_rng(args.seed).normal(size=N) * 1e-3
    x0[disc.num_inner(refs):] = 0.0
    xs = U.Vector(N)
    t_nz = []
    for it in range(args.warmup + args.steps):
        xs.upload(x0)
        U.synchronize()
        t0 = time.perf_counter()
        if not S32s.apply(xs, b):
            raise RuntimeError("f32 solve from a non-zero start did not converge")
        U.synchronize()
        if it >= args.warmup:
            t_nz.append(time.perf_counter() - t0)
    iters_nz = S32s.steps()
    log(f"[bench] f32-preconditioned from a non-zero start: {np.mean(t_nz)*1e3:.2f} ms, {iters_nz} CG steps")
    r = U.Vector(values=b.numpy())
    A.apply_sub(r, x["f32"])
    final_rel_res = float(np.linalg.norm(r.numpy()) / np.linalg.norm(b.numpy()))

    # ---- kernel timings with HIP events recorded on the library stream --------
    t_spmv = U.time_launches(lambda: A.apply(ys, xs), args.kernel_reps)
    spmv_bytes = nnz * 12 + (N + 1) * 4 + N * 8 + N * 8
    spmv_gbs = spmv_bytes / t_spmv / 1e9

    # the same operator with an OrderLex-style DoF numbering (lattice order):
    # isolates what the reference numbering costs the x-gathers
    lex = U.DomainDiscretization(0, refs, ordering="lex")
    lex.set_theta(theta)
    A_lex = U.SparseMatrix()
    b_lex = U.Vector(N)
    lex.assemble_linear(A_lex, b_lex)
    A_lex.set_spmv_layout(layout)
    t_spmv_lex = U.time_launches(lambda: A_lex.apply(ys, xs), args.kernel_reps)
    del A_lex, b_lex, lex
    log(f"[bench] SpMV f64 ({layout}) lexicographic numbering: {t_spmv_lex*1e3:.3f} ms")

    # measured STREAM-like device peak: triad c = a + 0.5 b (the library's
    # vector kernel) over three 2 GiB arrays, 24 B per element
    nst = 1 << 28
    bufs = [L.ugx_malloc(8 * nst) for _ in range(3)]
    try:
        assert all(bufs), "ugx_malloc failed for the triad buffers"
        for i, ptr in enumerate(bufs):
            assert L.ugx_vec_set_f64(nst, ptr, float(i + 1), None) == 0
        t_triad = U.time_launches(
            lambda: L.ugx_vec_scale_add2_f64(nst, bufs[2], 1.0, bufs[0], 0.5, bufs[1], None), args.kernel_reps)
    finally:
        for ptr in bufs:
            if ptr:
                L.ugx_free(ptr)
    triad_gbs = 24 * nst / t_triad / 1e9
    log(f"[bench] STREAM-like triad (3 x 2 GiB): {t_triad*1e3:.3f} ms  {triad_gbs:.0f} GB/s")

    g32 = solvers["f32"][0]
    nd = L.ugx_gmg_level_num_diags(g32.h, refs)
    nu_stored = nd // 2 + 1              # symmetric storage: centre + upper diagonals
    t_smooth = U.time_launches(lambda: g32.smooth_level(refs, 1), args.kernel_reps)
    # matrix-free top level: per-element coefficient + dinv + sd in/out + sc read/write
    smooth_bytes = (n ** 3 + N * 5) * 4

    # the operator the CG iterates with: surface operator in lattice order,
    # FP64 symmetric storage
    H = L.ugx_gmg_lattice_halo(g32.h)
    xl = U.DeviceArray(np.concatenate([np.zeros(H), np.random.default_rng(1).normal(size=N), np.zeros(H)]))
    yl = U.DeviceArray(np.zeros(N))
    xl_ptr = xl.ptr + 8 * H

    def lattice_spmv():
        assert L.ugx_gmg_surface_apply_lattice(g32.h, yl.ptr, xl_ptr) == 0

    t_lat = U.time_launches(lattice_spmv, args.kernel_reps)
    lat_bytes = n ** 3 * 8 + N * (8 + 8)   # matrix-free: per-element coefficient, x, y

    # the CG's direction update fused with its operator (the lattice CG's
    # k_q1mf_z<double,5>): p_new = z + beta p_old, q = A p_new, (q, p_new)
    zl = U.DeviceArray(np.concatenate([np.zeros(H), np.random.default_rng(2).normal(size=N), np.zeros(H)]))
    pl2 = U.DeviceArray(np.zeros(N + 2 * H))
    beta_d = U.DeviceArray(np.array([0.5]))
    qp_d = U.DeviceArray(np.zeros(1))
    part_d = U.DeviceArray(np.zeros(65536))

    def direction():
        assert L.ugx_gmg_surface_direction_lattice(g32.h, yl.ptr, xl_ptr, zl.ptr + 8 * H, beta_d.ptr,
                                                   pl2.ptr + 8 * H, qp_d.ptr, part_d.ptr) == 0

    t_dir = U.time_launches(direction, args.kernel_reps)
    dir_bytes = n ** 3 * 8 + N * 8 * 4   # coefficient; p_old, z read, p_new, q written
    del zl, pl2, beta_d, qp_d, part_d
    log(f"[bench] CG direction + operator f64 (matrix-free): {t_dir*1e3:.3f} ms  {dir_bytes/t_dir/1e9:.0f} GB/s")
    log(f"[bench] lattice surface SpMV f64 (symmetric DIA): {t_lat*1e3:.3f} ms  {lat_bytes/t_lat/1e9:.0f} GB/s")
    smooth_gbs = smooth_bytes / t_smooth / 1e9
    log(f"[bench] SpMV f64 ({layout}): {t_spmv*1e3:.3f} ms  {spmv_gbs:.0f} GB/s; "
        f"DIA Jacobi step f32: {t_smooth*1e3:.3f} ms  {smooth_gbs:.0f} GB/s")

    # the same FP32-preconditioned solve with the stored top-level operator
    # (LDS-staged symmetric diagonals) instead of the matrix-free one
    g_st = U.GeometricMultiGrid(disc, base_level=args.base_level, nu1=2, nu2=2, cycle=1, omega=0.66,
                                precision="f32", matrix_free=False)
    S_st = U.CG(g_st, max_steps=100, reduction=1e-10)
    S_st.init(A)
    x_st = U.Vector(N)

    def solve_stored():
        x_st.set(0.0)
        if not S_st.apply(x_st, b):
            raise RuntimeError("stored-operator solve did not converge")
    t32_stored, _ = _timed(U, solve_stored, args.steps, warmup=args.warmup)
    t_smooth_stored = U.time_launches(lambda: g_st.smooth_level(refs, 1), args.kernel_reps)
    del g_st, S_st, x_st
    log(f"[bench] stored top-level operator: solve {t32_stored*1e3:.2f} ms, smoother {t_smooth_stored*1e3:.3f} ms")

    t32 = float(np.mean(times["f32"]))
    t64 = float(np.mean(times["f64"]))
    # launches per FP32-preconditioned solve, one cycle per CG step plus the
    # initial one: of the 4 top-level smoothing steps of a V(2,2) cycle the
    # lattice path runs 1 as the plain fused Jacobi step timed here (the first
    # one reads the defect in its own kernel mode, the last full one is fused
    # with the correction update that ends the cycle), the surface path all
    # 4; the CG operator once per step, fused with the direction update after
    # the first step on the lattice path
    lattice = all(bool(solvers[p][1].used_lattice_path()) for p in solvers)
    smooth_launches = (1 if lattice else 4) * (iters["f32"] + 1)

    def roofline(kernel, nbytes, t, launches):
        return {"kernel": kernel, "bound": "hbm", "achieved": nbytes / t / 1e9, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": nbytes / t / 1e9 / HBM_PEAK_GBS, "algorithmic_bytes_per_launch": nbytes,
                "launch_ms": t * 1e3, "launches_per_solve": launches, "traffic": measured_traffic(kernel),
                "measured_stream_triad_GBps": triad_gbs}

    cands = [roofline("k_q1mf_z<float,0>", smooth_bytes, t_smooth, smooth_launches)]
    if lattice:
        cands.append(roofline("k_q1mf_z<double,5>", dir_bytes, t_dir, max(iters["f32"] - 1, 0)))
        cands.append(roofline("k_q1mf_z<double,2>", lat_bytes, t_lat, 1))
    else:
        op_kernel = {"dict": "k_dict_spmv", "sell": "k_sell_spmv<double,double>"}.get(
            layout, f"k_spmv_csr<double,double,{best_lanes}>")
        cands.append(roofline(op_kernel, spmv_bytes, t_spmv, iters["f32"]))
    # the kernel with the largest share of the solve
    roof = max(cands, key=lambda r: r["launch_ms"] * r["launches_per_solve"])
    roof_smoother = cands[0]

    secondary = None
    if not args.skip_extra:
        secondary = {}
        log("[bench] secondary: 2D FV1 Poisson")
        secondary["poisson_2d_fv1"] = bench_poisson_2d(U, args.seed, args.steps)
        for r_adj in (6, refs):
            log(f"[bench] secondary: adjoint gradient {2 ** r_adj}^3")
            secondary[f"adjoint_q1_{2 ** r_adj}"] = bench_adjoint(U, r_adj, args.seed, max(1, args.steps // 2))
        log("[bench] secondary: tetrahedra 96^3 x 6")
        secondary["tet_fv1_upwind_96"] = bench_tet(U, 5, args.seed, max(1, args.steps // 2))
        log(f"[bench] secondary: Galerkin coarse operators at {n}^3")
        secondary[f"galerkin_q1_{n}"] = bench_galerkin(U, disc, A, b, args.base_level, args.steps)
        log("[bench] secondary: Q2 element matrices")
        secondary["q2_hex_element_matrices_64"] = bench_q2_elements(U, args.seed, args.kernel_reps)

    cpu = None
    if not args.skip_cpu:
        log(f"[bench] CPU baseline (oracle, 1 core) on {2**args.cpu_refs}^3")
        cpu = cpu_baseline(args.cpu_refs, args.seed, iters["f64"], refs)
        log(f"[bench] CPU baseline est. {cpu['value']:.1f} s")

    out = {
        "metric": METRIC,
        "value": t32,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t32 * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 outer CG / f32 GMG levels",
        "data": "synthetic (seeded theta field, unit cube)",
        "config": {
            "workload": f"{n}^3 Q1 FE -div(exp(theta) grad u)=1, CG + GMG V(2,2) Jacobi 0.66, base LU 4^3, tol 1e-10",
            "dofs": N,
            "nnz": nnz,
            "fp32_precond_s": t32,
            "fp64_precond_s": t64,
            "speedup_fp32_vs_fp64_precond": t64 / t32,
            "cg_steps_fp32_precond": iters["f32"],
            "cg_steps_fp64_precond": iters["f64"],
            "rel_l2_diff_fp32_vs_fp64_precond_solution": err,
            "final_rel_residual_fp32": final_rel_res,
            "dofs_per_s_fp32": N / t32,
            "assembly_s": t_asm,
            "assembly_elements_per_s": n ** 3 / t_asm,
            "reassembly_s": t_reasm,
            "reassembly_elements_per_s": n ** 3 / t_reasm,
            "gmg_setup_s": setup_s,
            "mesh_and_numbering_s": t_mesh,
            "mesh_and_numbering_host_s": t_mesh_host,
            "mesh_numbering_device_equals_host": same_numbering,
            "spmv_layout": layout,
            "spmv_f64_ms": t_spmv * 1e3,
            "spmv_dictionary_ms": t_dict * 1e3,
            "spmv_csr_direct_ms": t_csr * 1e3,
            "spmv_sliced_ell_ms": t_sell * 1e3,
            "spmv_sliced_ell_slices_in_dof_order_ms": t_sell_dof_order * 1e3,
            "spmv_csr_lanes": best_lanes,
            "spmv_f64_lex_ordering_ms": t_spmv_lex * 1e3,
            "spmv_f64_lex_ordering_GBps": spmv_bytes / t_spmv_lex / 1e9,
            "spmv_csr_lanes_sweep_ms": {str(k): v * 1e3 for k, v in sweep.items()},
            "spmv_f64_GBps": spmv_gbs,
            "spmv_f64_frac_of_8TBps": spmv_gbs / HBM_PEAK_GBS,
            "measured_stream_triad_GBps": triad_gbs,
            "spmv_f64_lex_ordering_frac_of_stream_triad": spmv_bytes / t_spmv_lex / 1e9 / triad_gbs,
            "cg_lattice_path": {p: bool(solvers[p][1].used_lattice_path()) for p in solvers},
            "lattice_spmv_f64_matrix_free_ms": t_lat * 1e3,
            "lattice_spmv_f64_matrix_free_GBps": lat_bytes / t_lat / 1e9,
            "lattice_spmv_algorithmic_bytes": lat_bytes,
            "smoother_matrix_free_f32_ms": t_smooth * 1e3,
            "smoother_matrix_free_f32_GBps": smooth_gbs,
            "fp32_precond_stored_top_operator_s": t32_stored,
            "fp32_precond_nonzero_start_s": float(np.mean(t_nz)),
            "cg_steps_fp32_precond_nonzero_start": iters_nz,
            "smoother_stored_sym_dia_f32_ms": t_smooth_stored * 1e3,
            "smoother_stored_sym_dia_f32_GBps": N * 4 * (nu_stored + 5) / t_smooth_stored / 1e9,
            "secondary": secondary,
        },
        "roofline": roof,
        "roofline_fp32_smoother": roof_smoother,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
