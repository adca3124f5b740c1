"""Benchmark: DOF·steps/s of the FEM time stepper on one MI355X.

Default workload: Gray-Scott on an icosphere refined 6 times (40,962
vertices, 81,920 triangles, radius 1), two species, Du=2e-4, Dv=1e-4,
F=0.037, k=0.06, dt=0.5, u=1 and v=0 plus 50 seeded patches
(v = 0.25 + U(0, 0.05)), implicit diffusion and explicit reaction,
2,000 steps, FP32 CG with FP64 iterative refinement (tol 1e-10).  The
reaction uses lumped-mass weighting: the reference's nodal weighting
(CPUSolver.cpp:134,137) diverges at dt = 0.5 (u reaches -143 in one step).

A step is one implicit time step of both species; DOF = N x 2 species.
Prints one JSON line.  Extra keys: "roofline" (the persistent stepping
kernel against the 8 TB/s HBM peak), "cpu_baseline" (the FP32 port of the
reference CPU path, one core), "accuracy" (rel-L2 against FP64) and
"plain_fp64_gpu" (the plain FP64 GPU CG engine on the same workload).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "fea-visualizer_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
# HBM-side bytes measured with rocprofv3 FETCH_SIZE / WRITE_SIZE (calibrated
# against a plain copy kernel) for the dominant kernel of a workload; the
# summaries with their commands are in profiles/.
PROFILED_TRAFFIC = {
    "grayscott": {"bytes_per_step": 3.9685e7, "source": "profiles/k_step_grayscott_icosphere6.md"},
    "spmv": {"bytes_per_launch": 1.2087e9, "source": "profiles/k_spmv_stream_jittered4096.md"},
}
METRIC = "DOF·steps/s at rel-L2≤1e-8 vs FP64"


def gs_initial(mesh, n, seed_centres=0, seed_noise=1, npatch=50, radius=0.05):
    """u = 1, v = 0 plus npatch geodesic discs (radius in radians on the unit
    sphere) with v = 0.25 + U(0, 0.05); centres uniform on the sphere."""
    X = mesh["xyz"].astype(np.float64)
    X = X / np.linalg.norm(X, axis=1, keepdims=True)
    rc = np.random.Generator(np.random.PCG64(seed_centres))
    rn = np.random.Generator(np.random.PCG64(seed_noise))
    c = rc.normal(size=(npatch, 3))
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    inside = (np.arccos(np.clip(X @ c.T, -1.0, 1.0)) < radius).any(1)
    noise = rn.uniform(0.0, 0.05, size=X.shape[0])
    u = np.ones(n)
    v = np.zeros(n)
    v[inside] = 0.25 + noise[inside]
    return u, v


PREDICTORS = {"linear": 0, "previous": 1, "quadratic": 2, "cubic": 3, "quartic": 4}  # FEM_PREDICTOR_*


def step_bytes(n, nnz, k, kv, iters, sweeps, steps, ti=4, bicgstab=False, fused=False):
    """Algorithmic HBM bytes of a step sequence (int32 indices, inner vectors
    and values of ti bytes, FP64 refinement), every vector element and matrix
    entry touched once per phase (cache-perfect gathers).  A pipelined-CG
    iteration moves 11 vectors and the matrix once; a BiCGSTAB iteration 17
    vectors (p: r, p_old, v_old, rh read, p, v written; s: r, v read, s, t
    written; x: xi, p, s, t, rh read, xi, r written) and the matrix twice.
    fused: first-order equations on the persistent kernel, whose residual
    pass also reads M and the history combination and writes M x and S x0,
    so the update + right-hand side phase reads no matrix."""
    if bicgstab:
        it = 2 * nnz * (4 + ti * kv) + n * (17 * k * ti + ti * kv)
    else:
        it = nnz * (4 + ti * kv) + n * (11 * k * ti + ti * kv)     # one Krylov iteration
    sw = nnz * (4 + 8 * kv) + n * (3 * k * 8 + k * ti + ti * kv)  # FP64 residual
    st = nnz * (12 + 8 * kv) + n * (36 * k + 4 * kv) + n * 32 * k  # RHS + update
    if fused:
        sw += nnz * 8 + n * 3 * k * 8                                # + M, hvec read; xnext, hpre written
        st += -nnz * (12 + 8 * kv) + n * k * 8                       # no matrix; + hvec written
    return iters * it + sweeps * sw + steps * st


def cpu_info():
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def run_cpu_baseline(mesh, u0, v0, prm, nsteps):
    import oracle
    s = oracle.System(mesh, bc=1, dtype=np.float32)
    prep = oracle.port_prepare(s)  # one-off assembly/storage, outside the timed region
    t = time.perf_counter()
    its, _, _ = oracle.port_run(prep, oracle.GRAYSCOTT, prm, nsteps, u0, v0)
    dt = time.perf_counter() - t
    model, ncpu = cpu_info()
    return {
        "value": 2 * s.n * nsteps / dt,
        "unit": "DOF·steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{nsteps} steps of the same workload, FP32, CSC column-scatter SpMV, system "
                  f"re-formed every step, Eigen-default Jacobi-CG (tol=FLT_EPSILON, x0=0); "
                  f"{its} CG iterations; gcc -O3 -DNDEBUG, no -march/OpenMP; host: {model} ({ncpu} CPUs)",
    }


def spmv_measure(side=4096, reps=50):
    """Stand-alone CSR SpMV/SpMM (the product inside every CG iteration) on a
    synthetic CSR larger than the 256 MB Infinity Cache: the jittered
    side x side lattice (Delaunay split, Dirichlet), S = 1000 M + K."""
    import torch
    from femhip import FEMContext, Operator, _lib, meshes
    mesh = meshes.jittered_square(side, seed=2)
    ctx = FEMContext(mesh, 0)
    n = ctx.num_unknowns()
    nnz = ctx.info().nnz
    out = {"matrix": f"jittered {side}x{side} lattice, Delaunay split, Dirichlet: N={n}, nnz={nnz}, "
                     f"int32 columns, S = 1000 M + K",
           "bytes_formula": "(s_val + 4) nnz + 4 (N + 1) + 2 k N s_val", "results": []}
    for prec, vb, tdt in ((_lib.FEM_PREC_FP32_IR, 4, torch.float32), (_lib.FEM_PREC_FP64, 8, torch.float64)):
        op = Operator(ctx, (1000.0, 1.0, 0.0), prec)
        for k in (1, 16):
            if vb == 8 and k == 16:
                continue
            x = torch.randn(n * k, dtype=tdt, device="cuda")
            y = torch.empty_like(x)
            torch.cuda.synchronize()
            sched, _ = op.autotune(k, x.data_ptr(), y.data_ptr(), reps=5)
            ms = op.time(k, x.data_ptr(), y.data_ptr(), reps=reps)
            b = (vb + 4) * nnz + 4 * (n + 1) + 2 * k * n * vb
            gbs = b / (ms * 1e-3) / 1e9
            out["results"].append({"values": "f32" if vb == 4 else "f64", "nrhs": k, "schedule": sched,
                                   "ms": ms, "bytes": b, "achieved_GBs": gbs, "frac": gbs / HBM_PEAK_GBS})
            del x, y
        op.close()
    ctx.close()
    return out


def adjoint_measure(side=512, nsteps=100, every=10, fd_dirs=3):
    """Differentiable advection-diffusion: jittered side^2 lattice, rotating
    velocity (-(y-1/2), x-1/2) at element centroids (unnormalised), per-element
    diffusivity 1e-3 exp(0.3 z), dt = 2e-3, nsteps steps, L = |u_T - y|^2;
    adjoint with a checkpoint every `every` steps.  Timed: one forward and one
    backward of the FP32 + FP64-refinement solver (wall clock, inputs on the
    host as the C ABI takes them).  Accuracy: gradients against the FP64
    solver at tol 1e-13, and central differences (h = 1e-6, FP64) along
    random directions."""
    from femhip import Adjoint, AdvectionDiffusionParameters, FEMContext, SolverOptions, _lib, meshes
    mesh = meshes.jittered_square(side, seed=6)
    cen = mesh["xyz"].astype(np.float64)[mesh["tri"]].mean(axis=1)
    ve = np.stack([-(cen[:, 1] - 0.5), cen[:, 0] - 0.5, np.zeros(len(cen))], axis=1)
    rk = np.random.Generator(np.random.PCG64(7))
    kappa = 1e-3 * np.exp(0.3 * rk.standard_normal(mesh["tri"].shape[0]))
    ctx = FEMContext(mesh, 0, velocity_e=ve, normalize_velocity=False, kappa_e=kappa)
    n = ctx.num_unknowns()
    X = mesh["xyz"].astype(np.float64)
    keep = ctx.idx_map >= 0
    u0 = np.exp(-((X[:, 0] - 0.3) ** 2 + (X[:, 1] - 0.5) ** 2) / (2 * 0.05 ** 2))[keep]
    rt = np.random.Generator(np.random.PCG64(8))
    cc = rt.uniform(0.2, 0.8, 2)
    target = np.exp(-((X[:, 0] - cc[0]) ** 2 + (X[:, 1] - cc[1]) ** 2) / (2 * 0.05 ** 2))[keep]
    prm = AdvectionDiffusionParameters(2e-3, 1.0)
    res = {}
    timing = None
    for name, opts in (("fp32ir", SolverOptions(tol=1e-10)),
                       ("fp64", SolverOptions(precision=_lib.FEM_PREC_FP64, tol=1e-13))):
        adj = Adjoint(ctx, prm, opts)
        adj.forward(u0, nsteps, every)  # warm
        t = time.perf_counter()
        uT = adj.forward(u0, nsteps, every)
        tf = time.perf_counter() - t
        t = time.perf_counter()
        gu, gk = adj.backward(2.0 * (uT - target))
        tb = time.perf_counter() - t
        res[name] = (uT, gu, gk, adj)
        if name == "fp32ir":
            timing = {"forward_ms": 1e3 * tf, "backward_ms": 1e3 * tb, "adjoint_over_forward": tb / tf,
                      "forward_iterations": int(adj.forward_stats.iterations),
                      "backward_iterations": int(adj.backward_stats.iterations),
                      "forward_kernel_ms": adj.forward_stats.ms}

    def rel(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    uT64, gu64, gk64, adj64 = res["fp64"]
    out = {"workload": f"advection-diffusion, jittered {side}x{side} lattice (N={n}, {mesh['tri'].shape[0]} "
                       f"triangles), rotating velocity, kappa_e = 1e-3 exp(0.3 z), dt=2e-3, {nsteps} steps, "
                       f"checkpoint every {every}",
           **timing,
           "rel_l2_vs_fp64": {"u_T": rel(res["fp32ir"][0], uT64), "grad_u0": rel(res["fp32ir"][1], gu64),
                              "grad_kappa": rel(res["fp32ir"][2], gk64)}}
    # central differences of the FP64 forward along random directions
    rd = np.random.Generator(np.random.PCG64(9))
    h = 1e-6
    fd = []

    def loss(u, k=None):
        if k is not None:
            ctx.assemble_matrices(velocity_e=ve, normalize_velocity=False, kappa_e=k)
        v = adj64.forward(u, nsteps, every)
        return float(np.sum((v - target) ** 2))
    for _ in range(fd_dirs):
        d = rd.standard_normal(n)
        f = (loss(u0 + h * d) - loss(u0 - h * d)) / (2 * h)
        fd.append({"wrt": "u0", "fd": f, "adjoint": float(gu64 @ d), "rel": abs(f - gu64 @ d) / abs(f)})
        dk = rd.standard_normal(kappa.shape[0]) * kappa
        f = (loss(u0, kappa + h * dk) - loss(u0, kappa - h * dk)) / (2 * h)
        fd.append({"wrt": "kappa", "fd": f, "adjoint": float(gk64 @ dk), "rel": abs(f - gk64 @ dk) / abs(f)})
    ctx.assemble_matrices(velocity_e=ve, normalize_velocity=False, kappa_e=kappa)
    out["central_differences_fp64"] = fd
    for v in res.values():
        v[3].close()
    ctx.close()
    return out


def log(*a):
    """Progress on stderr (long workloads)."""
    print("[bench]", *a, file=sys.stderr, flush=True)


def contract_line(metric, value, unit, steps, warmup, ms_per_step, dtype, config, **extra):
    out = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
           "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": dtype, "data": "synthetic", "config": config}
    out.update(extra)
    return out


def gauss_sum(X, centres, sigmas):
    return np.exp(-((X[:, None, 0] - centres[None, :, 0]) ** 2 + (X[:, None, 1] - centres[None, :, 1]) ** 2)
                  / (2 * np.asarray(sigmas)[None, :] ** 2))


def run_heat_c1(args):
    """Heat equation on the structured 64 x 64 unit square (8,192 triangles,
    4,225 vertices, 3,969 unknowns), backward Euler dt = 1e-3, conductivity
    1, zero Dirichlet boundary, u0 = exp(-|x - (1/2,1/2)|^2 / 0.01), 100
    steps, FP64 CG to tol 1e-10 (the persistent kernel in FP64)."""
    from femhip import Equation, FEMContext, HeatParameters, HIPSolver, SolverOptions, _lib, meshes
    steps = args.steps or 100
    mesh = meshes.structured_square(64)
    ctx = FEMContext(mesh, 0)
    n, nnz = ctx.num_unknowns(), ctx.info().nnz
    X = mesh["xyz"].astype(np.float64)
    u0 = np.exp(-((X[:, 0] - 0.5) ** 2 + (X[:, 1] - 0.5) ** 2) / 0.01)[ctx.idx_map >= 0]
    ctx.equation = Equation.Heat
    ctx.parameters[Equation.Heat] = HeatParameters(1e-3, 1.0)
    sol = HIPSolver(ctx, SolverOptions(precision=_lib.FEM_PREC_FP64, tol=1e-10))
    sol.set_state_unknowns(_lib.FEM_FIELD_U, u0)
    tuned = sol.autotune(20) if not args.no_autotune else None
    sol.advance(args.warmup)
    sol.set_state_unknowns(_lib.FEM_FIELD_U, u0)
    sol.sync()
    t = time.perf_counter()
    status = sol.advance(steps)
    sol.sync()
    wall = time.perf_counter() - t
    st = sol.last_stats
    ach = step_bytes(n, nnz, 1, 1, st.iterations, st.sweeps, steps, ti=8, fused=True) / (st.ms * 1e-3) / 1e9
    ref = HIPSolver(ctx, SolverOptions(precision=_lib.FEM_PREC_FP64, tol=1e-13))
    ref.set_state_unknowns(_lib.FEM_FIELD_U, u0)
    ref.advance(steps)
    uu, ur = sol.get_state_unknowns(), ref.get_state_unknowns()
    out = contract_line(METRIC, n * steps / wall, "DOF·steps/s", steps, args.warmup, 1e3 * wall / steps, "f64",
                        {"workload": f"heat, structured 64x64 unit square (N={n}, nnz={nnz}), dt=1e-3, "
                                     f"conductivity 1, Dirichlet, Gaussian u0, FP64 CG tol 1e-10",
                         "global_batch": 1,
                         "parallelism": "single GPU, " + ("stream engine (autotuned)"
                                                          if sol.options.engine == _lib.FEM_ENGINE_STREAM
                                                          else "persistent cooperative kernel")},
                        status=int(status),
                        solver={"iterations": int(st.iterations), "sweeps": int(st.sweeps), "autotune": tuned,
                                "kernel_ms": st.ms},
                        roofline={"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": ach / HBM_PEAK_GBS, "traffic": None, "kernel": "k_step<1,double>",
                                  "note": "1.6 MB working set: cache-resident, latency-bound"},
                        accuracy={"trajectory_rel_l2_final": float(np.linalg.norm(uu - ur) / np.linalg.norm(ur)),
                                  "reference": "FP64 GPU CG, tol 1e-13"})
    if not args.no_cpu:
        import oracle
        s = oracle.System(mesh, bc=0, dtype=np.float32)
        prep = oracle.port_prepare(s)
        t = time.perf_counter()
        its, _, _ = oracle.port_run(prep, oracle.HEAT, [1e-3, 1.0], steps, u0)
        dt_ = time.perf_counter() - t
        model, ncpu = cpu_info()
        out["cpu_baseline"] = {"value": n * steps / dt_, "unit": "DOF·steps/s", "cores": 1, "kind": "port",
                               "sample": f"all {steps} steps, FP32 Eigen-default Jacobi-CG port ({its} CG "
                                         f"iterations), gcc -O3 -DNDEBUG; host: {model} ({ncpu} CPUs)"}
    return out


def run_wave_c3(args):
    """Wave equation on the jittered 1024 x 1024 lattice (interior vertices
    moved by U(+-h/4), seed 2, Delaunay split; 1,050,625 vertices, 2,097,152
    triangles), c = 1, Newmark beta = 1/4, gamma = 1/2, dt = 2e-4, u0 = sum of
    8 Gaussian bumps (sigma 0.02, centres U[0.1,0.9]^2, seed 3), v0 = 0, a0
    consistent (M a0 = -c^2 K u0), 500 steps, FP32 CG + FP64 refinement."""
    from femhip import Equation, FEMContext, HIPSolver, NewmarkParameters, SolverOptions, WaveParameters, _lib, meshes
    steps = args.steps or 500
    mesh = meshes.jittered_square(1024, seed=2)
    ctx = FEMContext(mesh, 0)
    n, nnz = ctx.num_unknowns(), ctx.info().nnz
    rng = np.random.Generator(np.random.PCG64(3))
    cen = rng.uniform(0.1, 0.9, size=(8, 2))
    X = mesh["xyz"].astype(np.float64)[ctx.idx_map >= 0]
    u0 = gauss_sum(X, cen, [0.02] * 8).sum(axis=1)
    ctx.equation = Equation.Wave_Newmark
    ctx.parameters[Equation.Wave_Newmark] = NewmarkParameters(2e-4, 1.0, 0.25, 0.5)

    def fresh(**kw):
        sol = HIPSolver(ctx, SolverOptions(**kw))
        sol.set_state_unknowns(_lib.FEM_FIELD_U, u0)
        sol.init_newmark_acceleration()
        return sol
    if args.engine == "stream":
        sol = fresh(tol=1e-10, engine=_lib.FEM_ENGINE_STREAM)
        tuned = None
    else:
        sol = fresh(tol=1e-10)
        tuned = sol.autotune(5) if not args.no_autotune else None
    a0 = sol.get_state_unknowns(_lib.FEM_FIELD_A)
    sol.advance(args.warmup)
    sol.set_state_unknowns(_lib.FEM_FIELD_U, u0)
    sol.set_state_unknowns(_lib.FEM_FIELD_V, np.zeros(n))
    sol.set_state_unknowns(_lib.FEM_FIELD_A, a0)
    sol.sync()
    t = time.perf_counter()
    status = sol.advance(steps)
    sol.sync()
    wall = time.perf_counter() - t
    st = sol.last_stats
    ach = step_bytes(n, nnz, 1, 1, st.iterations, st.sweeps, steps) / (st.ms * 1e-3) / 1e9
    ref = fresh(precision=_lib.FEM_PREC_FP64, tol=1e-13)
    ref.advance(steps)
    uu, ur = sol.get_state_unknowns(), ref.get_state_unknowns()
    out = contract_line(METRIC, n * steps / wall, "DOF·steps/s", steps, args.warmup, 1e3 * wall / steps,
                        "f32+f64 (FP32 CG, FP64 refinement residual and dots)",
                        {"workload": f"Newmark wave, jittered 1024x1024 lattice (N={n}, nnz={nnz}), c=1, "
                                     f"beta=1/4 gamma=1/2, dt=2e-4, 8 Gaussian bumps, tol 1e-10",
                         "global_batch": 1,
                         "parallelism": "single GPU, " + ("stream engine (autotuned)"
                                                          if sol.options.engine == _lib.FEM_ENGINE_STREAM
                                                          else "persistent cooperative kernel")},
                        status=int(status),
                        solver={"iterations": int(st.iterations), "sweeps": int(st.sweeps), "autotune": tuned,
                                "kernel_ms": st.ms, "iters_per_step": st.iterations / steps},
                        roofline={"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": ach / HBM_PEAK_GBS, "traffic": None, "kernel": "k_step<1,float>",
                                  "note": "~200 MB working set: mostly Infinity-Cache resident"},
                        accuracy={"trajectory_rel_l2_final": float(np.linalg.norm(uu - ur) / np.linalg.norm(ur)),
                                  "reference": "FP64 GPU CG, tol 1e-13"})
    # the reference's own wave scheme (CPUSolver.cpp:108-121) on the same mesh
    ctx.equation = Equation.Wave
    ctx.parameters[Equation.Wave] = WaveParameters(2e-4, 1.0)
    w = HIPSolver(ctx, SolverOptions(tol=1e-10))
    w.set_state_unknowns(_lib.FEM_FIELD_U, u0)
    w.advance(2)
    w.set_state_unknowns(_lib.FEM_FIELD_U, u0)
    w.set_state_unknowns(_lib.FEM_FIELD_V, np.zeros(n))
    w.sync()
    t = time.perf_counter()
    w.advance(steps)
    w.sync()
    out["reference_wave_scheme"] = {"value": n * steps / (time.perf_counter() - t), "unit": "DOF·steps/s",
                                    "iterations": int(w.last_stats.iterations)}
    if not args.no_cpu:
        import oracle
        s = oracle.System(mesh, bc=0, dtype=np.float32)
        prep = oracle.port_prepare(s)
        ns = 3
        t = time.perf_counter()
        its, _, _ = oracle.port_run(prep, oracle.WAVE, [2e-4, 1.0], ns, u0, np.zeros(n))
        dt_ = time.perf_counter() - t
        model, ncpu = cpu_info()
        out["cpu_baseline"] = {"value": n * ns / dt_, "unit": "DOF·steps/s", "cores": 1, "kind": "port",
                               "sample": f"{ns} steps of the reference's wave scheme (CPUSolver.cpp:108-121; "
                                         f"Newmark has no reference counterpart) on the same mesh, FP32 "
                                         f"Eigen-default CG port ({its} iterations); host: {model} ({ncpu} CPUs)"}
    return out


def c4_specs(ninst=256):
    """Parameters of the batched instances, without their meshes: side =
    round(exp(U(ln 32, ln 512))), kappa = exp(U(ln 1e-4, ln 1e-1)), v = U(0,1)
    (cos t, sin t, 0) (seed 4); 16 Gaussian bumps per mesh, centre U[0.2,0.8]^2
    and sigma U(0.02, 0.1) (seed 5).  Instance i's mesh is the jittered
    Delaunay-split lattice of that side with seed 1000 + i."""
    r4 = np.random.Generator(np.random.PCG64(4))
    r5 = np.random.Generator(np.random.PCG64(5))
    out = []
    for i in range(ninst):
        side = int(round(np.exp(r4.uniform(np.log(32), np.log(512)))))
        kap = float(np.exp(r4.uniform(np.log(1e-4), np.log(1e-1))))
        sp, th = r4.uniform(0, 1), r4.uniform(0, 2 * np.pi)
        cen = r5.uniform(0.2, 0.8, size=(16, 2))
        sig = r5.uniform(0.02, 0.1, size=16)
        out.append((side, kap, (sp * np.cos(th), sp * np.sin(th), 0.0), cen, sig))
    return out


def c4_instances(ninst=256, which=None):
    """The batched instances (c4_specs) with their meshes: (mesh, kappa,
    velocity, centres, sigmas) per instance; `which` selects indices."""
    from femhip import meshes
    specs = c4_specs(ninst)
    idx = range(ninst) if which is None else which
    return [(meshes.jittered_square(specs[i][0], seed=1000 + i),) + specs[i][1:] for i in idx]


def run_batched_c4(args):
    """256 advection-diffusion instances of uneven size (about 1.3e7
    unknowns in all), 16 right-hand sides each, backward Euler dt = 1e-3,
    unnormalised constant velocity per instance; one persistent launch for
    all instances and steps, FP32 BiCGSTAB + FP64 refinement (tol 1e-10)."""
    from femhip import AdvectionDiffusionParameters, Batch, FEMContext, SolverOptions, _lib
    steps = args.steps or 200
    inst = c4_instances(args.instances)
    log("batched: meshes generated")
    ctxs, fields, prms = [], [], []
    tot_n = tot_nnz = 0
    for m, kap, vel, cen, sig in inst:
        c = FEMContext(m, 0, velocity=vel, normalize_velocity=False)
        tot_n += c.num_unknowns()
        tot_nnz += c.info().nnz
        X = m["xyz"].astype(np.float64)[c.idx_map >= 0]
        ctxs.append(c)
        fields.append(np.ascontiguousarray(gauss_sum(X, cen, sig)))
        prms.append(AdvectionDiffusionParameters(1e-3, kap))

    def make(**kw):
        kw.setdefault("predictor", PREDICTORS[args.predictor])
        kw.setdefault("inner_tol", args.inner_tol or 1e-4)
        b = Batch(ctxs, SolverOptions(nrhs=16, **kw))
        for i in range(len(ctxs)):
            b.set_equation(i, prms[i])
            b.set_state_unknowns(i, _lib.FEM_FIELD_U, fields[i])
        return b
    log("batched: contexts assembled, N =", tot_n)
    if args.engine == "stream":
        b = make(tol=1e-10, engine=_lib.FEM_ENGINE_STREAM)
        tuned = None
    else:
        b = make(tol=1e-10)
        tuned = b.autotune(2) if not args.no_autotune else None
    log("batched: autotuned", tuned)
    engine = "stream" if b.options.engine == _lib.FEM_ENGINE_STREAM else "persistent"
    for i in range(len(ctxs)):
        b.set_state_unknowns(i, _lib.FEM_FIELD_U, fields[i])
    b.sync()
    t = time.perf_counter()
    status = b.advance(steps)
    b.sync()
    wall = time.perf_counter() - t
    st = b.last_stats
    log("batched: timed run done", wall)
    dof = 16 * tot_n
    # algorithmic bytes instance by instance: larger meshes iterate more
    alg = 0.0
    for i, c in enumerate(ctxs):
        si = b.instance_stats(i)
        alg += step_bytes(c.num_unknowns(), c.info().nnz, 16, 1, si.iterations, si.sweeps, steps, bicgstab=True)
    ach = alg / (st.ms * 1e-3) / 1e9
    finite = all(np.all(np.isfinite(b.get_state_unknowns(i))) for i in range(len(ctxs)))
    out = contract_line(METRIC, dof * steps / wall, "DOF·steps/s", steps, 0, 1e3 * wall / steps,
                        "f32+f64 (FP32 BiCGSTAB, FP64 refinement residual and dots)",
                        {"workload": f"batched advection-diffusion, {len(ctxs)} jittered meshes of side "
                                     f"log-U[32,512] (sum N={tot_n}, sum nnz={tot_nnz}), 16 RHS each, dt=1e-3, "
                                     f"tol 1e-10", "global_batch": len(ctxs) * 16,
                         "parallelism": f"single GPU, all instances per launch ({engine} engine"
                                         f"{', autotuned' if tuned is not None else ''})"},
                        status=int(status), all_finite=bool(finite),
                        solver={"iterations": int(st.iterations), "sweeps": int(st.sweeps), "autotune": tuned,
                                "kernel_ms": st.ms},
                        roofline={"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": ach / HBM_PEAK_GBS, "traffic": None,
                                  "kernel": "k_sphase<16,float,*> (stream engine phases)" if engine == "stream"
                                  else "k_step<16,float,BiCGSTAB>"})
    if not args.no_accuracy:
        ref = make(precision=_lib.FEM_PREC_FP64, tol=1e-13, engine=_lib.FEM_ENGINE_STREAM)
        for _ in range(steps // 10):
            ref.advance(10)
            log("batched: FP64 reference", ref.last_stats.steps)
        if steps % 10:
            ref.advance(steps % 10)
        errs = []
        for i in range(len(ctxs)):
            a, r = b.get_state_unknowns(i), ref.get_state_unknowns(i)
            errs.append(float(np.linalg.norm(a - r) / np.linalg.norm(r)))
        # one step teacher-forced from the FP64 state: both solvers start from it
        one = []
        for i in range(len(ctxs)):
            b.set_state_unknowns(i, _lib.FEM_FIELD_U, ref.get_state_unknowns(i))
        b.advance(1)
        ref.advance(1)
        for i in range(len(ctxs)):
            a, r = b.get_state_unknowns(i), ref.get_state_unknowns(i)
            one.append(float(np.linalg.norm(a - r) / np.linalg.norm(r)))
        out["accuracy"] = {"one_step_rel_l2_max_over_instances": float(np.max(one)),
                           "trajectory_rel_l2_final_max_over_instances": float(np.max(errs)),
                           "reference": "FP64 GPU (stream engine), tol 1e-13",
                           "gate_1e-8_one_step": bool(np.max(one) <= 1e-8)}
        ref.close()
    b.close()
    if not args.no_plain:
        # one step of each instance in turn on the plain FP64 engine, for up to
        # ~30 s of solver time (a sample when the batch does not fit)
        done_dof, spent, count = 0, 0.0, 0
        for i in range(len(ctxs)):
            pl = Batch([ctxs[i]], SolverOptions(nrhs=16, engine=_lib.FEM_ENGINE_PLAIN,
                                                precision=_lib.FEM_PREC_FP64))
            pl.set_equation(0, prms[i])
            pl.set_state_unknowns(0, _lib.FEM_FIELD_U, fields[i])
            pl.sync()
            t = time.perf_counter()
            pl.advance(1)
            pl.sync()
            spent += time.perf_counter() - t
            done_dof += 16 * ctxs[i].num_unknowns()
            count += 1
            pl.close()
            log("batched: plain FP64 engine, instance", i, "cumulative s", round(spent, 2))
            if spent > 30.0:
                break
        pv = done_dof / spent
        out["plain_fp64_gpu"] = {"value": pv, "unit": "DOF·steps/s", "steps": 1,
                                 "sample": f"1 step of instances 0..{count - 1} ({done_dof} DOF), one right-hand "
                                           f"side at a time with host-read reduction scalars",
                                 "speedup_of_value": out["value"] / pv}
    if not args.no_cpu:
        # the reference solves advection-diffusion with Eigen's CG, which can
        # stagnate up to its 2N iteration cap on these nonsymmetric systems:
        # sample the smallest instance, one column at a time, for up to ~20 s
        import oracle
        order = np.argsort([c.num_unknowns() for c in ctxs])[:8]
        spent, dofs, its, used = 0.0, 0, 0, []
        for i0 in order:
            m, kap, vel, cen, sig = inst[i0]
            s = oracle.System(m, bc=0, velocity=vel, normalize=False, dtype=np.float32)
            prep = oracle.port_prepare(s)
            for j in range(16):
                t = time.perf_counter()
                k_, _, _ = oracle.port_run(prep, oracle.ADVDIFF, [1e-3, kap], 1, fields[i0][:, j])
                spent += time.perf_counter() - t
                its += int(k_)
                dofs += s.n
            used.append(int(i0))
            log("batched: CPU port instance", int(i0), "N", s.n, "s", round(spent, 2))
            if spent > 20.0:
                break
        model, ncpu = cpu_info()
        out["cpu_baseline"] = {"value": dofs / spent, "unit": "DOF·steps/s", "cores": 1, "kind": "port",
                               "sample": f"the {len(used)} smallest instances {used}, 16 columns x 1 step each "
                                         f"({dofs} DOF), FP32 Eigen-default CG port ({its} iterations; the reference "
                                         f"solves advection-diffusion with CG); host: {model} ({ncpu} CPUs)"}
    for c in ctxs:
        c.close()
    return out


def run_spmv(args):
    """The SpMV HBM-bandwidth line: FP32 single right-hand side on the 1.14 GB
    CSR (other precisions and column counts under "results")."""
    sp = spmv_measure(reps=args.steps or 50)
    head = sp["results"][0]
    return contract_line("SpMV HBM GB/s", head["achieved_GBs"], "GB/s", args.steps or 50, 1, head["ms"], "f32",
                         {"workload": sp["matrix"], "global_batch": 1, "parallelism": "single GPU"},
                         roofline={"bound": "hbm", "achieved": head["achieved_GBs"], "peak": HBM_PEAK_GBS,
                                   "unit": "GB/s", "frac": head["frac"],
                                   "traffic": PROFILED_TRAFFIC["spmv"]["bytes_per_launch"],
                                   "traffic_source": PROFILED_TRAFFIC["spmv"]["source"],
                                   "kernel": "k_spmv_stream<float,1>", "bytes_per_launch": head["bytes"]},
                         spmv=sp)


def run_adjoint(args):
    """The adjoint/forward cost line (C5 workload)."""
    ad = adjoint_measure(nsteps=args.steps or 100)
    out = contract_line("adjoint/forward cost", ad["adjoint_over_forward"], "ratio", args.steps or 100, 1,
                        ad["backward_ms"] / (args.steps or 100), "f32+f64 (FP32 BiCGSTAB, FP64 refinement)",
                        {"workload": ad["workload"], "global_batch": 1, "parallelism": "single GPU"},
                        adjoint=ad)
    out["higher_is_better"] = False
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="grayscott",
                    choices=["grayscott", "heat", "wave", "batched", "spmv", "adjoint"],
                    help="grayscott: the headline line; the others: one workload each")
    ap.add_argument("--instances", type=int, default=256, help="batched: number of meshes")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--level", type=int, default=6)
    ap.add_argument("--cpu-steps", type=int, default=2000)
    ap.add_argument("--plain-steps", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--team-size", type=int, default=0)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--predictor", choices=sorted(PREDICTORS), default="cubic",
                    help="grayscott, batched: initial guess of each solve")
    ap.add_argument("--engine", choices=["auto", "stream"], default="auto",
                    help="batched, wave: 'stream' runs the streaming engine without autotuning")
    ap.add_argument("--inner-tol", type=float, default=None,
                    help="FP32 sweep reduction (grayscott: 1e-5, batched: the library's 1e-4)")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--no-spmv", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run only a 1 GiB read + 1 GiB write elementwise kernel (counter calibration)")
    args = ap.parse_args()
    if args.calib_copy:
        import torch
        x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            torch.add(x, 1.0, out=y)
        torch.cuda.synchronize()
        print(json.dumps({"calib_copy_bytes_read": x.numel() * 4, "calib_copy_bytes_written": y.numel() * 4}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "one GPU only (--gpus 1): sharding across GPUs is designed, not implemented",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.workload == "spmv":
        import torch  # noqa: F401  (before the library, as above)
    if args.workload != "grayscott":
        fn = {"heat": run_heat_c1, "wave": run_wave_c3, "batched": run_batched_c4, "spmv": run_spmv,
              "adjoint": run_adjoint}[args.workload]
        print(json.dumps(fn(args)))
        return
    if args.steps is None:
        args.steps = 2000

    if not args.no_spmv:
        # torch (device buffers of the SpMV leg) must load its HIP runtime
        # before the library does: a process holds one runtime
        import torch  # noqa: F401
    from femhip import Equation, FEMContext, HIPSolver, ReactionDiffusionParameters, SolverOptions, _lib, meshes

    mesh = meshes.icosphere(args.level)
    ctx = FEMContext(mesh, 1)
    n = ctx.num_unknowns()
    nnz = ctx.info().nnz
    u0, v0 = gs_initial(mesh, n)
    prm = ReactionDiffusionParameters(0.5, 2e-4, 1e-4, 0.037, 0.06, _lib.FEM_REACTION_LUMPED)
    ctx.equation = Equation.Reaction_Diffusion
    ctx.parameters[Equation.Reaction_Diffusion] = prm

    def solver(**kw):
        o = SolverOptions(tol=args.tol, inner_tol=args.inner_tol or 1e-5, team_size=args.team_size,
                          predictor=PREDICTORS[args.predictor])
        for k_, v_ in kw.items():
            setattr(o, k_, v_)
        s = HIPSolver(ctx, o)
        s.set_state_unknowns(_lib.FEM_FIELD_U, u0)
        s.set_state_unknowns(_lib.FEM_FIELD_V, v0)
        return s

    log("grayscott: N =", n)
    sol = solver()
    tuned = None
    if args.team_size == 0 and not args.no_autotune:
        tuned = sol.autotune(20)  # untimed; state restored
    if args.warmup > 0:
        sol.advance(args.warmup)
    sol.set_state_unknowns(_lib.FEM_FIELD_U, u0)
    sol.set_state_unknowns(_lib.FEM_FIELD_V, v0)
    sol.sync()
    t0 = time.perf_counter()
    status = sol.advance(args.steps)
    sol.sync()
    wall = time.perf_counter() - t0
    log("grayscott: timed run", wall)
    st = sol.last_stats
    kernel_ms = st.ms
    dof_steps = 2 * n * args.steps
    value = dof_steps / wall
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "DOF·steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64 (FP32 CG, FP64 refinement residual and dots)",
        "data": "synthetic",
        "config": {
            "workload": f"gray-scott icosphere L{args.level} (N={n}, nnz={nnz}, 2 species), dt=0.5, "
                        f"Du=2e-4 Dv=1e-4 F=0.037 k=0.06, 50 patches, lumped reaction weighting, "
                        f"tol {args.tol:g}",
            "global_batch": 1,
            "parallelism": "single GPU, persistent cooperative kernel",
        },
        "status": int(status),
        "solver": {"iterations": int(st.iterations), "sweeps": int(st.sweeps),
                   "team_size": sol.options.team_size, "autotune": tuned,
                   "iters_per_solve": st.iterations / max(1, 2 * args.steps),
                   "final_residual": st.residual, "kernel_ms": kernel_ms},
    }
    # roofline of the persistent stepping kernel (one launch = the timed region)
    bytes_alg = step_bytes(n, nnz, 2, 2, st.iterations, st.sweeps, args.steps, fused=True)
    ach = bytes_alg / (kernel_ms * 1e-3) / 1e9
    tr = PROFILED_TRAFFIC["grayscott"]
    out["roofline"] = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": ach / HBM_PEAK_GBS, "traffic": tr["bytes_per_step"] * args.steps,
                       "traffic_source": tr["source"] + " (per-step bytes x steps)",
                       "kernel": "k_step<2,float>", "bytes_per_launch": bytes_alg,
                       "kernel_ms": kernel_ms}

    if not args.no_accuracy:
        # FP64 reference trajectory (persistent FP64 CG, tol 1e-13) and
        # teacher-forced one-step errors at 5 checkpoints
        ref = solver(precision=_lib.FEM_PREC_FP64, tol=1e-13)
        one = solver()
        chk = max(1, args.steps // 5)
        done = 0
        one_step = []
        while done < args.steps:
            m = min(chk, args.steps - done)
            ref.advance(m)
            done += m
            uu = ref.get_state_unknowns(_lib.FEM_FIELD_U)
            vv = ref.get_state_unknowns(_lib.FEM_FIELD_V)
            r1 = solver(precision=_lib.FEM_PREC_FP64, tol=1e-13)
            for s_ in (one, r1):
                s_.set_state_unknowns(_lib.FEM_FIELD_U, uu)
                s_.set_state_unknowns(_lib.FEM_FIELD_V, vv)
                s_.advance(1)
            e = max(np.linalg.norm(one.get_state_unknowns(f) - r1.get_state_unknowns(f)) /
                    np.linalg.norm(r1.get_state_unknowns(f)) for f in (_lib.FEM_FIELD_U, _lib.FEM_FIELD_V))
            one_step.append(e)
            r1.close()
        traj = max(np.linalg.norm(sol.get_state_unknowns(f) - ref.get_state_unknowns(f)) /
                   np.linalg.norm(ref.get_state_unknowns(f)) for f in (_lib.FEM_FIELD_U, _lib.FEM_FIELD_V))
        out["accuracy"] = {"one_step_rel_l2_max": max(one_step), "one_step_rel_l2": one_step,
                           "trajectory_rel_l2_final": traj, "reference": "FP64 GPU CG, tol 1e-13",
                           "gate_1e-8_one_step": bool(max(one_step) <= 1e-8)}
        ref.close()
        one.close()

    if not args.no_plain and args.plain_steps > 0:
        pl = solver(engine=_lib.FEM_ENGINE_PLAIN, precision=_lib.FEM_PREC_FP64)
        pl.sync()
        t = time.perf_counter()
        pl.advance(args.plain_steps)
        pl.sync()
        tw = time.perf_counter() - t
        pv = 2 * n * args.plain_steps / tw
        out["plain_fp64_gpu"] = {"value": pv, "unit": "DOF·steps/s", "steps": args.plain_steps,
                                 "iterations": int(pl.last_stats.iterations),
                                 "speedup_of_value": value / pv}
        pl.close()

    if not args.no_spmv:
        log("spmv")
        out["spmv"] = spmv_measure()
    if not args.no_adjoint:
        log("adjoint")
        out["adjoint"] = adjoint_measure()

    if not args.no_cpu and args.cpu_steps > 0:
        log("cpu baseline")
        out["cpu_baseline"] = run_cpu_baseline(mesh, u0, v0, [0.5, 2e-4, 1e-4, 0.037, 0.06, 1], args.cpu_steps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
