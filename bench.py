"""Benchmark: DOF·steps/s of the FEM time stepper on one MI355X.

Default workload (the headline line): Gray-Scott on an icosphere refined 6
times (40,962 vertices, 81,920 triangles, radius 1), two species, Du=2e-4,
Dv=1e-4, F=0.037, k=0.06, dt=0.5, u=1 and v=0 plus 50 seeded patches
(v = 0.25 + U(0, 0.05)), implicit diffusion and explicit reaction, FP32 CG
with FP64 iterative refinement (tol 1e-10).  The reaction uses lumped-mass
weighting: the reference's nodal weighting (CPUSolver.cpp:134,137) diverges
at dt = 0.5 (u reaches -143 in one step).

A step is one implicit time step of both species; DOF = N x 2 species.
`--warmup W` untimed steps run from the initial state, then `--steps K`
timed steps continue the same trajectory in one launch (steps W+1..W+K).
Prints one JSON line.  Extra keys:
  roofline          the persistent stepping kernel against the 8 TB/s HBM peak
  accuracy          one-step (teacher-forced) and trajectory rel-L2 against the
                    oracle (C restatement of the reference CPU path in FP64)
  fp64_same_engine  the same kernel and steps with FP64 Krylov vectors
  plain_fp64_gpu    the plain FP64 engine (host-read reduction scalars)
  frames            per-frame cost of the reference's call pattern through the
                    surface-space Solver boundary (set_state, 1 step, get_state)
  spmv, adjoint     the SpMV bandwidth and adjoint/forward cost legs
  batched           the 256-mesh x 16-RHS batched line, K steps after W
  wave, heat        the Newmark 1024^2 and heat 64^2 lines, K steps each
  cpu_baseline      the FP32 port of the reference CPU path, one core
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
# Bytes of the dominant kernel's dispatch measured with rocprofv3
# FETCH_SIZE / WRITE_SIZE (calibrated against a plain copy kernel) for the
# exact command given; the summaries with their commands are in profiles/.
PROFILED_TRAFFIC = {
    "grayscott": {"steps": 20, "warmup": 5, "predictor": "sextic", "bytes_per_launch": 1.5290e9,
                  "source": "profiles/k_step_grayscott_icosphere6.md"},
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


PREDICTORS = {"linear": 0, "previous": 1, "quadratic": 2, "cubic": 3, "quartic": 4, "quintic": 5,
              "sextic": 6}  # FEM_PREDICTOR_*


def step_bytes(n, nnz, k, kv, iters, sweeps, steps, ti=4, bicgstab=False, fused=False, stream=False):
    """Algorithmic HBM bytes of a step sequence (int32 indices, inner vectors
    and values of ti bytes, FP64 refinement), every vector element and matrix
    entry touched once per phase (cache-perfect gathers).  A pipelined-CG
    iteration moves 11 vectors and the matrix once; a BiCGSTAB iteration the
    matrix twice and 17 vectors on the persistent kernel (p: r, p_old, v_old,
    rh read, p, v written; s: r, v read, s, t written; x: xi, p, s, t, rh
    read, xi, r written), 15 on the streaming engine's multi-column layout
    (stream: p: p, rh read, v written; s: r, v, rh read, t written; x: xi,
    r, t, p, v read, xi, r, p written).
    fused: first-order equations on the persistent kernel, whose residual
    pass also reads M and the history combination and writes M x and S x0,
    so the update + right-hand side phase reads no matrix."""
    if bicgstab:
        it = 2 * nnz * (4 + ti * kv) + n * ((15 if stream else 17) * k * ti + ti * kv)
    else:
        it = nnz * (4 + ti * kv) + n * (11 * k * ti + ti * kv)     # one Krylov iteration
    sw = nnz * (4 + 8 * kv) + n * (3 * k * 8 + k * ti + ti * kv)  # FP64 residual
    st = nnz * (12 + 8 * kv) + n * (36 * k + 4 * kv) + n * 32 * k  # RHS + update
    if fused:
        sw += nnz * 8 + n * 3 * k * 8                                # + M, hvec read; xnext, hpre written
        st += -nnz * (12 + 8 * kv) + n * k * 8                       # no matrix; + hvec written
    return iters * it + sweeps * sw + steps * st


def stats_copy(st):
    """A copy of a solver's last_stats: later advances (accuracy checks)
    overwrite the object in place."""
    return type(st).from_buffer_copy(st)


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
    st = stats_copy(sol.last_stats)
    ach = step_bytes(n, nnz, 1, 1, st.iterations, st.sweeps, steps, ti=8, fused=True) / (st.ms * 1e-3) / 1e9
    import oracle
    so = oracle.System(mesh, bc=0)
    _, ur, _, _ = so.run(oracle.HEAT, [1e-3, 1.0], steps, u0.copy(), tol=1e-14)
    uu = sol.get_state_unknowns()
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
                                  "reference": "oracle (C restatement, FP64 CG to 1e-14), same steps"})
    if not args.no_cpu:
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
    itol = args.inner_tol or 1e-6  # one FP32 sweep plus one refinement per step (1e-4: 2.8-3.0 sweeps)
    if args.engine == "stream":
        sol = fresh(tol=1e-10, inner_tol=itol, engine=_lib.FEM_ENGINE_STREAM)
        tuned = None
    else:
        sol = fresh(tol=1e-10, inner_tol=itol)
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
    st = stats_copy(sol.last_stats)
    ach = step_bytes(n, nnz, 1, 1, st.iterations, st.sweeps, steps) / (st.ms * 1e-3) / 1e9
    ref = fresh(precision=_lib.FEM_PREC_FP64, tol=1e-13)
    ref.advance(steps)
    uu, ur = sol.get_state_unknowns(), ref.get_state_unknowns()
    # one step teacher-forced from the final state against the oracle
    import oracle
    so = oracle.System(mesh, bc=0)
    fin = [sol.get_state_unknowns(f) for f in (_lib.FEM_FIELD_U, _lib.FEM_FIELD_V, _lib.FEM_FIELD_A)]
    sol.advance(1)
    _, u1, v1, a1 = so.step(oracle.NEWMARK, [2e-4, 1.0, 0.25, 0.5], fin[0].copy(), fin[1].copy(), fin[2].copy(),
                            tol=1e-14)
    one = max(float(np.linalg.norm(sol.get_state_unknowns(f) - r) / np.linalg.norm(r))
              for f, r in ((_lib.FEM_FIELD_U, u1), (_lib.FEM_FIELD_V, v1), (_lib.FEM_FIELD_A, a1)))
    out = contract_line(METRIC, n * steps / wall, "DOF·steps/s", steps, args.warmup, 1e3 * wall / steps,
                        "f32+f64 (FP32 CG, FP64 refinement residual and dots)",
                        {"workload": f"Newmark wave, jittered 1024x1024 lattice (N={n}, nnz={nnz}), c=1, "
                                     f"beta=1/4 gamma=1/2, dt=2e-4, 8 Gaussian bumps, tol 1e-10",
                         "global_batch": 1,
                         "parallelism": "single GPU, " + ("stream engine (autotuned)"
                                                          if sol.options.engine == _lib.FEM_ENGINE_STREAM
                                                          else "persistent cooperative kernel")},
                        status=int(status),
                        solver={"iterations": int(st.iterations), "sweeps": int(st.sweeps), "inner_tol": itol, "autotune": tuned,
                                "kernel_ms": st.ms, "iters_per_step": st.iterations / steps},
                        roofline={"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": ach / HBM_PEAK_GBS, "traffic": None,
                                  "kernel": "k_sphase<1,float,*> (streaming engine phases)"
                                  if sol.options.engine == _lib.FEM_ENGINE_STREAM else "k_step<1,float>",
                                  "note": "~200 MB working set: mostly Infinity-Cache resident"},
                        accuracy={"one_step_rel_l2": one,
                                  "one_step_reference": "oracle (C restatement, FP64 CG to 1e-14), one step from "
                                                        "the final state (u, v, a)",
                                  "trajectory_rel_l2_final": float(np.linalg.norm(uu - ur) / np.linalg.norm(ur)),
                                  "trajectory_reference": "FP64 GPU CG, tol 1e-13 (the oracle takes ~0.7 s per "
                                                          "step at this size)"})
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
                               "sample": f"{ns} steps of the reference's wave scheme (CPUSolver.cpp:108-121) on the "
                                         f"same mesh, FP32 Eigen-default CG port ({its} iterations); a different "
                                         f"scheme from the GPU line's Newmark, which has no reference "
                                         f"counterpart; host: {model} ({ncpu} CPUs)"}
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


def batched_measure(args, steps, warmup, plain_seconds=8.0, cpu_seconds=8.0, graph_replay=False):
    """256 advection-diffusion instances of uneven size (about 1.3e7
    unknowns in all), 16 right-hand sides each, backward Euler dt = 1e-3,
    unnormalised constant velocity per instance, FP32 BiCGSTAB + FP64
    refinement (tol 1e-10), autotuned between the persistent team plans and
    the streaming engine.  `warmup` untimed steps, then `steps` timed steps
    continuing the same trajectories.  Outside the timed region: one step of
    a sample of instances and columns against the oracle (teacher-forced from
    the final state), the same steps on the same engine in FP64, a sample of
    the plain FP64 engine and of the FP32 CPU port.  graph_replay (the
    batched workload on its own): the next K steps as one replay of a
    recorded streaming advance.  The default line leaves it out: rocprofv3
    (kernel trace, and counters with a kernel filter that excludes the
    graph's kernels) crashed in the replay of a 20-step graph, so profiled
    commands run without one (--no-graph for the batched line)."""
    from femhip import AdvectionDiffusionParameters, Batch, FEMContext, SolverOptions, _lib
    ninst = args.instances
    t_gen = time.perf_counter()
    inst = c4_instances(ninst)
    log("batched: meshes generated", round(time.perf_counter() - t_gen, 1), "s")
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
    log("batched: contexts assembled, N =", tot_n)

    def make(**kw):
        kw.setdefault("predictor", PREDICTORS[args.predictor])
        kw.setdefault("inner_tol", args.inner_tol or 1e-6)
        b = Batch(ctxs, SolverOptions(nrhs=16, **kw))
        for i in range(len(ctxs)):
            b.set_equation(i, prms[i])
            b.set_state_unknowns(i, _lib.FEM_FIELD_U, fields[i])
        return b

    def timed(b):
        if warmup > 0:
            b.advance(warmup)
        b.sync()
        t = time.perf_counter()
        status = b.advance(steps)
        b.sync()
        return status, time.perf_counter() - t

    if args.engine == "stream":
        b = make(tol=args.tol, engine=_lib.FEM_ENGINE_STREAM)
        tuned = None
    else:
        b = make(tol=args.tol)
        tuned = b.autotune(2) if not args.no_autotune else None
        for i in range(len(ctxs)):  # the autotuner restores the state; start from the initial fields
            b.set_state_unknowns(i, _lib.FEM_FIELD_U, fields[i])
    log("batched: autotuned", tuned)
    engine = "stream" if b.options.engine == _lib.FEM_ENGINE_STREAM else "persistent"
    status, wall = timed(b)
    st = stats_copy(b.last_stats)
    log("batched: timed run", round(wall, 3), "s")
    dof = 16 * tot_n
    # algorithmic bytes instance by instance: larger meshes iterate more
    def alg_bytes(nsteps):
        tot = 0.0
        for i, c in enumerate(ctxs):
            si = b.instance_stats(i)
            tot += step_bytes(c.num_unknowns(), c.info().nnz, 16, 1, si.iterations, si.sweeps, nsteps,
                              bicgstab=True, stream=engine == "stream")
        return tot

    alg = alg_bytes(steps)
    st_ms, st_its, st_sweeps = st.ms, int(st.iterations), int(st.sweeps)
    ach = alg / (st_ms * 1e-3) / 1e9
    graph = None
    if engine == "stream" and graph_replay:
        # the next K steps as replays of a recorded advance, continuing the
        # trajectory; a recording is sized from the last eager advance, so a
        # few eager steps first size it for the steps that follow (the long
        # advance above includes the costly early steps); a recording holds at
        # most the library's launch limit, so G <= K steps are replayed K // G
        # times
        b.advance(4)
        gsteps = steps
        while True:
            try:
                b.graph_capture(gsteps)
                break
            except _lib.FemError:
                if gsteps == 1:
                    raise
                gsteps = max(1, gsteps // 2)
        reps = steps // gsteps
        b.sync()
        t = time.perf_counter()
        gstatus, gits, gms, galg = 0, 0, 0.0, 0.0
        for _ in range(reps):
            gstatus = max(gstatus, int(b.graph_launch()))
            gits += int(b.last_stats.iterations)
            gms += b.last_stats.ms
            galg += alg_bytes(gsteps)
        b.sync()
        gw = time.perf_counter() - t
        graph = {"value": dof * gsteps * reps / gw, "unit": "DOF·steps/s", "ms_per_step": 1e3 * gw / (gsteps * reps),
                 "status": gstatus, "iterations": gits, "steps_per_graph": gsteps, "replays": reps,
                 "algorithmic_GBs": galg / (gms * 1e-3) / 1e9,
                 "note": "after 4 more eager steps, the next steps as replays of a HIP graph of every phase "
                         "launch, with the sweeps and iterations of those 4 steps plus a margin and no host reads "
                         "(later steps need fewer iterations: compare algorithmic_GBs with roofline.achieved); wall "
                         "time includes the host's per-replay statistics reads"}
        log("batched: graph replays", reps, "x", gsteps, "steps", round(gw, 3), "s")
    out = {"value": dof * steps / wall, "unit": "DOF·steps/s", "steps": steps, "warmup": warmup,
           "ms_per_step": 1e3 * wall / steps, "status": int(status),
           "workload": f"batched advection-diffusion, {len(ctxs)} jittered meshes of side log-U[32,512] "
                       f"(sum N={tot_n}, sum nnz={tot_nnz}), 16 RHS each, dt=1e-3, tol {args.tol:g}; timed steps "
                       f"{warmup + 1}..{warmup + steps}",
           "engine": engine + (" (autotuned)" if tuned is not None else ""),
           "solver": {"iterations": st_its, "sweeps": st_sweeps, "autotune": tuned, "kernel_ms": st_ms},
           "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": ach / HBM_PEAK_GBS, "traffic": None,
                        "kernel": "k_sphase<16,float,*> (streaming engine phases)" if engine == "stream"
                        else "k_step<16,float,BiCGSTAB>",
                        "bytes_per_launch": alg,
                        "note": "algorithmic bytes summed instance by instance (iterations and sweeps of each) "
                                "over the device time of the advance"}}
    if graph is not None:
        out["graph_replay"] = graph
    if not args.no_accuracy:
        # one step teacher-forced from the final state, against the oracle
        # (FP64 BiCGSTAB to 1e-14) on the 2 smallest and the largest
        # instances, 4 columns each
        import oracle
        order = np.argsort([c.num_unknowns() for c in ctxs], kind="stable")
        sample = [int(order[0]), int(order[1]), int(order[-1])]
        starts = {i: b.get_state_unknowns(i).reshape(-1, 16) for i in sample}
        b.advance(1)
        errs = []
        for i in sample:
            m, kap, vel, cen, sig = inst[i]
            s = oracle.System(m, bc=0, velocity=vel, normalize=False)
            got = b.get_state_unknowns(i).reshape(-1, 16)
            for j in (0, 5, 10, 15):
                _, ref, _, _ = s.step(oracle.ADVDIFF, [1e-3, kap], starts[i][:, j].copy(), krylov=oracle.BICGSTAB,
                                      tol=1e-14)
                errs.append(float(np.linalg.norm(got[:, j] - ref) / np.linalg.norm(ref)))
        out["accuracy"] = {"one_step_rel_l2_max": max(errs), "gate_1e-8": bool(max(errs) <= 1e-8),
                           "sample": f"instances {sample} (2 smallest, largest), columns 0,5,10,15, one step from "
                                     f"the final state",
                           "reference": "oracle (C restatement, FP64 BiCGSTAB, tol 1e-14)"}
    b.close()
    if not args.no_fp64:
        ref = make(tol=args.tol, precision=_lib.FEM_PREC_FP64,
                   engine=b.options.engine if engine == "stream" else _lib.FEM_ENGINE_PERSISTENT)
        _, w64 = timed(ref)
        v64 = dof * steps / w64
        out["fp64_same_engine"] = {"value": v64, "unit": "DOF·steps/s", "ms_per_step": 1e3 * w64 / steps,
                                   "speedup_of_value": out["value"] / v64,
                                   "note": f"the same {engine} engine and steps with FP64 Krylov vectors and matrix"}
        ref.close()
    if not args.no_plain and plain_seconds > 0:
        # one step of each instance in turn on the plain FP64 engine (one
        # right-hand side at a time, host-read reduction scalars), for up to
        # plain_seconds of solver time
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
            if spent > plain_seconds:
                break
        pv = done_dof / spent
        out["plain_fp64_gpu"] = {"value": pv, "unit": "DOF·steps/s", "steps": 1,
                                 "sample": f"1 step of instances 0..{count - 1} ({done_dof} DOF), one right-hand "
                                           f"side at a time with host-read reduction scalars",
                                 "speedup_of_value": out["value"] / pv}
    if not args.no_cpu and cpu_seconds > 0:
        # the reference solves advection-diffusion with Eigen's CG, which can
        # stagnate up to its 2N iteration cap on these nonsymmetric systems:
        # sample the smallest instances, one column at a time
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
            if spent > cpu_seconds:
                break
        model, ncpu = cpu_info()
        out["cpu_baseline"] = {"value": dofs / spent, "unit": "DOF·steps/s", "cores": 1, "kind": "port",
                               "sample": f"the {len(used)} smallest instances {used}, 16 columns x 1 step each "
                                         f"({dofs} DOF), FP32 Eigen-default CG port ({its} iterations; the reference "
                                         f"solves advection-diffusion with CG); host: {model} ({ncpu} CPUs)"}
    for c in ctxs:
        c.close()
    return out


def run_batched_c4(args):
    """The batched line on its own (C4: 200 steps by default)."""
    steps = args.steps or 200
    r = batched_measure(args, steps, args.warmup, plain_seconds=30.0, cpu_seconds=20.0,
                        graph_replay=not args.no_graph)
    extra = {k: v for k, v in r.items() if k not in ("value", "unit", "steps", "warmup", "ms_per_step",
                                                      "workload", "engine")}
    return contract_line(METRIC, r["value"], r["unit"], steps, args.warmup, r["ms_per_step"],
                         "f32+f64 (FP32 BiCGSTAB, FP64 refinement residual and dots)",
                         {"workload": r["workload"], "global_batch": args.instances * 16,
                          "parallelism": f"single GPU, all instances per launch ({r['engine']})"}, **extra)


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
    ap.add_argument("--predictor", choices=sorted(PREDICTORS), default="sextic",
                    help="grayscott, batched: initial guess of each solve")
    ap.add_argument("--engine", choices=["auto", "stream"], default="auto",
                    help="batched, wave: 'stream' runs the streaming engine without autotuning")
    ap.add_argument("--inner-tol", type=float, default=None,
                    help="FP32 sweep reduction (default 1e-6; batched 256 x 16 RHS: 1e-4 -> 1e-6 took 30%% of the "
                         "refinement sweeps off at the same iterations)")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--no-spmv", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="batched: leave out the graph replay")
    ap.add_argument("--no-wave", action="store_true", help="default line: leave out the Newmark wave leg")
    ap.add_argument("--no-heat", action="store_true", help="default line: leave out the heat leg")
    ap.add_argument("--frames", type=int, default=200,
                    help="grayscott: frames timed through the surface-space Solver path")
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
    out = grayscott_line(args)
    if not args.no_spmv:
        log("spmv")
        out["spmv"] = spmv_measure()
    if not args.no_adjoint:
        log("adjoint")
        out["adjoint"] = adjoint_measure()
    if not args.no_batched:
        log("batched")
        out["batched"] = batched_measure(args, args.steps if args.steps <= 200 else 200, args.warmup)
    # the other single-instance benchmark workloads, K steps each (at most 200)
    short = argparse.Namespace(**{**vars(args), "steps": min(args.steps, 200)})
    if not args.no_wave:
        log("wave")
        out["wave"] = run_wave_c3(short)
    if not args.no_heat:
        log("heat")
        out["heat"] = run_heat_c1(short)
    if not args.no_cpu and args.cpu_steps > 0:
        log("cpu baseline")
        out["cpu_baseline"] = run_cpu_baseline(out.pop("_mesh"), out.pop("_u0"), out.pop("_v0"),
                                               [0.5, 2e-4, 1e-4, 0.037, 0.06, 1], args.cpu_steps)
    for k in ("_mesh", "_u0", "_v0"):
        out.pop(k, None)
    print(json.dumps(out))


def grayscott_line(args):
    """The headline line: Gray-Scott on the icosphere refined args.level
    times.  W = args.warmup untimed steps from the initial state, then
    K = args.steps timed steps continuing the same trajectory in one launch.
    Outside the timed region: the trajectory after W + K steps and one step
    teacher-forced from the oracle's states at three points against the
    oracle; the same steps on the same engine in FP64; the plain FP64 engine;
    and frames through the surface-space Solver boundary."""
    import oracle
    from femhip import Equation, FEMContext, HIPSolver, ReactionDiffusionParameters, SolverOptions, _lib, meshes

    mesh = meshes.icosphere(args.level)
    ctx = FEMContext(mesh, 1)
    n = ctx.num_unknowns()
    nnz = ctx.info().nnz
    u0, v0 = gs_initial(mesh, n)
    prm = ReactionDiffusionParameters(0.5, 2e-4, 1e-4, 0.037, 0.06, _lib.FEM_REACTION_LUMPED)
    oprm = [0.5, 2e-4, 1e-4, 0.037, 0.06, 1]
    ctx.equation = Equation.Reaction_Diffusion
    ctx.parameters[Equation.Reaction_Diffusion] = prm
    W, K = args.warmup, args.steps

    def solver(**kw):
        o = SolverOptions(tol=args.tol, inner_tol=args.inner_tol or 1e-6, team_size=args.team_size,
                          predictor=PREDICTORS[args.predictor])
        for k_, v_ in kw.items():
            setattr(o, k_, v_)
        s = HIPSolver(ctx, o)
        s.set_state_unknowns(_lib.FEM_FIELD_U, u0)
        s.set_state_unknowns(_lib.FEM_FIELD_V, v0)
        return s

    def timed(s):
        if W > 0:
            s.advance(W)
        s.sync()
        t = time.perf_counter()
        status = s.advance(K)
        s.sync()
        return status, time.perf_counter() - t

    log("grayscott: N =", n)
    sol = solver()
    tuned = None
    if args.team_size == 0 and not args.no_autotune:
        tuned = sol.autotune(20)  # untimed; the state is restored
    status, wall = timed(sol)
    log("grayscott: timed run", wall)
    st = stats_copy(sol.last_stats)
    kernel_ms = st.ms
    value = 2 * n * K / wall
    out = contract_line(
        METRIC, value, "DOF·steps/s", K, W, 1e3 * wall / K, "f32+f64 (FP32 CG, FP64 refinement residual and dots)",
        {"workload": f"gray-scott icosphere L{args.level} (N={n}, nnz={nnz}, 2 species), dt=0.5, Du=2e-4 Dv=1e-4 "
                     f"F=0.037 k=0.06, 50 patches, lumped reaction weighting, tol {args.tol:g}; timed steps "
                     f"{W + 1}..{W + K} of the trajectory in one launch",
         "global_batch": 1, "parallelism": "single GPU, persistent cooperative kernel"},
        status=int(status),
        solver={"krylov_iterations": int(st.iterations), "sweeps": int(st.sweeps),
                "krylov_iterations_per_step": st.iterations / K, "sweeps_per_step": st.sweeps / K,
                "note": "one Krylov iteration advances the u and v solves together",
                "team_size": sol.options.team_size, "autotune": tuned, "final_residual": st.residual,
                "kernel_ms": kernel_ms})
    # roofline of the persistent stepping kernel: the timed region is one launch
    bytes_alg = step_bytes(n, nnz, 2, 2, st.iterations, st.sweeps, K, fused=True)
    ach = bytes_alg / (kernel_ms * 1e-3) / 1e9
    prof = PROFILED_TRAFFIC["grayscott"]
    same_cmd = ((args.steps, args.warmup, args.level, args.tol, args.predictor, args.inner_tol, args.team_size)
                == (prof["steps"], prof["warmup"], 6, 1e-10, prof["predictor"], None, 0))
    out["roofline"] = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": ach / HBM_PEAK_GBS,
                       "traffic": prof["bytes_per_launch"] if same_cmd else None,
                       "traffic_source": prof["source"] if same_cmd else "not profiled for these arguments",
                       "traffic_note": "rocprofv3 FETCH_SIZE + WRITE_SIZE of the timed dispatch: bytes between L2 "
                                       "and the fabric (Infinity Cache hits included), scaled by a copy-kernel "
                                       "calibration; the 8 MB working set stays cache-resident",
                       "kernel": "k_step<2,float>", "bytes_per_launch": bytes_alg, "kernel_ms": kernel_ms}
    uT = sol.get_state_unknowns(_lib.FEM_FIELD_U)
    vT = sol.get_state_unknowns(_lib.FEM_FIELD_V)

    if not args.no_accuracy:
        # the oracle trajectory (C restatement, FP64 CG to 1e-14) and one step
        # teacher-forced from its states after W, W + K/2 and W + K steps
        s = oracle.System(mesh, bc=1)
        marks = sorted({W, W + K // 2, W + K})
        u, v = u0.copy(), v0.copy()
        one = []
        done = 0
        probe = solver()
        for mk in marks:
            if mk > done:
                _, u, v, _ = s.run(oracle.GRAYSCOTT, oprm, mk - done, u, v, tol=1e-14)
                done = mk
            if mk == W + K:
                traj = max(float(np.linalg.norm(uT - u) / np.linalg.norm(u)),
                           float(np.linalg.norm(vT - v) / np.linalg.norm(v)))
            probe.set_state_unknowns(_lib.FEM_FIELD_U, u)
            probe.set_state_unknowns(_lib.FEM_FIELD_V, v)
            probe.advance(1)
            _, u1, v1, _ = s.step(oracle.GRAYSCOTT, oprm, u.copy(), v.copy(), tol=1e-14)
            one.append(max(float(np.linalg.norm(probe.get_state_unknowns(_lib.FEM_FIELD_U) - u1) / np.linalg.norm(u1)),
                           float(np.linalg.norm(probe.get_state_unknowns(_lib.FEM_FIELD_V) - v1) / np.linalg.norm(v1))))
        probe.close()
        out["accuracy"] = {"one_step_rel_l2_max": max(one), "one_step_rel_l2": one, "one_step_after_steps": marks,
                           "trajectory_rel_l2_final": traj, "trajectory_steps": W + K,
                           "reference": "oracle (C restatement of CPUSolver, FP64 CG to 1e-14)",
                           "gate_1e-8_one_step": bool(max(one) <= 1e-8),
                           "gate_1e-8_trajectory": bool(traj <= 1e-8)}

    if not args.no_fp64:
        f64 = solver(precision=_lib.FEM_PREC_FP64)
        _, w64 = timed(f64)
        v64 = 2 * n * K / w64
        out["fp64_same_engine"] = {"value": v64, "unit": "DOF·steps/s", "ms_per_step": 1e3 * w64 / K,
                                   "kernel": "k_step<2,double>", "speedup_of_value": value / v64,
                                   "note": "the same persistent kernel and steps with FP64 Krylov vectors and matrix"}
        f64.close()

    if not args.no_plain and args.plain_steps > 0:
        pl = solver(engine=_lib.FEM_ENGINE_PLAIN, precision=_lib.FEM_PREC_FP64)
        pl.sync()
        t = time.perf_counter()
        pl.advance(args.plain_steps)
        pl.sync()
        tw = time.perf_counter() - t
        pv = 2 * n * args.plain_steps / tw
        out["plain_fp64_gpu"] = {"value": pv, "unit": "DOF·steps/s", "steps": args.plain_steps,
                                 "iterations": int(pl.last_stats.iterations), "speedup_of_value": value / pv,
                                 "note": "plain FP64 engine: one right-hand side at a time, separate kernels, "
                                         "host-read reduction scalars"}
        pl.close()

    if args.frames > 0:
        # the reference's per-frame call pattern (Application.cpp:530-540):
        # the surface values (v for Gray-Scott, float per vertex) go in, one
        # step, the values come back; u stays on the device
        fr = solver()
        fr.advance(W)
        surf = fr.get_state(_lib.FEM_FIELD_V)
        fr.sync()
        t = time.perf_counter()
        its = 0
        for _ in range(args.frames):
            fr.set_state(_lib.FEM_FIELD_V, surf)
            fr.advance(1)
            its += fr.last_stats.iterations
            surf = fr.get_state(_lib.FEM_FIELD_V)
        tf = time.perf_counter() - t
        out["frames"] = {"frames": args.frames, "ms_per_frame": 1e3 * tf / args.frames,
                         "value": 2 * n * args.frames / tf, "unit": "DOF·steps/s",
                         "krylov_iterations_per_frame": its / args.frames,
                         "path": "HIPSolver.set_state(v) -> advance(1) -> get_state(v) per frame, through the C ABI"}
        fr.close()
    out["_mesh"], out["_u0"], out["_v0"] = mesh, u0, v0
    sol.close()
    ctx.close()
    return out


if __name__ == "__main__":
    main()
