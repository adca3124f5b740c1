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


def step_bytes(n, nnz, k, kv, iters, sweeps, steps):
    """Algorithmic HBM bytes of the persistent stepping kernel (int32 indices,
    FP32 inner vectors and values, FP64 refinement), every vector element and
    matrix entry touched once per phase (cache-perfect gathers)."""
    it = nnz * (4 + 4 * kv) + n * (11 * k * 4 + 4 * kv)       # CG phases A + B
    sw = nnz * (4 + 8 * kv) + n * (28 * k + 4 * kv)           # FP64 residual
    st = nnz * (12 + 8 * kv) + n * (36 * k + 4 * kv) + n * 32 * k  # RHS + update
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--level", type=int, default=6)
    ap.add_argument("--cpu-steps", type=int, default=2000)
    ap.add_argument("--plain-steps", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--team-size", type=int, default=0)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--inner-tol", type=float, default=1e-4)
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
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is design-only "
                                   "(see DESIGN.md)", "n_gpus": args.gpus}))
        sys.exit(2)

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
        o = SolverOptions(tol=args.tol, inner_tol=args.inner_tol, team_size=args.team_size)
        for k_, v_ in kw.items():
            setattr(o, k_, v_)
        s = HIPSolver(ctx, o)
        s.set_state_unknowns(_lib.FEM_FIELD_U, u0)
        s.set_state_unknowns(_lib.FEM_FIELD_V, v0)
        return s

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
    bytes_alg = step_bytes(n, nnz, 2, 2, st.iterations, st.sweeps, args.steps)
    ach = bytes_alg / (kernel_ms * 1e-3) / 1e9
    out["roofline"] = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": ach / HBM_PEAK_GBS, "traffic": None,
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
        out["spmv"] = spmv_measure()
    if not args.no_adjoint:
        out["adjoint"] = adjoint_measure()

    if not args.no_cpu and args.cpu_steps > 0:
        out["cpu_baseline"] = run_cpu_baseline(mesh, u0, v0, [0.5, 2e-4, 1e-4, 0.037, 0.06, 1], args.cpu_steps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
