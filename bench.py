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

    if not args.no_cpu and args.cpu_steps > 0:
        out["cpu_baseline"] = run_cpu_baseline(mesh, u0, v0, [0.5, 2e-4, 1e-4, 0.037, 0.06, 1], args.cpu_steps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
