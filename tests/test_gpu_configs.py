"""GPU parity on the benchmark workloads themselves, at their full sizes.

Each test runs the HIP library (through the C ABI) on one benchmark
configuration's mesh, parameters and initial data, and checks it against the
FP64 oracle solved to 1e-14: every compared field, column and instance to
rel-L2 <= 1e-8, the accuracy bar of the benchmark metric.  The golden
fixtures written by tests/golden/make_golden.py (an independent direct-solve
implementation) are checked on the GPU path as well.
"""
import concurrent.futures as cf
import os
import sys

import numpy as np
import pytest

import oracle
from femhip import (AdvectionDiffusionParameters, Adjoint, Batch, Equation, FEMContext, HeatParameters, HIPSolver,
                    NewmarkParameters, ReactionDiffusionParameters, SolverOptions, _lib, meshes)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
import bench  # noqa: E402  (workload definitions: initial data, instance parameters)

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
RTOL = 1e-8


def rel(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / (nb if nb > 0 else 1.0))


def load_gold(name):
    return dict(np.load(os.path.join(GOLD, name)))


def mesh_of(g):
    return {"xyz": g["xyz"], "tri": g["tri"], "on_boundary": g["on_boundary"]}


def test_c1_heat_square64_100_steps(gpu):
    """Heat, 64x64 structured square, dt = 1e-3, c = 1, Dirichlet, Gaussian
    u0, 100 steps, FP64 CG tol 1e-10 (as benchmarked), against the oracle."""
    print("case C1 heat 64x64, 100 steps", flush=True)
    m = meshes.structured_square(64)
    s = oracle.System(m, bc=0)
    X = m["xyz"].astype(np.float64)
    u0 = s.gather(np.exp(-((X[:, 0] - 0.5) ** 2 + (X[:, 1] - 0.5) ** 2) / 0.01))
    ctx = FEMContext(m, 0)
    ctx.equation = Equation.Heat
    ctx.parameters[Equation.Heat] = HeatParameters(1e-3, 1.0)
    sol = HIPSolver(ctx, SolverOptions(precision=_lib.FEM_PREC_FP64, tol=1e-10))
    sol.set_state_unknowns(_lib.FEM_FIELD_U, u0)
    _, ref, _, _ = s.run(oracle.HEAT, [1e-3, 1.0], 100, u0.copy(), tol=1e-14)
    assert sol.advance(100) == _lib.FEM_OK
    assert rel(sol.get_state_unknowns(), ref) <= RTOL


def test_c2_grayscott_icosphere6_lumped(gpu):
    """Gray-Scott on the icosphere refined 6 times (N = 40,962), dt = 0.5,
    the benchmark's 50 patches, lumped reaction weighting, FP32 CG + FP64
    refinement at the benchmark's tolerances; 10 single-step launches, then 10
    steps in one launch, each state against the oracle trajectory."""
    print("case C2 gray-scott icosphere L6", flush=True)
    m = meshes.icosphere(6)
    s = oracle.System(m, bc=1)
    assert (s.n, s.nnz) == (40962, 286722)
    u0, v0 = bench.gs_initial(m, s.n)
    prm = ReactionDiffusionParameters(0.5, 2e-4, 1e-4, 0.037, 0.06, _lib.FEM_REACTION_LUMPED)
    ctx = FEMContext(m, 1)
    ctx.equation = Equation.Reaction_Diffusion
    ctx.parameters[Equation.Reaction_Diffusion] = prm
    sol = HIPSolver(ctx, SolverOptions(tol=1e-10, inner_tol=1e-5))
    sol.set_state_unknowns(_lib.FEM_FIELD_U, u0)
    sol.set_state_unknowns(_lib.FEM_FIELD_V, v0)
    oprm = [0.5, 2e-4, 1e-4, 0.037, 0.06, 1]
    u, v = u0.copy(), v0.copy()
    for k in range(10):
        _, u, v, _ = s.step(oracle.GRAYSCOTT, oprm, u, v, tol=1e-14)
        assert sol.advance(1) == _lib.FEM_OK
        assert rel(sol.get_state_unknowns(_lib.FEM_FIELD_U), u) <= RTOL, k
        assert rel(sol.get_state_unknowns(_lib.FEM_FIELD_V), v) <= RTOL, k
    _, u, v, _ = s.run(oracle.GRAYSCOTT, oprm, 10, u, v, tol=1e-14)
    assert sol.advance(10) == _lib.FEM_OK
    assert rel(sol.get_state_unknowns(_lib.FEM_FIELD_U), u) <= RTOL
    assert rel(sol.get_state_unknowns(_lib.FEM_FIELD_V), v) <= RTOL
    # the patches have started to react: v differs from pure diffusion
    assert np.max(np.abs(v - v0)) > 1e-3


def test_c3_newmark_jittered1024(gpu):
    """Newmark-beta wave on the jittered 1024x1024 lattice (N = 1,046,529),
    c = 1, beta = 1/4, gamma = 1/2, dt = 2e-4, 8 Gaussian bumps, consistent
    initial acceleration; 3 steps of FP32 CG + FP64 refinement (tol 1e-10) on
    the persistent and the streaming engine: u, v, a against the oracle."""
    print("case C3 newmark 1024^2", flush=True)
    m = meshes.jittered_square(1024, seed=2)
    s = oracle.System(m, bc=0)
    rng = np.random.Generator(np.random.PCG64(3))
    cen = rng.uniform(0.1, 0.9, size=(8, 2))
    X = m["xyz"].astype(np.float64)[s.idx_map >= 0]
    u0 = bench.gauss_sum(X, cen, [0.02] * 8).sum(axis=1)
    a0 = s.newmark_init(1.0, u0)
    _, ur, vr, ar = s.run(oracle.NEWMARK, [2e-4, 1.0, 0.25, 0.5], 3, u0.copy(), np.zeros(s.n), a0.copy(),
                          tol=1e-14)
    ctx = FEMContext(m, 0)
    ctx.equation = Equation.Wave_Newmark
    ctx.parameters[Equation.Wave_Newmark] = NewmarkParameters(2e-4, 1.0, 0.25, 0.5)
    for engine in (_lib.FEM_ENGINE_PERSISTENT, _lib.FEM_ENGINE_STREAM):
        print("  engine", engine, flush=True)
        sol = HIPSolver(ctx, SolverOptions(tol=1e-10, engine=engine))
        sol.set_state_unknowns(_lib.FEM_FIELD_U, u0)
        a = sol.init_newmark_acceleration()
        assert rel(a, a0) <= RTOL
        assert sol.advance(3) == _lib.FEM_OK
        assert rel(sol.get_state_unknowns(_lib.FEM_FIELD_U), ur) <= RTOL
        assert rel(sol.get_state_unknowns(_lib.FEM_FIELD_V), vr) <= RTOL
        assert rel(sol.get_state_unknowns(_lib.FEM_FIELD_A), ar) <= RTOL
        sol.close()


C4_STEPS = 2


@pytest.fixture(scope="module")
def c4_cases():
    """The 4 smallest and 4 largest of the 256 batched instances, with the
    oracle's 2-step trajectories of all 16 columns (BiCGSTAB to 1e-14; the
    oracle releases the GIL, so columns run in parallel threads)."""
    specs = bench.c4_specs()
    order = np.argsort([sp[0] for sp in specs], kind="stable")
    which = [int(i) for i in list(order[:4]) + list(order[-4:])]
    cases = []
    jobs = []
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        for i, (m, kap, vel, cen, sig) in zip(which, bench.c4_instances(which=which)):
            s = oracle.System(m, bc=0, velocity=vel, normalize=False)
            X = m["xyz"].astype(np.float64)[s.idx_map >= 0]
            F = np.ascontiguousarray(bench.gauss_sum(X, cen, sig))
            cases.append({"index": i, "mesh": m, "kappa": kap, "vel": vel, "F": F, "n": s.n})
            for j in range(16):
                jobs.append((len(cases) - 1, j, ex.submit(
                    lambda s=s, kap=kap, f=F[:, j].copy(): s.run(oracle.ADVDIFF, [1e-3, kap], C4_STEPS, f,
                                                                 krylov=oracle.BICGSTAB, tol=1e-14)[1])))
        for c in cases:
            c["ref"] = np.zeros_like(c["F"])
        for ci, j, fut in jobs:
            cases[ci]["ref"][:, j] = fut.result()
    return cases


@pytest.mark.parametrize("engine", ["stream", "stream_graph", "persistent"])
def test_c4_batched_advdiff_16rhs(gpu, c4_cases, engine):
    """Batched advection-diffusion: the 4 smallest and 4 largest benchmark
    instances (sides 32-505) in one batch, 16 right-hand sides each,
    unnormalised velocity, dt = 1e-3, FP32 BiCGSTAB + FP64 refinement
    (tol 1e-10): every column of every instance after 2 steps against the
    oracle.  stream_graph: one eager step, then the second as a replay of a
    recorded streaming advance."""
    print("case C4 batched, engine", engine, [c["index"] for c in c4_cases], flush=True)
    ctxs = [FEMContext(c["mesh"], 0, velocity=c["vel"], normalize_velocity=False) for c in c4_cases]
    eng = _lib.FEM_ENGINE_PERSISTENT if engine == "persistent" else _lib.FEM_ENGINE_STREAM
    b = Batch(ctxs, SolverOptions(nrhs=16, tol=1e-10, engine=eng))
    for i, c in enumerate(c4_cases):
        b.set_equation(i, AdvectionDiffusionParameters(1e-3, c["kappa"]))
        b.set_state_unknowns(i, _lib.FEM_FIELD_U, c["F"])
    if engine == "stream_graph":
        assert b.advance(C4_STEPS - 1) == _lib.FEM_OK
        b.graph_capture(1)
        assert b.graph_launch() == _lib.FEM_OK
        assert b.last_stats.steps == 1
    else:
        assert b.advance(C4_STEPS) == _lib.FEM_OK
    for i, c in enumerate(c4_cases):
        got = b.get_state_unknowns(i).reshape(c["n"], 16)
        for j in range(16):
            assert rel(got[:, j], c["ref"][:, j]) <= RTOL, (c["index"], j)
    b.close()
    for c in ctxs:
        c.close()


def test_c5_adjoint_jittered512(gpu):
    """Differentiable advection-diffusion on the jittered 512x512 lattice
    (N = 261,121): rotating per-element velocity (unnormalised), per-element
    diffusivity 1e-3 exp(0.3 z), dt = 2e-3, 10 steps with a checkpoint every
    5, loss |u_T - target|^2.  u_T and the gradients w.r.t. u0 and kappa
    against the oracle's FP64 adjoint; the FP64 GPU gradient against a central
    difference (h = 1e-6) along one random direction of each."""
    print("case C5 adjoint 512^2", flush=True)
    side, nsteps, every, dt = 512, 10, 5, 2e-3
    m = meshes.jittered_square(side, seed=6)
    cen = m["xyz"].astype(np.float64)[m["tri"]].mean(axis=1)
    ve = np.stack([-(cen[:, 1] - 0.5), cen[:, 0] - 0.5, np.zeros(len(cen))], axis=1)
    rk = np.random.Generator(np.random.PCG64(7))
    kappa = 1e-3 * np.exp(0.3 * rk.standard_normal(m["tri"].shape[0]))
    s = oracle.System(m, bc=0, velocity_e=ve, normalize=False, kappa_e=kappa)
    X = m["xyz"].astype(np.float64)
    keep = s.idx_map >= 0
    u0 = np.exp(-((X[:, 0] - 0.3) ** 2 + (X[:, 1] - 0.5) ** 2) / (2 * 0.05 ** 2))[keep]
    rt = np.random.Generator(np.random.PCG64(8))
    cc = rt.uniform(0.2, 0.8, 2)
    target = np.exp(-((X[:, 0] - cc[0]) ** 2 + (X[:, 1] - cc[1]) ** 2) / (2 * 0.05 ** 2))[keep]
    loss_r, uT_r, gu_r, gk_r = s.adjoint(dt, nsteps, u0, target, tol=1e-14)
    ctx = FEMContext(m, 0, velocity_e=ve, normalize_velocity=False, kappa_e=kappa)
    prm = AdvectionDiffusionParameters(dt, 1.0)
    for name, opts, gtol in (("fp32ir", SolverOptions(tol=1e-10), 1e-7),
                             ("fp64", SolverOptions(precision=_lib.FEM_PREC_FP64, tol=1e-13), 1e-9)):
        print("  precision", name, flush=True)
        adj = Adjoint(ctx, prm, opts)
        uT = adj.forward(u0, nsteps, every)
        assert rel(uT, uT_r) <= RTOL
        gu, gk = adj.backward(2.0 * (uT - target))
        assert rel(gu, gu_r) <= gtol
        assert rel(gk, gk_r) <= gtol
        if name == "fp64":
            rd = np.random.Generator(np.random.PCG64(9))
            h = 1e-6
            d = rd.standard_normal(s.n)

            def loss(u, k=None):
                if k is not None:
                    ctx.assemble_matrices(velocity_e=ve, normalize_velocity=False, kappa_e=k)
                return float(np.sum((adj.forward(u, nsteps, every) - target) ** 2))
            fd = (loss(u0 + h * d) - loss(u0 - h * d)) / (2 * h)
            assert abs(fd - gu @ d) <= 1e-6 * abs(fd)
            dk = rd.standard_normal(kappa.shape[0]) * kappa
            fk = (loss(u0, kappa + h * dk) - loss(u0, kappa - h * dk)) / (2 * h)
            ctx.assemble_matrices(velocity_e=ve, normalize_velocity=False, kappa_e=kappa)
            assert abs(fk - gk @ dk) <= 1e-4 * abs(fk)
        adj.close()


# ---- the independent golden fixtures on the GPU path ----------------------

def test_golden_icosphere_advection_matrix(gpu):
    """Curved-surface advection (velocity projected on each element and
    normalised) of the shipped icosphere against the fixture."""
    g = load_gold("icosphere_shipped.npz")
    ctx = FEMContext(mesh_of(g), 1)
    for which, key in ((_lib.FEM_MATRIX_K, "K"), (_lib.FEM_MATRIX_M, "M"), (_lib.FEM_MATRIX_A, "A")):
        assert np.max(np.abs(ctx.matrix(which) - g[key])) <= 1e-12 * np.max(np.abs(g[key]))


@pytest.mark.parametrize("precision", [_lib.FEM_PREC_FP32_IR, _lib.FEM_PREC_FP64])
def test_golden_icosphere_lumped_grayscott(gpu, precision):
    """Three Gray-Scott steps at dt = 0.5 with lumped weighting on the shipped
    icosphere against the direct-solve fixture."""
    g = load_gold("icosphere_shipped.npz")
    ctx = FEMContext(mesh_of(g), 1)
    ctx.equation = Equation.Reaction_Diffusion
    ctx.parameters[Equation.Reaction_Diffusion] = ReactionDiffusionParameters(
        0.5, 2e-4, 1e-4, 0.037, 0.06, _lib.FEM_REACTION_LUMPED)
    sol = HIPSolver(ctx, SolverOptions(precision=precision, tol=1e-12))
    sol.set_state_unknowns(_lib.FEM_FIELD_U, np.ones(g["lumped_v0"].shape[0]))
    sol.set_state_unknowns(_lib.FEM_FIELD_V, g["lumped_v0"])
    for k in range(3):
        assert sol.advance(1) == _lib.FEM_OK
        assert rel(sol.get_state_unknowns(_lib.FEM_FIELD_U), g["lumped_u"][k]) <= RTOL
        assert rel(sol.get_state_unknowns(_lib.FEM_FIELD_V), g["lumped_v"][k]) <= RTOL


@pytest.mark.parametrize("engine", [_lib.FEM_ENGINE_PERSISTENT, _lib.FEM_ENGINE_STREAM])
def test_golden_jittered_kappa_velocity(gpu, engine):
    """Per-element diffusivity and unnormalised per-element velocity:
    pattern, K, M, A and three advection-diffusion steps against the
    fixture."""
    g = load_gold("jittered_kappa.npz")
    ctx = FEMContext(mesh_of(g), 0, velocity_e=g["velocity_e"], normalize_velocity=False, kappa_e=g["kappa_e"])
    rp, ci = ctx.pattern()
    assert np.array_equal(ctx.idx_map, g["idx_map"])
    assert np.array_equal(rp, g["row_ptr"]) and np.array_equal(ci, g["col_idx"])
    for which, key in ((_lib.FEM_MATRIX_K, "K"), (_lib.FEM_MATRIX_M, "M"), (_lib.FEM_MATRIX_A, "A")):
        assert np.max(np.abs(ctx.matrix(which) - g[key])) <= 1e-12 * np.max(np.abs(g[key]))
    ctx.equation = Equation.Advection_Diffusion
    ctx.parameters[Equation.Advection_Diffusion] = AdvectionDiffusionParameters(2e-3, 1.0)
    sol = HIPSolver(ctx, SolverOptions(tol=1e-12, engine=engine))
    sol.set_state_unknowns(_lib.FEM_FIELD_U, g["u0"])
    for k in range(3):
        assert sol.advance(1) == _lib.FEM_OK
        assert rel(sol.get_state_unknowns(), g["advdiff"][k]) <= RTOL
