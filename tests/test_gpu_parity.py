"""GPU parity tests: the HIP library (through the C ABI) against the oracle
and the golden fixtures.  Pattern and idx_map are compared bit for bit;
FP64 matrices to 1e-12 of their largest entry; every time-step result to
rel-L2 <= 1e-8 against the FP64 oracle solved to 1e-14 (the tolerance of
the benchmark's accuracy gate)."""
import os

import numpy as np
import pytest

import oracle
from femhip import (AdvectionDiffusionParameters, Batch, Equation, FEMContext, HeatParameters, HIPSolver,
                    NewmarkParameters, ReactionDiffusionParameters, SolverOptions, WaveParameters, _lib, meshes)

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
RTOL = 1e-8


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / (nb if nb > 0 else 1.0)


def gauss(mesh, c=(0.5, 0.5), w=0.01):
    x = mesh["xyz"].astype(np.float64)
    return np.exp(-((x[:, 0] - c[0]) ** 2 + (x[:, 1] - c[1]) ** 2) / w)


MESHES = {
    "square16": (lambda: meshes.structured_square(16), 0),
    "jitter20": (lambda: meshes.jittered_square(20, seed=4), 0),
    "jitter20N": (lambda: meshes.jittered_square(20, seed=4), 1),
    "ico3": (lambda: meshes.icosphere(3), 1),
}


@pytest.mark.parametrize("name", sorted(MESHES))
def test_pattern_and_matrices(gpu, name):
    mk, bc = MESHES[name]
    m = mk()
    print("case", name, flush=True)
    ctx = FEMContext(m, bc, velocity=(0.3, 1.0, 0.2))
    s = oracle.System(m, bc=bc, velocity=(0.3, 1.0, 0.2))
    assert np.array_equal(ctx.idx_map, s.idx_map)
    rp, ci = ctx.pattern()
    assert np.array_equal(rp, s.row_ptr)
    assert np.array_equal(ci, s.col_idx)
    info = ctx.info()
    assert info.num_unknowns == s.n and info.nnz == s.nnz
    assert info.max_row_nonzeros == int(np.diff(s.row_ptr).max())
    for which, ref in ((_lib.FEM_MATRIX_K, s.K), (_lib.FEM_MATRIX_M, s.M), (_lib.FEM_MATRIX_A, s.A)):
        got = ctx.matrix(which)
        assert np.max(np.abs(got - ref)) <= 1e-12 * max(1e-300, np.max(np.abs(ref)))


def test_golden_square_matrices(gpu):
    g = dict(np.load(os.path.join(GOLD, "square64.npz")))
    m = {"xyz": g["xyz"], "tri": g["tri"], "on_boundary": g["on_boundary"]}
    ctx = FEMContext(m, 0)
    rp, ci = ctx.pattern()
    assert np.array_equal(ctx.idx_map, g["idx_map"])
    assert np.array_equal(rp, g["row_ptr"]) and np.array_equal(ci, g["col_idx"])
    for which, key in ((_lib.FEM_MATRIX_K, "K"), (_lib.FEM_MATRIX_M, "M"), (_lib.FEM_MATRIX_A, "A")):
        assert np.max(np.abs(ctx.matrix(which) - g[key])) <= 1e-12 * np.max(np.abs(g[key]))


def make_solver(ctx, eq, params, **opts):
    ctx.equation = eq
    ctx.parameters[eq] = params
    return HIPSolver(ctx, SolverOptions(**opts))


@pytest.mark.parametrize("precision,engine", [
    (_lib.FEM_PREC_FP32_IR, _lib.FEM_ENGINE_PERSISTENT),
    (_lib.FEM_PREC_FP64, _lib.FEM_ENGINE_PERSISTENT),
    (_lib.FEM_PREC_FP64, _lib.FEM_ENGINE_PLAIN),
])
def test_heat_golden_steps(gpu, precision, engine):
    """64x64 square, dt=1e-3, c=1: three steps against the direct-solve golden."""
    g = dict(np.load(os.path.join(GOLD, "square64.npz")))
    m = {"xyz": g["xyz"], "tri": g["tri"], "on_boundary": g["on_boundary"]}
    print("case heat golden", precision, engine, flush=True)
    ctx = FEMContext(m, 0)
    sol = make_solver(ctx, Equation.Heat, HeatParameters(1e-3, 1.0), precision=precision, engine=engine, tol=1e-12)
    sol.set_state_unknowns(_lib.FEM_FIELD_U, g["u0"])
    for k in range(3):
        st = sol.advance_time(1)
        assert st == _lib.FEM_OK
        assert rel(sol.get_state_unknowns(), g["heat"][k]) < RTOL


@pytest.mark.parametrize("precision", [_lib.FEM_PREC_FP32_IR, _lib.FEM_PREC_FP64])
def test_heat_trajectory_vs_oracle(gpu, precision):
    m = meshes.jittered_square(40, seed=1)
    s = oracle.System(m, bc=0)
    u0 = s.gather(gauss(m))
    ctx = FEMContext(m, 0)
    sol = make_solver(ctx, Equation.Heat, HeatParameters(1e-3, 1.0), precision=precision)
    sol.set_state_unknowns(_lib.FEM_FIELD_U, u0)
    u = u0.copy()
    for k in range(10):
        _, u, _, _ = s.step(oracle.HEAT, [1e-3, 1.0], u, tol=1e-14)
        sol.advance_time(1)
        assert rel(sol.get_state_unknowns(), u) < RTOL, k
    # many steps in one launch
    _, u, _, _ = s.run(oracle.HEAT, [1e-3, 1.0], 20, u, tol=1e-14)
    assert sol.advance_time(20) == _lib.FEM_OK
    assert rel(sol.get_state_unknowns(), u) < RTOL
    assert sol.last_stats.steps == 20 and sol.last_stats.iterations > 0


@pytest.mark.parametrize("engine", [_lib.FEM_ENGINE_PERSISTENT, _lib.FEM_ENGINE_PLAIN])
def test_advection_diffusion_vs_oracle(gpu, engine):
    m = meshes.jittered_square(32, seed=2)
    vel = (0.6, 0.8, 0.0)
    s = oracle.System(m, bc=0, velocity=vel)
    ctx = FEMContext(m, 0, velocity=vel)
    p = AdvectionDiffusionParameters(1e-3, 0.25, vel)
    sol = make_solver(ctx, Equation.Advection_Diffusion, p, engine=engine,
                      precision=_lib.FEM_PREC_FP32_IR if engine == 0 else _lib.FEM_PREC_FP64)
    u = s.gather(gauss(m, (0.4, 0.45), 0.02))
    sol.set_state_unknowns(_lib.FEM_FIELD_U, u)
    for k in range(5):
        _, u, _, _ = s.step(oracle.ADVDIFF, [1e-3, 0.25], u, krylov=oracle.BICGSTAB, tol=1e-14)
        assert sol.advance_time(1) == _lib.FEM_OK
        assert rel(sol.get_state_unknowns(), u) < RTOL, k


def test_wave_reference_scheme_vs_golden(gpu):
    g = dict(np.load(os.path.join(GOLD, "square64.npz")))
    m = {"xyz": g["xyz"], "tri": g["tri"], "on_boundary": g["on_boundary"]}
    ctx = FEMContext(m, 0)
    sol = make_solver(ctx, Equation.Wave, WaveParameters(0.05, 0.05))
    sol.set_state_unknowns(_lib.FEM_FIELD_U, g["u0"])
    for k in range(3):
        assert sol.advance_time(1) == _lib.FEM_OK
        assert rel(sol.get_state_unknowns(_lib.FEM_FIELD_U), g["wave_u"][k]) < RTOL
        assert rel(sol.get_state_unknowns(_lib.FEM_FIELD_V), g["wave_v"][k]) < RTOL


def test_grayscott_nodal_shipped_icosphere_golden(gpu):
    """Reference defaults on the reference's own icosphere.obj (u starts at 0)."""
    g = dict(np.load(os.path.join(GOLD, "icosphere_shipped.npz")))
    m = {"xyz": g["xyz"], "tri": g["tri"], "on_boundary": g["on_boundary"]}
    ctx = FEMContext(m, 1)
    sol = make_solver(ctx, Equation.Reaction_Diffusion, ReactionDiffusionParameters())
    sol.set_state_unknowns(_lib.FEM_FIELD_V, g["v0"])
    for k in range(3):
        assert sol.advance_time(1) == _lib.FEM_OK
        assert rel(sol.get_state_unknowns(_lib.FEM_FIELD_U), g["gs_u"][k]) < RTOL
        assert rel(sol.get_state_unknowns(_lib.FEM_FIELD_V), g["gs_v"][k]) < RTOL


@pytest.mark.parametrize("precision", [_lib.FEM_PREC_FP32_IR, _lib.FEM_PREC_FP64])
def test_grayscott_lumped_vs_oracle(gpu, precision):
    m = meshes.icosphere(3)
    s = oracle.System(m, bc=1)
    rng = np.random.default_rng(0)
    u = np.ones(s.n)
    v = np.zeros(s.n)
    idx = rng.choice(s.n, 30, replace=False)
    v[idx] = 0.25 + rng.uniform(0, 0.05, 30)
    prm = ReactionDiffusionParameters(0.5, 2e-4, 1e-4, 0.037, 0.06, _lib.FEM_REACTION_LUMPED)
    ctx = FEMContext(m, 1)
    sol = make_solver(ctx, Equation.Reaction_Diffusion, prm, precision=precision)
    sol.set_state_unknowns(_lib.FEM_FIELD_U, u)
    sol.set_state_unknowns(_lib.FEM_FIELD_V, v)
    op = [0.5, 2e-4, 1e-4, 0.037, 0.06, 1]
    for k in range(20):
        _, u, v, _ = s.step(oracle.GRAYSCOTT, op, u, v, tol=1e-14)
    assert sol.advance_time(20) == _lib.FEM_OK
    assert rel(sol.get_state_unknowns(_lib.FEM_FIELD_U), u) < RTOL
    assert rel(sol.get_state_unknowns(_lib.FEM_FIELD_V), v) < RTOL


def test_newmark_vs_oracle_and_energy(gpu):
    m = meshes.jittered_square(32, seed=6)
    s = oracle.System(m, bc=0)
    u = s.gather(np.exp(-((m["xyz"][:, 0] - 0.5) ** 2 + (m["xyz"][:, 1] - 0.5) ** 2) / (2 * 0.05 ** 2)))
    v = np.zeros(s.n)
    a = s.newmark_init(1.0, u)
    ctx = FEMContext(m, 0)
    sol = make_solver(ctx, Equation.Wave_Newmark, NewmarkParameters(2e-3, 1.0, 0.25, 0.5))
    sol.set_state_unknowns(_lib.FEM_FIELD_U, u)
    sol.set_state_unknowns(_lib.FEM_FIELD_A, a)
    for _ in range(15):
        _, u, v, a = s.step(oracle.NEWMARK, [2e-3, 1.0, 0.25, 0.5], u, v, a, tol=1e-14)
    assert sol.advance_time(15) == _lib.FEM_OK
    assert rel(sol.get_state_unknowns(_lib.FEM_FIELD_U), u) < RTOL
    assert rel(sol.get_state_unknowns(_lib.FEM_FIELD_V), v) < RTOL


def test_batch_uneven_meshes_multi_rhs(gpu):
    specs = [(12, 1), (20, 2), (7, 3)]
    ctxs, systems, states = [], [], []
    rng = np.random.default_rng(5)
    k = 4
    for side, seed in specs:
        m = meshes.jittered_square(side, seed=seed)
        ctxs.append(FEMContext(m, 0))
        s = oracle.System(m, bc=0)
        systems.append(s)
        states.append(rng.normal(size=(s.n, k)))
    b = Batch(ctxs, SolverOptions(nrhs=k))
    kappas = [1e-1, 1e-2, 3e-3]
    for i, kap in enumerate(kappas):
        b.set_equation(i, HeatParameters(1e-3, kap))
        b.set_state_unknowns(i, _lib.FEM_FIELD_U, states[i])
    assert b.advance(3) == _lib.FEM_OK
    # per-instance counters add up to the call's totals
    per = [b.instance_stats(i) for i in range(len(specs))]
    assert sum(p.iterations for p in per) == b.last_stats.iterations
    assert sum(p.sweeps for p in per) == b.last_stats.sweeps
    assert all(p.steps == 3 and p.iterations > 0 for p in per)
    for i, s in enumerate(systems):
        got = b.get_state_unknowns(i).reshape(s.n, k)
        for j in range(k):
            _, u, _, _ = s.run(oracle.HEAT, [1e-3, kappas[i]], 3, states[i][:, j].copy(), tol=1e-14)
            assert rel(got[:, j], u) < RTOL, (i, j)


def test_batch_advection_diffusion_16_rhs(gpu):
    """16 right-hand sides per instance with BiCGSTAB on uneven meshes with
    per-instance unnormalised velocity and diffusivity."""
    specs = [(10, 1, (0.6, 0.2, 0.0), 3e-2), (16, 2, (-0.1, 0.9, 0.0), 1e-3), (6, 3, (0.3, -0.3, 0.0), 1e-1)]
    k, dt, T = 16, 1e-3, 3
    rng = np.random.default_rng(17)
    ctxs, systems, states = [], [], []
    for side, seed, vel, _ in specs:
        m = meshes.jittered_square(side, seed=seed)
        ctxs.append(FEMContext(m, 0, velocity=vel, normalize_velocity=False))
        s = oracle.System(m, bc=0, velocity=vel, normalize=False)
        systems.append(s)
        states.append(rng.normal(size=(s.n, k)))
    b = Batch(ctxs, SolverOptions(nrhs=k))
    for i, spec in enumerate(specs):
        b.set_equation(i, AdvectionDiffusionParameters(dt, spec[3]))
        b.set_state_unknowns(i, _lib.FEM_FIELD_U, states[i])
    print("case batch advdiff 16 rhs", flush=True)
    assert b.advance(T) == _lib.FEM_OK
    for i, s in enumerate(systems):
        got = b.get_state_unknowns(i).reshape(s.n, k)
        for j in range(k):
            _, u, _, _ = s.run(oracle.ADVDIFF, [dt, specs[i][3]], T, states[i][:, j].copy(), krylov=oracle.BICGSTAB,
                               tol=1e-14)
            assert rel(got[:, j], u) < RTOL, (i, j)


@pytest.mark.parametrize("krylov", [_lib.FEM_KRYLOV_CG, _lib.FEM_KRYLOV_BICGSTAB, _lib.FEM_KRYLOV_CG_CLASSIC])
def test_fenced_exchange_large_row_blocks(gpu, krylov):
    """Three workgroups over ~2,200 rows x 4 columns: each workgroup streams
    more rows than its LDS row cache holds, so the launch uses fenced
    barriers with cached data accesses."""
    m = meshes.jittered_square(48, seed=6)
    vel = (0.4, 0.7, 0.0)
    k, dt, c, T = 4, 1e-3, 2e-2, 4
    ctx = FEMContext(m, 0, velocity=vel, normalize_velocity=False)
    s = oracle.System(m, bc=0, velocity=vel, normalize=False)
    rng = np.random.default_rng(23)
    x0 = rng.normal(size=(s.n, k))
    b = Batch([ctx], SolverOptions(nrhs=k, team_size=3, krylov=krylov))
    b.set_equation(0, AdvectionDiffusionParameters(dt, c))
    b.set_state_unknowns(0, _lib.FEM_FIELD_U, x0)
    print("case fenced", krylov, flush=True)
    assert b.advance(T) == _lib.FEM_OK
    got = b.get_state_unknowns(0).reshape(s.n, k)
    for j in range(k):
        _, u, _, _ = s.run(oracle.ADVDIFF, [dt, c], T, x0[:, j].copy(), krylov=oracle.BICGSTAB, tol=1e-14)
        assert rel(got[:, j], u) < RTOL, j


def test_surface_state_and_instability(gpu):
    m = meshes.structured_square(16)
    ctx = FEMContext(m, 0)
    sol = make_solver(ctx, Equation.Heat, HeatParameters(1e-3, 1.0))
    vals = gauss(m).astype(np.float32)
    sol.set_state(_lib.FEM_FIELD_U, vals)
    got = sol.get_state(_lib.FEM_FIELD_U)
    onb = m["on_boundary"].astype(bool)
    assert np.all(got[onb] == 0.0)  # CPUSolver::map_vector_to_surface writes 0 on the boundary
    assert np.array_equal(got[~onb], vals[~onb])
    assert sol.advance_time(1) == _lib.FEM_OK and not sol.has_numerical_instability()
    vals[40] = 5e6  # the reference rule: u > 1e4
    sol.set_state(_lib.FEM_FIELD_U, vals)
    assert sol.advance_time(1) == _lib.FEM_DIVERGED
    assert sol.has_numerical_instability()
    sol.clear_values()
    assert not np.any(sol.get_state(_lib.FEM_FIELD_U))


def test_deterministic_bitwise(gpu):
    m = meshes.jittered_square(48, seed=3)
    out = []
    for _ in range(2):
        ctx = FEMContext(m, 0)
        sol = make_solver(ctx, Equation.Heat, HeatParameters(1e-3, 1.0))
        sol.set_state_unknowns(_lib.FEM_FIELD_U, np.linspace(0, 1, ctx.num_unknowns()))
        sol.advance_time(5)
        out.append(sol.get_state_unknowns())
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("k,krylov", [(4, _lib.FEM_KRYLOV_CG), (16, _lib.FEM_KRYLOV_BICGSTAB)])
def test_batch_graph_replays_match_eager(gpu, k, krylov):
    """A recorded advance (HIP graph: slot-reset kernel + stepping kernel)
    replayed several times, with eager work and host reads between the
    replays, matches an identical batch advanced eagerly bit for bit after
    every replay, and against the oracle; a rebuilt plan makes it stale."""
    print("case graph replay", k, krylov, flush=True)
    rng = np.random.default_rng(11)
    specs = [(14, 1), (22, 2), (9, 3)]
    vel = (0.4, -0.2, 0.0)
    ctxs, systems, states = [], [], []
    for side, seed in specs:
        m = meshes.jittered_square(side, seed=seed)
        ctxs.append(FEMContext(m, 0, velocity=vel, normalize_velocity=False))
        systems.append(oracle.System(m, bc=0, velocity=vel, normalize=False))
        states.append(rng.normal(size=(systems[-1].n, k)))
    adv = krylov == _lib.FEM_KRYLOV_BICGSTAB
    prm = AdvectionDiffusionParameters(1e-3, 2e-2) if adv else HeatParameters(1e-3, 2e-2)
    eq, oprm = (oracle.ADVDIFF, [1e-3, 2e-2])

    def make():
        b = Batch(ctxs, SolverOptions(nrhs=k, krylov=krylov))
        for i in range(len(ctxs)):
            b.set_equation(i, prm)
            b.set_state_unknowns(i, _lib.FEM_FIELD_U, states[i])
        return b
    g, e = make(), make()
    g.graph_capture(2)
    ref = [s.copy() for s in states]
    for rep in range(3):
        assert g.graph_launch() == _lib.FEM_OK
        assert g.last_stats.steps == 2
        assert e.advance(2) == _lib.FEM_OK
        for i, s in enumerate(systems):
            a, b_ = g.get_state_unknowns(i), e.get_state_unknowns(i)
            assert np.array_equal(a, b_), (rep, i)
            for j in range(k):
                _, ref_j, _, _ = s.run(eq if adv else oracle.HEAT, oprm, 2, ref[i][:, j].copy(),
                                       krylov=oracle.BICGSTAB if adv else oracle.CG, tol=1e-14)
                ref[i][:, j] = ref_j
            assert rel(a.reshape(s.n, k), ref[i]) < RTOL, (rep, i)
    # a state write between replays is picked up by the next replay
    for b in (g, e):
        b.set_state_unknowns(1, _lib.FEM_FIELD_U, states[1])
    assert g.graph_launch() == _lib.FEM_OK and e.advance(2) == _lib.FEM_OK
    for i in range(len(ctxs)):
        assert np.array_equal(g.get_state_unknowns(i), e.get_state_unknowns(i))
    # new parameters rebuild the matrices and the plan: the graph is stale
    g.set_equation(0, AdvectionDiffusionParameters(2e-3, 2e-2) if adv else HeatParameters(2e-3, 2e-2))
    with pytest.raises(_lib.FemError):
        g.graph_launch()
    g.graph_capture(1)
    assert g.graph_launch() == _lib.FEM_OK
    g.close()
    e.close()


def test_batch_mixed_team_sizes(gpu):
    """One large and several tiny meshes: the plan mixes multi-workgroup teams
    with single-workgroup teams that each walk several instances."""
    sides = [64, 4, 5, 6, 4, 5, 6, 7]
    ctxs, systems, states = [], [], []
    rng = np.random.default_rng(9)
    for i, side in enumerate(sides):
        m = meshes.jittered_square(side, seed=20 + i)
        ctxs.append(FEMContext(m, 0))
        s = oracle.System(m, bc=0)
        systems.append(s)
        states.append(rng.normal(size=s.n))
    b = Batch(ctxs, SolverOptions())
    for i in range(len(sides)):
        b.set_equation(i, HeatParameters(2e-3, 0.05 * (i + 1)))
        b.set_state_unknowns(i, _lib.FEM_FIELD_U, states[i])
    assert b.advance(4) == _lib.FEM_OK
    for i, s in enumerate(systems):
        _, u, _, _ = s.run(oracle.HEAT, [2e-3, 0.05 * (i + 1)], 4, states[i].copy(), tol=1e-14)
        assert rel(b.get_state_unknowns(i), u) < RTOL, i


@pytest.mark.parametrize("team", [3, 37, 400])
def test_team_sizes_and_reduction_modes(gpu, team):
    """Explicit team sizes, including one large enough that rank 0 reduces and
    publishes the sums (Gray-Scott: 2 columns x 3 partials x 400 > 2048)."""
    m = meshes.icosphere(4)
    s = oracle.System(m, bc=1)
    rng = np.random.default_rng(2)
    u = np.ones(s.n)
    v = np.where(rng.uniform(size=s.n) < 0.05, 0.25, 0.0)
    prm = ReactionDiffusionParameters(0.5, 2e-4, 1e-4, 0.037, 0.06, _lib.FEM_REACTION_LUMPED)
    ctx = FEMContext(m, 1)
    sol = make_solver(ctx, Equation.Reaction_Diffusion, prm, team_size=team)
    sol.set_state_unknowns(_lib.FEM_FIELD_U, u)
    sol.set_state_unknowns(_lib.FEM_FIELD_V, v)
    for _ in range(6):
        _, u, v, _ = s.step(oracle.GRAYSCOTT, [0.5, 2e-4, 1e-4, 0.037, 0.06, 1], u, v, tol=1e-14)
    assert sol.advance_time(6) == _lib.FEM_OK
    assert rel(sol.get_state_unknowns(_lib.FEM_FIELD_U), u) < RTOL
    assert rel(sol.get_state_unknowns(_lib.FEM_FIELD_V), v) < RTOL


STREAM_CASES = ["heat_cg", "heat_classic", "advdiff_batch16", "advdiff_k1", "grayscott", "newmark", "wave_fp64"]


@pytest.mark.parametrize("case", STREAM_CASES)
def test_stream_engine_matches_oracle(gpu, case):
    """The streaming engine (one launch per phase over all row blocks) against
    the oracle, for every Krylov method, several equations and a batch."""
    print("case stream", case, flush=True)
    rng = np.random.default_rng(31)
    if case == "advdiff_batch16":
        specs = [(18, 1, (0.6, 0.2, 0.0), 3e-2), (9, 2, (-0.1, 0.9, 0.0), 1e-3), (25, 3, (0.3, -0.3, 0.0), 1e-1)]
        k, dt, T = 16, 1e-3, 3
        ctxs, systems, states = [], [], []
        for side, seed, vel, _ in specs:
            m = meshes.jittered_square(side, seed=seed)
            ctxs.append(FEMContext(m, 0, velocity=vel, normalize_velocity=False))
            systems.append(oracle.System(m, bc=0, velocity=vel, normalize=False))
            states.append(rng.normal(size=(systems[-1].n, k)))
        b = Batch(ctxs, SolverOptions(nrhs=k, engine=_lib.FEM_ENGINE_STREAM))
        for i, spec in enumerate(specs):
            b.set_equation(i, AdvectionDiffusionParameters(dt, spec[3]))
            b.set_state_unknowns(i, _lib.FEM_FIELD_U, states[i])
        assert b.advance(T) == _lib.FEM_OK
        for i, s in enumerate(systems):
            got = b.get_state_unknowns(i).reshape(s.n, k)
            for j in range(k):
                _, u, _, _ = s.run(oracle.ADVDIFF, [dt, specs[i][3]], T, states[i][:, j].copy(),
                                   krylov=oracle.BICGSTAB, tol=1e-14)
                assert rel(got[:, j], u) < RTOL, (i, j)
        return
    if case == "advdiff_k1":  # one column: FP32 BiCGSTAB on the staged one-column layout
        m = meshes.jittered_square(40, seed=8)
        vel = (0.6, 0.8, 0.0)
        s = oracle.System(m, bc=0, velocity=vel)
        ctx = FEMContext(m, 0, velocity=vel)
        sol = make_solver(ctx, Equation.Advection_Diffusion, AdvectionDiffusionParameters(1e-3, 0.05, vel),
                          engine=_lib.FEM_ENGINE_STREAM)
        u = s.gather(gauss(m, (0.4, 0.45), 0.02))
        sol.set_state_unknowns(_lib.FEM_FIELD_U, u)
        _, u, _, _ = s.run(oracle.ADVDIFF, [1e-3, 0.05], 5, u.copy(), krylov=oracle.BICGSTAB, tol=1e-14)
        assert sol.advance_time(5) == _lib.FEM_OK
        assert rel(sol.get_state_unknowns(_lib.FEM_FIELD_U), u) < RTOL
        return
    if case == "grayscott":
        m = meshes.icosphere(3)
        s = oracle.System(m, bc=1)
        u = np.ones(s.n)
        v = np.where(rng.uniform(size=s.n) < 0.05, 0.25, 0.0)
        prm = ReactionDiffusionParameters(0.5, 2e-4, 1e-4, 0.037, 0.06, _lib.FEM_REACTION_LUMPED)
        ctx = FEMContext(m, 1)
        sol = make_solver(ctx, Equation.Reaction_Diffusion, prm, engine=_lib.FEM_ENGINE_STREAM)
        sol.set_state_unknowns(_lib.FEM_FIELD_U, u)
        sol.set_state_unknowns(_lib.FEM_FIELD_V, v)
        for _ in range(5):
            _, u, v, _ = s.step(oracle.GRAYSCOTT, [0.5, 2e-4, 1e-4, 0.037, 0.06, 1], u, v, tol=1e-14)
        assert sol.advance_time(5) == _lib.FEM_OK
        assert rel(sol.get_state_unknowns(_lib.FEM_FIELD_U), u) < RTOL
        assert rel(sol.get_state_unknowns(_lib.FEM_FIELD_V), v) < RTOL
        return
    m = meshes.jittered_square(40, seed=8)
    s = oracle.System(m, bc=0)
    u = s.gather(gauss(m, (0.45, 0.5), 0.02))
    ctx = FEMContext(m, 0)
    if case == "newmark":
        a = s.newmark_init(1.0, u)
        sol = make_solver(ctx, Equation.Wave_Newmark, NewmarkParameters(2e-3, 1.0, 0.25, 0.5),
                          engine=_lib.FEM_ENGINE_STREAM)
        sol.set_state_unknowns(_lib.FEM_FIELD_U, u)
        sol.set_state_unknowns(_lib.FEM_FIELD_A, a)
        v = np.zeros(s.n)
        for _ in range(6):
            _, u, v, a = s.step(oracle.NEWMARK, [2e-3, 1.0, 0.25, 0.5], u, v, a, tol=1e-14)
        assert sol.advance_time(6) == _lib.FEM_OK
    elif case == "wave_fp64":
        sol = make_solver(ctx, Equation.Wave, WaveParameters(5e-3, 0.8), engine=_lib.FEM_ENGINE_STREAM,
                          precision=_lib.FEM_PREC_FP64)
        sol.set_state_unknowns(_lib.FEM_FIELD_U, u)
        v = np.zeros(s.n)
        for _ in range(6):
            _, u, v, _ = s.step(oracle.WAVE, [5e-3, 0.8], u, v, tol=1e-14)
        assert sol.advance_time(6) == _lib.FEM_OK
    else:
        kr = _lib.FEM_KRYLOV_CG if case == "heat_cg" else _lib.FEM_KRYLOV_CG_CLASSIC
        sol = make_solver(ctx, Equation.Heat, HeatParameters(1e-3, 1.0), engine=_lib.FEM_ENGINE_STREAM, krylov=kr)
        sol.set_state_unknowns(_lib.FEM_FIELD_U, u)
        _, u, _, _ = s.run(oracle.HEAT, [1e-3, 1.0], 6, u.copy(), tol=1e-14)
        assert sol.advance_time(6) == _lib.FEM_OK
    assert rel(sol.get_state_unknowns(_lib.FEM_FIELD_U), u) < RTOL


STREAM_MULTI = [
    ("heat", _lib.FEM_KRYLOV_CG, 4, _lib.FEM_PREC_FP32_IR),
    ("heat", _lib.FEM_KRYLOV_CG_CLASSIC, 16, _lib.FEM_PREC_FP64),
    ("heat", _lib.FEM_KRYLOV_CG, 16, _lib.FEM_PREC_FP64),
    ("wave", _lib.FEM_KRYLOV_CG, 4, _lib.FEM_PREC_FP64),
    ("newmark", _lib.FEM_KRYLOV_CG_CLASSIC, 16, _lib.FEM_PREC_FP32_IR),
    ("advdiff", _lib.FEM_KRYLOV_BICGSTAB, 4, _lib.FEM_PREC_FP64),
]


@pytest.mark.parametrize("eq,krylov,k,prec", STREAM_MULTI)
def test_stream_engine_multi_rhs(gpu, eq, krylov, k, prec):
    """The streaming engine's multi-column phase bodies (each lane owns 16
    bytes of adjacent columns) for every Krylov method, both inner
    precisions and the heat, wave, Newmark and advection-diffusion
    right-hand sides, column by column against the oracle on two uneven
    meshes."""
    print("case stream multi", eq, krylov, k, prec, flush=True)
    rng = np.random.default_rng(41 + k)
    vel = (0.5, -0.3, 0.0)
    dt, T = 2e-3, 4
    ctxs, systems, u0s, a0s = [], [], [], []
    for side, seed in ((15, 1), (23, 2)):
        m = meshes.jittered_square(side, seed=seed)
        ctxs.append(FEMContext(m, 0, velocity=vel, normalize_velocity=False))
        s = oracle.System(m, bc=0, velocity=vel, normalize=False)
        systems.append(s)
        u0 = rng.normal(size=(s.n, k))
        u0s.append(u0)
        if eq == "newmark":
            a0s.append(np.stack([s.newmark_init(1.0, u0[:, j]) for j in range(k)], axis=1))
    b = Batch(ctxs, SolverOptions(nrhs=k, engine=_lib.FEM_ENGINE_STREAM, krylov=krylov, precision=prec))
    params = {"heat": (HeatParameters(dt, 0.05), oracle.HEAT, [dt, 0.05]),
              "wave": (WaveParameters(dt, 0.9), oracle.WAVE, [dt, 0.9]),
              "newmark": (NewmarkParameters(dt, 1.0, 0.25, 0.5), oracle.NEWMARK, [dt, 1.0, 0.25, 0.5]),
              "advdiff": (AdvectionDiffusionParameters(dt, 2e-2), oracle.ADVDIFF, [dt, 2e-2])}[eq]
    for i in range(len(ctxs)):
        b.set_equation(i, params[0])
        b.set_state_unknowns(i, _lib.FEM_FIELD_U, u0s[i])
        if eq == "newmark":
            b.set_state_unknowns(i, _lib.FEM_FIELD_A, a0s[i])
    assert b.advance(T) == _lib.FEM_OK
    okr = oracle.BICGSTAB if krylov == _lib.FEM_KRYLOV_BICGSTAB else oracle.CG
    for i, s in enumerate(systems):
        got = b.get_state_unknowns(i).reshape(s.n, k)
        gotv = b.get_state_unknowns(i, _lib.FEM_FIELD_V).reshape(s.n, k)
        for j in range(k):
            a0 = a0s[i][:, j].copy() if eq == "newmark" else None
            _, u, v, _ = s.run(params[1], params[2], T, u0s[i][:, j].copy(), None, a0, krylov=okr, tol=1e-14)
            assert rel(got[:, j], u) < RTOL, (i, j)
            if eq in ("wave", "newmark"):
                assert rel(gotv[:, j], v) < RTOL, (i, j)
    b.close()


@pytest.mark.parametrize("precision", [_lib.FEM_PREC_FP32_IR, _lib.FEM_PREC_FP64])
def test_fused_steps_with_and_without_sweeps(gpu, precision):
    """Inside one launch, a step whose initial guess already meets the
    tolerance runs no refinement sweep; the next step's update then forms
    its right-hand side by gathering the solution, while after a sweep it
    takes b and r from the last residual pass (and the pipelined CG starts
    in that phase).  Heat under Neumann conditions from 1 + 1e-2 noise,
    with the previous solution as the initial guess: the fast modes die out
    within a few steps, after which that guess of the nearly constant state
    meets a loose tolerance, so one launch runs both kinds of step.  (An
    extrapolating guess amplifies the tolerance-level error of its history
    and nearly always needs a sweep.)  The launch must match single-step
    launches, which form every right-hand side from the state, and the
    oracle."""
    print("case fused sweeps mix", precision, flush=True)
    m = meshes.jittered_square(40, seed=8)
    s = oracle.System(m, bc=1)
    u0 = 1.0 + 1e-2 * np.random.default_rng(3).uniform(-1.0, 1.0, s.n)
    nsteps = 24

    def run(tol, split):
        sol = make_solver(FEMContext(m, 1), Equation.Heat, HeatParameters(1e-3, 1.0), precision=precision, tol=tol,
                          predictor=_lib.FEM_PREDICTOR_PREVIOUS)
        sol.set_state_unknowns(_lib.FEM_FIELD_U, u0)
        sweeps = []
        for _ in range(nsteps if split else 1):
            assert sol.advance_time(1 if split else nsteps) == _lib.FEM_OK
            sweeps.append(sol.last_stats.sweeps)
        return sol.get_state_unknowns(), sweeps

    # per-step sweeps of single-step launches (same initial guesses as one launch)
    for tol in (2e-5, 1e-5, 3e-5, 5e-6, 5e-5):
        steps1, sweeps = run(tol, True)
        print("tol", tol, "sweeps per step", sweeps, flush=True)
        if 0 in sweeps[1:-1] and any(sweeps[1:-1]):
            break
    else:
        pytest.fail("no tolerance gave steps with and without sweeps")
    whole, wsweeps = run(tol, False)
    print("one launch: sweeps", wsweeps, flush=True)
    assert 0 < wsweeps[0] < nsteps
    # the perturbation about 1, which the solves resolve to tol |b| (|b| ~ |M 1| / dt)
    assert rel(whole - 1.0, steps1 - 1.0) < 1e-3
    # the oracle's CG from the same guess with the same stopping rule: it also
    # stops at once where the guess meets the tolerance
    _, ref, _, _ = s.run(oracle.HEAT, [1e-3, 1.0], nsteps, u0.copy(), tol=tol, warm=True)
    assert rel(whole, ref) < 10 * tol


@pytest.mark.parametrize("pred", [_lib.FEM_PREDICTOR_LINEAR, _lib.FEM_PREDICTOR_QUADRATIC, _lib.FEM_PREDICTOR_CUBIC,
                                  _lib.FEM_PREDICTOR_QUARTIC, _lib.FEM_PREDICTOR_QUINTIC, _lib.FEM_PREDICTOR_SEXTIC])
@pytest.mark.parametrize("case", ["heat_persistent", "grayscott_persistent", "advdiff16_stream", "heat_stream",
                                  "newmark_persistent", "newmark_stream"])
def test_predictors(gpu, case, pred):
    """The linear (2 x_n - x_{n-1}) and quadratic (3 x_n - 3 x_{n-1} +
    x_{n-2}, x_{n-2} kept in the v field) initial guesses change only where
    each solve starts: results stay within the tolerance of the oracle, over
    launches of one and of several steps (the history carries across them)."""
    print("case predictor", case, pred, flush=True)
    Q = pred
    eng = _lib.FEM_ENGINE_STREAM if case.endswith("stream") else _lib.FEM_ENGINE_PERSISTENT
    if case == "grayscott_persistent":
        m = meshes.icosphere(3)
        s = oracle.System(m, bc=1)
        rng = np.random.default_rng(2)
        u, v = np.ones(s.n), np.where(rng.uniform(size=s.n) < 0.05, 0.25, 0.0)
        prm = ReactionDiffusionParameters(0.5, 2e-4, 1e-4, 0.037, 0.06, _lib.FEM_REACTION_LUMPED)
        sol = make_solver(FEMContext(m, 1), Equation.Reaction_Diffusion, prm, predictor=Q, engine=eng)
        sol.set_state_unknowns(_lib.FEM_FIELD_U, u)
        sol.set_state_unknowns(_lib.FEM_FIELD_V, v)
        op = [0.5, 2e-4, 1e-4, 0.037, 0.06, 1]
        for n in (1, 1, 4, 6):
            for _ in range(n):
                _, u, v, _ = s.step(oracle.GRAYSCOTT, op, u, v, tol=1e-14)
            assert sol.advance_time(n) == _lib.FEM_OK
            assert rel(sol.get_state_unknowns(_lib.FEM_FIELD_U), u) < RTOL
            assert rel(sol.get_state_unknowns(_lib.FEM_FIELD_V), v) < RTOL
        return
    if case.startswith("newmark"):  # the acceleration is extrapolated (linear or quadratic)
        m = meshes.jittered_square(36, seed=5)
        s = oracle.System(m, bc=0)
        u = s.gather(gauss(m, (0.5, 0.45), 0.02))
        a = s.newmark_init(1.0, u)
        v = np.zeros(s.n)
        prm = [1e-3, 1.0, 0.25, 0.5]
        sol = make_solver(FEMContext(m, 0), Equation.Wave_Newmark, NewmarkParameters(*prm), predictor=Q, engine=eng)
        sol.set_state_unknowns(_lib.FEM_FIELD_U, u)
        sol.set_state_unknowns(_lib.FEM_FIELD_A, a)
        for n in (1, 2, 6):
            for _ in range(n):
                _, u, v, a = s.step(oracle.NEWMARK, prm, u, v, a, tol=1e-14)
            assert sol.advance_time(n) == _lib.FEM_OK
            assert rel(sol.get_state_unknowns(_lib.FEM_FIELD_U), u) < RTOL, n
            assert rel(sol.get_state_unknowns(_lib.FEM_FIELD_V), v) < RTOL, n
            assert rel(sol.get_state_unknowns(_lib.FEM_FIELD_A), a) < RTOL, n
        return
    if case == "advdiff16_stream":
        k, vel = 16, (0.5, 0.3, 0.0)
        rng = np.random.default_rng(9)
        ctxs, systems, states = [], [], []
        for side, seed in ((17, 1), (11, 2)):
            m = meshes.jittered_square(side, seed=seed)
            ctxs.append(FEMContext(m, 0, velocity=vel, normalize_velocity=False))
            systems.append(oracle.System(m, bc=0, velocity=vel, normalize=False))
            states.append(rng.normal(size=(systems[-1].n, k)))
        b = Batch(ctxs, SolverOptions(nrhs=k, engine=eng, predictor=Q))
        for i in range(2):
            b.set_equation(i, AdvectionDiffusionParameters(1e-3, 3e-2))
            b.set_state_unknowns(i, _lib.FEM_FIELD_U, states[i])
        for n in (1, 5):
            assert b.advance(n) == _lib.FEM_OK
            for i, s in enumerate(systems):
                for j in range(k):
                    _, states[i][:, j], _, _ = s.run(oracle.ADVDIFF, [1e-3, 3e-2], n, states[i][:, j].copy(),
                                                      krylov=oracle.BICGSTAB, tol=1e-14)
                assert rel(b.get_state_unknowns(i).reshape(s.n, k), states[i]) < RTOL, (n, i)
        b.close()
        return
    m = meshes.jittered_square(40, seed=8)
    s = oracle.System(m, bc=0)
    u = s.gather(gauss(m, (0.45, 0.5), 0.02))
    sol = make_solver(FEMContext(m, 0), Equation.Heat, HeatParameters(1e-3, 1.0), predictor=Q, engine=eng)
    sol.set_state_unknowns(_lib.FEM_FIELD_U, u)
    for n in (1, 2, 7):
        _, u, _, _ = s.run(oracle.HEAT, [1e-3, 1.0], n, u.copy(), tol=1e-14)
        assert sol.advance_time(n) == _lib.FEM_OK
        assert rel(sol.get_state_unknowns(), u) < RTOL, n


@pytest.mark.parametrize("pred", [_lib.FEM_PREDICTOR_CUBIC, _lib.FEM_PREDICTOR_SEXTIC])
def test_history_across_engines_and_autotune(gpu, pred):
    """The streaming engine turns the predictor history by pointer, the
    persistent kernel shifts it by copies: a trajectory that starts on the
    persistent kernel, goes through the autotuner's probes of both engines
    (each restoring the state) and continues on the chosen one stays on the
    oracle's trajectory, and so do stream launches of three, one and two steps
    (turn counts that the history slots do not divide).  Sextic: the
    persistent kernel uses quartic at most and leaves the two oldest slots to
    the streaming engine."""
    print("case history across engines", pred, flush=True)
    k, vel, prm = 16, (0.5, 0.3, 0.0), [1e-3, 3e-2]
    rng = np.random.default_rng(21)
    ctxs, systems, states = [], [], []
    for side, seed in ((17, 3), (11, 4)):
        m = meshes.jittered_square(side, seed=seed)
        ctxs.append(FEMContext(m, 0, velocity=vel, normalize_velocity=False))
        systems.append(oracle.System(m, bc=0, velocity=vel, normalize=False))
        states.append(rng.normal(size=(systems[-1].n, k)))

    def oracle_steps(n):
        for i, s in enumerate(systems):
            for j in range(k):
                _, states[i][:, j], _, _ = s.run(oracle.ADVDIFF, prm, n, states[i][:, j].copy(),
                                                  krylov=oracle.BICGSTAB, tol=1e-14)

    def check(b, tag):
        for i, s in enumerate(systems):
            assert rel(b.get_state_unknowns(i).reshape(s.n, k), states[i]) < RTOL, (tag, i)

    def batch(engine):
        b = Batch(ctxs, SolverOptions(nrhs=k, engine=engine, predictor=pred))
        for i in range(2):
            b.set_equation(i, AdvectionDiffusionParameters(*prm))
        return b

    b = batch(_lib.FEM_ENGINE_PERSISTENT)
    for i in range(2):
        b.set_state_unknowns(i, _lib.FEM_FIELD_U, states[i])
    assert b.advance(3) == _lib.FEM_OK
    oracle_steps(3)
    check(b, "persistent")
    b.autotune(2)
    assert b.advance(4) == _lib.FEM_OK
    oracle_steps(4)
    check(b, "after autotune")
    b.close()
    # stream launches of 3 and 1 steps (history of 3 slots turned 3 and 4 times)
    b = batch(_lib.FEM_ENGINE_STREAM)
    for i in range(2):
        b.set_state_unknowns(i, _lib.FEM_FIELD_U, states[i])
    for n in (3, 1, 2):
        assert b.advance(n) == _lib.FEM_OK
        oracle_steps(n)
        check(b, f"stream {n}")
    b.close()


def test_stream_engine_adjoint_matches_persistent(gpu):
    from femhip import Adjoint
    m = meshes.jittered_square(24, seed=12)
    ctx = FEMContext(m, 0, velocity=(0.4, 0.2, 0.0), normalize_velocity=False)
    rng = np.random.default_rng(4)
    u0 = rng.normal(size=ctx.num_unknowns())
    g = rng.normal(size=ctx.num_unknowns())
    out = []
    for eng in (_lib.FEM_ENGINE_PERSISTENT, _lib.FEM_ENGINE_STREAM):
        adj = Adjoint(ctx, AdvectionDiffusionParameters(2e-3, 1e-2),
                      SolverOptions(engine=eng, precision=_lib.FEM_PREC_FP64, tol=1e-13))
        uT = adj.forward(u0, 7, 3)
        gu, gk = adj.backward(g)
        out.append((uT, gu, gk))
        adj.close()
    for a_, b_ in zip(out[0], out[1]):
        assert rel(b_, a_) < 1e-10


@pytest.mark.parametrize("eq", ["heat", "advdiff", "wave", "grayscott"])
def test_gl_compat_engine_vs_oracle(gpu, eq):
    """The GL-compatible engine against the float32 restatement of the
    reference's OpenGL solver (10 CG iterations per solve, warm start)."""
    print("case gl", eq, flush=True)
    if eq == "grayscott":
        m = meshes.icosphere(3)
        s = oracle.System(m, bc=1)
        rng = np.random.default_rng(3)
        u = np.zeros(s.n)
        v = np.where(rng.uniform(size=s.n) < 0.1, 0.25, 0.0)
        prm = ReactionDiffusionParameters()  # reference defaults
        ctx = FEMContext(m, 1)
        sol = make_solver(ctx, Equation.Reaction_Diffusion, prm, engine=_lib.FEM_ENGINE_GL_COMPAT)
        sol.set_state_unknowns(_lib.FEM_FIELD_V, v)
        op = [prm.time_step, prm.Du, prm.Dv, prm.feed_rate, prm.kill_rate]
        code = oracle.GRAYSCOTT
    else:
        m = meshes.jittered_square(30, seed=5)
        s = oracle.System(m, bc=0)
        u = s.gather(gauss(m, (0.45, 0.5), 0.02))
        v = np.zeros(s.n)
        ctx = FEMContext(m, 0)
        if eq == "heat":
            sol = make_solver(ctx, Equation.Heat, HeatParameters(), engine=_lib.FEM_ENGINE_GL_COMPAT)
            op, code = [0.01, 0.05], oracle.HEAT
        elif eq == "advdiff":
            sol = make_solver(ctx, Equation.Advection_Diffusion, AdvectionDiffusionParameters(),
                              engine=_lib.FEM_ENGINE_GL_COMPAT)
            op, code = [0.001, 0.25], oracle.ADVDIFF
        else:
            sol = make_solver(ctx, Equation.Wave, WaveParameters(), engine=_lib.FEM_ENGINE_GL_COMPAT)
            op, code = [0.05, 0.05], oracle.WAVE
        sol.set_state_unknowns(_lib.FEM_FIELD_U, u)
    for _ in range(4):
        u, v, rr = oracle.gl_step(s, code, op, u, v)
    assert sol.advance_time(4) == _lib.FEM_OK
    if eq == "grayscott":
        assert rel(sol.get_state_unknowns(_lib.FEM_FIELD_U), u) < 1e-4
        assert rel(sol.get_state_unknowns(_lib.FEM_FIELD_V), v) < 1e-4
    else:
        assert rel(sol.get_state_unknowns(_lib.FEM_FIELD_U), u) < 1e-4


def fan_mesh(m):
    """A disc fanned from its centre: vertex 0 is adjacent to all m rim
    vertices, so matrix row 0 holds m + 1 nonzeros."""
    th = np.linspace(0.0, 2.0 * np.pi, m, endpoint=False)
    xyz = np.zeros((m + 1, 3), dtype=np.float32)
    xyz[1:, 0], xyz[1:, 1] = np.cos(th), np.sin(th)
    tri = np.array([(0, 1 + i, 1 + (i + 1) % m) for i in range(m)], dtype=np.uint32)
    return {"xyz": xyz, "tri": tri, "on_boundary": meshes.boundary_from_edges(tri, m + 1).astype(np.uint8)}


@pytest.mark.parametrize("k", [1, 4])
def test_stream_engine_row_longer_than_block_cap(gpu, k):
    """A row of 2201 nonzeros exceeds the 2048-nonzero row-block cap: it gets
    a block of its own and a stage sized to it; results match the oracle."""
    print("case fan 2200", k, flush=True)
    m = fan_mesh(2200)
    s = oracle.System(m, bc=1)
    assert int(np.diff(s.row_ptr).max()) == 2201
    rng = np.random.default_rng(13)
    u0 = rng.normal(size=(s.n, k))
    b = Batch([FEMContext(m, 1)], SolverOptions(nrhs=k, engine=_lib.FEM_ENGINE_STREAM))
    b.set_equation(0, HeatParameters(1e-2, 0.3))
    b.set_state_unknowns(0, _lib.FEM_FIELD_U, u0)
    assert b.advance(3) == _lib.FEM_OK
    got = b.get_state_unknowns(0).reshape(s.n, k)
    for j in range(k):
        _, u, _, _ = s.run(oracle.HEAT, [1e-2, 0.3], 3, u0[:, j].copy(), tol=1e-14)
        assert rel(got[:, j], u) < RTOL, j
    b.close()


def test_stream_engine_refuses_row_beyond_lds(gpu):
    """A 5401-nonzero row in FP64 needs a 65 KB stage, beyond a workgroup's
    64 KiB of LDS: the streaming engine refuses with a clear error, and the
    persistent engine solves the same system."""
    print("case fan 5400", flush=True)
    m = fan_mesh(5400)
    s = oracle.System(m, bc=1)
    u0 = np.random.default_rng(14).normal(size=s.n)
    ctx = FEMContext(m, 1)
    st = make_solver(ctx, Equation.Heat, HeatParameters(1e-2, 0.3), engine=_lib.FEM_ENGINE_STREAM,
                     precision=_lib.FEM_PREC_FP64)
    st.set_state_unknowns(_lib.FEM_FIELD_U, u0)
    with pytest.raises(_lib.FemError, match="LDS"):
        st.advance_time(1)
    pe = make_solver(ctx, Equation.Heat, HeatParameters(1e-2, 0.3), precision=_lib.FEM_PREC_FP64)
    pe.set_state_unknowns(_lib.FEM_FIELD_U, u0)
    assert pe.advance_time(1) == _lib.FEM_OK
    _, u, _, _ = s.run(oracle.HEAT, [1e-2, 0.3], 1, u0.copy(), tol=1e-14)
    assert rel(pe.get_state_unknowns(), u) < RTOL


def test_graph_capture_engines(gpu):
    """The plain engine is host-driven and is not recorded; the streaming
    engine is, once an eager advance has shown the sweeps and iterations
    its steps need."""
    m = meshes.jittered_square(8, seed=1)
    b = Batch([FEMContext(m, 0)], SolverOptions(engine=_lib.FEM_ENGINE_PLAIN))
    b.set_equation(0, HeatParameters(1e-3, 1.0))
    with pytest.raises(_lib.FemError, match="host-driven"):
        b.graph_capture(2)
    with pytest.raises(_lib.FemError):
        b.graph_launch()
    b.close()
    ctx = FEMContext(m, 0)
    b = Batch([ctx], SolverOptions(engine=_lib.FEM_ENGINE_STREAM))
    b.set_equation(0, HeatParameters(1e-3, 1.0))
    b.set_state_unknowns(0, _lib.FEM_FIELD_U, np.ones(ctx.num_unknowns()))
    with pytest.raises(_lib.FemError, match="advance the batch once first"):
        b.graph_capture(2)
    assert b.advance(1) == _lib.FEM_OK
    b.graph_capture(2)
    assert b.graph_launch() == _lib.FEM_OK
    with pytest.raises(_lib.FemError, match="launches"):  # refused on the host, nothing recorded
        b.graph_capture(100000)
    b.close()


@pytest.mark.parametrize("k,krylov", [(16, _lib.FEM_KRYLOV_BICGSTAB), (1, _lib.FEM_KRYLOV_CG),
                                      (4, _lib.FEM_KRYLOV_CG_CLASSIC)])
def test_stream_graph_replays_match_eager(gpu, k, krylov):
    """A recorded streaming advance (every phase, reduction and compaction of
    two steps, with the sweeps and iterations of the eager runs plus a
    margin, nothing read on the host) replayed several times, with eager
    work and host reads between the replays, matches an identical batch
    advanced eagerly bit for bit after every replay, and the oracle; state
    writes and new equation parameters between replays are picked up."""
    print("case stream graph replay", k, krylov, flush=True)
    rng = np.random.default_rng(13)
    specs = [(14, 1), (22, 2), (9, 3)]
    vel = (0.4, -0.2, 0.0)
    ctxs, systems, states = [], [], []
    for side, seed in specs:
        m = meshes.jittered_square(side, seed=seed)
        ctxs.append(FEMContext(m, 0, velocity=vel, normalize_velocity=False))
        systems.append(oracle.System(m, bc=0, velocity=vel, normalize=False))
        states.append(rng.normal(size=(systems[-1].n, k)))
    adv = krylov == _lib.FEM_KRYLOV_BICGSTAB
    kap = 2e-2

    def params(dt):
        return AdvectionDiffusionParameters(dt, kap) if adv else HeatParameters(dt, kap)

    def make():
        b = Batch(ctxs, SolverOptions(nrhs=k, krylov=krylov, engine=_lib.FEM_ENGINE_STREAM))
        for i in range(len(ctxs)):
            b.set_equation(i, params(1e-3))
            b.set_state_unknowns(i, _lib.FEM_FIELD_U, states[i])
        return b

    def oracle_run(ref, dt, n):
        for i, s in enumerate(systems):
            for j in range(k):
                _, ref[i][:, j], _, _ = s.run(oracle.ADVDIFF if adv else oracle.HEAT, [dt, kap], n,
                                              ref[i][:, j].copy(), krylov=oracle.BICGSTAB if adv else oracle.CG,
                                              tol=1e-14)

    g, e = make(), make()
    ref = [s.copy() for s in states]
    assert g.advance(2) == _lib.FEM_OK and e.advance(2) == _lib.FEM_OK  # sizes the recorded iterations
    oracle_run(ref, 1e-3, 2)
    g.graph_capture(2)
    for rep in range(3):
        assert g.graph_launch() == _lib.FEM_OK
        assert g.last_stats.steps == 2
        assert e.advance(2) == _lib.FEM_OK
        oracle_run(ref, 1e-3, 2)
        for i, s in enumerate(systems):
            a, b_ = g.get_state_unknowns(i), e.get_state_unknowns(i)
            assert np.array_equal(a, b_), (rep, i)
            assert rel(a.reshape(s.n, k), ref[i]) < RTOL, (rep, i)
    # a state write and new parameters between replays are read by the next one
    for b in (g, e):
        b.set_state_unknowns(1, _lib.FEM_FIELD_U, states[1])
        b.set_equation(0, params(1.5e-3))
    assert g.graph_launch() == _lib.FEM_OK and e.advance(2) == _lib.FEM_OK
    for i in range(len(ctxs)):
        assert np.array_equal(g.get_state_unknowns(i), e.get_state_unknowns(i)), i
    g.close()
    e.close()
