"""GPU tests of the solver boundary: launch plans against the co-resident
limit of the kernel variant they launch, and the surface-space frame path
(set_state -> advance(1) -> get_state, the reference's per-frame call
pattern of Application.cpp:530-540 and CPUSolver.cpp:13-37)."""
import itertools

import numpy as np
import pytest

import oracle
from femhip import (AdvectionDiffusionParameters, Batch, Equation, FEMContext, HeatParameters, HIPSolver,
                    ReactionDiffusionParameters, SolverOptions, _lib, meshes)

pytestmark = pytest.mark.gpu


def rel(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / (nb if nb > 0 else 1.0))


@pytest.fixture(scope="module")
def many_contexts():
    # enough instances that a multi-instance plan spreads over the whole
    # co-resident grid, and two large ones that get multi-workgroup teams
    ms = [meshes.jittered_square(s, seed=20 + i) for i, s in enumerate([160, 128] + [24] * 40)]
    return [FEMContext(m, 0, velocity=(0.5, 0.3, 0.0), normalize_velocity=False) for m in ms]


@pytest.mark.parametrize("nrhs,precision,krylov,exchange", list(itertools.product(
    [1, 2, 4, 16], [_lib.FEM_PREC_FP32_IR, _lib.FEM_PREC_FP64],
    [_lib.FEM_KRYLOV_CG, _lib.FEM_KRYLOV_BICGSTAB, _lib.FEM_KRYLOV_CG_CLASSIC],
    [_lib.FEM_EXCHANGE_AUTO, _lib.FEM_EXCHANGE_COHERENT, _lib.FEM_EXCHANGE_FENCED])))
def test_plan_fits_coresident_limit(gpu, many_contexts, nrhs, precision, krylov, exchange):
    """Every (columns, precision, Krylov method, exchange mode) plan stays
    within the co-resident workgroups of the variant it launches (the spin
    barriers need every workgroup resident)."""
    print("case plan", nrhs, precision, krylov, exchange, flush=True)
    b = Batch(many_contexts, SolverOptions(nrhs=nrhs, precision=precision, krylov=krylov, exchange=exchange))
    for i in range(len(many_contexts)):
        b.set_equation(i, AdvectionDiffusionParameters(1e-3, 0.01))
    info = b.plan_info()
    assert 1 <= info.grid <= info.coresident, (info.grid, info.coresident)
    assert info.teams >= 1
    if exchange != _lib.FEM_EXCHANGE_AUTO and info.exchange != 0:
        assert info.exchange == exchange
    b.close()


@pytest.mark.parametrize("nrhs", [4, 16])
def test_fp64_bicgstab_fenced_batch_advances(gpu, many_contexts, nrhs):
    """The variant with the fewest co-resident workgroups (FP64 BiCGSTAB,
    fenced exchange, 4 or 16 columns) runs a full-grid plan to completion and
    matches the oracle on the largest instance."""
    ctxs = many_contexts[:6]
    b = Batch(ctxs, SolverOptions(nrhs=nrhs, precision=_lib.FEM_PREC_FP64, krylov=_lib.FEM_KRYLOV_BICGSTAB,
                                  exchange=_lib.FEM_EXCHANGE_FENCED, tol=1e-12))
    m0 = ctxs[0].mesh
    s = oracle.System(m0, bc=0, velocity=(0.5, 0.3, 0.0), normalize=False)
    rng = np.random.default_rng(3)
    fields = []
    for i, c in enumerate(ctxs):
        b.set_equation(i, AdvectionDiffusionParameters(1e-3, 0.01))
        f = rng.standard_normal((c.num_unknowns(), nrhs))
        fields.append(f)
        b.set_state_unknowns(i, _lib.FEM_FIELD_U, f)
    info = b.plan_info()
    assert info.grid <= info.coresident
    print("case fenced fp64 bicgstab", nrhs, info.grid, info.coresident, flush=True)
    assert b.advance(2) == _lib.FEM_OK
    got = b.get_state_unknowns(0).reshape(-1, nrhs)
    for j in range(nrhs):
        _, ref, _, _ = s.run(oracle.ADVDIFF, [1e-3, 0.01], 2, fields[0][:, j].copy(), krylov=oracle.BICGSTAB,
                             tol=1e-14)
        assert rel(got[:, j], ref) <= 1e-8
    b.close()


def heat_solver(m, **opts):
    ctx = FEMContext(m, 0)
    ctx.equation = Equation.Heat
    ctx.parameters[Equation.Heat] = HeatParameters(1e-3, 1.0)
    return ctx, HIPSolver(ctx, SolverOptions(**opts))


def test_frame_loop_keeps_state_and_predictor(gpu):
    """Writing back what get_state returned changes nothing: the frame loop
    set_state -> advance(1) -> get_state gives bit for bit the states of
    advance(1) calls without any state write, with the same iteration
    counts (the predictor keeps its history), and stays within 1e-8 of the
    oracle's FP64 trajectory."""
    m = meshes.jittered_square(40, seed=1)
    s = oracle.System(m, bc=0)
    x = m["xyz"].astype(np.float64)
    surf0 = np.exp(-((x[:, 0] - 0.5) ** 2 + (x[:, 1] - 0.5) ** 2) / 0.01).astype(np.float32)
    ca, a = heat_solver(m)
    cb, b = heat_solver(m)
    a.set_state(_lib.FEM_FIELD_U, surf0)
    b.set_state(_lib.FEM_FIELD_U, surf0)
    u = s.gather(surf0.astype(np.float64))
    surf = surf0.copy()
    its_a, its_b = [], []
    for k in range(12):
        a.set_state(_lib.FEM_FIELD_U, surf)
        assert a.advance(1) == _lib.FEM_OK
        surf = a.get_state(_lib.FEM_FIELD_U)
        its_a.append(a.last_stats.iterations)
        assert b.advance(1) == _lib.FEM_OK
        its_b.append(b.last_stats.iterations)
        ua, ub = a.get_state_unknowns(), b.get_state_unknowns()
        assert np.array_equal(ua, ub), k
        _, u, _, _ = s.step(oracle.HEAT, [1e-3, 1.0], u, tol=1e-14)
        assert rel(ua, u) <= 1e-8, k
        onb = m["on_boundary"].astype(bool)
        assert np.all(surf[onb] == 0.0)
        assert np.array_equal(surf[~onb], ua[s.idx_map[~onb]].astype(np.float32))
    assert its_a == its_b
    # with history the later solves start closer: fewer iterations than the first
    assert its_a[-1] < its_a[0]


def test_frame_write_replaces_only_changed_values(gpu):
    """A brush write (one surface value changed) replaces exactly that
    unknown with the float value; every other unknown keeps its FP64 state."""
    m = meshes.jittered_square(24, seed=3)
    ctx, sol = heat_solver(m)
    x = m["xyz"].astype(np.float64)
    surf = np.exp(-((x[:, 0] - 0.5) ** 2 + (x[:, 1] - 0.5) ** 2) / 0.01).astype(np.float32)
    sol.set_state(_lib.FEM_FIELD_U, surf)
    sol.advance(3)
    before = sol.get_state_unknowns()
    surf = sol.get_state(_lib.FEM_FIELD_U)
    idx = ctx.idx_map
    v = int(np.flatnonzero(idx >= 0)[7])
    surf[v] = np.float32(0.75)
    sol.set_state(_lib.FEM_FIELD_U, surf)
    after = sol.get_state_unknowns()
    changed = np.flatnonzero(after != before)
    assert list(changed) == [idx[v]]
    assert after[idx[v]] == np.float64(np.float32(0.75))


def test_grayscott_surface_is_v(gpu):
    """Gray-Scott's surface field is v (CPUSolver.cpp:130,144): set_state /
    get_state of FEM_FIELD_V touch the v column only, u stays as set."""
    m = meshes.icosphere(2)
    ctx = FEMContext(m, 1)
    ctx.equation = Equation.Reaction_Diffusion
    ctx.parameters[Equation.Reaction_Diffusion] = ReactionDiffusionParameters(
        0.5, 2e-4, 1e-4, 0.037, 0.06, _lib.FEM_REACTION_LUMPED)
    sol = HIPSolver(ctx, SolverOptions())
    n = ctx.num_unknowns()
    u = np.linspace(0.5, 1.0, n)
    sol.set_state_unknowns(_lib.FEM_FIELD_U, u)
    vs = (np.arange(m["xyz"].shape[0]) % 5 == 0).astype(np.float32) * 0.25
    sol.set_state(_lib.FEM_FIELD_V, vs)
    assert np.array_equal(sol.get_state_unknowns(_lib.FEM_FIELD_U), u)
    assert np.array_equal(sol.get_state(_lib.FEM_FIELD_V), vs)


def fan_disc(rings=40, per_ring=16):
    """Disc of concentric rings of per_ring vertices around one centre vertex
    whose row holds per_ring + 1 entries (longer than the LDS row cache's
    10), every other row 7."""
    xyz = [(0.0, 0.0, 0.0)]
    for i in range(1, rings + 1):
        for j in range(per_ring):
            t = 2 * np.pi * (j + 0.5 * (i % 2)) / per_ring
            xyz.append((i / rings * np.cos(t), i / rings * np.sin(t), 0.0))
    tri = [(0, 1 + j, 1 + (j + 1) % per_ring) for j in range(per_ring)]
    for i in range(1, rings):
        a, b = 1 + (i - 1) * per_ring, 1 + i * per_ring
        for j in range(per_ring):
            j1 = (j + 1) % per_ring
            tri.append((a + j, b + j, b + j1) if i % 2 else (a + j, b + j, a + j1))
            tri.append((a + j, b + j1, a + j1) if i % 2 else (a + j1, b + j, b + j1))
    return meshes._pack(np.array(xyz), np.array(tri))


@pytest.mark.parametrize("exchange", [_lib.FEM_EXCHANGE_COHERENT, _lib.FEM_EXCHANGE_FENCED])
def test_team_where_only_last_workgroup_fits_row_cache(gpu, exchange):
    """Gray-Scott on the icosphere refined 6 times with a team of 128
    workgroups: every workgroup is dealt 321 rows (more than the LDS row
    cache holds) but the last only 195.  The whole team must take the same
    uncached path (the cached one starts the CG one barrier earlier); steps
    inside one launch use the fused update, where the difference shows."""
    import bench
    m = meshes.icosphere(6)
    s = oracle.System(m, bc=1)
    u0, v0 = bench.gs_initial(m, s.n)
    ctx = FEMContext(m, 1)
    ctx.equation = Equation.Reaction_Diffusion
    ctx.parameters[Equation.Reaction_Diffusion] = ReactionDiffusionParameters(
        0.5, 2e-4, 1e-4, 0.037, 0.06, _lib.FEM_REACTION_LUMPED)
    sol = HIPSolver(ctx, SolverOptions(team_size=128, exchange=exchange, inner_tol=1e-6))
    sol.set_state_unknowns(_lib.FEM_FIELD_U, u0)
    sol.set_state_unknowns(_lib.FEM_FIELD_V, v0)
    print("case team 128 on ico6, exchange", exchange, flush=True)
    assert sol.advance(4) == _lib.FEM_OK
    _, u, v, _ = s.run(oracle.GRAYSCOTT, [0.5, 2e-4, 1e-4, 0.037, 0.06, 1], 4, u0.copy(), v0.copy(), tol=1e-14)
    assert rel(sol.get_state_unknowns(_lib.FEM_FIELD_U), u) <= 1e-8
    assert rel(sol.get_state_unknowns(_lib.FEM_FIELD_V), v) <= 1e-8


def test_long_row_in_one_workgroup(gpu):
    """A fan mesh whose centre row (17 entries) exceeds the row cache while
    every other row fits: the team of several workgroups runs uncached as a
    whole.  Heat (pipelined CG), several steps per launch, against the
    oracle."""
    m = fan_disc()
    s = oracle.System(m, bc=0)
    assert int(np.diff(s.row_ptr).max()) == 17
    x = m["xyz"].astype(np.float64)
    u0 = s.gather(np.exp(-((x[:, 0] - 0.1) ** 2 + x[:, 1] ** 2) / 0.05))
    ctx, sol = heat_solver(m)
    assert ctx.num_max_nonzeros_per_row() == 17
    sol.set_state_unknowns(_lib.FEM_FIELD_U, u0)
    print("case fan mesh, n", s.n, flush=True)
    assert sol.advance(5) == _lib.FEM_OK
    _, u, _, _ = s.run(oracle.HEAT, [1e-3, 1.0], 5, u0.copy(), tol=1e-14)
    assert rel(sol.get_state_unknowns(), u) <= 1e-8


@pytest.mark.parametrize("mesh_type", [_lib.FEM_MESH_OPEN, _lib.FEM_MESH_MIRRORED])
def test_solver_export_ply_from_device(gpu, tmp_path, mesh_type):
    """fem_solver_export_ply writes the solver's current surface field (v for
    Gray-Scott) exactly as the clipping restatement does on the values
    get_state returns."""
    import test_surface_io as sio
    m = meshes.icosphere(3)
    ctx = FEMContext(m, 1)
    ctx.equation = Equation.Reaction_Diffusion
    ctx.parameters[Equation.Reaction_Diffusion] = ReactionDiffusionParameters(
        0.5, 2e-4, 1e-4, 0.037, 0.06, _lib.FEM_REACTION_LUMPED)
    sol = HIPSolver(ctx, SolverOptions())
    n = ctx.num_unknowns()
    sol.set_state_unknowns(_lib.FEM_FIELD_U, np.ones(n))
    sol.set_state_unknowns(_lib.FEM_FIELD_V, np.where(np.arange(n) % 9 == 0, 0.25, 0.0))
    assert sol.advance(3) == _lib.FEM_OK
    x = m["xyz"].astype(np.float32)
    nrm = x / np.linalg.norm(x, axis=1, keepdims=True)
    vals = sol.get_state(_lib.FEM_FIELD_V)
    thr = float(np.quantile(vals, 0.6))
    path = str(tmp_path / "solver.ply")
    sol.export_ply(path, nrm, 2.0, thr, mesh_type)
    head, rows, faces = sio.read_ply_exact(path)
    pos, pval, fref = oracle.ply_clip(x, m["tri"], nrm, vals, 2.0, thr, mesh_type)
    assert len(pos) > 0 and len(fref) > 0
    assert sio.same(rows[:, :3], pos) and np.array_equal(faces[:, 1:], fref)


def test_barrier_timeout_path_forced_from_host(gpu):
    """The abort path of the persistent kernel's bounded barrier waits,
    forced from the host with a zero wait limit: every workgroup leaves at
    its first unsatisfied wait, the call reports FEM_ERR_TIMEOUT and the
    batch recovers with the default limit (state written again) to within
    1e-8 of the oracle."""
    m = meshes.icosphere(4)
    s = oracle.System(m, bc=1)
    ctx = FEMContext(m, 1)
    prm = ReactionDiffusionParameters(0.5, 2e-4, 1e-4, 0.037, 0.06, _lib.FEM_REACTION_LUMPED)
    b = Batch([ctx], SolverOptions(exchange=_lib.FEM_EXCHANGE_COHERENT))
    b.set_equation(0, prm)
    u = np.ones(s.n)
    v = np.where(np.arange(s.n) % 13 == 0, 0.25, 0.0)
    b.set_state_unknowns(0, _lib.FEM_FIELD_U, u)
    b.set_state_unknowns(0, _lib.FEM_FIELD_V, v)
    assert b.plan_info().teams == 1 and b.plan_info().grid > 1
    b.set_wait_limit(0.0)
    print("case forced barrier timeout", flush=True)
    with pytest.raises(_lib.FemError) as err:
        b.advance(3)
    assert err.value.code == _lib.FEM_ERR_TIMEOUT
    b.set_wait_limit(4.0)
    b.set_state_unknowns(0, _lib.FEM_FIELD_U, u)
    b.set_state_unknowns(0, _lib.FEM_FIELD_V, v)
    assert b.advance(3) == _lib.FEM_OK
    _, ur, vr, _ = s.run(oracle.GRAYSCOTT, [0.5, 2e-4, 1e-4, 0.037, 0.06, 1], 3, u.copy(), v.copy(), tol=1e-14)
    assert rel(b.get_state_unknowns(0, _lib.FEM_FIELD_U)[: s.n], ur) <= 1e-8
    assert rel(b.get_state_unknowns(0, _lib.FEM_FIELD_V)[: s.n], vr) <= 1e-8
    with pytest.raises(_lib.FemError):
        b.set_wait_limit(-1.0)
    b.close()
