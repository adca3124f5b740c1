"""GPU tests of the stand-alone CSR operator and of the discrete adjoint,
through the C ABI, against the oracle (FP64 CPU restatement) and against
finite differences of the GPU forward itself."""
import numpy as np
import pytest

import oracle
from femhip import (AdvectionDiffusionParameters, Adjoint, FEMContext, HeatParameters, Operator, SolverOptions,
                    _lib, meshes)

pytestmark = pytest.mark.gpu


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / (nb if nb > 0 else 1.0)


def rotating_velocity(mesh):
    """(-(y - 1/2), x - 1/2, 0) at the element centroids."""
    c = mesh["xyz"].astype(np.float64)[mesh["tri"]].mean(axis=1)
    return np.stack([-(c[:, 1] - 0.5), c[:, 0] - 0.5, np.zeros(len(c))], axis=1)


def c5_like(n, seed):
    m = meshes.jittered_square(n, seed=seed)
    rng = np.random.Generator(np.random.PCG64(seed + 1))
    kappa = 1e-3 * np.exp(0.3 * rng.standard_normal(m["tri"].shape[0]))
    return m, rotating_velocity(m), kappa


@pytest.mark.parametrize("precision", [_lib.FEM_PREC_FP64, _lib.FEM_PREC_FP32_IR])
@pytest.mark.parametrize("transpose", [False, True])
def test_operator_vs_oracle_all_schedules(gpu, precision, transpose):
    m, ve, kappa = c5_like(24, 3)
    print("case operator", precision, transpose, flush=True)
    ctx = FEMContext(m, 0, velocity_e=ve, normalize_velocity=False, kappa_e=kappa)
    s = oracle.System(m, bc=0, velocity_e=ve, normalize=False, kappa_e=kappa)
    coef = (1.0 / 2e-3, 1.0, -1.0)
    S = coef[0] * s.scipy("M") + coef[1] * s.scipy("K") + coef[2] * s.scipy("A")
    if transpose:
        S = S.T.tocsr()
    op = Operator(ctx, coef, precision, transpose)
    rng = np.random.Generator(np.random.PCG64(11))
    tol = 1e-14 if precision == _lib.FEM_PREC_FP64 else 2e-6
    for k in (1, 2, 4, 16):
        x = rng.standard_normal((s.n, k)) if k > 1 else rng.standard_normal(s.n)
        ref = S @ x
        for g in _lib.SPMV_SCHEDULES:
            if g < k:
                continue
            op.set_schedule(k, g)
            y = op.apply(x)
            assert rel(y, ref) <= tol, (k, g, rel(y, ref))
    op.close()


@pytest.mark.parametrize("precision", [_lib.FEM_PREC_FP64, _lib.FEM_PREC_FP32_IR])
def test_operator_pipe_many_blocks_per_workgroup(gpu, precision):
    """The pipelined stream schedule on a mesh with more row blocks than its
    resident grid (600 x 600 jittered, ~358K rows: 1,400 blocks of 256 rows
    at one column, 5,600 of 64 at 16 FP32 columns, 11,200 of 32 at 16 FP64),
    so every workgroup walks several blocks through both LDS buffers.  Each
    row's arithmetic is the plain stream schedule's, so the results are
    bit-identical to it, and they match scipy."""
    m = meshes.jittered_square(600, seed=4)
    print("case operator pipe", precision, flush=True)
    ctx = FEMContext(m, 0)
    s = oracle.System(m, bc=0)
    coef = (1.0 / 2e-3, 1.0, -1.0)
    S = coef[0] * s.scipy("M") + coef[1] * s.scipy("K") + coef[2] * s.scipy("A")
    op = Operator(ctx, coef, precision)
    rng = np.random.Generator(np.random.PCG64(12))
    tol = 1e-14 if precision == _lib.FEM_PREC_FP64 else 2e-6
    for k in (1, 16):
        x = rng.standard_normal((s.n, k)) if k > 1 else rng.standard_normal(s.n)
        op.set_schedule(k, _lib.FEM_SPMV_STREAM)
        y0 = op.apply(x)
        for g in (_lib.FEM_SPMV_PIPE, _lib.FEM_SPMV_PIPE_NT):
            op.set_schedule(k, g)
            y = op.apply(x)
            assert np.array_equal(y, y0), (k, g)
        assert rel(y0, S @ x) <= tol, (k, rel(y0, S @ x))
    op.close()


def test_operator_transpose_dot_product(gpu):
    m, ve, kappa = c5_like(20, 5)
    ctx = FEMContext(m, 0, velocity_e=ve, normalize_velocity=False, kappa_e=kappa)
    coef = (500.0, 1.0, -1.0)
    op = Operator(ctx, coef, _lib.FEM_PREC_FP64, False)
    opT = Operator(ctx, coef, _lib.FEM_PREC_FP64, True)
    rng = np.random.Generator(np.random.PCG64(2))
    x = rng.standard_normal(ctx.num_unknowns())
    y = rng.standard_normal(ctx.num_unknowns())
    a, b = np.dot(op.apply(x), y), np.dot(x, opT.apply(y))
    assert abs(a - b) <= 1e-13 * abs(a)


@pytest.mark.parametrize("every", [1, 3, 7])
def test_adjoint_vs_oracle_advection_diffusion(gpu, every):
    m, ve, kappa = c5_like(16, 7)
    dt, T = 2e-3, 7
    print("case adjoint every", every, flush=True)
    s = oracle.System(m, bc=0, velocity_e=ve, normalize=False, kappa_e=kappa)
    rng = np.random.Generator(np.random.PCG64(8))
    u0 = rng.standard_normal(s.n)
    target = rng.standard_normal(s.n)
    loss, uT, gu, gk = s.adjoint(dt, T, u0, target, tol=1e-14)
    ctx = FEMContext(m, 0, velocity_e=ve, normalize_velocity=False, kappa_e=kappa)
    adj = Adjoint(ctx, AdvectionDiffusionParameters(dt, 1.0), SolverOptions(precision=_lib.FEM_PREC_FP64,
                                                                           tol=1e-14))
    uT_g = adj.forward(u0, T, every)
    assert rel(uT_g, uT) <= 1e-11
    gu_g, gk_g = adj.backward(2.0 * (uT_g - target))
    assert rel(gu_g, gu) <= 1e-9, rel(gu_g, gu)
    assert rel(gk_g, gk) <= 1e-9, rel(gk_g, gk)
    adj.close()


def test_adjoint_checkpoint_interval_invariance_fp32(gpu):
    m, ve, kappa = c5_like(24, 9)
    ctx = FEMContext(m, 0, velocity_e=ve, normalize_velocity=False, kappa_e=kappa)
    rng = np.random.Generator(np.random.PCG64(3))
    n = ctx.num_unknowns()
    u0 = rng.standard_normal(n)
    g = rng.standard_normal(n)
    out = []
    for every in (1, 4, 10):
        adj = Adjoint(ctx, AdvectionDiffusionParameters(2e-3, 1.0), SolverOptions(tol=1e-10))
        adj.forward(u0, 10, every)
        out.append(adj.backward(g))
        adj.close()
    for gu, gk in out[1:]:
        # the recomputed forward is the same deterministic launch sequence
        assert rel(gu, out[0][0]) <= 1e-12 and rel(gk, out[0][1]) <= 1e-12


def test_adjoint_fp32_refined_vs_oracle(gpu):
    m, ve, kappa = c5_like(32, 11)
    dt, T = 2e-3, 12
    s = oracle.System(m, bc=0, velocity_e=ve, normalize=False, kappa_e=kappa)
    rng = np.random.Generator(np.random.PCG64(12))
    u0 = rng.standard_normal(s.n)
    target = rng.standard_normal(s.n)
    loss, uT, gu, gk = s.adjoint(dt, T, u0, target, tol=1e-14)
    ctx = FEMContext(m, 0, velocity_e=ve, normalize_velocity=False, kappa_e=kappa)
    adj = Adjoint(ctx, AdvectionDiffusionParameters(dt, 1.0), SolverOptions(tol=1e-10))
    uT_g = adj.forward(u0, T, 5)
    gu_g, gk_g = adj.backward(2.0 * (uT_g - target))
    assert rel(uT_g, uT) <= 1e-8
    assert rel(gu_g, gu) <= 1e-7 and rel(gk_g, gk) <= 1e-7, (rel(gu_g, gu), rel(gk_g, gk))
    adj.close()


def test_adjoint_central_differences_heat_fp64(gpu):
    """Gradients of L = |u_T - y|^2 against central differences (h = 1e-6)
    of the GPU forward itself, heat equation with conductivity 0.7."""
    m = meshes.jittered_square(12, seed=2)
    rng = np.random.Generator(np.random.PCG64(9))
    ntri = m["tri"].shape[0]
    kappa = np.exp(0.3 * rng.standard_normal(ntri))
    ctx = FEMContext(m, 0, kappa_e=kappa)
    n = ctx.num_unknowns()
    u0 = rng.standard_normal(n)
    y = rng.standard_normal(n)
    opts = SolverOptions(precision=_lib.FEM_PREC_FP64, tol=1e-14)
    prm = HeatParameters(1e-2, 0.7)
    adj = Adjoint(ctx, prm, opts)
    T = 6
    uT = adj.forward(u0, T, 4)
    gu, gk = adj.backward(2.0 * (uT - y))

    def loss_u(u):
        return float(np.sum((adj.forward(u, T, 4) - y) ** 2))

    def loss_k(k):
        ctx.assemble_matrices(kappa_e=k)
        v = float(np.sum((adj.forward(u0, T, 4) - y) ** 2))
        return v

    h = 1e-6
    for q in range(3):
        d = rng.standard_normal(n)
        fd = (loss_u(u0 + h * d) - loss_u(u0 - h * d)) / (2 * h)
        assert abs(fd - gu @ d) <= 1e-6 * abs(fd) + 1e-9, (fd, gu @ d)
        dk = rng.standard_normal(ntri) * kappa
        fd = (loss_k(kappa + h * dk) - loss_k(kappa - h * dk)) / (2 * h)
        assert abs(fd - gk @ dk) <= 1e-6 * abs(fd) + 1e-9, (fd, gk @ dk)
    ctx.assemble_matrices(kappa_e=kappa)
    adj.close()


@pytest.mark.parametrize("device", ["cpu", "cuda"])
def test_torch_autograd_rollout(gpu, device):
    """Rollout through torch.autograd against the oracle's FP64 adjoint, with
    host tensors and with device tensors (device pointers cross the ABI);
    a second call with a changed kappa re-assembles."""
    import torch
    from femhip.autograd import Rollout

    m, ve, kappa = c5_like(16, 13)
    dt, T = 2e-3, 9
    rng = np.random.Generator(np.random.PCG64(14))
    s = oracle.System(m, bc=0, velocity_e=ve, normalize=False, kappa_e=kappa)
    u0 = rng.standard_normal(s.n)
    target = rng.standard_normal(s.n)
    roll = Rollout(m, AdvectionDiffusionParameters(dt, 1.0), T, checkpoint_every=4, velocity_e=ve,
                   normalize_velocity=False, options=SolverOptions(precision=_lib.FEM_PREC_FP64, tol=1e-14))
    for kap in (kappa, kappa * 1.5):
        so = oracle.System(m, bc=0, velocity_e=ve, normalize=False, kappa_e=kap)
        loss, uT, gu, gk = so.adjoint(dt, T, u0, target, tol=1e-14)
        tu0 = torch.tensor(u0, dtype=torch.float64, device=device, requires_grad=True)
        tk = torch.tensor(kap, dtype=torch.float64, device=device, requires_grad=True)
        out = roll(tu0, tk)
        assert out.device.type == device
        L = ((out - torch.tensor(target, dtype=torch.float64, device=device)) ** 2).sum()
        L.backward()
        assert tu0.grad.device.type == device and tk.grad.device.type == device
        assert abs(L.item() - loss) <= 1e-10 * loss
        assert rel(tu0.grad.cpu().numpy(), gu) <= 1e-9
        assert rel(tk.grad.cpu().numpy(), gk) <= 1e-9
    roll.close()


@pytest.mark.parametrize("coef", [(1.0, 0.0, 0.0), (500.0, 1.0, -1.0)])
def test_ctx_solve_vs_direct(gpu, coef):
    """fem_ctx_solve against a direct FP64 solve of the oracle's matrices."""
    import scipy.sparse.linalg as sla
    m, ve, kappa = c5_like(28, 15)
    ctx = FEMContext(m, 0, velocity_e=ve, normalize_velocity=False, kappa_e=kappa)
    s = oracle.System(m, bc=0, velocity_e=ve, normalize=False, kappa_e=kappa)
    S = (coef[0] * s.scipy("M") + coef[1] * s.scipy("K") + coef[2] * s.scipy("A")).tocsc()
    rng = np.random.Generator(np.random.PCG64(16))
    b = rng.standard_normal((s.n, 4))
    ref = sla.spsolve(S, b)
    x = ctx.solve(coef, b, SolverOptions(tol=1e-12))
    assert rel(x, ref) <= 1e-10, rel(x, ref)
    x1 = ctx.solve(coef, b[:, 0], SolverOptions(precision=_lib.FEM_PREC_FP64, tol=1e-13))
    assert rel(x1, ref[:, 0]) <= 1e-11


def test_newmark_initial_acceleration(gpu):
    from femhip import Equation, HIPSolver, NewmarkParameters
    m = meshes.jittered_square(30, seed=4)
    s = oracle.System(m, bc=0)
    u = s.gather(np.exp(-((m["xyz"][:, 0] - 0.4) ** 2 + (m["xyz"][:, 1] - 0.5) ** 2) / (2 * 0.05 ** 2)))
    a_ref = s.newmark_init(1.3, u)
    ctx = FEMContext(m, 0)
    ctx.equation = Equation.Wave_Newmark
    ctx.parameters[Equation.Wave_Newmark] = NewmarkParameters(2e-4, 1.3, 0.25, 0.5)
    sol = HIPSolver(ctx)
    sol.set_state_unknowns(_lib.FEM_FIELD_U, u)
    a = sol.init_newmark_acceleration()
    assert rel(a, a_ref) <= 1e-10
    assert rel(sol.get_state_unknowns(_lib.FEM_FIELD_A), a_ref) <= 1e-10


def test_smooth_normals_vs_oracle(gpu):
    from femhip import Equation, HIPSolver, ReactionDiffusionParameters
    m = meshes.icosphere(4)
    p = m["xyz"].astype(np.float32)
    nrm = p / np.linalg.norm(p, axis=1, keepdims=True)
    rng = np.random.Generator(np.random.PCG64(21))
    vals = rng.uniform(-0.5, 1.5, p.shape[0]).astype(np.float32)
    ctx = FEMContext(m, 1)
    got = ctx.smooth_normals(nrm, vals, 0.3)
    ref = oracle.smooth_normals(p, m["tri"], nrm, vals, 0.3)
    scale = np.linalg.norm(ref, axis=1).max()
    assert np.abs(got - ref).max() <= 1e-5 * scale
    # from the solver state (Gray-Scott: the surface carries v)
    ctx.equation = Equation.Reaction_Diffusion
    ctx.parameters[Equation.Reaction_Diffusion] = ReactionDiffusionParameters()
    sol = HIPSolver(ctx)
    sol.set_state(_lib.FEM_FIELD_V, vals)
    got2 = sol.smooth_normals(nrm, 0.3)
    assert np.array_equal(got2, got)
