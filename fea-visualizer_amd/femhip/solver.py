"""Host-side mirror of the reference's FEM interface over the C ABI.

``FEMContext`` mirrors include/FEM/FEMContext.hpp:24-59 (mesh, boundary
condition, idx_map, assembled matrices, per-equation parameters) and
``HIPSolver`` mirrors the abstract ``Solver`` (include/FEM/Solver.hpp:8-15):
``advance_time()``, ``clear_values()`` and ``has_numerical_instability()``.
Parameter classes mirror include/FEM/EquationParameters.hpp:4-49, with the
reference's defaults.  All arithmetic runs in the HIP library; this module
only moves arrays across the boundary.
"""
from __future__ import annotations

import atexit
import ctypes as C
import itertools
import sys
import weakref
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _lib as L

# Every live handle, keyed by creation order.  At interpreter exit they are
# closed explicitly, newest first (solvers before the contexts they use),
# while the HIP runtime is still fully up; finalizers that run later during
# interpreter teardown make no library call.
_LIVE: "weakref.WeakValueDictionary[int, object]" = weakref.WeakValueDictionary()
_ORDER = itertools.count()


def _register(obj):
    _LIVE[next(_ORDER)] = obj


@atexit.register
def close_all():
    """Close every live handle, newest first."""
    for key in sorted(list(_LIVE.keys()), reverse=True):
        obj = _LIVE.get(key)
        if obj is not None:
            try:
                obj.close()
            except Exception:
                pass


def _finalize(obj):
    if sys.is_finalizing():
        return
    try:
        obj.close()
    except Exception:
        pass

__all__ = [
    "HeatParameters", "AdvectionDiffusionParameters", "WaveParameters",
    "ReactionDiffusionParameters", "NewmarkParameters", "SolverOptions",
    "FEMContext", "HIPSolver", "Batch", "Equation", "BoundaryCondition", "Operator", "Adjoint",
]


class Equation:
    """enum class Equation (FEMContext.hpp:9-14) plus Newmark-beta."""
    Heat = L.FEM_HEAT
    Wave = L.FEM_WAVE
    Advection_Diffusion = L.FEM_ADVECTION_DIFFUSION
    Reaction_Diffusion = L.FEM_REACTION_DIFFUSION
    Wave_Newmark = L.FEM_WAVE_NEWMARK


class BoundaryCondition:
    """enum class BoundaryCondition (FEMContext.hpp:16-19)."""
    Dirichlet = L.FEM_DIRICHLET
    Neumann = L.FEM_NEUMANN


@dataclass
class HeatParameters:  # EquationParameters.hpp:8-15
    time_step: float = 0.01
    conductivity: float = 0.05

    def to_c(self):
        return L.FEM_HEAT, L.fem_heat_params(self.time_step, self.conductivity)


@dataclass
class AdvectionDiffusionParameters:  # EquationParameters.hpp:17-26
    time_step: float = 0.001
    c: float = 0.25
    velocity: Sequence[float] = (1.0, 0.0, 0.0)

    def to_c(self):
        return L.FEM_ADVECTION_DIFFUSION, L.fem_advdiff_params(self.time_step, self.c)


@dataclass
class WaveParameters:  # EquationParameters.hpp:28-35
    time_step: float = 0.05
    c: float = 0.05

    def to_c(self):
        return L.FEM_WAVE, L.fem_wave_params(self.time_step, self.c)


@dataclass
class ReactionDiffusionParameters:  # EquationParameters.hpp:37-49
    time_step: float = 0.001
    Du: float = 0.08
    Dv: float = 0.04
    feed_rate: float = 0.035
    kill_rate: float = 0.06
    reaction_weighting: int = L.FEM_REACTION_NODAL  # reference: nodal (CPUSolver.cpp:134,137)

    def to_c(self):
        return L.FEM_REACTION_DIFFUSION, L.fem_rd_params(
            self.time_step, self.Du, self.Dv, self.feed_rate, self.kill_rate, self.reaction_weighting)


@dataclass
class NewmarkParameters:  # no reference counterpart
    time_step: float = 2e-4
    c: float = 1.0
    beta: float = 0.25
    gamma: float = 0.5

    def to_c(self):
        return L.FEM_WAVE_NEWMARK, L.fem_newmark_params(self.time_step, self.c, self.beta, self.gamma)


@dataclass
class SolverOptions:
    engine: int = L.FEM_ENGINE_PERSISTENT
    precision: int = L.FEM_PREC_FP32_IR
    krylov: int = L.FEM_KRYLOV_AUTO
    nrhs: int = 1
    warm_start: int = 1
    max_iters: int = 0
    max_inner: int = 0
    max_sweeps: int = 0
    team_size: int = 0
    tol: float = 1e-10
    inner_tol: float = 1e-4
    exchange: int = L.FEM_EXCHANGE_AUTO
    predictor: int = L.FEM_PREDICTOR_SEXTIC

    def to_c(self):
        o = L.fem_solver_opts()
        for name, _ in L.fem_solver_opts._fields_:
            setattr(o, name, getattr(self, name))
        return o


class FEMContext:
    """Mesh + boundary condition + assembled K, M, A on the GPU."""

    def __init__(self, mesh: dict, boundary_condition: int = BoundaryCondition.Dirichlet,
                 velocity=(1.0, 0.0, 0.0), velocity_e=None, normalize_velocity=True, kappa_e=None):
        self._lib = L.load()
        self.mesh = mesh
        self.boundary_condition = boundary_condition
        self.equation = Equation.Wave  # FEMContext.hpp:31 default
        self.parameters = {
            Equation.Heat: HeatParameters(),
            Equation.Advection_Diffusion: AdvectionDiffusionParameters(),
            Equation.Wave: WaveParameters(),
            Equation.Reaction_Diffusion: ReactionDiffusionParameters(),
            Equation.Wave_Newmark: NewmarkParameters(),
        }
        self._keep = {}
        xyz = np.ascontiguousarray(mesh["xyz"], dtype=np.float32)
        tri = np.ascontiguousarray(mesh["tri"], dtype=np.uint32)
        onb = np.ascontiguousarray(mesh["on_boundary"], dtype=np.uint8)
        m = L.fem_mesh(xyz.shape[0], tri.shape[0], L.ptr(xyz, C.c_float), L.ptr(tri, C.c_uint32),
                       L.ptr(onb, C.c_uint8))
        h = C.c_void_p()
        L.check(self._lib.fem_ctx_create(C.byref(m), int(boundary_condition), C.byref(h)))
        self.handle = h
        _register(self)
        self.assemble_matrices(velocity, velocity_e, normalize_velocity, kappa_e)

    def assemble_matrices(self, velocity=None, velocity_e=None, normalize_velocity=True, kappa_e=None):
        """FEMContext::assemble_matrices (FEMContext.cpp:57-64) on the GPU."""
        if velocity is None:
            velocity = self.parameters[Equation.Advection_Diffusion].velocity
        p = L.fem_assembly_params()
        for q in range(3):
            p.velocity[q] = float(velocity[q])
        p.normalize_velocity = 1 if normalize_velocity else 0
        if velocity_e is not None:
            ve = np.ascontiguousarray(velocity_e, dtype=np.float64).reshape(-1)
            self._keep["ve"] = ve
            p.velocity_e = L.ptr(ve, C.c_double)
        if kappa_e is not None:
            ke = np.ascontiguousarray(kappa_e, dtype=np.float64).reshape(-1)
            self._keep["ke"] = ke
            p.kappa_e = L.ptr(ke, C.c_double)
        L.check(self._lib.fem_ctx_assemble(self.handle, C.byref(p)))

    def info(self):
        i = L.fem_ctx_info()
        L.check(self._lib.fem_ctx_info_get(self.handle, C.byref(i)))
        return i

    def num_nodes(self):
        return self.info().num_nodes

    def num_unknowns(self):
        return self.info().num_unknowns

    def num_max_nonzeros_per_row(self):
        return self.info().max_row_nonzeros

    @property
    def idx_map(self):
        out = np.empty(self.num_nodes(), dtype=np.int32)
        L.check(self._lib.fem_ctx_get_idx_map(self.handle, L.ptr(out, C.c_int32)))
        return out

    def pattern(self):
        i = self.info()
        rp = np.empty(i.num_unknowns + 1, dtype=np.int32)
        ci = np.empty(max(1, i.nnz), dtype=np.int32)
        L.check(self._lib.fem_ctx_get_pattern(self.handle, L.ptr(rp, C.c_int32), L.ptr(ci, C.c_int32)))
        return rp, ci[: i.nnz]

    def matrix(self, which: int):
        i = self.info()
        out = np.empty(max(1, i.nnz), dtype=np.float64)
        L.check(self._lib.fem_ctx_get_matrix(self.handle, int(which), L.ptr(out, C.c_double)))
        return out[: i.nnz]

    def solve(self, coef, b, options: Optional[SolverOptions] = None, x0=None):
        """x with (coef[0] M + coef[1] K + coef[2] A) x = b (b: [N] or [N, nrhs]),
        solved on the GPU with the given solver options."""
        o = options or SolverOptions()
        bb = np.ascontiguousarray(b, dtype=np.float64)
        k = 1 if bb.ndim == 1 else bb.shape[1]
        co = o.to_c()
        co.nrhs = k
        co.warm_start = 1 if x0 is not None else 0
        x = np.zeros_like(bb) if x0 is None else np.array(x0, dtype=np.float64).reshape(bb.shape)
        cf = np.ascontiguousarray(coef, dtype=np.float64)
        st = L.fem_step_stats()
        L.check(self._lib.fem_ctx_solve(self.handle, L.ptr(cf, C.c_double), C.byref(co), L.ptr(bb, C.c_double),
                                        L.ptr(x, C.c_double), C.byref(st)))
        self.last_solve_stats = st
        return x

    def smooth_normals(self, normals, values, extrusion: float):
        """Summed (unnormalised) normals of the surface extruded by `values`
        (smooth_normals.glsl:22-57), computed on the GPU."""
        nrm = np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 3)
        val = np.ascontiguousarray(values, dtype=np.float32).reshape(-1)
        out = np.empty_like(nrm)
        L.check(self._lib.fem_ctx_smooth_normals(self.handle, L.ptr(nrm, C.c_float), L.ptr(val, C.c_float),
                                                 float(extrusion), L.ptr(out, C.c_float)))
        return out

    def close(self):
        if getattr(self, "handle", None):
            self._lib.fem_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        _finalize(self)


class HIPSolver:
    """The third backend behind the reference's Solver interface."""

    def __init__(self, fem_ctx: FEMContext, options: Optional[SolverOptions] = None, **kw):
        self._lib = L.load()
        self.fem_ctx = fem_ctx
        self.options = options or SolverOptions(**kw)
        h = C.c_void_p()
        o = self.options.to_c()
        L.check(self._lib.fem_solver_create(fem_ctx.handle, C.byref(o), C.byref(h)))
        self.handle = h
        _register(self)
        self.last_stats = L.fem_step_stats()
        self._eq = None
        self.set_equation()

    # -- parameters -------------------------------------------------------
    def set_equation(self, equation: Optional[int] = None, params=None):
        if equation is None:
            equation = self.fem_ctx.equation
        if params is None:
            params = self.fem_ctx.parameters[equation]
        eq, cp = params.to_c()
        if eq != equation:
            raise ValueError("parameter class does not match the equation")
        L.check(self._lib.fem_solver_set_equation(self.handle, eq, C.byref(cp), C.sizeof(cp)))
        self._eq = equation

    # -- Solver interface -------------------------------------------------
    def advance_time(self, nsteps: int = 1) -> int:
        """Solver::advance_time; returns FEM_OK / FEM_NOT_CONVERGED / FEM_DIVERGED."""
        self.set_equation(self.fem_ctx.equation)
        return L.check(self._lib.fem_solver_advance(self.handle, int(nsteps), C.byref(self.last_stats)))

    def advance(self, nsteps: int = 1) -> int:
        """Advance with the equation already set (no parameter re-upload)."""
        return L.check(self._lib.fem_solver_advance(self.handle, int(nsteps), C.byref(self.last_stats)))

    def clear_values(self):
        L.check(self._lib.fem_solver_clear(self.handle))

    def has_numerical_instability(self) -> bool:
        out = C.c_int32()
        L.check(self._lib.fem_solver_unstable(self.handle, C.byref(out)))
        return bool(out.value)

    # -- state --------------------------------------------------------------
    def _ncols(self, field_id):
        return 1 if self._eq == Equation.Reaction_Diffusion else self.options.nrhs

    def set_state(self, field_id: int, values_nv):
        v = np.ascontiguousarray(values_nv, dtype=np.float32).reshape(-1)
        L.check(self._lib.fem_solver_set_state(self.handle, int(field_id), L.ptr(v, C.c_float)))

    def get_state(self, field_id: int = L.FEM_FIELD_U):
        nv = self.fem_ctx.num_nodes()
        out = np.empty(nv * self._ncols(field_id), dtype=np.float32)
        L.check(self._lib.fem_solver_get_state(self.handle, int(field_id), L.ptr(out, C.c_float)))
        return out

    def set_state_unknowns(self, field_id: int, x):
        v = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        L.check(self._lib.fem_solver_set_state_unknowns(self.handle, int(field_id), L.ptr(v, C.c_double)))

    def get_state_unknowns(self, field_id: int = L.FEM_FIELD_U):
        n = self.fem_ctx.num_unknowns()
        out = np.empty(n * self._ncols(field_id), dtype=np.float64)
        L.check(self._lib.fem_solver_get_state_unknowns(self.handle, int(field_id), L.ptr(out, C.c_double)))
        return out

    def device_ptr(self, field_id: int = L.FEM_FIELD_U) -> int:
        p = C.c_void_p()
        L.check(self._lib.fem_solver_state_device_ptr(self.handle, int(field_id), C.byref(p)))
        return p.value

    def sync(self):
        L.check(self._lib.fem_solver_sync(self.handle))

    def smooth_normals(self, normals, extrusion: float):
        """Normals of the surface extruded by the current field, from the
        device state (Surface::calculate_normals after advance_time)."""
        nrm = np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 3)
        out = np.empty_like(nrm)
        L.check(self._lib.fem_solver_smooth_normals(self.handle, L.ptr(nrm, C.c_float), float(extrusion),
                                                    L.ptr(out, C.c_float)))
        return out

    def export_ply(self, path: str, normals, vertex_extrusion: float = 0.25, threshold: float = 0.0,
                   mesh_type: int = L.FEM_MESH_OPEN):
        """Surface::export_to_ply of the current field, read from the device
        (fem_solver_export_ply); normals: the mesh's vertex normals."""
        nrm = np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 3)
        L.check(self._lib.fem_solver_export_ply(self.handle, L.ptr(nrm, C.c_float), float(vertex_extrusion),
                                                float(threshold), int(mesh_type), str(path).encode()))

    def init_newmark_acceleration(self):
        """Consistent initial acceleration of the Newmark wave from the current
        u: M a = -c^2 K u (both products and the solve on the GPU)."""
        prm = self.fem_ctx.parameters[Equation.Wave_Newmark]
        u = self.get_state_unknowns(L.FEM_FIELD_U)
        op = Operator(self.fem_ctx, (0.0, -prm.c * prm.c, 0.0), L.FEM_PREC_FP64)
        rhs = op.apply(u)
        op.close()
        o = SolverOptions(precision=L.FEM_PREC_FP64, tol=min(self.options.tol, 1e-12))
        a = self.fem_ctx.solve((1.0, 0.0, 0.0), rhs, o)
        self.set_state_unknowns(L.FEM_FIELD_A, a)
        return a

    def autotune(self, probe_steps: int = 20):
        """Pick the team size by timing probe_steps under a few launch plans
        (the state is restored); returns (team_size, best probe ms)."""
        c = C.c_int32()
        ms = C.c_double()
        L.check(self._lib.fem_solver_autotune(self.handle, int(probe_steps), C.byref(c), C.byref(ms)))
        if c.value == 0:
            self.options.engine = L.FEM_ENGINE_STREAM
        self.options.team_size = c.value
        return c.value, ms.value

    def close(self):
        if getattr(self, "handle", None):
            self._lib.fem_solver_destroy(self.handle)
            self.handle = None

    def __del__(self):
        _finalize(self)


class Batch:
    """Many contexts advanced together in one persistent launch."""

    def __init__(self, contexts: Sequence[FEMContext], options: Optional[SolverOptions] = None, **kw):
        self._lib = L.load()
        self.contexts = list(contexts)
        self.options = options or SolverOptions(**kw)
        arr = (C.c_void_p * len(self.contexts))(*[c.handle.value for c in self.contexts])
        h = C.c_void_p()
        o = self.options.to_c()
        L.check(self._lib.fem_batch_create(arr, len(self.contexts), C.byref(o), C.byref(h)))
        self.handle = h
        _register(self)
        self.last_stats = L.fem_step_stats()

    def set_equation(self, instance: int, params):
        eq, cp = params.to_c()
        L.check(self._lib.fem_batch_set_equation(self.handle, int(instance), eq, C.byref(cp), C.sizeof(cp)))

    def set_state_unknowns(self, instance: int, field_id: int, x):
        v = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        L.check(self._lib.fem_batch_set_state_unknowns(self.handle, int(instance), int(field_id),
                                                       L.ptr(v, C.c_double)))

    def get_state_unknowns(self, instance: int, field_id: int = L.FEM_FIELD_U):
        n = self.contexts[instance].num_unknowns()
        out = np.empty(n * self.options.nrhs, dtype=np.float64)
        L.check(self._lib.fem_batch_get_state_unknowns(self.handle, int(instance), int(field_id),
                                                       L.ptr(out, C.c_double)))
        return out

    def advance(self, nsteps: int = 1) -> int:
        return L.check(self._lib.fem_batch_advance(self.handle, int(nsteps), C.byref(self.last_stats)))

    def sync(self):
        L.check(self._lib.fem_batch_sync(self.handle))

    def instance_stats(self, instance: int):
        """Counters of one instance over the last advance (fem_step_stats)."""
        st = L.fem_step_stats()
        L.check(self._lib.fem_batch_instance_stats(self.handle, int(instance), C.byref(st)))
        return st

    def graph_capture(self, nsteps: int):
        """Record one advance(nsteps) as a HIP graph (persistent or streaming
        engine; a streaming batch needs one eager advance first, see
        fem_batch_graph_capture)."""
        L.check(self._lib.fem_batch_graph_capture(self.handle, int(nsteps)))

    def graph_launch(self) -> int:
        """Replay the recorded advance; returns the step status as advance()."""
        return L.check(self._lib.fem_batch_graph_launch(self.handle, C.byref(self.last_stats)))

    def set_wait_limit(self, seconds: float):
        """Longest barrier wait of the persistent kernel before the launch
        aborts with FEM_ERR_TIMEOUT (default 4 s; 0: no waiting at all)."""
        L.check(self._lib.fem_batch_set_wait_limit(self.handle, float(seconds)))

    def plan_info(self):
        """Persistent-engine launch plan: grid, teams, exchange mode and the
        co-resident limit of the launched variant (fem_plan_info)."""
        info = L.fem_plan_info()
        L.check(self._lib.fem_batch_plan_info(self.handle, C.byref(info)))
        return info

    def autotune(self, probe_steps: int = 5):
        """Pick the workgroup budget of the team plan (state restored);
        returns (budget, best probe ms)."""
        c = C.c_int32()
        ms = C.c_double()
        L.check(self._lib.fem_batch_autotune(self.handle, int(probe_steps), C.byref(c), C.byref(ms)))
        if c.value == 0:
            self.options.engine = L.FEM_ENGINE_STREAM
        return c.value, ms.value

    def close(self):
        if getattr(self, "handle", None):
            self._lib.fem_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        _finalize(self)


class Operator:
    """S = cM M + cK K + cA A of an assembled context (or its transpose) as a
    stand-alone CSR operator: the sparse product of the reference's solvers
    (CPUSolver.cpp:78,96,114,134,137; cgm.glsl:75-102)."""

    def __init__(self, fem_ctx: FEMContext, coef=(0.0, 1.0, 0.0), precision: int = L.FEM_PREC_FP64,
                 transpose: bool = False):
        self._lib = L.load()
        self.fem_ctx = fem_ctx
        self.precision = precision
        self.n = fem_ctx.num_unknowns()
        c = np.ascontiguousarray(coef, dtype=np.float64)
        h = C.c_void_p()
        L.check(self._lib.fem_operator_create(fem_ctx.handle, L.ptr(c, C.c_double), int(precision),
                                              1 if transpose else 0, C.byref(h)))
        self.handle = h
        _register(self)

    def apply(self, x):
        """y = S x for x of shape [N] or [N, nrhs] (host arrays, FP64 in and out)."""
        xa = np.ascontiguousarray(x, dtype=np.float64)
        k = 1 if xa.ndim == 1 else xa.shape[1]
        y = np.empty_like(xa)
        L.check(self._lib.fem_operator_apply(self.handle, k, L.ptr(xa, C.c_double), L.ptr(y, C.c_double)))
        return y

    def apply_device(self, nrhs: int, x_ptr: int, y_ptr: int):
        """Enqueue y = S x on device pointers of the operator's value type."""
        L.check(self._lib.fem_operator_apply_device(self.handle, int(nrhs), C.c_void_p(x_ptr), C.c_void_p(y_ptr)))

    def time(self, nrhs: int, x_ptr: int, y_ptr: int, reps: int = 20) -> float:
        """Average ms per application (HIP events on the operator's stream)."""
        ms = C.c_double()
        L.check(self._lib.fem_operator_time(self.handle, int(nrhs), C.c_void_p(x_ptr), C.c_void_p(y_ptr),
                                            int(reps), C.byref(ms)))
        return ms.value

    def autotune(self, nrhs: int, x_ptr: int, y_ptr: int, reps: int = 10):
        """Pick the SpMV schedule for nrhs; returns (schedule, ms)."""
        g = C.c_int32()
        ms = C.c_double()
        L.check(self._lib.fem_operator_autotune(self.handle, int(nrhs), C.c_void_p(x_ptr), C.c_void_p(y_ptr),
                                                int(reps), C.byref(g), C.byref(ms)))
        return g.value, ms.value

    def set_schedule(self, nrhs: int, schedule: int):
        L.check(self._lib.fem_operator_set_schedule(self.handle, int(nrhs), int(schedule)))

    def sync(self):
        L.check(self._lib.fem_operator_sync(self.handle))

    def close(self):
        if getattr(self, "handle", None):
            self._lib.fem_operator_destroy(self.handle)
            self.handle = None

    def __del__(self):
        _finalize(self)


class Adjoint:
    """Backward-Euler heat / advection-diffusion rollout with its discrete
    adjoint: gradients of L(u_T) w.r.t. u_0 and the per-element diffusivity,
    by checkpoint-and-recompute on the GPU."""

    def __init__(self, fem_ctx: FEMContext, params, options: Optional[SolverOptions] = None, **kw):
        self._lib = L.load()
        self.fem_ctx = fem_ctx
        self.options = options or SolverOptions(**kw)
        h = C.c_void_p()
        o = self.options.to_c()
        L.check(self._lib.fem_adjoint_create(fem_ctx.handle, C.byref(o), C.byref(h)))
        self.handle = h
        _register(self)
        self.forward_stats = L.fem_step_stats()
        self.backward_stats = L.fem_step_stats()
        self.set_equation(params)

    def set_equation(self, params):
        eq, cp = params.to_c()
        L.check(self._lib.fem_adjoint_set_equation(self.handle, eq, C.byref(cp), C.sizeof(cp)))
        self.params = params

    def forward(self, u0, nsteps: int, checkpoint_every: int = 10):
        """u_T of the rollout from u_0 (unknown space); keeps checkpoints."""
        u = np.ascontiguousarray(u0, dtype=np.float64)
        uT = np.empty_like(u)
        L.check(self._lib.fem_adjoint_forward(self.handle, L.ptr(u, C.c_double), int(nsteps),
                                              int(checkpoint_every), L.ptr(uT, C.c_double),
                                              C.byref(self.forward_stats)))
        return uT

    def backward(self, grad_uT):
        """(dL/du_0, dL/dkappa_e) for dL/du_T of the last forward."""
        g = np.ascontiguousarray(grad_uT, dtype=np.float64)
        gu = np.empty_like(g)
        gk = np.empty(self.fem_ctx.mesh["tri"].shape[0], dtype=np.float64)
        L.check(self._lib.fem_adjoint_backward(self.handle, L.ptr(g, C.c_double), L.ptr(gu, C.c_double),
                                               L.ptr(gk, C.c_double), C.byref(self.backward_stats)))
        return gu, gk

    def forward_device(self, u0_ptr: int, uT_ptr: int, nsteps: int, checkpoint_every: int = 10):
        """forward on device pointers of FP64 [N] arrays (inputs complete)."""
        L.check(self._lib.fem_adjoint_forward_device(self.handle, C.c_void_p(u0_ptr), int(nsteps),
                                                     int(checkpoint_every), C.c_void_p(uT_ptr),
                                                     C.byref(self.forward_stats)))

    def backward_device(self, grad_uT_ptr: int, grad_u0_ptr: int, grad_kappa_ptr: int):
        """backward on device pointers (FP64 [N], [N], [num_triangles])."""
        L.check(self._lib.fem_adjoint_backward_device(self.handle, C.c_void_p(grad_uT_ptr), C.c_void_p(grad_u0_ptr),
                                                      C.c_void_p(grad_kappa_ptr), C.byref(self.backward_stats)))

    def close(self):
        if getattr(self, "handle", None):
            self._lib.fem_adjoint_destroy(self.handle)
            self.handle = None

    def __del__(self):
        _finalize(self)
