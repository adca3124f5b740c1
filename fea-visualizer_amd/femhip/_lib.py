"""ctypes binding of the C ABI declared in include/femhip.h.

The shared library ``libfemhip.so`` is built in-tree (see ``femhip.build``).
Loading fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfemhip.so")

# status codes
FEM_OK = 0
FEM_NOT_CONVERGED = 1
FEM_DIVERGED = 2
FEM_ERR_INVALID = -1
FEM_ERR_HIP = -2
FEM_ERR_OOM = -3
FEM_ERR_STATE = -4
FEM_ERR_TIMEOUT = -5

# equations (0-3 as the reference's enum class Equation)
FEM_HEAT = 0
FEM_WAVE = 1
FEM_ADVECTION_DIFFUSION = 2
FEM_REACTION_DIFFUSION = 3
FEM_WAVE_NEWMARK = 4

FEM_DIRICHLET = 0
FEM_NEUMANN = 1

FEM_FIELD_U = 0
FEM_FIELD_V = 1
FEM_FIELD_A = 2

FEM_MATRIX_K = 0
FEM_MATRIX_M = 1
FEM_MATRIX_A = 2

FEM_ENGINE_PERSISTENT = 0
FEM_ENGINE_PLAIN = 1
FEM_ENGINE_STREAM = 2
FEM_ENGINE_GL_COMPAT = 3
FEM_PREC_FP64 = 0
FEM_PREC_FP32_IR = 1
FEM_KRYLOV_AUTO = 0
FEM_KRYLOV_CG = 1
FEM_KRYLOV_BICGSTAB = 2
FEM_KRYLOV_CG_CLASSIC = 3
FEM_REACTION_NODAL = 0
FEM_REACTION_LUMPED = 1
FEM_EXCHANGE_AUTO = 0
FEM_EXCHANGE_COHERENT = 1
FEM_EXCHANGE_FENCED = 2
FEM_PREDICTOR_LINEAR = 0
FEM_PREDICTOR_PREVIOUS = 1
FEM_PREDICTOR_QUADRATIC = 2
FEM_PREDICTOR_CUBIC = 3
FEM_PREDICTOR_QUARTIC = 4
FEM_PREDICTOR_QUINTIC = 5
FEM_PREDICTOR_SEXTIC = 6
FEM_MESH_OPEN = 0
FEM_MESH_CLOSED = 1
FEM_MESH_MIRRORED = 2
FEM_SPMV_STREAM = 32
FEM_SPMV_STREAM_NT = 33
FEM_SPMV_PIPE = 34
FEM_SPMV_PIPE_NT = 35
SPMV_SCHEDULES = (1, 2, 4, 8, 16, FEM_SPMV_STREAM, FEM_SPMV_STREAM_NT, FEM_SPMV_PIPE, FEM_SPMV_PIPE_NT)


class fem_mesh(C.Structure):
    _fields_ = [
        ("num_vertices", C.c_int32),
        ("num_triangles", C.c_int32),
        ("xyz", C.POINTER(C.c_float)),
        ("triangles", C.POINTER(C.c_uint32)),
        ("on_boundary", C.POINTER(C.c_uint8)),
    ]


class fem_assembly_params(C.Structure):
    _fields_ = [
        ("velocity", C.c_double * 3),
        ("velocity_e", C.POINTER(C.c_double)),
        ("normalize_velocity", C.c_int32),
        ("kappa_e", C.POINTER(C.c_double)),
    ]


class fem_heat_params(C.Structure):
    _fields_ = [("time_step", C.c_double), ("conductivity", C.c_double)]


class fem_advdiff_params(C.Structure):
    _fields_ = [("time_step", C.c_double), ("c", C.c_double)]


class fem_wave_params(C.Structure):
    _fields_ = [("time_step", C.c_double), ("c", C.c_double)]


class fem_rd_params(C.Structure):
    _fields_ = [
        ("time_step", C.c_double),
        ("Du", C.c_double),
        ("Dv", C.c_double),
        ("feed_rate", C.c_double),
        ("kill_rate", C.c_double),
        ("reaction_weighting", C.c_int32),
    ]


class fem_newmark_params(C.Structure):
    _fields_ = [("time_step", C.c_double), ("c", C.c_double), ("beta", C.c_double), ("gamma", C.c_double)]


class fem_solver_opts(C.Structure):
    _fields_ = [
        ("engine", C.c_int32),
        ("precision", C.c_int32),
        ("krylov", C.c_int32),
        ("nrhs", C.c_int32),
        ("warm_start", C.c_int32),
        ("max_iters", C.c_int32),
        ("max_inner", C.c_int32),
        ("max_sweeps", C.c_int32),
        ("team_size", C.c_int32),
        ("tol", C.c_double),
        ("inner_tol", C.c_double),
        ("exchange", C.c_int32),
        ("predictor", C.c_int32),
    ]


class fem_step_stats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int64),
        ("sweeps", C.c_int64),
        ("steps", C.c_int64),
        ("residual", C.c_double),
        ("ms", C.c_double),
        ("status", C.c_int32),
    ]


class fem_ctx_info(C.Structure):
    _fields_ = [
        ("num_nodes", C.c_int32),
        ("num_unknowns", C.c_int32),
        ("num_triangles", C.c_int32),
        ("max_row_nonzeros", C.c_int32),
        ("nnz", C.c_int64),
    ]


class fem_plan_info(C.Structure):
    _fields_ = [
        ("grid", C.c_int32),
        ("teams", C.c_int32),
        ("exchange", C.c_int32),
        ("coresident", C.c_int32),
    ]


P = C.c_void_p
I32 = C.c_int32
PD = C.POINTER(C.c_double)
PF = C.POINTER(C.c_float)
PI = C.POINTER(C.c_int32)

# (name, restype, argtypes) for every function declared in include/femhip.h
SIGNATURES = [
    ("fem_last_error", C.c_char_p, []),
    ("fem_abi_version", C.c_int, []),
    ("fem_device_count", C.c_int, [PI]),
    ("fem_opts_default", None, [C.POINTER(fem_solver_opts)]),
    ("fem_ctx_create", C.c_int, [C.POINTER(fem_mesh), I32, C.POINTER(P)]),
    ("fem_ctx_destroy", C.c_int, [P]),
    ("fem_ctx_assemble", C.c_int, [P, C.POINTER(fem_assembly_params)]),
    ("fem_ctx_info_get", C.c_int, [P, C.POINTER(fem_ctx_info)]),
    ("fem_ctx_get_idx_map", C.c_int, [P, PI]),
    ("fem_ctx_get_pattern", C.c_int, [P, PI, PI]),
    ("fem_ctx_get_matrix", C.c_int, [P, I32, PD]),
    ("fem_ctx_solve", C.c_int, [P, PD, C.POINTER(fem_solver_opts), PD, PD, C.POINTER(fem_step_stats)]),
    ("fem_ctx_solve_noalloc", C.c_int, [P, PD, C.POINTER(fem_solver_opts), PD, PD, C.POINTER(fem_step_stats)]),
    ("fem_ctx_smooth_normals", C.c_int, [P, PF, PF, C.c_float, PF]),
    ("fem_surface_from_obj", C.c_int, [C.c_char_p, C.POINTER(P)]),
    ("fem_surface_get", C.c_int, [P, C.POINTER(fem_mesh), C.POINTER(PF), PI]),
    ("fem_surface_destroy", C.c_int, [P]),
    ("fem_export_ply", C.c_int, [C.POINTER(fem_mesh), PF, PF, C.c_float, C.c_float, I32, C.c_char_p]),
    ("fem_solver_create", C.c_int, [P, C.POINTER(fem_solver_opts), C.POINTER(P)]),
    ("fem_solver_destroy", C.c_int, [P]),
    ("fem_solver_set_equation", C.c_int, [P, I32, P, C.c_size_t]),
    ("fem_solver_set_state", C.c_int, [P, I32, PF]),
    ("fem_solver_get_state", C.c_int, [P, I32, PF]),
    ("fem_solver_set_state_unknowns", C.c_int, [P, I32, PD]),
    ("fem_solver_get_state_unknowns", C.c_int, [P, I32, PD]),
    ("fem_solver_advance", C.c_int, [P, I32, C.POINTER(fem_step_stats)]),
    ("fem_solver_clear", C.c_int, [P]),
    ("fem_solver_unstable", C.c_int, [P, PI]),
    ("fem_solver_state_device_ptr", C.c_int, [P, I32, C.POINTER(P)]),
    ("fem_solver_sync", C.c_int, [P]),
    ("fem_solver_smooth_normals", C.c_int, [P, PF, C.c_float, PF]),
    ("fem_solver_export_ply", C.c_int, [P, PF, C.c_float, C.c_float, I32, C.c_char_p]),
    ("fem_solver_autotune", C.c_int, [P, I32, PI, PD]),
    ("fem_batch_create", C.c_int, [C.POINTER(P), I32, C.POINTER(fem_solver_opts), C.POINTER(P)]),
    ("fem_batch_destroy", C.c_int, [P]),
    ("fem_batch_set_equation", C.c_int, [P, I32, I32, P, C.c_size_t]),
    ("fem_batch_set_state_unknowns", C.c_int, [P, I32, I32, PD]),
    ("fem_batch_get_state_unknowns", C.c_int, [P, I32, I32, PD]),
    ("fem_batch_advance", C.c_int, [P, I32, C.POINTER(fem_step_stats)]),
    ("fem_batch_sync", C.c_int, [P]),
    ("fem_batch_instance_stats", C.c_int, [P, I32, C.POINTER(fem_step_stats)]),
    ("fem_batch_graph_capture", C.c_int, [P, I32]),
    ("fem_batch_graph_launch", C.c_int, [P, C.POINTER(fem_step_stats)]),
    ("fem_batch_plan_info", C.c_int, [P, C.POINTER(fem_plan_info)]),
    ("fem_batch_set_wait_limit", C.c_int, [P, C.c_double]),
    ("fem_batch_autotune", C.c_int, [P, I32, PI, PD]),
    ("fem_operator_create", C.c_int, [P, PD, I32, I32, C.POINTER(P)]),
    ("fem_operator_destroy", C.c_int, [P]),
    ("fem_operator_apply_device", C.c_int, [P, I32, P, P]),
    ("fem_operator_apply", C.c_int, [P, I32, PD, PD]),
    ("fem_operator_time", C.c_int, [P, I32, P, P, I32, PD]),
    ("fem_operator_autotune", C.c_int, [P, I32, P, P, I32, PI, PD]),
    ("fem_operator_set_schedule", C.c_int, [P, I32, I32]),
    ("fem_operator_sync", C.c_int, [P]),
    ("fem_adjoint_create", C.c_int, [P, C.POINTER(fem_solver_opts), C.POINTER(P)]),
    ("fem_adjoint_destroy", C.c_int, [P]),
    ("fem_adjoint_set_equation", C.c_int, [P, I32, P, C.c_size_t]),
    ("fem_adjoint_forward", C.c_int, [P, PD, I32, I32, PD, C.POINTER(fem_step_stats)]),
    ("fem_adjoint_backward", C.c_int, [P, PD, PD, PD, C.POINTER(fem_step_stats)]),
    ("fem_adjoint_forward_device", C.c_int, [P, P, I32, I32, P, C.POINTER(fem_step_stats)]),
    ("fem_adjoint_backward_device", C.c_int, [P, P, P, P, C.POINTER(fem_step_stats)]),
]

_lib = None


class FemError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"femhip error {code}: {msg}")
        self.code = code


def load(path: str | None = None):
    """Load libfemhip.so and bind every declared symbol (raises if absent)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FileNotFoundError(
            f"{p} is missing: build the HIP library first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(p)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def hip_runtimes():
    """Paths of the HIP runtime libraries mapped into this process.  A
    library loaded before torch binds /opt/rocm's runtime and torch then maps
    its own copy: two runtimes, whose device pointers the other does not
    know.  Loaded after torch, the library binds torch's copy (same soname)."""
    try:
        with open("/proc/self/maps") as f:
            return sorted({l.split()[-1] for l in f if "libamdhip64" in l})
    except OSError:
        return []


def last_error() -> str:
    return load().fem_last_error().decode()


def check(code: int, allow=(FEM_OK,)):
    """Raise FemError for negative codes; positive solver outcomes pass through."""
    if code < 0:
        raise FemError(code, last_error())
    return code


def ptr(arr, ctype):
    return arr.ctypes.data_as(C.POINTER(ctype))
