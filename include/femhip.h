/*
 * femhip.h -- C ABI of the MI355X (gfx950) FEM time-stepping backend.
 *
 * The library assembles P1-triangle stiffness / mass / advection matrices into
 * CSR on the GPU and advances the heat, wave, advection-diffusion and
 * Gray-Scott equations with implicit time steps, behind the reference's
 * solver interface (include/FEM/Solver.hpp:8-15 of tvumcc/fea-visualizer).
 *
 * Conventions
 *   - Every function returns an int status: FEM_OK (0), a positive solver
 *     outcome (FEM_NOT_CONVERGED, FEM_DIVERGED) or a negative error.  No C++
 *     exception crosses this boundary; fem_last_error() returns a thread-local
 *     message for the last failure.
 *   - Host arrays are owned by the caller and copied in or out.  The library
 *     owns its device memory and one HIP stream per handle.
 *   - A handle is used from one host thread at a time (the reference drives
 *     its solvers from the single GL thread).
 */
#ifndef FEMHIP_H
#define FEMHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FEMHIP_ABI_VERSION 1

/* Status codes. */
enum {
  FEM_OK = 0,
  FEM_NOT_CONVERGED = 1,   /* a solve hit its iteration cap */
  FEM_DIVERGED = 2,        /* instability rule tripped or non-finite state */
  FEM_ERR_INVALID = -1,    /* bad argument */
  FEM_ERR_HIP = -2,        /* HIP runtime error (no GPU, launch failure, ...) */
  FEM_ERR_OOM = -3,
  FEM_ERR_STATE = -4,      /* call out of order (e.g. advance before set_equation) */
  FEM_ERR_TIMEOUT = -5     /* a device-side wait exceeded its spin limit */
};

/* Equation ids: 0-3 as enum class Equation (include/FEM/FEMContext.hpp:9-14). */
enum {
  FEM_HEAT = 0,
  FEM_WAVE = 1,                  /* the reference's velocity-form backward Euler */
  FEM_ADVECTION_DIFFUSION = 2,
  FEM_REACTION_DIFFUSION = 3,    /* Gray-Scott */
  FEM_WAVE_NEWMARK = 4           /* no reference counterpart: Newmark-beta */
};

/* Boundary conditions: enum class BoundaryCondition (FEMContext.hpp:16-19). */
enum { FEM_DIRICHLET = 0, FEM_NEUMANN = 1 };

/* State fields. */
enum { FEM_FIELD_U = 0, FEM_FIELD_V = 1, FEM_FIELD_A = 2 };

/* Matrices held by a context. */
enum { FEM_MATRIX_K = 0, FEM_MATRIX_M = 1, FEM_MATRIX_A = 2 };

/* Solver engines and arithmetic. */
/* PERSISTENT: one cooperative launch per advance (latency-bound instances);
 * PLAIN: FP64 baseline, one right-hand side at a time with host-read scalars;
 * STREAM: each phase a full-occupancy launch over the row blocks of all
 * instances (bandwidth-bound instances: large meshes, many right-hand sides). */
enum { FEM_ENGINE_PERSISTENT = 0, FEM_ENGINE_PLAIN = 1, FEM_ENGINE_STREAM = 2, FEM_ENGINE_GL_COMPAT = 3 };
/* GL_COMPAT reproduces the reference's OpenGL compute-shader solver for A/B
 * comparison (src/FEM/GPUSolver.cpp:74-172): FP32, warm start, unpreconditioned
 * CG with at most max_iters (default 10) iterations, Gray-Scott reaction summed
 * per row slot and the v solve after the u solve, instability when the last
 * r.r > 1e4.  Not the accurate path. */
enum { FEM_PREC_FP64 = 0, FEM_PREC_FP32_IR = 1 };
/* FEM_KRYLOV_CG is single-reduction (Chronopoulos-Gear) PCG;
 * FEM_KRYLOV_CG_CLASSIC the two-reduction form of Eigen's conjugate_gradient. */
enum { FEM_KRYLOV_AUTO = 0, FEM_KRYLOV_CG = 1, FEM_KRYLOV_BICGSTAB = 2, FEM_KRYLOV_CG_CLASSIC = 3 };
/* Gray-Scott reaction weighting: the reference adds the nodal reaction
 * without mass weighting (CPUSolver.cpp:134,137); LUMPED multiplies it by the
 * row sums of M (a documented deviation that is stable at large time steps). */
enum { FEM_REACTION_NODAL = 0, FEM_REACTION_LUMPED = 1 };

/* Mesh: Surface::{vertices, triangles, on_boundary} (Surface.hpp:36-40). */
typedef struct fem_mesh {
  int32_t num_vertices;
  int32_t num_triangles;
  const float* xyz;            /* 3 * num_vertices */
  const uint32_t* triangles;   /* 3 * num_triangles */
  const uint8_t* on_boundary;  /* num_vertices, 0/1 */
} fem_mesh;

/* Assembly inputs: AdvectionDiffusionParameters::velocity
 * (EquationParameters.hpp:17-26, read at FEMContext.cpp:58) plus extensions. */
typedef struct fem_assembly_params {
  double velocity[3];          /* constant velocity */
  const double* velocity_e;    /* optional: 3 per triangle (overrides velocity) */
  int32_t normalize_velocity;  /* 1 = reference (unit tangential velocity, FEMContext.cpp:180) */
  const double* kappa_e;       /* optional: per-triangle diffusivity multiplying K */
} fem_assembly_params;

/* Equation parameters: EquationParameters.hpp:4-49 (widened to double). */
typedef struct fem_heat_params { double time_step; double conductivity; } fem_heat_params;
typedef struct fem_advdiff_params { double time_step; double c; } fem_advdiff_params;
typedef struct fem_wave_params { double time_step; double c; } fem_wave_params;
typedef struct fem_rd_params {
  double time_step;
  double Du;
  double Dv;
  double feed_rate;
  double kill_rate;
  int32_t reaction_weighting;  /* FEM_REACTION_NODAL (reference) or FEM_REACTION_LUMPED */
} fem_rd_params;
typedef struct fem_newmark_params { double time_step; double c; double beta; double gamma; } fem_newmark_params;

typedef struct fem_solver_opts {
  int32_t engine;        /* FEM_ENGINE_PERSISTENT (default) or FEM_ENGINE_PLAIN (FP64 baseline) */
  int32_t precision;     /* FEM_PREC_FP32_IR (default) or FEM_PREC_FP64 */
  int32_t krylov;        /* FEM_KRYLOV_AUTO: CG, BiCGSTAB for advection-diffusion */
  int32_t nrhs;          /* right-hand sides per instance: 1, 2, 4 or 16 (Gray-Scott: 1) */
  int32_t warm_start;    /* 1: start from the previous state; 0: x0 = 0 as Eigen's solve() */
  int32_t max_iters;     /* Krylov iterations per solve; 0 = 2N (Eigen's default) */
  int32_t max_inner;     /* FP32 iterations per refinement sweep; 0 = 200 */
  int32_t max_sweeps;    /* refinement sweeps per solve; 0 = 30 */
  int32_t team_size;     /* persistent engine: workgroups per instance; stream engine: rows per
                            workgroup; 0 = autotuned / by column count */
  double tol;            /* relative residual |b - Ax| <= tol |b| in FP64; 0 = 1e-10 */
  double inner_tol;      /* FP32 sweep reduction; 0 = 1e-4 */
  int32_t exchange;      /* multi-workgroup teams: FEM_EXCHANGE_AUTO (row-cache rule or
                            autotuner), _COHERENT or _FENCED */
  int32_t predictor;     /* initial guess with warm_start, from the previous solutions:
                            FEM_PREDICTOR_SEXTIC (the fem_opts_default choice; see below),
                            FEM_PREDICTOR_QUARTIC (5 x_n - 10 x_{n-1} + 10 x_{n-2} -
                            5 x_{n-3} + x_{n-4}; two more state-sized history vectors),
                            FEM_PREDICTOR_CUBIC (4 x_n -
                            6 x_{n-1} + 4 x_{n-2} - x_{n-3}; one more),
                            FEM_PREDICTOR_QUADRATIC (3 x_n - 3 x_{n-1} + x_{n-2}),
                            FEM_PREDICTOR_LINEAR (0: 2 x_n - x_{n-1}),
                            FEM_PREDICTOR_PREVIOUS (x_n), or on the streaming engine
                            FEM_PREDICTOR_QUINTIC / _SEXTIC (orders 5 and 6, weights
                            (-1)^j C(order + 1, j + 1) on x_{n-j}; three and four more
                            vectors; the persistent engine and the first step of each
                            streaming launch use quartic at most).  Heat, advection-diffusion and
                            Gray-Scott take any order; the wave equation is linear at most;
                            Newmark extrapolates its acceleration, quadratic at most. */
} fem_solver_opts;

enum {
  FEM_PREDICTOR_LINEAR = 0,
  FEM_PREDICTOR_PREVIOUS = 1,
  FEM_PREDICTOR_QUADRATIC = 2,
  FEM_PREDICTOR_CUBIC = 3,
  FEM_PREDICTOR_QUARTIC = 4,
  FEM_PREDICTOR_QUINTIC = 5,  /* streaming engine; the persistent engine uses quartic at most */
  FEM_PREDICTOR_SEXTIC = 6    /* streaming engine; the persistent engine uses quartic at most */
};

/* How the workgroups of one team exchange vectors between barriers:
 * COHERENT: device-coherent loads/stores, fence-free barriers (latency-bound
 * phases); FENCED: cached accesses, L2 write-back/invalidate at barriers
 * (bandwidth-bound phases). */
enum { FEM_EXCHANGE_AUTO = 0, FEM_EXCHANGE_COHERENT = 1, FEM_EXCHANGE_FENCED = 2 };

typedef struct fem_step_stats {
  int64_t iterations;    /* Krylov iterations summed over solves and steps */
  int64_t sweeps;        /* FP64 refinement sweeps */
  int64_t steps;
  double residual;       /* largest final relative FP64 residual of the last step */
  double ms;             /* device time of the call */
  int32_t status;        /* OR of the per-instance outcome bits */
} fem_step_stats;

typedef struct fem_ctx_info {
  int32_t num_nodes;        /* FEMContext::num_nodes */
  int32_t num_unknowns;     /* FEMContext::num_unknowns */
  int32_t num_triangles;
  int32_t max_row_nonzeros; /* FEMContext::num_max_nonzeros_per_row */
  int64_t nnz;
} fem_ctx_info;

typedef struct fem_ctx fem_ctx;
typedef struct fem_solver fem_solver;
typedef struct fem_batch fem_batch;
typedef struct fem_adjoint fem_adjoint;
typedef struct fem_operator fem_operator;

const char* fem_last_error(void);
int fem_abi_version(void);
/* Number of visible GPUs (0 when none); never throws. */
int fem_device_count(int32_t* count);
void fem_opts_default(fem_solver_opts* opts);

/* ---- context: mesh, idx_map, CSR pattern, assembled matrices ----
 * fem_ctx_create replaces FEMContext::init_from_surface +
 * update_boundary_conditions (FEMContext.cpp:21-27, 224-240);
 * fem_ctx_assemble replaces FEMContext::assemble_matrices
 * (FEMContext.cpp:57-64, 70-219).  create also assembles with the reference
 * defaults (velocity (1,0,0), normalised). */
int fem_ctx_create(const fem_mesh* mesh, int32_t bc, fem_ctx** out);
int fem_ctx_destroy(fem_ctx* ctx);
int fem_ctx_assemble(fem_ctx* ctx, const fem_assembly_params* params);
/* FEMContext::num_nodes / num_unknowns / num_max_nonzeros_per_row (FEMContext.cpp:33-52). */
int fem_ctx_info_get(const fem_ctx* ctx, fem_ctx_info* info);
int fem_ctx_get_idx_map(const fem_ctx* ctx, int32_t* idx_map);
int fem_ctx_get_pattern(const fem_ctx* ctx, int32_t* row_ptr, int32_t* col_idx);
int fem_ctx_get_matrix(const fem_ctx* ctx, int32_t which, double* values);
/* One solve of (coef[0] M + coef[1] K + coef[2] A) x = b with the solver
 * options (precision, Krylov method, tolerance; nrhs columns, arrays
 * [N][nrhs]): the cg.compute / cg.solve pair of CPUSolver.cpp:80-81 for a
 * caller-given matrix combination.  x is the initial guess when
 * opts->warm_start, else x0 = 0 as Eigen's solve(). */
int fem_ctx_solve(fem_ctx* ctx, const double coef[3], const fem_solver_opts* opts, const double* b, double* x,
                  fem_step_stats* stats);
/* fem_ctx_solve without any allocation: no reference counterpart; for callers
 * that cannot allocate device memory (tight serving loops, code that must not
 * call hipMalloc).  Same arguments, semantics and results as fem_ctx_solve,
 * on the persistent engine only.  Every device buffer the solve needs beyond
 * the caller's b and x (formed matrix, Krylov vectors, device copies of b and
 * x, barrier slots) lives in a static 32 MiB array per GPU that the library
 * holds until the process ends; systems that need more return
 * FEM_ERR_INVALID (about 180,000 unknowns fit at one right-hand side in
 * FP32 + refinement).  Calls on one GPU share that array: calls that overlap
 * on one GPU race.  Runs on the context's stream; stats->ms is the call's
 * wall time. */
int fem_ctx_solve_noalloc(fem_ctx* ctx, const double coef[3], const fem_solver_opts* opts, const double* b,
                          double* x, fem_step_stats* stats);
/* Smooth vertex normals of the surface extruded by per-vertex values
 * (shaders/FEM/smooth_normals.glsl:22-57): every vertex moves along its
 * normal by min(1, extrusion * max(0, value)); each triangle's unnormalised
 * normal cross(b - a, c - a) of the moved positions is summed into its three
 * vertices.  Each vertex sums its incident triangles in ascending order, so
 * the result is deterministic (the shader accumulates with float
 * compare-and-swap in arbitrary order).  Host arrays: normals and out
 * [nv][3], values [nv]. */
int fem_ctx_smooth_normals(fem_ctx* ctx, const float* normals_nv3, const float* values_nv, float extrusion,
                           float* out_nv3);

/* ---- surface input and output: host code around the GPU path ----
 * fem_surface_from_obj <- Surface::init_from_obj (src/Utils/Surface.cpp:57-133):
 *   faces triangulated (triangles kept; quads split along the shorter
 *   diagonal, squared lengths in float; larger polygons fanned from their
 *   first corner); every object or group appends the whole vertex list again
 *   while its face indices keep pointing at the first copy (Surface.cpp:77-80);
 *   vertex normals = glm::normalize of the summed corner normals
 *   (Surface.cpp:89-102, NaN for a vertex no face uses); boundary = endpoints
 *   of edges used by one triangle (Surface.cpp:104-127).  A missing file, a
 *   file without vertices or without normals is refused with FEM_ERR_INVALID
 *   and the reference's message.
 * fem_export_ply <- Surface::export_to_ply (src/Utils/Surface.cpp:138-285):
 *   the surface extruded along its normals by value * max(0, extrusion),
 *   clipped where values fall below threshold + 1e-9, closed by a projected
 *   copy for FEM_MESH_CLOSED / _MIRRORED, positions de-duplicated in first-use
 *   order with glm::vec3 equality, coloured by the Viridis polynomial
 *   (AssetStorage.cpp:141-152), written as ASCII PLY with the reference's
 *   header and shortest round-trip numbers. */
enum { FEM_MESH_OPEN = 0, FEM_MESH_CLOSED = 1, FEM_MESH_MIRRORED = 2 };  /* MeshType (Surface.hpp:25-29) */
typedef struct fem_surface fem_surface;
int fem_surface_from_obj(const char* path, fem_surface** out);
/* Borrowed views of the surface's arrays, valid until fem_surface_destroy:
 * the mesh to pass to fem_ctx_create, vertex normals [nv][3], boundary count
 * (Surface::num_boundary_points).  Any output pointer may be NULL. */
int fem_surface_get(const fem_surface* s, fem_mesh* mesh, const float** normals_nv3, int32_t* num_boundary_points);
int fem_surface_destroy(fem_surface* s);
/* mesh: positions and triangles (on_boundary unused); normals and values per
 * vertex ([nv][3], [nv]). */
int fem_export_ply(const fem_mesh* mesh, const float* normals_nv3, const float* values_nv, float vertex_extrusion,
                   float threshold, int32_t mesh_type, const char* path);

/* ---- solver: the Solver interface (Solver.hpp:8-15) ----
 * fem_solver_advance  <- Solver::advance_time        (CPUSolver.cpp:64-147)
 * fem_solver_clear    <- Solver::clear_values        (CPUSolver.cpp:54-59)
 * fem_solver_unstable <- Solver::has_numerical_instability (CPUSolver.cpp:41-48)
 * set/get_state work in surface space (nv values, gather/scatter through
 * idx_map as CPUSolver.cpp:13-37, on the device; boundary values read back
 * as 0).  set_state keeps every unknown whose FP64 state, rounded to float,
 * equals the value given, so writing back what get_state returned (the
 * reference's gather/step/scatter frame loop) leaves the FP64 state and the
 * predictor history as they were; any other value replaces the state and
 * restarts the predictor.  With nrhs > 1 the arrays hold nv*nrhs values,
 * right-hand side fastest. */
int fem_solver_create(fem_ctx* ctx, const fem_solver_opts* opts, fem_solver** out);
int fem_solver_destroy(fem_solver* s);
int fem_solver_set_equation(fem_solver* s, int32_t equation, const void* params, size_t size);
int fem_solver_set_state(fem_solver* s, int32_t field, const float* values_nv);
int fem_solver_get_state(fem_solver* s, int32_t field, float* values_nv);
int fem_solver_set_state_unknowns(fem_solver* s, int32_t field, const double* x);
int fem_solver_get_state_unknowns(fem_solver* s, int32_t field, double* x);
int fem_solver_advance(fem_solver* s, int32_t nsteps, fem_step_stats* stats);
int fem_solver_clear(fem_solver* s);
int fem_solver_unstable(fem_solver* s, int32_t* unstable);
/* Device pointer of a state field ([N][nrhs] doubles) for zero-copy callers. */
int fem_solver_state_device_ptr(fem_solver* s, int32_t field, void** dptr);
/* Synchronise the solver stream. */
int fem_solver_sync(fem_solver* s);
/* Smooth normals of the surface extruded by the solver's current field (the
 * visualisation step after advance_time: Surface::calculate_normals,
 * src/Utils/Surface.cpp:287-307, and shaders/FEM/smooth_normals.glsl:22-57),
 * with the values taken from the state on the device: u, or v for Gray-Scott,
 * boundary vertices 0.  normals_nv3: the mesh's vertex normals [nv][3];
 * out_nv3 receives the unnormalised sums [nv][3]. */
int fem_solver_smooth_normals(fem_solver* s, const float* normals_nv3, float extrusion, float* out_nv3);
/* PLY export of the solver's current surface field (u, or v for Gray-Scott,
 * read from the device; boundary vertices 0 as after the frame's scatter,
 * CPUSolver.cpp:28-37) on the context's mesh: fem_export_ply with the
 * caller's vertex normals [nv][3]. */
int fem_solver_export_ply(fem_solver* s, const float* normals_nv3, float vertex_extrusion, float threshold,
                          int32_t mesh_type, const char* path);
/* Autotuner (no reference counterpart): runs probe_steps under a few launch
 * plans from the current state -- persistent team sizes in both exchange
 * modes, and the streaming engine -- restores the state after each, and keeps
 * the fastest.  *chosen receives the team size (workgroups per instance), or
 * 0 when the streaming engine won. */
int fem_solver_autotune(fem_solver* s, int32_t probe_steps, int32_t* chosen, double* best_ms);

/* ---- batch: many contexts (meshes of uneven size) advanced in one
 * persistent launch; no reference counterpart ---- */
int fem_batch_create(fem_ctx* const* ctxs, int32_t count, const fem_solver_opts* opts, fem_batch** out);
int fem_batch_destroy(fem_batch* b);
int fem_batch_set_equation(fem_batch* b, int32_t instance, int32_t equation, const void* params, size_t size);
int fem_batch_set_state_unknowns(fem_batch* b, int32_t instance, int32_t field, const double* x);
int fem_batch_get_state_unknowns(fem_batch* b, int32_t instance, int32_t field, double* x);
int fem_batch_advance(fem_batch* b, int32_t nsteps, fem_step_stats* stats);
int fem_batch_sync(fem_batch* b);
/* Counters of one instance over the last advance: its Krylov iterations,
 * sweeps and steps, its final residual and outcome bits (ms is 0: device
 * time is reported per call). */
int fem_batch_instance_stats(fem_batch* b, int32_t instance, fem_step_stats* stats);
/* Record one fem_batch_advance(b, nsteps) as a HIP graph owned by the batch
 * (no memset nodes); fem_batch_graph_launch replays it and returns what
 * fem_batch_advance would.  The state may change between replays (it is read
 * from device memory).
 * Persistent engine: the barrier-slot reset kernel and the stepping kernel.
 * A rebuilt launch plan (new equation parameters or matrices, autotuning)
 * makes the graph stale: the launch then returns FEM_ERR_STATE and the batch
 * must be captured again.
 * Streaming engine: every phase, reduction and compaction launch of nsteps
 * steps, with nothing read on the host.  Each step issues the sweeps the
 * batch's last eager advance needed (its costliest step) plus one, each of
 * the iterations that sweep needed plus a quarter (at least 2); launches past
 * an instance's end exit on the device.  Capture therefore needs one eager
 * fem_batch_advance first (else FEM_ERR_STATE); a few eager steps just before
 * the capture size it for the steps that follow.  An instance that would need more iterations than
 * recorded stops there and the replay reports FEM_NOT_CONVERGED; new
 * equation parameters are read by the next replay; a rebuilt row-block plan
 * (autotuning) makes the graph stale.  A recording holds at most 8192
 * launches (FEM_ERR_INVALID beyond, naming the steps that fit): longer
 * advances replay a shorter recording several times.
 * The plain and GL-compatible engines are host-driven and are not captured. */
int fem_batch_graph_capture(fem_batch* b, int32_t nsteps);
int fem_batch_graph_launch(fem_batch* b, fem_step_stats* stats);
/* Launch plan of the persistent engine (built here if it is not yet):
 * workgroups in the grid, teams, the exchange mode of the launched variant
 * (0: one-workgroup teams, FEM_EXCHANGE_COHERENT or FEM_EXCHANGE_FENCED) and
 * how many workgroups of that variant the device holds at once.  The
 * stepping kernel's barriers need grid <= coresident; every launch checks
 * it.  No reference counterpart. */
typedef struct fem_plan_info {
  int32_t grid;
  int32_t teams;
  int32_t exchange;
  int32_t coresident;
} fem_plan_info;
int fem_batch_plan_info(fem_batch* b, fem_plan_info* info);
/* Watchdog of the persistent kernel's barriers (no reference counterpart):
 * the longest a workgroup waits at a barrier before the launch aborts and
 * the call returns FEM_ERR_TIMEOUT with the state invalid.  Default 4 s; 0
 * aborts at the first barrier wait that is not already satisfied.  Takes
 * effect at the next launch plan (a captured graph must be captured again). */
int fem_batch_set_wait_limit(fem_batch* b, double seconds);
/* Autotuner over the number of workgroups the batch plan spreads over
 * (*chosen; 0 when the streaming engine won); state restored as for
 * fem_solver_autotune. */
int fem_batch_autotune(fem_batch* b, int32_t probe_steps, int32_t* chosen, double* best_ms);

/* ---- stand-alone system operator: S = coef[0] M + coef[1] K + coef[2] A of
 * an assembled context (or its transpose), stored once in CSR with FP32
 * (FEM_PREC_FP32_IR) or FP64 (FEM_PREC_FP64) values.  Replaces the sparse
 * products of the reference's solvers: Eigen's S * x at
 * src/FEM/CPUSolver.cpp:78,96,114,134,137 and the ELL row loop Ad_i of
 * shaders/FEM/cgm.glsl:75-102.  Columns: nrhs in {1, 2, 4, 16}, vectors
 * stored [N][nrhs]. ---- */
int fem_operator_create(fem_ctx* ctx, const double coef[3], int32_t precision, int32_t transpose,
                        fem_operator** out);
int fem_operator_destroy(fem_operator* op);
/* y = S x on device pointers of the operator's value type (float or double),
 * enqueued on the operator's stream (returns before completion). */
int fem_operator_apply_device(fem_operator* op, int32_t nrhs, const void* x, void* y);
/* y = S x on host arrays of doubles (converted to the value type), synchronous. */
int fem_operator_apply(fem_operator* op, int32_t nrhs, const double* x, double* y);
/* Average device time of `reps` back-to-back applications, HIP events on
 * the operator's stream. */
int fem_operator_time(fem_operator* op, int32_t nrhs, const void* x, void* y, int32_t reps, double* ms);
/* SpMV schedules: 1, 2, 4, 8 or 16 lanes per row (>= nrhs), or the row-block
 * stream (a workgroup stages its rows' column/value stream in LDS with 16-byte
 * loads), optionally with non-temporal stream loads; the pipelined stream
 * walks the row blocks with a resident grid and loads the next blocks' row
 * pointers and stream while it gathers the current one. */
enum { FEM_SPMV_STREAM = 32, FEM_SPMV_STREAM_NT = 33, FEM_SPMV_PIPE = 34, FEM_SPMV_PIPE_NT = 35 };
/* Schedule autotuner: times every schedule valid for nrhs and keeps the
 * fastest (*schedule receives it). */
int fem_operator_autotune(fem_operator* op, int32_t nrhs, const void* x, void* y, int32_t reps,
                          int32_t* schedule, double* ms);
/* Fix the schedule used for nrhs columns. */
int fem_operator_set_schedule(fem_operator* op, int32_t nrhs, int32_t schedule);
int fem_operator_sync(fem_operator* op);

/* ---- adjoint of a backward-Euler rollout (no reference counterpart; the
 * forward is the heat / advection-diffusion step of
 * src/FEM/CPUSolver.cpp:72-102 with per-element diffusivity).  The forward
 * keeps every `checkpoint_every`-th state; the backward recomputes each
 * segment from its checkpoint and runs transposed solves.  Gradients are
 * returned for a loss L(u_T) given dL/du_T: dL/du_0 (unknown space) and
 * dL/dkappa_e (one per triangle, for K = sum_e kappa_e K_e as assembled with
 * fem_assembly_params.kappa_e). ---- */
int fem_adjoint_create(fem_ctx* ctx, const fem_solver_opts* opts, fem_adjoint** out);
int fem_adjoint_destroy(fem_adjoint* a);
/* FEM_HEAT or FEM_ADVECTION_DIFFUSION with their parameter structs. */
int fem_adjoint_set_equation(fem_adjoint* a, int32_t equation, const void* params, size_t size);
/* u0 and uT: [N] unknowns.  stats->ms is device time of the rollout. */
int fem_adjoint_forward(fem_adjoint* a, const double* u0, int32_t nsteps, int32_t checkpoint_every, double* uT,
                        fem_step_stats* stats);
/* grad_uT: dL/du_T [N]; grad_u0 [N] and grad_kappa_e [num_triangles] may be
 * NULL.  stats->ms is the wall time of the whole backward pass. */
int fem_adjoint_backward(fem_adjoint* a, const double* grad_uT, double* grad_u0, double* grad_kappa_e,
                         fem_step_stats* stats);
/* The same two calls on device pointers (FP64 arrays in device memory, for
 * example torch tensors): no host copy.  Work that produces the inputs on
 * another stream must be complete before the call; the outputs are complete
 * when it returns. */
int fem_adjoint_forward_device(fem_adjoint* a, const double* u0, int32_t nsteps, int32_t checkpoint_every,
                               double* uT, fem_step_stats* stats);
int fem_adjoint_backward_device(fem_adjoint* a, const double* grad_uT, double* grad_u0, double* grad_kappa_e,
                                fem_step_stats* stats);

#ifdef __cplusplus
}
#endif
#endif /* FEMHIP_H */
