// api.hip -- the extern "C" boundary (include/femhip.h).  Every entry point
// validates its arguments, converts C++ failures into status codes and
// records a message for fem_last_error(); no exception crosses the ABI.
#include "femhip_adjoint.h"

#include <cstdio>
#include <cstring>
#include <exception>
#include <new>

namespace femhip {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int hip_fail(hipError_t e, const char* what) {
  char buf[512];
  std::snprintf(buf, sizeof(buf), "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
  g_last_error = buf;
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return FEM_ERR_OOM;
  return FEM_ERR_HIP;
}

}  // namespace femhip

using namespace femhip;

#define FEM_GUARD(...)                               \
  try {                                              \
    __VA_ARGS__                                      \
  } catch (const std::bad_alloc&) {                  \
    set_error("host allocation failed");             \
    return FEM_ERR_OOM;                              \
  } catch (const std::exception& ex) {               \
    set_error(ex.what());                            \
    return FEM_ERR_INVALID;                          \
  } catch (...) {                                    \
    set_error("unknown C++ exception");              \
    return FEM_ERR_INVALID;                          \
  }

extern "C" {

const char* fem_last_error(void) { return g_last_error.c_str(); }

int fem_abi_version(void) { return FEMHIP_ABI_VERSION; }

int fem_device_count(int32_t* count) {
  if (!count) return FEM_ERR_INVALID;
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  *count = (e == hipSuccess) ? c : 0;
  return FEM_OK;
}

void fem_opts_default(fem_solver_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->engine = FEM_ENGINE_PERSISTENT;
  o->precision = FEM_PREC_FP32_IR;
  o->krylov = FEM_KRYLOV_AUTO;
  o->nrhs = 1;
  o->warm_start = 1;
  o->tol = 1e-10;
  o->inner_tol = 1e-4;
  o->predictor = FEM_PREDICTOR_SEXTIC;
}

// ------------------------------------------------------------------ context

int fem_ctx_create(const fem_mesh* mesh, int32_t bc, fem_ctx** out) {
  FEM_GUARD({
    if (!mesh || !out || mesh->num_vertices < 0 || mesh->num_triangles < 0 ||
        (mesh->num_vertices > 0 && (!mesh->xyz || !mesh->on_boundary)) ||
        (mesh->num_triangles > 0 && !mesh->triangles)) {
      set_error("fem_ctx_create: invalid mesh");
      return FEM_ERR_INVALID;
    }
    if (bc != FEM_DIRICHLET && bc != FEM_NEUMANN) {
      set_error("fem_ctx_create: unknown boundary condition");
      return FEM_ERR_INVALID;
    }
    *out = nullptr;
    return ctx_create(mesh, bc, out);
  })
}

int fem_ctx_destroy(fem_ctx* ctx) {
  FEM_GUARD({
    ctx_free(ctx);
    return FEM_OK;
  })
}

int fem_ctx_assemble(fem_ctx* ctx, const fem_assembly_params* params) {
  FEM_GUARD({
    if (!ctx) {
      set_error("fem_ctx_assemble: null context");
      return FEM_ERR_INVALID;
    }
    return ctx_assemble(ctx, params);
  })
}

int fem_ctx_info_get(const fem_ctx* ctx, fem_ctx_info* info) {
  if (!ctx || !info) {
    set_error("fem_ctx_info_get: null argument");
    return FEM_ERR_INVALID;
  }
  info->num_nodes = ctx->nv;
  info->num_unknowns = ctx->n;
  info->num_triangles = ctx->ntri;
  info->max_row_nonzeros = ctx->max_row;
  info->nnz = ctx->nnz;
  return FEM_OK;
}

int fem_ctx_get_idx_map(const fem_ctx* ctx, int32_t* idx_map) {
  if (!ctx || !idx_map) {
    set_error("fem_ctx_get_idx_map: null argument");
    return FEM_ERR_INVALID;
  }
  if (ctx->nv > 0) FEM_HIP(hipMemcpy(idx_map, ctx->d_idx, sizeof(int) * ctx->nv, hipMemcpyDeviceToHost));
  return FEM_OK;
}

int fem_ctx_get_pattern(const fem_ctx* ctx, int32_t* row_ptr, int32_t* col_idx) {
  if (!ctx || !row_ptr || (ctx->nnz > 0 && !col_idx)) {
    set_error("fem_ctx_get_pattern: null argument");
    return FEM_ERR_INVALID;
  }
  FEM_HIP(hipMemcpy(row_ptr, ctx->d_rp, sizeof(int) * (ctx->n + 1), hipMemcpyDeviceToHost));
  if (ctx->nnz > 0) FEM_HIP(hipMemcpy(col_idx, ctx->d_ci, sizeof(int) * ctx->nnz, hipMemcpyDeviceToHost));
  return FEM_OK;
}

int fem_ctx_get_matrix(const fem_ctx* ctx, int32_t which, double* values) {
  if (!ctx || (!values && ctx->nnz > 0)) {
    set_error("fem_ctx_get_matrix: null argument");
    return FEM_ERR_INVALID;
  }
  const double* src = which == FEM_MATRIX_K ? ctx->d_K : which == FEM_MATRIX_M ? ctx->d_M : which == FEM_MATRIX_A ? ctx->d_A : nullptr;
  if (!src) {
    set_error("fem_ctx_get_matrix: unknown matrix");
    return FEM_ERR_INVALID;
  }
  if (ctx->nnz > 0) FEM_HIP(hipMemcpy(values, src, sizeof(double) * ctx->nnz, hipMemcpyDeviceToHost));
  return FEM_OK;
}

// ------------------------------------------------------------------ solver

int fem_solver_create(fem_ctx* ctx, const fem_solver_opts* opts, fem_solver** out) {
  FEM_GUARD({
    if (!ctx || !out) {
      set_error("fem_solver_create: null argument");
      return FEM_ERR_INVALID;
    }
    fem_batch* b = nullptr;
    int rc = batch_create(&ctx, 1, opts, &b);
    if (rc) return rc;
    fem_solver* s = new fem_solver();
    s->batch = b;
    *out = s;
    return FEM_OK;
  })
}

int fem_solver_destroy(fem_solver* s) {
  FEM_GUARD({
    if (s) {
      batch_free(s->batch);
      delete s;
    }
    return FEM_OK;
  })
}

int fem_solver_set_equation(fem_solver* s, int32_t equation, const void* params, size_t size) {
  FEM_GUARD({
    if (!s) {
      set_error("null solver");
      return FEM_ERR_INVALID;
    }
    return batch_set_equation(s->batch, 0, equation, params, size);
  })
}

// Surface-space state: gather nv -> N through idx_map (CPUSolver.cpp:13-21)
// on the device; entries equal to the state rounded to float are kept.
int fem_solver_set_state(fem_solver* s, int32_t field, const float* values_nv) {
  FEM_GUARD({
    if (!s || !values_nv) {
      set_error("null argument");
      return FEM_ERR_INVALID;
    }
    return batch_set_surface(s->batch, 0, field, values_nv);
  })
}

// Scatter N -> nv on the device, boundary entries written as 0 (CPUSolver.cpp:28-37).
int fem_solver_get_state(fem_solver* s, int32_t field, float* values_nv) {
  FEM_GUARD({
    if (!s || !values_nv) {
      set_error("null argument");
      return FEM_ERR_INVALID;
    }
    return batch_get_surface(s->batch, 0, field, values_nv);
  })
}

int fem_solver_set_state_unknowns(fem_solver* s, int32_t field, const double* x) {
  FEM_GUARD({
    if (!s || !x) {
      set_error("null argument");
      return FEM_ERR_INVALID;
    }
    return batch_set_state_unknowns(s->batch, 0, field, x);
  })
}

int fem_solver_get_state_unknowns(fem_solver* s, int32_t field, double* x) {
  FEM_GUARD({
    if (!s || !x) {
      set_error("null argument");
      return FEM_ERR_INVALID;
    }
    return batch_get_state_unknowns(s->batch, 0, field, x);
  })
}

int fem_solver_advance(fem_solver* s, int32_t nsteps, fem_step_stats* stats) {
  FEM_GUARD({
    if (!s) {
      set_error("null solver");
      return FEM_ERR_INVALID;
    }
    return batch_advance(s->batch, nsteps, stats);
  })
}

int fem_solver_clear(fem_solver* s) {
  FEM_GUARD({
    if (!s) {
      set_error("null solver");
      return FEM_ERR_INVALID;
    }
    return batch_clear(s->batch, -1);
  })
}

int fem_solver_unstable(fem_solver* s, int32_t* unstable) {
  FEM_GUARD({
    if (!s || !unstable) {
      set_error("null argument");
      return FEM_ERR_INVALID;
    }
    int st = 0;
    const auto& h = s->batch->insts[0];
    if (h.d_status) FEM_HIP(hipMemcpy(&st, h.d_status, sizeof(int), hipMemcpyDeviceToHost));
    *unstable = (st & (ST_UNSTABLE | ST_NONFINITE)) ? 1 : 0;
    return FEM_OK;
  })
}

int fem_solver_state_device_ptr(fem_solver* s, int32_t field, void** dptr) {
  FEM_GUARD({
    if (!s || !dptr) {
      set_error("null argument");
      return FEM_ERR_INVALID;
    }
    double* d;
    int col, nc;
    int rc = batch_field_ptr(s->batch, 0, field, &d, &col, &nc);
    if (rc) return rc;
    *dptr = (void*)d;
    rc = batch_reset_history(s->batch, 0);  // the caller may write the state: restart the predictor
    if (rc) return rc;
    return FEM_OK;
  })
}

int fem_solver_sync(fem_solver* s) {
  if (!s) return FEM_ERR_INVALID;
  FEM_HIP(hipStreamSynchronize(s->batch->stream));
  return FEM_OK;
}

int fem_solver_autotune(fem_solver* s, int32_t probe_steps, int32_t* chosen, double* best_ms) {
  FEM_GUARD({
    if (!s) {
      set_error("null solver");
      return FEM_ERR_INVALID;
    }
    int c = 0;
    const int rc = batch_autotune(s->batch, probe_steps, &c, best_ms);
    if (chosen) *chosen = c;
    return rc;
  })
}

// ------------------------------------------------------------------ batch

int fem_batch_create(fem_ctx* const* ctxs, int32_t count, const fem_solver_opts* opts, fem_batch** out) {
  FEM_GUARD({
    if (!out) {
      set_error("null argument");
      return FEM_ERR_INVALID;
    }
    return batch_create(ctxs, count, opts, out);
  })
}

int fem_batch_destroy(fem_batch* b) {
  FEM_GUARD({
    batch_free(b);
    return FEM_OK;
  })
}

int fem_batch_set_equation(fem_batch* b, int32_t instance, int32_t equation, const void* params, size_t size) {
  FEM_GUARD({
    if (!b) {
      set_error("null batch");
      return FEM_ERR_INVALID;
    }
    return batch_set_equation(b, instance, equation, params, size);
  })
}

int fem_batch_set_state_unknowns(fem_batch* b, int32_t instance, int32_t field, const double* x) {
  FEM_GUARD({
    if (!b || !x) {
      set_error("null argument");
      return FEM_ERR_INVALID;
    }
    return batch_set_state_unknowns(b, instance, field, x);
  })
}

int fem_batch_get_state_unknowns(fem_batch* b, int32_t instance, int32_t field, double* x) {
  FEM_GUARD({
    if (!b || !x) {
      set_error("null argument");
      return FEM_ERR_INVALID;
    }
    return batch_get_state_unknowns(b, instance, field, x);
  })
}

int fem_batch_advance(fem_batch* b, int32_t nsteps, fem_step_stats* stats) {
  FEM_GUARD({
    if (!b) {
      set_error("null batch");
      return FEM_ERR_INVALID;
    }
    return batch_advance(b, nsteps, stats);
  })
}

int fem_batch_instance_stats(fem_batch* b, int32_t instance, fem_step_stats* stats) {
  FEM_GUARD({
    if (!b || !stats) {
      set_error("null batch or stats");
      return FEM_ERR_INVALID;
    }
    return batch_instance_stats(b, instance, stats);
  })
}

int fem_batch_graph_capture(fem_batch* b, int32_t nsteps) {
  FEM_GUARD({
    if (!b) {
      set_error("null batch");
      return FEM_ERR_INVALID;
    }
    return batch_graph_capture(b, nsteps);
  })
}

int fem_batch_graph_launch(fem_batch* b, fem_step_stats* stats) {
  FEM_GUARD({
    if (!b) {
      set_error("null batch");
      return FEM_ERR_INVALID;
    }
    return batch_graph_launch(b, stats);
  })
}

int fem_batch_plan_info(fem_batch* b, fem_plan_info* info) {
  FEM_GUARD({
    if (!b || !info) {
      set_error("null batch or info");
      return FEM_ERR_INVALID;
    }
    return batch_plan_info(b, info);
  })
}

int fem_batch_set_wait_limit(fem_batch* b, double seconds) {
  FEM_GUARD({
    if (!b) {
      set_error("null batch");
      return FEM_ERR_INVALID;
    }
    return batch_set_wait_limit(b, seconds);
  })
}

int fem_batch_sync(fem_batch* b) {
  if (!b) return FEM_ERR_INVALID;
  FEM_HIP(hipStreamSynchronize(b->stream));
  return FEM_OK;
}

int fem_batch_autotune(fem_batch* b, int32_t probe_steps, int32_t* chosen, double* best_ms) {
  FEM_GUARD({
    if (!b) {
      set_error("null batch");
      return FEM_ERR_INVALID;
    }
    int c = 0;
    const int rc = batch_autotune(b, probe_steps, &c, best_ms);
    if (chosen) *chosen = c;
    return rc;
  })
}

int fem_ctx_solve(fem_ctx* ctx, const double coef[3], const fem_solver_opts* opts, const double* b, double* x,
                  fem_step_stats* stats) {
  FEM_GUARD({
    if (!ctx || !coef || (ctx->n > 0 && (!b || !x))) {
      set_error("fem_ctx_solve: null argument");
      return FEM_ERR_INVALID;
    }
    if (opts && (opts->engine == FEM_ENGINE_PLAIN || opts->engine == FEM_ENGINE_GL_COMPAT)) {
      set_error("fem_ctx_solve runs on the persistent or streaming engine");
      return FEM_ERR_INVALID;
    }
    return linear_solve(ctx, coef, opts, b, x, stats);
  })
}

int fem_ctx_solve_noalloc(fem_ctx* ctx, const double coef[3], const fem_solver_opts* opts, const double* b,
                          double* x, fem_step_stats* stats) {
  FEM_GUARD({
    if (!ctx || !coef || (ctx->n > 0 && (!b || !x))) {
      set_error("fem_ctx_solve_noalloc: null argument");
      return FEM_ERR_INVALID;
    }
    return linear_solve_noalloc(ctx, coef, opts, b, x, stats);
  })
}

int fem_ctx_smooth_normals(fem_ctx* ctx, const float* normals_nv3, const float* values_nv, float extrusion,
                           float* out_nv3) {
  FEM_GUARD({
    if (!ctx || (ctx->nv > 0 && (!normals_nv3 || !values_nv || !out_nv3))) {
      set_error("fem_ctx_smooth_normals: null argument");
      return FEM_ERR_INVALID;
    }
    return ctx_smooth_normals(ctx, normals_nv3, values_nv, extrusion, out_nv3);
  })
}

int fem_solver_smooth_normals(fem_solver* s, const float* normals_nv3, float extrusion, float* out_nv3) {
  FEM_GUARD({
    if (!s || !normals_nv3 || !out_nv3) {
      set_error("fem_solver_smooth_normals: null argument");
      return FEM_ERR_INVALID;
    }
    fem_batch* b = s->batch;
    fem_ctx* c = b->insts[0].ctx;
    const size_t nv = (size_t)c->nv;
    if (nv == 0) return FEM_OK;
    int rc = 0;
    float* d = dalloc<float>(7 * nv, &rc);
    if (!d) return rc;
    float *dn = d, *dv = d + 3 * nv, *dout = d + 4 * nv;
    hipError_t e = hipMemcpyAsync(dn, normals_nv3, sizeof(float) * 3 * nv, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) {
      rc = batch_surface_values(b, 0, dv);
      if (!rc) rc = ctx_smooth_normals_device(c, dn, dv, extrusion, dout, b->stream);
      if (!rc) e = hipMemcpyAsync(out_nv3, dout, sizeof(float) * 3 * nv, hipMemcpyDeviceToHost, b->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(e, "fem_solver_smooth_normals");
    return rc;
  })
}

// ------------------------------------------------------------------ surface

int fem_surface_from_obj(const char* path, fem_surface** out) {
  FEM_GUARD({
    if (!path || !out) {
      set_error("fem_surface_from_obj: null argument");
      return FEM_ERR_INVALID;
    }
    *out = nullptr;
    return surface_from_obj(path, out);
  })
}

int fem_surface_get(const fem_surface* s, fem_mesh* mesh, const float** normals_nv3, int32_t* num_boundary_points) {
  FEM_GUARD({
    if (!s) {
      set_error("fem_surface_get: null surface");
      return FEM_ERR_INVALID;
    }
    return surface_get(s, mesh, normals_nv3, num_boundary_points);
  })
}

int fem_surface_destroy(fem_surface* s) {
  FEM_GUARD({
    surface_free(s);
    return FEM_OK;
  })
}

int fem_export_ply(const fem_mesh* mesh, const float* normals_nv3, const float* values_nv, float vertex_extrusion,
                   float threshold, int32_t mesh_type, const char* path) {
  FEM_GUARD({
    if (!mesh || !path || mesh->num_vertices < 0 || mesh->num_triangles < 0 ||
        (mesh->num_vertices > 0 && (!mesh->xyz || !normals_nv3 || !values_nv)) ||
        (mesh->num_triangles > 0 && !mesh->triangles)) {
      set_error("fem_export_ply: invalid argument");
      return FEM_ERR_INVALID;
    }
    return ply_export(mesh->xyz, mesh->num_vertices, mesh->triangles, mesh->num_triangles, normals_nv3, values_nv,
                      vertex_extrusion, threshold, mesh_type, path);
  })
}

int fem_solver_export_ply(fem_solver* s, const float* normals_nv3, float vertex_extrusion, float threshold,
                          int32_t mesh_type, const char* path) {
  FEM_GUARD({
    if (!s || !path || !normals_nv3) {
      set_error("fem_solver_export_ply: null argument");
      return FEM_ERR_INVALID;
    }
    fem_batch* b = s->batch;
    fem_ctx* c = b->insts[0].ctx;
    const size_t nv = (size_t)c->nv, nt = (size_t)c->ntri;
    std::vector<float> xyz(3 * nv), vals(nv);
    std::vector<uint32_t> tri(3 * nt);
    if (nv > 0) {
      int rc = 0;
      float* d = dalloc<float>(nv, &rc);
      if (!d) return rc;
      rc = batch_surface_values(b, 0, d);
      hipError_t e = hipSuccess;
      if (!rc) e = hipMemcpyAsync(vals.data(), d, sizeof(float) * nv, hipMemcpyDeviceToHost, b->stream);
      if (!rc && e == hipSuccess)
        e = hipMemcpyAsync(xyz.data(), c->d_xyz, sizeof(float) * 3 * nv, hipMemcpyDeviceToHost, b->stream);
      if (!rc && e == hipSuccess && nt > 0)
        e = hipMemcpyAsync(tri.data(), c->d_tri, sizeof(uint32_t) * 3 * nt, hipMemcpyDeviceToHost, b->stream);
      if (!rc && e == hipSuccess) e = hipStreamSynchronize(b->stream);
      (void)hipFree(d);
      if (rc) return rc;
      if (e != hipSuccess) return hip_fail(e, "fem_solver_export_ply");
    }
    return ply_export(xyz.data(), (int32_t)nv, tri.data(), (int32_t)nt, normals_nv3, vals.data(), vertex_extrusion,
                      threshold, mesh_type, path);
  })
}

// ----------------------------------------------------------------- operator

int fem_operator_create(fem_ctx* ctx, const double coef[3], int32_t precision, int32_t transpose,
                        fem_operator** out) {
  FEM_GUARD({
    if (!ctx || !coef || !out) {
      set_error("fem_operator_create: null argument");
      return FEM_ERR_INVALID;
    }
    return operator_create(ctx, coef, precision, transpose ? 1 : 0, out);
  })
}

int fem_operator_destroy(fem_operator* op) {
  FEM_GUARD({
    operator_free(op);
    return FEM_OK;
  })
}

int fem_operator_apply_device(fem_operator* op, int32_t nrhs, const void* x, void* y) {
  FEM_GUARD({
    if (!op || (op->n > 0 && (!x || !y))) {
      set_error("fem_operator_apply_device: null argument");
      return FEM_ERR_INVALID;
    }
    return operator_apply_device(op, nrhs, x, y);
  })
}

int fem_operator_apply(fem_operator* op, int32_t nrhs, const double* x, double* y) {
  FEM_GUARD({
    if (!op || (op->n > 0 && (!x || !y))) {
      set_error("fem_operator_apply: null argument");
      return FEM_ERR_INVALID;
    }
    return operator_apply_host(op, nrhs, x, y);
  })
}

int fem_operator_time(fem_operator* op, int32_t nrhs, const void* x, void* y, int32_t reps, double* ms) {
  FEM_GUARD({
    if (!op || !ms || (op->n > 0 && (!x || !y))) {
      set_error("fem_operator_time: null argument");
      return FEM_ERR_INVALID;
    }
    return operator_time(op, nrhs, x, y, reps, ms);
  })
}

int fem_operator_autotune(fem_operator* op, int32_t nrhs, const void* x, void* y, int32_t reps, int32_t* lanes,
                          double* ms) {
  FEM_GUARD({
    if (!op || (op->n > 0 && (!x || !y))) {
      set_error("fem_operator_autotune: null argument");
      return FEM_ERR_INVALID;
    }
    int g = 0;
    const int rc = operator_autotune(op, nrhs, x, y, reps, &g, ms);
    if (lanes) *lanes = g;
    return rc;
  })
}

int fem_operator_set_schedule(fem_operator* op, int32_t nrhs, int32_t schedule) {
  FEM_GUARD({
    if (!op) {
      set_error("null operator");
      return FEM_ERR_INVALID;
    }
    return operator_set_lanes(op, nrhs, schedule);
  })
}

int fem_operator_sync(fem_operator* op) {
  FEM_GUARD({
    if (!op) {
      set_error("null operator");
      return FEM_ERR_INVALID;
    }
    FEM_HIP(hipStreamSynchronize(op->stream));
    return FEM_OK;
  })
}

// ------------------------------------------------------------------ adjoint

int fem_adjoint_create(fem_ctx* ctx, const fem_solver_opts* opts, fem_adjoint** out) {
  FEM_GUARD({
    if (!ctx || !out) {
      set_error("fem_adjoint_create: null argument");
      return FEM_ERR_INVALID;
    }
    return adjoint_create(ctx, opts, out);
  })
}

int fem_adjoint_destroy(fem_adjoint* a) {
  FEM_GUARD({
    adjoint_free(a);
    return FEM_OK;
  })
}

int fem_adjoint_set_equation(fem_adjoint* a, int32_t equation, const void* params, size_t size) {
  FEM_GUARD({
    if (!a || !params) {
      set_error("fem_adjoint_set_equation: null argument");
      return FEM_ERR_INVALID;
    }
    return adjoint_set_equation(a, equation, params, size);
  })
}

int fem_adjoint_forward(fem_adjoint* a, const double* u0, int32_t nsteps, int32_t checkpoint_every, double* uT,
                        fem_step_stats* stats) {
  FEM_GUARD({
    if (!a) {
      set_error("null adjoint");
      return FEM_ERR_INVALID;
    }
    return adjoint_forward(a, u0, nsteps, checkpoint_every, uT, stats, false);
  })
}

int fem_adjoint_forward_device(fem_adjoint* a, const double* u0, int32_t nsteps, int32_t checkpoint_every,
                               double* uT, fem_step_stats* stats) {
  FEM_GUARD({
    if (!a) {
      set_error("null adjoint");
      return FEM_ERR_INVALID;
    }
    return adjoint_forward(a, u0, nsteps, checkpoint_every, uT, stats, true);
  })
}

int fem_adjoint_backward(fem_adjoint* a, const double* grad_uT, double* grad_u0, double* grad_kappa_e,
                         fem_step_stats* stats) {
  FEM_GUARD({
    if (!a) {
      set_error("null adjoint");
      return FEM_ERR_INVALID;
    }
    return adjoint_backward(a, grad_uT, grad_u0, grad_kappa_e, stats, false);
  })
}

int fem_adjoint_backward_device(fem_adjoint* a, const double* grad_uT, double* grad_u0, double* grad_kappa_e,
                                fem_step_stats* stats) {
  FEM_GUARD({
    if (!a) {
      set_error("null adjoint");
      return FEM_ERR_INVALID;
    }
    return adjoint_backward(a, grad_uT, grad_u0, grad_kappa_e, stats, true);
  })
}

}  // extern "C"
