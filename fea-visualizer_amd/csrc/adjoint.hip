// adjoint.hip -- discrete adjoint of a backward-Euler heat / advection-
// diffusion rollout, with checkpoint-and-recompute.
//
// Forward (CPUSolver.cpp:77-81, 95-99 with per-element diffusivity):
//   S u_{t+1} = B u_t,  B = M/dt,  S = M/dt + c K(kappa) - A,  t = 0..T-1,
// where K(kappa) = sum_e kappa_e K_e.  For a loss L(u_T) with g = dL/du_T:
//   mu_T = S^-T g;  mu_t = S^-T B mu_{t+1};  dL/du_0 = B mu_1;
//   dL/dkappa_e = -c sum_{t=1..T} mu_t^T K_e u_t.
// The recursion mu_t = S^-T B mu_{t+1} is a forward step of the same equation
// with the transposed system matrix, so it runs in the persistent stepping
// kernel (a second instance whose matrix is formed transposed).  The forward
// keeps u every `every` steps; the backward walks the segments from the last,
// recomputes each segment's states from its checkpoint, runs the segment's
// adjoint steps with their trajectory recorded, and accumulates the
// per-element gradient in one element-parallel pass per segment (each element
// owns its output, summed in a fixed order: deterministic, no atomics).
#include "femhip_adjoint.h"
#include "femhip_geom.h"

#include <algorithm>
#include <chrono>
#include <cstring>

namespace femhip {

namespace {

__global__ void k_grad_kappa(int ntri, const float* xyz, const uint32_t* tri, const int* idx, int L,
                             const double* carry, const double* MU, const double* U, long long n, double coef,
                             double* gk) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ntri) return;
  const uint32_t* tv = tri + 3 * (size_t)e;
  int r[3];
  bool any = false;
  for (int a = 0; a < 3; a++) {
    r[a] = idx[tv[a]];
    any |= r[a] >= 0;
  }
  if (!any) return;
  Geo G;
  elem_geom(xyz, tv, G);
  double kab[3][3];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) kab[a][b] = elem_kab(G, a, b);
  double g = gk[e];
  // pair s: mu_{t1-s} (s = 0: carry, else MU[s-1]) with u_{t1-s} = U[L-1-s]
  for (int s = 0; s < L; s++) {
    const double* mu = (s == 0) ? carry : MU + (size_t)(s - 1) * n;
    const double* u = U + (size_t)(L - 1 - s) * n;
    double acc = 0.0;
    for (int a = 0; a < 3; a++) {
      if (r[a] < 0) continue;
      for (int b = 0; b < 3; b++) {
        if (r[b] < 0) continue;
        acc += mu[r[a]] * kab[a][b] * u[r[b]];
      }
    }
    g -= coef * acc;
  }
  gk[e] = g;
}

// y = (M / dt) x
__global__ void k_mass_mv(int n, const int* rp, const int* ci, const double* M, double dt, const double* x,
                          double* y) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double acc = 0.0;
  for (int s = rp[r]; s < rp[r + 1]; s++) acc += (M[s] / dt) * x[ci[s]];
  y[r] = acc;
}

double* state_u(fem_batch* b) { return b->insts[0].d_u; }

void free_bufs(fem_adjoint* a) {
  void* ptrs[] = {a->d_ckpt, a->d_U, a->d_MU, a->d_carry, a->d_gk, a->d_tmp};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  a->d_ckpt = a->d_U = a->d_MU = a->d_carry = a->d_gk = a->d_tmp = nullptr;
  a->ckpt_cap = 0;
  a->seg_cap = 0;
}

}  // namespace

void adjoint_free(fem_adjoint* a) {
  if (!a) return;
  free_bufs(a);
  if (a->fwd) batch_free(a->fwd);
  if (a->adj) batch_free(a->adj);
  delete a;
}

int adjoint_create(fem_ctx* ctx, const fem_solver_opts* opts, fem_adjoint** out) {
  fem_solver_opts o;
  fem_opts_default(&o);
  if (opts) o = *opts;
  if (o.engine == FEM_ENGINE_PLAIN || o.engine == FEM_ENGINE_GL_COMPAT) {
    set_error("the adjoint runs on the persistent or streaming engine");
    return FEM_ERR_INVALID;
  }
  if (o.nrhs > 1) {
    set_error("the adjoint takes one right-hand side");
    return FEM_ERR_INVALID;
  }
  o.nrhs = 1;
  // recomputed segments must reproduce the forward bit for bit: every step
  // starts from the previous solution alone (no extrapolation history to
  // carry across checkpoints)
  o.predictor = FEM_PREDICTOR_PREVIOUS;
  fem_adjoint* a = new fem_adjoint();
  a->ctx = ctx;
  int rc = batch_create(&ctx, 1, &o, &a->fwd);
  if (!rc) rc = batch_create(&ctx, 1, &o, &a->adj);
  if (rc) {
    adjoint_free(a);
    return rc;
  }
  a->adj->insts[0].transpose = 1;
  *out = a;
  return 0;
}

int adjoint_set_equation(fem_adjoint* a, int eq, const void* params, size_t size) {
  if (eq != FEM_HEAT && eq != FEM_ADVECTION_DIFFUSION) {
    set_error("the adjoint supports the heat and advection-diffusion equations");
    return FEM_ERR_INVALID;
  }
  int rc = batch_set_equation(a->fwd, 0, eq, params, size);
  if (!rc) rc = batch_set_equation(a->adj, 0, eq, params, size);
  if (rc) return rc;
  // both parameter structs are {time_step, coefficient of K}
  const double* p = (const double*)params;
  a->eq = eq;
  a->dt = p[0];
  a->coef = p[1];
  a->have_forward = false;
  return 0;
}

// dev: u0 and uT are device pointers (copied device to device), else host.
int adjoint_forward(fem_adjoint* a, const double* u0, int nsteps, int every, double* uT, fem_step_stats* stats,
                    bool dev) {
  if (a->eq < 0) {
    set_error("adjoint: set_equation first");
    return FEM_ERR_STATE;
  }
  if (nsteps < 1 || every < 1 || !u0) {
    set_error("adjoint_forward: need nsteps >= 1, checkpoint interval >= 1 and u0");
    return FEM_ERR_INVALID;
  }
  const long long n = a->ctx->n;
  const long long nck = nsteps / every + 1;
  int rc = 0;
  hipStream_t st = a->fwd->stream;
  if (a->ckpt_cap < nck) {
    if (a->d_ckpt) (void)hipFree(a->d_ckpt);
    a->d_ckpt = dalloc<double>((size_t)(nck * std::max(1LL, n)), &rc);
    if (rc) {
      a->ckpt_cap = 0;
      return rc;
    }
    a->ckpt_cap = nck;
  }
  a->have_forward = false;
  if (n > 0) {
    FEM_HIP(hipMemcpyAsync(a->d_ckpt, u0, sizeof(double) * n, dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                           st));
    FEM_HIP(hipMemcpyAsync(state_u(a->fwd), a->d_ckpt, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    FEM_HIP(hipStreamSynchronize(st));
  }
  int hr = batch_reset_history(a->fwd, 0);
  if (hr) return hr;
  InstHost& h = a->fwd->insts[0];
  h.traj = a->d_ckpt + n;
  h.traj_every = every;
  const int r = batch_advance(a->fwd, nsteps, stats);
  h.traj = nullptr;
  if (r < 0) return r;
  if (uT && n > 0) {
    FEM_HIP(hipMemcpyAsync(uT, state_u(a->fwd), sizeof(double) * n,
                           dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st));
    FEM_HIP(hipStreamSynchronize(st));
  }
  a->nsteps = nsteps;
  a->every = every;
  a->fwd_version = a->ctx->version;
  a->have_forward = true;
  return r;
}

// dev: gT, gu0 and gk are device pointers, else host.
int adjoint_backward(fem_adjoint* a, const double* gT, double* gu0, double* gk, fem_step_stats* stats, bool dev) {
  if (!a->have_forward || a->fwd_version != a->ctx->version) {
    set_error("adjoint_backward: run adjoint_forward (after the last assembly) first");
    return FEM_ERR_STATE;
  }
  if (!gT) {
    set_error("adjoint_backward: null dL/du_T");
    return FEM_ERR_INVALID;
  }
  const auto t_start = std::chrono::steady_clock::now();
  fem_ctx* c = a->ctx;
  const long long n = c->n;
  const int T = a->nsteps, C = a->every;
  int rc = 0;
  if (a->seg_cap < C || !a->d_carry) {
    void* ptrs[] = {a->d_U, a->d_MU, a->d_carry, a->d_gk, a->d_tmp};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
    a->d_U = dalloc<double>((size_t)C * std::max(1LL, n), &rc);
    a->d_MU = dalloc<double>((size_t)C * std::max(1LL, n), &rc);
    a->d_carry = dalloc<double>((size_t)std::max(1LL, n), &rc);
    a->d_gk = dalloc<double>((size_t)std::max(1, c->ntri), &rc);
    a->d_tmp = dalloc<double>((size_t)std::max(1LL, n), &rc);
    if (rc) {
      a->seg_cap = 0;
      return rc;
    }
    a->seg_cap = C;
  }
  hipStream_t st = a->fwd->stream;
  long long its = 0, sweeps = 0, steps = 0;
  double res = 0.0;
  int worst = FEM_OK, status = 0;
  // adjoint states are not physical fields: the magnitude rule of the
  // forward (|u| > 1e4) does not apply, only non-finite values do
  auto acc_stats = [&](const fem_step_stats& s, int r) {
    its += s.iterations;
    sweeps += s.sweeps;
    steps += s.steps;
    res = std::max(res, s.residual);
    status |= s.status & ~ST_UNSTABLE;
    if (r == FEM_DIVERGED)
      r = (s.status & ST_NONFINITE) ? FEM_DIVERGED : (s.status & ST_NOT_CONVERGED) ? FEM_NOT_CONVERGED : FEM_OK;
    if (r > worst) worst = r;
  };
  if (c->ntri > 0) FEM_HIP(hipMemsetAsync(a->d_gk, 0, sizeof(double) * c->ntri, st));
  InstHost& ha = a->adj->insts[0];
  InstHost& hf = a->fwd->insts[0];
  if (n > 0) {
    // mu_T = S^-T g: one plain solve on the transposed instance
    FEM_HIP(hipMemcpyAsync(ha.d_b, gT, sizeof(double) * n, dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                           st));
    FEM_HIP(hipMemsetAsync(ha.d_x, 0, sizeof(double) * n, st));
    FEM_HIP(hipStreamSynchronize(st));
    ha.eq_override = EQ_SOLVE;
    fem_step_stats s{};
    int r = batch_advance(a->adj, 1, &s);
    ha.eq_override = -1;
    if (r < 0) return r;
    acc_stats(s, r);
    FEM_HIP(hipMemcpyAsync(ha.d_u, ha.d_x, sizeof(double) * n, hipMemcpyDeviceToDevice, a->adj->stream));
    FEM_HIP(hipStreamSynchronize(a->adj->stream));
    if ((r = batch_reset_history(a->adj, 0))) return r;
    const int nseg = (T + C - 1) / C;
    for (int j = nseg - 1; j >= 0; j--) {
      const int t0 = j * C, t1 = std::min(T, t0 + C), L = t1 - t0;
      // recompute u_{t0+1} .. u_{t1} from the checkpoint u_{t0}
      FEM_HIP(hipMemcpyAsync(hf.d_u, a->d_ckpt + (size_t)j * n, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
      FEM_HIP(hipStreamSynchronize(st));
      if ((r = batch_reset_history(a->fwd, 0))) return r;
      hf.traj = a->d_U;
      hf.traj_every = 1;
      r = batch_advance(a->fwd, L, &s);
      hf.traj = nullptr;
      if (r < 0) return r;
      acc_stats(s, r);
      // mu_{t1} is the carry; the adjoint steps record mu_{t1-1}, mu_{t1-2}, ...
      FEM_HIP(hipMemcpyAsync(a->d_carry, ha.d_u, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
      FEM_HIP(hipStreamSynchronize(st));
      const int m = (t0 > 0) ? L : L - 1;  // the last segment stops at mu_1
      if (m > 0) {
        ha.traj = a->d_MU;
        ha.traj_every = 1;
        r = batch_advance(a->adj, m, &s);
        ha.traj = nullptr;
        if (r < 0) return r;
        acc_stats(s, r);
      }
      if (c->ntri > 0) {
        hipLaunchKernelGGL(k_grad_kappa, dim3(div_up(c->ntri, kBlock)), dim3(kBlock), 0, st, c->ntri, c->d_xyz,
                           c->d_tri, c->d_idx, L, a->d_carry, a->d_MU, a->d_U, n, a->coef, a->d_gk);
        FEM_HIP(hipGetLastError());
      }
    }
    // dL/du_0 = B mu_1
    hipLaunchKernelGGL(k_mass_mv, dim3(div_up(n, kBlock)), dim3(kBlock), 0, st, (int)n, c->d_rp, c->d_ci, c->d_M,
                       a->dt, ha.d_u, a->d_tmp);
    FEM_HIP(hipGetLastError());
    if (gu0)
      FEM_HIP(hipMemcpyAsync(gu0, a->d_tmp, sizeof(double) * n, dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                             st));
  }
  if (gk && c->ntri > 0)
    FEM_HIP(hipMemcpyAsync(gk, a->d_gk, sizeof(double) * c->ntri,
                           dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st));
  FEM_HIP(hipStreamSynchronize(st));
  const double ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
  if (stats) {
    stats->iterations = its;
    stats->sweeps = sweeps;
    stats->steps = steps;
    stats->residual = res;
    stats->ms = ms;
    stats->status = status;
  }
  return worst;
}

}  // namespace femhip
