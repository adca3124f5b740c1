// gl_compat.hip -- the reference's OpenGL compute-shader solver semantics
// (src/FEM/GPUSolver.cpp:74-172, shaders/FEM/cgm_helper.glsl:74-117,
// shaders/FEM/cgm.glsl:75-148) as an engine of this library, for A/B
// comparison with that backend's behaviour.  It is not the accurate path:
//   - everything in FP32, warm start from the current state;
//   - unpreconditioned CG, at most `max_iterations` (the reference's 10)
//     iterations, loop while r.r > 1e-10 and stop once r.r <= 1e-10 r0.r0;
//   - Gray-Scott: the reaction terms are summed over every slot of the row
//     with the neighbours' values (cgm_helper.glsl:97,101), u is solved
//     first and the v solve uses the new u (GPUSolver.cpp:140-168);
//   - instability: the last r.r > 1e4 (GPUSolver.cpp:56-59).
// The shader's direction update reads and rewrites r.r from every invocation
// (a race); here every row uses the intended ratio r_{i+1}.r_{i+1} / r_i.r_i.
// Rows are the CSR rows of the context, whose ascending columns are the ELL
// slot order the reference builds (GPUSolver.cpp:243-266).
#include "femhip_step.h"

#include <algorithm>
#include <cmath>

namespace femhip {

namespace {

enum { GL_HEAT = 0, GL_ADVDIFF = 1, GL_WAVE = 2, GL_GS_U = 3, GL_GS_V = 4 };

struct GlPrm {
  int eq;
  float dt, c, Du, Dv, F, kr;
};

__device__ inline float gl_coef(const GlPrm& p, float m, float k, float a) {
  switch (p.eq) {
    case GL_HEAT: return m / p.dt + p.c * k;
    case GL_ADVDIFF: return m / p.dt + p.c * k - a;
    case GL_WAVE: return m / p.dt + p.c * p.c * k * p.dt;
    case GL_GS_U: return m / p.dt + p.Du * k;
    default: return m / p.dt + p.Dv * k;
  }
}

// cgm_helper.glsl stage 1: b, r = d = b - A x (x = u, or v for the wave and
// the Gray-Scott v solve).
__global__ void k_gl_init(GlPrm p, int n, const int* rp, const int* ci, const float* M, const float* K,
                          const float* A, const float* u, const float* v, float* b, float* r, float* d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float bi = 0.0f, ax = 0.0f;
  for (int s = rp[i]; s < rp[i + 1]; s++) {
    const int c = ci[s];
    const float m = M[s], k = K[s], a = A[s];
    switch (p.eq) {
      case GL_HEAT:
      case GL_ADVDIFF:
        bi += u[c] * (m / p.dt);
        ax += u[c] * gl_coef(p, m, k, a);
        break;
      case GL_WAVE:
        bi += v[c] * (m / p.dt) - p.c * p.c * u[c] * k;
        ax += v[c] * gl_coef(p, m, k, a);
        break;
      case GL_GS_U:
        bi += u[c] * (m / p.dt) - u[c] * v[c] * v[c] + p.F * (1.0f - u[c]);
        ax += u[c] * gl_coef(p, m, k, a);
        break;
      default:
        bi += v[c] * (m / p.dt) + u[c] * v[c] * v[c] - v[c] * (p.F + p.kr);
        ax += v[c] * gl_coef(p, m, k, a);
        break;
    }
  }
  b[i] = bi;
  r[i] = bi - ax;
  d[i] = bi - ax;
}

__device__ inline float gl_ad(const GlPrm& p, int i, const int* rp, const int* ci, const float* M, const float* K,
                              const float* A, const float* d) {
  float s = 0.0f;
  for (int q = rp[i]; q < rp[i + 1]; q++) s += d[ci[q]] * gl_coef(p, M[q], K[q], A[q]);
  return s;
}

// Per-block partial sums of x.y (what = 0: r.r, 1: d.(A d)).
__global__ void k_gl_dot(GlPrm p, int n, const int* rp, const int* ci, const float* M, const float* K,
                         const float* A, const float* r, const float* d, int what, float* part) {
  __shared__ float sh[kBlock];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  float v = 0.0f;
  if (i < n) v = what == 0 ? r[i] * r[i] : d[i] * gl_ad(p, i, rp, ci, M, K, A, d);
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int h = kBlock / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_gl_sum(const float* part, int nb, float* out) {
  __shared__ float sh[kBlock];
  float v = 0.0f;
  for (int q = threadIdx.x; q < nb; q += kBlock) v += part[q];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int h = kBlock / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

// cgm.glsl stage 2: r -= alpha A d, x += alpha d.
__global__ void k_gl_update(GlPrm p, int n, const int* rp, const int* ci, const float* M, const float* K,
                            const float* A, float alpha, const float* d, float* r, float* x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float ad = gl_ad(p, i, rp, ci, M, K, A, d);
  r[i] = r[i] - alpha * ad;
  x[i] = x[i] + alpha * d[i];
}

// cgm.glsl stage 4: d = r + beta d.
__global__ void k_gl_dir(int n, float beta, const float* r, float* d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = r[i] + beta * d[i];
}

__global__ void k_gl_wave(int n, float dt, const float* v, float* u) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) u[i] = u[i] + v[i] * dt;
}

__global__ void k_to32(long long n, const double* a, float* b) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}
__global__ void k_to64(long long n, const float* a, double* b) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (double)a[i];
}

struct GlBufs {
  float *M, *K, *A, *u, *v, *b, *r, *d, *part, *scal;
};

// One CG solve (GPUSolver::cgm, GPUSolver.cpp:325-356); returns r.r at exit.
int gl_cgm(const GlPrm& p, fem_ctx* c, GlBufs& g, float* x, int max_it, hipStream_t st, float* rr_out, int* its) {
  const int n = c->n;
  const unsigned nb = div_up(n, kBlock);
  float h[2];
  auto dot = [&](int what, float* out) -> int {
    hipLaunchKernelGGL(k_gl_dot, dim3(nb), dim3(kBlock), 0, st, p, n, c->d_rp, c->d_ci, g.M, g.K, g.A, g.r, g.d,
                       what, g.part);
    hipLaunchKernelGGL(k_gl_sum, dim3(1), dim3(kBlock), 0, st, g.part, (int)nb, g.scal);
    FEM_HIP(hipGetLastError());
    FEM_HIP(hipMemcpyAsync(h, g.scal, sizeof(float), hipMemcpyDeviceToHost, st));
    FEM_HIP(hipStreamSynchronize(st));
    *out = h[0];
    return 0;
  };
  float r0 = 0.f, ri = 0.f, dad = 0.f, ri1 = 0.f;
  int rc = dot(0, &ri);
  if (rc) return rc;
  r0 = ri;
  const float eps = 1e-10f;
  int it = 0;
  while (ri > eps && it < max_it) {
    if ((rc = dot(1, &dad))) return rc;
    const float alpha = ri / dad;
    hipLaunchKernelGGL(k_gl_update, dim3(nb), dim3(kBlock), 0, st, p, n, c->d_rp, c->d_ci, g.M, g.K, g.A, alpha,
                       g.d, g.r, x);
    if ((rc = dot(0, &ri1))) return rc;
    hipLaunchKernelGGL(k_gl_dir, dim3(nb), dim3(kBlock), 0, st, n, ri1 / ri, g.r, g.d);
    FEM_HIP(hipGetLastError());
    ri = ri1;
    it++;
    if (ri <= eps * r0) break;
  }
  *rr_out = ri;
  *its += it;
  return 0;
}

}  // namespace

int gl_advance(fem_batch* b, int nsteps, fem_step_stats* stats) {
  const auto t0 = hipEventRecord(b->ev0, b->stream);
  if (t0 != hipSuccess) return hip_fail(t0, "gl_advance");
  long long its_all = 0;
  int status_all = 0;
  double rr_max = 0.0;
  for (auto& h : b->insts) {
    if (h.eq < 0 || !h.d_u) {
      set_error("advance: set_equation was not called for every instance");
      return FEM_ERR_STATE;
    }
    if (h.k != (h.eq == FEM_REACTION_DIFFUSION ? 2 : 1)) {
      set_error("the GL-compatible engine takes one right-hand side");
      return FEM_ERR_INVALID;
    }
    if (h.eq == FEM_WAVE_NEWMARK) {
      set_error("the GL-compatible engine has no Newmark scheme (the reference GL solver has none)");
      return FEM_ERR_INVALID;
    }
    fem_ctx* c = h.ctx;
    const long long n = c->n, nnz = c->nnz;
    if (n == 0) continue;
    int rc = 0;
    float* slab = dalloc<float>((size_t)(3 * nnz + 6 * n + div_up(n, kBlock) + 4), &rc);
    if (!slab) return rc;
    GlBufs g;
    g.M = slab;
    g.K = g.M + nnz;
    g.A = g.K + nnz;
    g.u = g.A + nnz;
    g.v = g.u + n;
    g.b = g.v + n;
    g.r = g.b + n;
    g.d = g.r + n;
    g.part = g.d + n;
    g.scal = g.part + div_up(n, kBlock) + 1;
    const unsigned nbz = div_up(nnz, kBlock), nbn = div_up(n, kBlock);
    hipStream_t st = b->stream;
    hipLaunchKernelGGL(k_to32, dim3(nbz), dim3(kBlock), 0, st, nnz, c->d_M, g.M);
    hipLaunchKernelGGL(k_to32, dim3(nbz), dim3(kBlock), 0, st, nnz, c->d_K, g.K);
    hipLaunchKernelGGL(k_to32, dim3(nbz), dim3(kBlock), 0, st, nnz, c->d_A, g.A);
    // state: u, v (Gray-Scott: the interleaved [n][2] array holds u and v)
    std::vector<double> hs((size_t)n * h.k);
    std::vector<float> hu(n), hv(n, 0.f);
    FEM_HIP(hipMemcpyAsync(hs.data(), h.d_u, sizeof(double) * n * h.k, hipMemcpyDeviceToHost, st));
    FEM_HIP(hipStreamSynchronize(st));
    if (h.eq == FEM_REACTION_DIFFUSION) {
      for (long long i = 0; i < n; i++) {
        hu[i] = (float)hs[2 * i];
        hv[i] = (float)hs[2 * i + 1];
      }
    } else {
      for (long long i = 0; i < n; i++) hu[i] = (float)hs[i];
      if (h.eq == FEM_WAVE) {
        FEM_HIP(hipMemcpyAsync(hs.data(), h.d_v, sizeof(double) * n, hipMemcpyDeviceToHost, st));
        FEM_HIP(hipStreamSynchronize(st));
        for (long long i = 0; i < n; i++) hv[i] = (float)hs[i];
      }
    }
    FEM_HIP(hipMemcpyAsync(g.u, hu.data(), sizeof(float) * n, hipMemcpyHostToDevice, st));
    FEM_HIP(hipMemcpyAsync(g.v, hv.data(), sizeof(float) * n, hipMemcpyHostToDevice, st));
    GlPrm p{};
    p.dt = (float)h.dt;
    switch (h.eq) {
      case FEM_HEAT: p.eq = GL_HEAT; p.c = (float)h.p1; break;
      case FEM_ADVECTION_DIFFUSION: p.eq = GL_ADVDIFF; p.c = (float)h.p1; break;
      case FEM_WAVE: p.eq = GL_WAVE; p.c = (float)h.p1; break;
      default:
        p.Du = (float)h.p1;
        p.Dv = (float)h.p2;
        p.F = (float)h.p3;
        p.kr = (float)h.p4;
        break;
    }
    const int max_it = b->opts.max_iters > 0 ? b->opts.max_iters : 10;  // GPUSolver.hpp:31
    float rr = 0.f;
    int its = 0;
    for (int s = 0; s < nsteps && !rc; s++) {
      if (h.eq == FEM_REACTION_DIFFUSION) {
        p.eq = GL_GS_U;
        hipLaunchKernelGGL(k_gl_init, dim3(nbn), dim3(kBlock), 0, st, p, (int)n, c->d_rp, c->d_ci, g.M, g.K, g.A,
                           g.u, g.v, g.b, g.r, g.d);
        rc = gl_cgm(p, c, g, g.u, max_it, st, &rr, &its);
        if (rc) break;
        p.eq = GL_GS_V;
        hipLaunchKernelGGL(k_gl_init, dim3(nbn), dim3(kBlock), 0, st, p, (int)n, c->d_rp, c->d_ci, g.M, g.K, g.A,
                           g.u, g.v, g.b, g.r, g.d);
        rc = gl_cgm(p, c, g, g.v, max_it, st, &rr, &its);
      } else {
        hipLaunchKernelGGL(k_gl_init, dim3(nbn), dim3(kBlock), 0, st, p, (int)n, c->d_rp, c->d_ci, g.M, g.K, g.A,
                           g.u, g.v, g.b, g.r, g.d);
        rc = gl_cgm(p, c, g, p.eq == GL_WAVE ? g.v : g.u, max_it, st, &rr, &its);
        if (!rc && p.eq == GL_WAVE)
          hipLaunchKernelGGL(k_gl_wave, dim3(nbn), dim3(kBlock), 0, st, (int)n, p.dt, g.v, g.u);
      }
    }
    if (!rc) {
      FEM_HIP(hipMemcpyAsync(hu.data(), g.u, sizeof(float) * n, hipMemcpyDeviceToHost, st));
      FEM_HIP(hipMemcpyAsync(hv.data(), g.v, sizeof(float) * n, hipMemcpyDeviceToHost, st));
      FEM_HIP(hipStreamSynchronize(st));
      if (h.eq == FEM_REACTION_DIFFUSION) {
        for (long long i = 0; i < n; i++) {
          hs[2 * i] = hu[i];
          hs[2 * i + 1] = hv[i];
        }
        FEM_HIP(hipMemcpyAsync(h.d_u, hs.data(), sizeof(double) * n * 2, hipMemcpyHostToDevice, st));
      } else {
        for (long long i = 0; i < n; i++) hs[i] = hu[i];
        FEM_HIP(hipMemcpyAsync(h.d_u, hs.data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
        FEM_HIP(hipStreamSynchronize(st));
        if (h.eq == FEM_WAVE) {
          for (long long i = 0; i < n; i++) hs[i] = hv[i];
          FEM_HIP(hipMemcpyAsync(h.d_v, hs.data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
        }
      }
      FEM_HIP(hipStreamSynchronize(st));
      // GPUSolver::has_numerical_instability: the last r.r above 1e4 (NaN too)
      const int flag = (!(rr <= 1e4f)) ? ST_UNSTABLE : 0;
      FEM_HIP(hipMemcpyAsync(h.d_status, &flag, sizeof(int), hipMemcpyHostToDevice, st));
      FEM_HIP(hipStreamSynchronize(st));
      status_all |= flag;
      rr_max = std::max(rr_max, (double)rr);
      its_all += its;
      // no extrapolation history is kept here
      if ((rc = batch_reset_history(b, (int)(&h - b->insts.data())))) break;
    }
    (void)hipFree(slab);
    if (rc) return rc;
  }
  FEM_HIP(hipEventRecord(b->ev1, b->stream));
  FEM_HIP(hipEventSynchronize(b->ev1));
  float ms = 0.f;
  FEM_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  if (stats) {
    stats->iterations = its_all;
    stats->sweeps = 0;
    stats->steps = nsteps;
    stats->residual = rr_max;
    stats->ms = ms;
    stats->status = status_all;
  }
  return (status_all & ST_UNSTABLE) ? FEM_DIVERGED : FEM_OK;
}

}  // namespace femhip
