// plain.hip -- the plain FP64 GPU engine: one right-hand side at a time,
// one kernel per CG stage and a host read of the reduced scalars after each
// reduction, the way the reference's GL solver drives its compute shaders
// (src/FEM/GPUSolver.cpp:326-361 reads the norms after every stage).  It is
// the FP64 GPU baseline the persistent engine is compared against, and it
// shares the right-hand side / update semantics of the persistent kernel.
#include "femhip_step.h"

#include <cfloat>
#include <cmath>
#include <cstring>

namespace femhip {

namespace {

constexpr int kDotBlocks = 512;

// r = b - S x0 for column j after building b and x0 (same formulas as the
// persistent kernel's phase R).
__global__ void k_p_rhs(Inst I, int j, double* r, double* z, double* p) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= I.n) return;
  const int K = I.k, kv = I.kv, vj = (kv == 1) ? 0 : j;
  const int lo = I.rp[row], hi = I.rp[row + 1];
  const size_t o = (size_t)row * K + j;
  const bool warm = I.warm != 0;
  const double inv_dt = 1.0 / I.dt;
  double b = 0.0, sx = 0.0, x0 = 0.0;
  switch (I.eq) {
    case EQ_HEAT:
    case EQ_ADVDIFF: {
      double bs = 0.0;
      for (int s = lo; s < hi; s++) {
        const double uc = I.u[(size_t)I.ci[s] * K + j];
        bs += I.M64[s] * uc;
        if (warm) sx += I.S64[(size_t)s * kv + vj] * uc;
      }
      b = bs * inv_dt;
      x0 = warm ? I.u[o] : 0.0;
    } break;
    case EQ_WAVE: {
      double bs = 0.0, ks = 0.0;
      for (int s = lo; s < hi; s++) {
        const size_t oc = (size_t)I.ci[s] * K + j;
        bs += I.M64[s] * I.v[oc];
        ks += I.K64[s] * I.u[oc];
        if (warm) sx += I.S64[(size_t)s * kv + vj] * I.v[oc];
      }
      b = bs * inv_dt - I.p1 * I.p1 * ks;
      x0 = warm ? I.v[o] : 0.0;
    } break;
    case EQ_GS: {
      double bs = 0.0;
      for (int s = lo; s < hi; s++) {
        const double yc = I.u[(size_t)I.ci[s] * K + j];
        bs += I.M64[s] * yc;
        if (warm) sx += I.S64[(size_t)s * kv + vj] * yc;
      }
      const double uu = I.u[(size_t)row * K], vv = I.u[(size_t)row * K + 1];
      const double uvv = uu * (vv * vv);
      const double f = (j == 0) ? (-uvv + I.p3 * (1.0 - uu)) : (uvv - (I.p3 + I.p4) * vv);
      const double w = I.weighting ? I.lumped[row] : 1.0;
      b = bs * inv_dt + w * f;
      x0 = warm ? I.u[o] : 0.0;
    } break;
    case EQ_NEWMARK: {
      const double h = (0.5 - I.p2) * I.dt * I.dt;
      double ks = 0.0;
      for (int s = lo; s < hi; s++) {
        const size_t oc = (size_t)I.ci[s] * K + j;
        ks += I.K64[s] * (I.u[oc] + I.dt * I.v[oc] + h * I.a[oc]);
        if (warm) sx += I.S64[(size_t)s * kv + vj] * I.a[oc];
      }
      b = -I.p1 * I.p1 * ks;
      x0 = warm ? I.a[o] : 0.0;
    } break;
    default: {
      b = I.b[o];
      if (warm)
        for (int s = lo; s < hi; s++) sx += I.S64[(size_t)s * kv + vj] * I.x[(size_t)I.ci[s] * K + j];
      x0 = warm ? I.x[o] : 0.0;
    } break;
  }
  I.b[o] = b;
  I.x[o] = x0;
  const double rr = b - sx;
  r[row] = rr;
  const double zz = I.dinv64[(size_t)row * kv + vj] * rr;
  z[row] = zz;
  p[row] = zz;
}

// y = S x (column j of the value table)
__global__ void k_p_spmv(Inst I, int j, const double* x, double* y) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= I.n) return;
  const int kv = I.kv, vj = (kv == 1) ? 0 : j;
  double acc = 0.0;
  for (int s = I.rp[row]; s < I.rp[row + 1]; s++) acc += I.S64[(size_t)s * kv + vj] * x[I.ci[s]];
  y[row] = acc;
}

// r = b - S x for column j; z = D^-1 r; p = z
__global__ void k_p_resid(Inst I, int j, double* r, double* z, double* p) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= I.n) return;
  const int K = I.k, kv = I.kv, vj = (kv == 1) ? 0 : j;
  double sx = 0.0;
  for (int s = I.rp[row]; s < I.rp[row + 1]; s++) sx += I.S64[(size_t)s * kv + vj] * I.x[(size_t)I.ci[s] * K + j];
  const double rr = I.b[(size_t)row * K + j] - sx;
  r[row] = rr;
  const double zz = I.dinv64[(size_t)row * kv + vj] * rr;
  z[row] = zz;
  p[row] = zz;
}

// partial[blk] = sum a[i*sa+oa]*b[i*sb+ob] over a grid-stride range
__global__ void k_p_dot(const double* a, int sa, int oa, const double* b, int sb, int ob, int n, double* partial) {
  __shared__ double lds[kBlock];
  double acc = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    acc += a[(size_t)i * sa + oa] * b[(size_t)i * sb + ob];
  lds[threadIdx.x] = acc;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) lds[threadIdx.x] += lds[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = lds[0];
}
__global__ void k_p_dot_final(const double* partial, int m, double* out) {
  __shared__ double lds[kBlock];
  double acc = 0.0;
  for (int i = threadIdx.x; i < m; i += kBlock) acc += partial[i];
  lds[threadIdx.x] = acc;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) lds[threadIdx.x] += lds[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = lds[0];
}

// x += alpha p (strided into column j); r -= alpha q; z = D^-1 r
__global__ void k_p_cg_update(Inst I, int j, double alpha, const double* p, const double* q, double* r, double* z) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= I.n) return;
  const int kv = I.kv, vj = (kv == 1) ? 0 : j;
  I.x[(size_t)row * I.k + j] += alpha * p[row];
  const double rr = r[row] - alpha * q[row];
  r[row] = rr;
  z[row] = I.dinv64[(size_t)row * kv + vj] * rr;
}

__global__ void k_p_axpby(int n, const double* x, double beta, double* y) {  // y = x + beta y
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = x[i] + beta * y[i];
}

// BiCGSTAB helpers (right Jacobi preconditioning)
__global__ void k_p_bi_p(int n, const double* r, double beta, double omega, const double* v, double* p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = r[i] + beta * (p[i] - omega * v[i]);
}
__global__ void k_p_scale(Inst I, int j, const double* x, double* y) {  // y = D^-1 x
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < I.n) y[i] = I.dinv64[(size_t)i * I.kv + ((I.kv == 1) ? 0 : j)] * x[i];
}
__global__ void k_p_axpy(int n, double a, const double* x, const double* y, double* out) {  // out = y + a x
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = y[i] + a * x[i];
}
__global__ void k_p_bi_x(Inst I, int j, double alpha, const double* ph, double omega, const double* sh) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < I.n) I.x[(size_t)i * I.k + j] += alpha * ph[i] + omega * sh[i];
}

__global__ void k_p_update(Inst I, int* flags) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= I.n * I.k) return;
  const size_t o = (size_t)i;
  const double xf = I.x[o];
  double c1 = xf, c2 = 0.0;
  switch (I.eq) {
    case EQ_HEAT:
    case EQ_ADVDIFF:
    case EQ_GS:
      I.u[o] = xf;
      break;
    case EQ_WAVE: {
      I.v[o] = xf;
      const double un = I.u[o] + xf * I.dt;
      I.u[o] = un;
      c2 = un;
    } break;
    case EQ_NEWMARK: {
      const double dt = I.dt, be = I.p2, ga = I.p3;
      const double ao = I.a[o], vo = I.v[o];
      const double ut = I.u[o] + dt * vo + (0.5 - be) * dt * dt * ao;
      const double un = ut + be * dt * dt * xf;
      const double vn = vo + dt * ((1.0 - ga) * ao + ga * xf);
      I.u[o] = un;
      I.v[o] = vn;
      I.a[o] = xf;
      c1 = un;
      c2 = vn;
    } break;
    default:
      break;
  }
  int f = 0;
  if (c1 > 1e4 || c2 > 1e4) f |= ST_UNSTABLE;
  if (!isfinite(c1) || !isfinite(c2)) f |= ST_NONFINITE;
  if (f) atomicOr(flags, f);
}

__global__ void k_p_zero_col(Inst I, int j) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < I.n) I.x[(size_t)i * I.k + j] = 0.0;
}

struct Plain {
  fem_batch* b;
  hipStream_t st;
  double* part;
  double* res;   // 8 result slots
  double* host;  // pinned-free host mirror
};

int dot(Plain& P, const double* a, int sa, int oa, const double* bb, int sb, int ob, int n, int slot) {
  hipLaunchKernelGGL(k_p_dot, dim3(kDotBlocks), dim3(kBlock), 0, P.st, a, sa, oa, bb, sb, ob, n, P.part);
  hipLaunchKernelGGL(k_p_dot_final, dim3(1), dim3(kBlock), 0, P.st, P.part, kDotBlocks, P.res + slot);
  FEM_HIP(hipGetLastError());
  return 0;
}

int fetch(Plain& P, int nslots, double* out) {
  FEM_HIP(hipMemcpyAsync(out, P.res, sizeof(double) * nslots, hipMemcpyDeviceToHost, P.st));
  FEM_HIP(hipStreamSynchronize(P.st));
  return 0;
}

#define PCK(x)          \
  do {                  \
    int rc_ = (x);      \
    if (rc_) return rc_; \
  } while (0)

}  // namespace

int batch_prepare(fem_batch* b, bool need_plan);
int collect_stats(fem_batch* b, fem_step_stats* stats, float ms);

int plain_advance(fem_batch* b, int nsteps, fem_step_stats* stats) {
  int rc = batch_prepare(b, false);
  if (rc) return rc;
  long long nmax = 1;
  for (auto& h : b->insts) nmax = std::max(nmax, (long long)h.ctx->n);
  if (b->plain_n < nmax) {
    if (b->d_plain) (void)hipFree(b->d_plain);
    b->d_plain = dalloc<double>((size_t)nmax * 10 + kDotBlocks + 16, &rc);
    if (!b->d_plain) return rc;
    b->plain_n = nmax;
  }
  Plain P;
  P.b = b;
  P.st = b->stream;
  double* base = b->d_plain;
  double *r = base, *z = base + nmax, *p = base + 2 * nmax, *q = base + 3 * nmax;
  double *rh = base + 4 * nmax, *vv = base + 5 * nmax, *sv = base + 6 * nmax, *tv = base + 7 * nmax,
         *ph = base + 8 * nmax, *shv = base + 9 * nmax;
  P.part = base + 10 * nmax;
  P.res = P.part + kDotBlocks;
  long long its_all = 0, sweeps_all = 0;
  int status = 0;
  double res_all = 0.0;
  FEM_HIP(hipEventRecord(b->ev0, b->stream));
  for (auto& h : b->insts) {
    Inst& I = h.dev;
    FEM_HIP(hipMemsetAsync(h.d_status, 0, sizeof(int), P.st));
    const int n = I.n;
    if (n == 0) continue;
    const unsigned g = div_up(n, kBlock);
    const double tol2 = I.tol2;
    for (int step = 0; step < nsteps; step++) {
      // every right-hand side from the old state first (Gray-Scott couples u and v)
      for (int j = 0; j < I.k; j++)
        hipLaunchKernelGGL(k_p_rhs, dim3(g), dim3(kBlock), 0, P.st, I, j, r, z, p);
      FEM_HIP(hipGetLastError());
      for (int j = 0; j < I.k; j++) {
        double hv[4];
        PCK(dot(P, I.b, I.k, j, I.b, I.k, j, n, 0));
        hipLaunchKernelGGL(k_p_resid, dim3(g), dim3(kBlock), 0, P.st, I, j, r, z, p);
        PCK(dot(P, r, 1, 0, r, 1, 0, n, 1));
        PCK(dot(P, r, 1, 0, z, 1, 0, n, 2));
        PCK(fetch(P, 3, hv));
        const double rhs2 = hv[0];
        double rr = hv[1], gam = hv[2];
        if (rhs2 == 0.0) {  // Eigen: x = 0
          hipLaunchKernelGGL(k_p_zero_col, dim3(g), dim3(kBlock), 0, P.st, I, j);
          continue;
        }
        const double thr = std::max(tol2 * rhs2, DBL_MIN);
        int it = 0;
        if (I.krylov == 0) {
          while (rr >= thr && it < I.max_iters) {
            hipLaunchKernelGGL(k_p_spmv, dim3(g), dim3(kBlock), 0, P.st, I, j, p, q);
            PCK(dot(P, p, 1, 0, q, 1, 0, n, 0));
            PCK(fetch(P, 1, hv));
            const double alpha = gam / hv[0];
            hipLaunchKernelGGL(k_p_cg_update, dim3(g), dim3(kBlock), 0, P.st, I, j, alpha, p, q, r, z);
            PCK(dot(P, r, 1, 0, r, 1, 0, n, 0));
            PCK(dot(P, r, 1, 0, z, 1, 0, n, 1));
            PCK(fetch(P, 2, hv));
            it++;
            rr = hv[0];
            if (rr < thr || !(rr == rr)) break;
            const double beta = hv[1] / gam;
            gam = hv[1];
            hipLaunchKernelGGL(k_p_axpby, dim3(g), dim3(kBlock), 0, P.st, n, z, beta, p);
          }
        } else {
          double rho = 1.0, alpha = 1.0, omega = 1.0;
          FEM_HIP(hipMemcpyAsync(rh, r, sizeof(double) * n, hipMemcpyDeviceToDevice, P.st));
          FEM_HIP(hipMemsetAsync(p, 0, sizeof(double) * n, P.st));
          FEM_HIP(hipMemsetAsync(vv, 0, sizeof(double) * n, P.st));
          double r0n2 = rr;
          while (rr >= thr && it < I.max_iters) {
            PCK(dot(P, rh, 1, 0, r, 1, 0, n, 0));
            PCK(fetch(P, 1, hv));
            double rho_new = hv[0];
            if (std::fabs(rho_new) < 1e-30 * r0n2 || omega == 0.0) {
              FEM_HIP(hipMemcpyAsync(rh, r, sizeof(double) * n, hipMemcpyDeviceToDevice, P.st));
              FEM_HIP(hipMemsetAsync(p, 0, sizeof(double) * n, P.st));
              FEM_HIP(hipMemsetAsync(vv, 0, sizeof(double) * n, P.st));
              rho = alpha = omega = 1.0;
              rho_new = rr;
              r0n2 = rr;
            }
            const double beta = (rho_new / rho) * (alpha / omega);
            rho = rho_new;
            hipLaunchKernelGGL(k_p_bi_p, dim3(g), dim3(kBlock), 0, P.st, n, r, beta, omega, vv, p);
            hipLaunchKernelGGL(k_p_scale, dim3(g), dim3(kBlock), 0, P.st, I, j, p, ph);
            hipLaunchKernelGGL(k_p_spmv, dim3(g), dim3(kBlock), 0, P.st, I, j, ph, vv);
            PCK(dot(P, rh, 1, 0, vv, 1, 0, n, 0));
            PCK(fetch(P, 1, hv));
            alpha = rho / hv[0];
            hipLaunchKernelGGL(k_p_axpy, dim3(g), dim3(kBlock), 0, P.st, n, -alpha, vv, r, sv);
            PCK(dot(P, sv, 1, 0, sv, 1, 0, n, 0));
            PCK(fetch(P, 1, hv));
            it++;
            if (hv[0] < thr) {
              hipLaunchKernelGGL(k_p_bi_x, dim3(g), dim3(kBlock), 0, P.st, I, j, alpha, ph, 0.0, ph);
              FEM_HIP(hipMemcpyAsync(r, sv, sizeof(double) * n, hipMemcpyDeviceToDevice, P.st));
              rr = hv[0];
              break;
            }
            hipLaunchKernelGGL(k_p_scale, dim3(g), dim3(kBlock), 0, P.st, I, j, sv, shv);
            hipLaunchKernelGGL(k_p_spmv, dim3(g), dim3(kBlock), 0, P.st, I, j, shv, tv);
            PCK(dot(P, tv, 1, 0, sv, 1, 0, n, 0));
            PCK(dot(P, tv, 1, 0, tv, 1, 0, n, 1));
            PCK(fetch(P, 2, hv));
            omega = hv[1] > 0.0 ? hv[0] / hv[1] : 0.0;
            hipLaunchKernelGGL(k_p_bi_x, dim3(g), dim3(kBlock), 0, P.st, I, j, alpha, ph, omega, shv);
            hipLaunchKernelGGL(k_p_axpy, dim3(g), dim3(kBlock), 0, P.st, n, -omega, tv, sv, r);
            PCK(dot(P, r, 1, 0, r, 1, 0, n, 0));
            PCK(fetch(P, 1, hv));
            rr = hv[0];
            if (!(rr == rr)) break;
          }
        }
        its_all += it;
        sweeps_all += 1;
        if (rr >= thr) status |= ST_NOT_CONVERGED;
        res_all = std::max(res_all, std::sqrt(rr / rhs2));
      }
      hipLaunchKernelGGL(k_p_update, dim3(div_up((long long)n * I.k, kBlock)), dim3(kBlock), 0, P.st, I,
                         h.d_status);
      FEM_HIP(hipGetLastError());
    }
  }
  // this engine keeps no extrapolation history: a later launch restarts the predictor
  rc = batch_reset_history(b, -1);
  if (rc) return rc;
  FEM_HIP(hipEventRecord(b->ev1, b->stream));
  FEM_HIP(hipEventSynchronize(b->ev1));
  float ms = 0.f;
  FEM_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  int dev_status = 0;
  for (auto& h : b->insts) {
    int st = 0;
    FEM_HIP(hipMemcpy(&st, h.d_status, sizeof(int), hipMemcpyDeviceToHost));
    dev_status |= st;
  }
  status |= dev_status;
  if (stats) {
    stats->iterations = its_all;
    stats->sweeps = sweeps_all;
    stats->steps = nsteps;
    stats->residual = res_all;
    stats->ms = ms;
    stats->status = status;
  }
  if (status & (ST_UNSTABLE | ST_NONFINITE)) return FEM_DIVERGED;
  if (status & ST_NOT_CONVERGED) return FEM_NOT_CONVERGED;
  return FEM_OK;
}

}  // namespace femhip
