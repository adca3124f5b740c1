// step.hip -- persistent, cooperative time-stepping kernel.
//
// One launch advances every instance of a batch by `nsteps` implicit time
// steps.  The grid is split into teams (contiguous workgroup ranges); a team
// owns a list of instances and walks them one after the other.  Inside a team
// each workgroup owns a contiguous block of CSR rows and the team meets at a
// device-scope barrier between phases.  The last workgroup to arrive reduces
// the per-workgroup partial dot products in rank order (so every reduction is
// deterministic) and publishes the result with the barrier release.
//
// Per step and instance (the reference's CPUSolver::advance_time,
// src/FEM/CPUSolver.cpp:64-147, re-expressed as fused phases):
//   R    right-hand side (FP64) incl. the Gray-Scott reaction, the initial
//        residual r = b - S x0 (FP64), r -> inner precision, z = D^-1 r
//   A/B  preconditioned CG in the inner precision (FP32 by default):
//        A: p = z + beta p (double-buffered, recomputed for gathered
//           neighbours), q = S p, <p,q>;  B: alpha, xi += alpha p,
//           r -= alpha q, z = D^-1 r, <r,z>, <r,r>
//        (BiCGSTAB for advection-diffusion: phases P, S, X)
//   Res  FP64 refinement residual r = b - S64 (x + xi), restart of the inner
//        solve when |r| >= tol |b|
//   U    state update (heat: u = x; wave: v = x, u += dt v; Gray-Scott: u,v;
//        Newmark: u, v, a) and the instability rule (CPUSolver.cpp:41-48)
// Vectors are [rows][k] with the right-hand side fastest, so the k lanes of a
// row gather one contiguous span per neighbour.
#include "femhip_internal.h"
#include "femhip_step.h"

#include <cfloat>

namespace femhip {

namespace {

// wall_clock64 ticks at 100 MHz: a barrier wait longer than 4 s aborts the
// launch (it cannot happen in a correct run; it bounds every wait).
constexpr long long kSpinTimeoutTicks = 400000000LL;

struct Shared {
  double wred[kBlock / kWave][kMaxRed];
  double red[kMaxRed];     // reduced values of the last barrier
  double gamma[kMaxK];     // CG: <r,z>;  BiCGSTAB: rho
  double alpha[kMaxK];
  double beta[kMaxK];
  double omega[kMaxK];
  double thr[kMaxK];       // tol^2 |b|^2
  double ithr[kMaxK];      // inner stopping threshold of the sweep
  double bb[kMaxK];
  int active[kMaxK];
  int flag;
  int uni;                 // uniform control value broadcast
  unsigned gen;
};

__device__ inline double ald(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Rank of this workgroup inside its team, ordered so that ranks on the same
// XCD (workgroups are dealt round-robin over the 8 XCDs) own adjacent row
// blocks and share halo rows in their XCD's L2.
__device__ inline int xcd_rank(int l, int g) {
  int x = l & 7, base = 0;
  for (int y = 0; y < x; y++) base += (g - y + 7) >> 3;
  return base + (l >> 3);
}

// Per-lane partial sums -> this workgroup's row of the team partial table.
template <int K, int NS>
__device__ inline void block_partials(const double (&acc)[NS], Shared& sh, double* dst) {
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
  for (int q = 0; q < NS; q++) {
    double v = acc[q];
#pragma unroll
    for (int m = kWave / 2; m >= K; m >>= 1) v += __shfl_xor(v, m, kWave);
    if (lane < K) sh.wred[wave][q * K + lane] = v;
  }
  __syncthreads();
  if (threadIdx.x < NS * K) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kBlock / kWave; w++) s += sh.wred[w][threadIdx.x];
    dst[threadIdx.x] = s;
  }
}

// Team barrier.  nvals > 0: the partial table is reduced and sh.red filled.
// Returns false when the launch must stop (a wait timed out somewhere).
__device__ bool team_sync(const Team& T, int nvals, Shared& sh) {
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned g = __hip_atomic_load(&T.bar->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    unsigned prev = __hip_atomic_fetch_add(&T.bar->count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    sh.flag = (prev == (unsigned)(T.size - 1)) ? 1 : 0;
    sh.gen = g;
  }
  __syncthreads();
  if (sh.flag == 1) {
    if (nvals > 0) {
      // chunks of ranks per value, summed in fixed rank order
      const int chunks = kBlock / nvals;
      const int q = threadIdx.x % nvals, c = threadIdx.x / nvals;
      double s = 0.0;
      if (c < chunks)
        for (int w = c; w < T.size; w += chunks) s += ald(T.partial + (size_t)w * kMaxRed + q);
      __syncthreads();
      double* buf = &sh.wred[0][0];  // kBlock doubles of scratch
      if (c < chunks) buf[c * nvals + q] = s;
      __syncthreads();
      if (threadIdx.x < nvals) {
        double t = 0.0;
        for (int cc = 0; cc < chunks; cc++) t += buf[cc * nvals + threadIdx.x];
        T.scal[threadIdx.x] = t;
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      __hip_atomic_store(&T.bar->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __threadfence();
      __hip_atomic_store(&T.bar->gen, sh.gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
  } else if (threadIdx.x == 0) {
    long long t0 = wall_clock64();
    while (__hip_atomic_load(&T.bar->gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == sh.gen) {
      __builtin_amdgcn_s_sleep(1);
      if (__hip_atomic_load(T.abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
          wall_clock64() - t0 > kSpinTimeoutTicks) {
        __hip_atomic_fetch_or(T.abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        sh.flag = -1;
        break;
      }
    }
    __threadfence();
  }
  __syncthreads();
  if (sh.flag < 0) return false;
  if (nvals > 0 && threadIdx.x < nvals) sh.red[threadIdx.x] = ald(T.scal + threadIdx.x);
  __syncthreads();
  return true;
}

template <typename TI>
__device__ inline TI to_ti(double v) { return (TI)v; }

// ---------------------------------------------------------------------------
// Phases.  Lane layout: column jr = tid % K, rows tid / K + m * (kBlock / K).
// ---------------------------------------------------------------------------

template <int K, typename TI>
__device__ void phase_rhs(const Inst& I, int row_lo, int row_hi, double (&acc)[3]) {
  constexpr int RPP = kBlock / K;
  const int jr = threadIdx.x % K;
  const int kv = I.kv, vj = (kv == 1) ? 0 : jr;
  TI* r = (TI*)I.r;
  TI* z = (TI*)I.z;
  TI* xi = (TI*)I.xi;
  const TI* dinv = (const TI*)I.dinvi;
  const double inv_dt = 1.0 / I.dt;
  const bool warm = I.warm != 0;
  const int* __restrict__ rp = I.rp;
  const int* __restrict__ ci = I.ci;
  for (int row = row_lo + (int)threadIdx.x / K; row < row_hi; row += RPP) {
    const int lo = rp[row], hi = rp[row + 1];
    const size_t o = (size_t)row * K + jr;
    double b = 0.0, sx = 0.0, x0 = 0.0;
    switch (I.eq) {
      case EQ_HEAT:
      case EQ_ADVDIFF: {  // b = (M/dt) u                       CPUSolver.cpp:78,96
        double bs = 0.0;
        for (int s = lo; s < hi; s++) {
          const int c = ci[s];
          const double uc = I.u[(size_t)c * K + jr];
          bs += I.M64[s] * uc;
          if (warm) sx += I.S64[(size_t)s * kv + vj] * uc;
        }
        b = bs * inv_dt;
        x0 = warm ? I.u[o] : 0.0;
      } break;
      case EQ_WAVE: {  // b = (M/dt) v - c^2 K u                    CPUSolver.cpp:114
        double bs = 0.0, ks = 0.0;
        for (int s = lo; s < hi; s++) {
          const int c = ci[s];
          const double vc = I.v[(size_t)c * K + jr];
          bs += I.M64[s] * vc;
          ks += I.K64[s] * I.u[(size_t)c * K + jr];
          if (warm) sx += I.S64[(size_t)s * kv + vj] * vc;
        }
        b = bs * inv_dt - I.p1 * I.p1 * ks;
        x0 = warm ? I.v[o] : 0.0;
      } break;
      case EQ_GS: {  // b_u = (M/dt)u - uvv + F(1-u); b_v = (M/dt)v + uvv - (F+k)v   CPUSolver.cpp:134,137
        double bs = 0.0;
        for (int s = lo; s < hi; s++) {
          const int c = ci[s];
          const double yc = I.u[(size_t)c * K + jr];
          bs += I.M64[s] * yc;
          if (warm) sx += I.S64[(size_t)s * kv + vj] * yc;
        }
        const double uu = I.u[(size_t)row * K], vv = I.u[(size_t)row * K + 1];
        const double uvv = uu * (vv * vv);
        const double f = (jr == 0) ? (-uvv + I.p3 * (1.0 - uu)) : (uvv - (I.p3 + I.p4) * vv);
        const double w = I.weighting ? I.lumped[row] : 1.0;
        b = bs * inv_dt + w * f;
        x0 = warm ? I.u[o] : 0.0;
      } break;
      case EQ_NEWMARK: {  // b = -c^2 K (u + dt v + (1/2 - beta) dt^2 a)
        const double h = (0.5 - I.p2) * I.dt * I.dt;
        double ks = 0.0;
        for (int s = lo; s < hi; s++) {
          const int c = ci[s];
          const size_t oc = (size_t)c * K + jr;
          ks += I.K64[s] * (I.u[oc] + I.dt * I.v[oc] + h * I.a[oc]);
          if (warm) sx += I.S64[(size_t)s * kv + vj] * I.a[oc];
        }
        b = -I.p1 * I.p1 * ks;
        x0 = warm ? I.a[o] : 0.0;
      } break;
      default: {  // EQ_SOLVE: b supplied, x holds the guess
        b = I.b[o];
        if (warm)
          for (int s = lo; s < hi; s++) sx += I.S64[(size_t)s * kv + vj] * I.x[(size_t)ci[s] * K + jr];
        x0 = warm ? I.x[o] : 0.0;
      } break;
    }
    const double r64 = b - sx;
    I.b[o] = b;
    I.x[o] = x0;
    const TI rr = to_ti<TI>(r64);
    const TI zz = dinv[(size_t)row * kv + vj] * rr;
    r[o] = rr;
    z[o] = zz;
    xi[o] = (TI)0;
    acc[0] += b * b;
    acc[1] += r64 * r64;
    acc[2] += (I.krylov == 1) ? (double)rr * (double)rr : (double)rr * (double)zz;
  }
}

// FP64 refinement residual r = b - S64 (x + xi); restarts the inner solve.
template <int K, typename TI>
__device__ void phase_res(const Inst& I, int row_lo, int row_hi, double (&acc)[2]) {
  constexpr int RPP = kBlock / K;
  const int jr = threadIdx.x % K;
  const int kv = I.kv, vj = (kv == 1) ? 0 : jr;
  TI* r = (TI*)I.r;
  TI* z = (TI*)I.z;
  const TI* xi = (const TI*)I.xi;
  const TI* dinv = (const TI*)I.dinvi;
  const int* __restrict__ rp = I.rp;
  const int* __restrict__ ci = I.ci;
  for (int row = row_lo + (int)threadIdx.x / K; row < row_hi; row += RPP) {
    const size_t o = (size_t)row * K + jr;
    double sx = 0.0;
    for (int s = rp[row]; s < rp[row + 1]; s++) {
      const size_t oc = (size_t)ci[s] * K + jr;
      sx += I.S64[(size_t)s * kv + vj] * (I.x[oc] + (double)xi[oc]);
    }
    const double r64 = I.b[o] - sx;
    const TI rr = to_ti<TI>(r64);
    const TI zz = dinv[(size_t)row * kv + vj] * rr;
    r[o] = rr;
    z[o] = zz;
    acc[0] += r64 * r64;
    acc[1] += (I.krylov == 1) ? (double)rr * (double)rr : (double)rr * (double)zz;
  }
}

// CG phase A: p = z + beta p_old, q = S p, <p, q>.  Neighbour p values are
// recomputed from z and p_old, so p needs no barrier of its own.  On the
// first iteration of a sweep the pending correction is folded: x += xi.
template <int K, typename TI>
__device__ void phase_cg_a(const Inst& I, int row_lo, int row_hi, const TI* pold, TI* pnew, bool first,
                           const Shared& sh, double (&acc)[1]) {
  constexpr int RPP = kBlock / K;
  const int jr = threadIdx.x % K;
  const int kv = I.kv, vj = (kv == 1) ? 0 : jr;
  const TI* z = (const TI*)I.z;
  TI* q = (TI*)I.q;
  TI* xi = (TI*)I.xi;
  const TI* S = (const TI*)I.Si;
  const TI bj = first ? (TI)0 : to_ti<TI>(sh.beta[jr]);
  const int* __restrict__ rp = I.rp;
  const int* __restrict__ ci = I.ci;
  for (int row = row_lo + (int)threadIdx.x / K; row < row_hi; row += RPP) {
    const size_t o = (size_t)row * K + jr;
    if (first) {
      I.x[o] += (double)xi[o];
      xi[o] = (TI)0;
    }
    TI qa = (TI)0;
    if (first) {
      for (int s = rp[row]; s < rp[row + 1]; s++) qa += S[(size_t)s * kv + vj] * z[(size_t)ci[s] * K + jr];
    } else {
      for (int s = rp[row]; s < rp[row + 1]; s++) {
        const size_t oc = (size_t)ci[s] * K + jr;
        qa += S[(size_t)s * kv + vj] * (z[oc] + bj * pold[oc]);
      }
    }
    const TI pn = first ? z[o] : (z[o] + bj * pold[o]);
    pnew[o] = pn;
    q[o] = qa;
    acc[0] += (double)pn * (double)qa;
  }
}

// CG phase B: xi += alpha p, r -= alpha q, z = D^-1 r, <r,z>, <r,r>.
template <int K, typename TI>
__device__ void phase_cg_b(const Inst& I, int row_lo, int row_hi, const TI* p, const Shared& sh,
                           double (&acc)[2]) {
  constexpr int RPP = kBlock / K;
  const int jr = threadIdx.x % K;
  const int kv = I.kv, vj = (kv == 1) ? 0 : jr;
  TI* r = (TI*)I.r;
  TI* z = (TI*)I.z;
  TI* xi = (TI*)I.xi;
  const TI* q = (const TI*)I.q;
  const TI* dinv = (const TI*)I.dinvi;
  const TI al = to_ti<TI>(sh.alpha[jr]);
  for (int row = row_lo + (int)threadIdx.x / K; row < row_hi; row += RPP) {
    const size_t o = (size_t)row * K + jr;
    xi[o] += al * p[o];
    const TI rr = r[o] - al * q[o];
    const TI zz = dinv[(size_t)row * kv + vj] * rr;
    r[o] = rr;
    z[o] = zz;
    acc[0] += (double)rr * (double)zz;
    acc[1] += (double)rr * (double)rr;
  }
}

// BiCGSTAB (right Jacobi preconditioning) phase P:
// p = r + beta (p_old - omega v_old); v = S D^-1 p; <rh, v>.
template <int K, typename TI>
__device__ void phase_bi_p(const Inst& I, int row_lo, int row_hi, const TI* pold, TI* pnew, const TI* vold,
                           TI* vnew, bool first, const Shared& sh, double (&acc)[1]) {
  constexpr int RPP = kBlock / K;
  const int jr = threadIdx.x % K;
  const int kv = I.kv, vj = (kv == 1) ? 0 : jr;
  const TI* r = (const TI*)I.r;
  TI* rh = (TI*)I.rh;
  TI* xi = (TI*)I.xi;
  const TI* S = (const TI*)I.Si;
  const TI* dinv = (const TI*)I.dinvi;
  const TI be = first ? (TI)0 : to_ti<TI>(sh.beta[jr]);
  const TI om = first ? (TI)0 : to_ti<TI>(sh.omega[jr]);
  const int* __restrict__ rp = I.rp;
  const int* __restrict__ ci = I.ci;
  for (int row = row_lo + (int)threadIdx.x / K; row < row_hi; row += RPP) {
    const size_t o = (size_t)row * K + jr;
    if (first) {
      I.x[o] += (double)xi[o];
      xi[o] = (TI)0;
      rh[o] = r[o];
    }
    TI va = (TI)0;
    for (int s = rp[row]; s < rp[row + 1]; s++) {
      const int c = ci[s];
      const size_t oc = (size_t)c * K + jr;
      const TI pc = first ? r[oc] : (r[oc] + be * (pold[oc] - om * vold[oc]));
      va += S[(size_t)s * kv + vj] * (dinv[(size_t)c * kv + vj] * pc);
    }
    const TI pn = first ? r[o] : (r[o] + be * (pold[o] - om * vold[o]));
    pnew[o] = pn;
    vnew[o] = va;
    acc[0] += (double)(first ? r[o] : rh[o]) * (double)va;
  }
}

// BiCGSTAB phase S: s = r - alpha v; t = S D^-1 s; <t,s>, <t,t>, <s,s>.
template <int K, typename TI>
__device__ void phase_bi_s(const Inst& I, int row_lo, int row_hi, const TI* vnew, const Shared& sh,
                           double (&acc)[3]) {
  constexpr int RPP = kBlock / K;
  const int jr = threadIdx.x % K;
  const int kv = I.kv, vj = (kv == 1) ? 0 : jr;
  const TI* r = (const TI*)I.r;
  TI* sv = (TI*)I.s;
  TI* tv = (TI*)I.t;
  const TI* S = (const TI*)I.Si;
  const TI* dinv = (const TI*)I.dinvi;
  const TI al = to_ti<TI>(sh.alpha[jr]);
  const int* __restrict__ rp = I.rp;
  const int* __restrict__ ci = I.ci;
  for (int row = row_lo + (int)threadIdx.x / K; row < row_hi; row += RPP) {
    const size_t o = (size_t)row * K + jr;
    TI ta = (TI)0;
    for (int s = rp[row]; s < rp[row + 1]; s++) {
      const int c = ci[s];
      const size_t oc = (size_t)c * K + jr;
      ta += S[(size_t)s * kv + vj] * (dinv[(size_t)c * kv + vj] * (r[oc] - al * vnew[oc]));
    }
    const TI so = r[o] - al * vnew[o];
    sv[o] = so;
    tv[o] = ta;
    acc[0] += (double)ta * (double)so;
    acc[1] += (double)ta * (double)ta;
    acc[2] += (double)so * (double)so;
  }
}

// BiCGSTAB phase X: xi += D^-1 (alpha p + omega s); r = s - omega t; <rh,r>, <r,r>.
template <int K, typename TI>
__device__ void phase_bi_x(const Inst& I, int row_lo, int row_hi, const TI* p, const Shared& sh,
                           double (&acc)[2]) {
  constexpr int RPP = kBlock / K;
  const int jr = threadIdx.x % K;
  const int kv = I.kv, vj = (kv == 1) ? 0 : jr;
  TI* r = (TI*)I.r;
  TI* xi = (TI*)I.xi;
  const TI* rh = (const TI*)I.rh;
  const TI* sv = (const TI*)I.s;
  const TI* tv = (const TI*)I.t;
  const TI* dinv = (const TI*)I.dinvi;
  const TI al = to_ti<TI>(sh.alpha[jr]);
  const TI om = to_ti<TI>(sh.omega[jr]);
  for (int row = row_lo + (int)threadIdx.x / K; row < row_hi; row += RPP) {
    const size_t o = (size_t)row * K + jr;
    const TI d = dinv[(size_t)row * kv + vj];
    xi[o] += d * (al * p[o] + om * sv[o]);
    const TI rr = sv[o] - om * tv[o];
    r[o] = rr;
    acc[0] += (double)rh[o] * (double)rr;
    acc[1] += (double)rr * (double)rr;
  }
}

// State update and the instability rule (CPUSolver.cpp:41-48, widened to
// flag NaN/Inf separately).
template <int K, typename TI>
__device__ void phase_update(const Inst& I, int row_lo, int row_hi, const Shared& sh, int* flags) {
  constexpr int RPP = kBlock / K;
  const int jr = threadIdx.x % K;
  TI* xi = (TI*)I.xi;
  int f = 0;
  for (int row = row_lo + (int)threadIdx.x / K; row < row_hi; row += RPP) {
    const size_t o = (size_t)row * K + jr;
    double xf = I.x[o] + (double)xi[o];
    if (sh.bb[jr] == 0.0) xf = 0.0;  // Eigen: rhs == 0 -> x = 0
    I.x[o] = xf;
    xi[o] = (TI)0;
    double c1 = xf, c2 = 0.0;
    switch (I.eq) {
      case EQ_HEAT:
      case EQ_ADVDIFF:
        I.u[o] = xf;
        break;
      case EQ_WAVE: {
        I.v[o] = xf;
        const double un = I.u[o] + xf * I.dt;  // CPUSolver.cpp:118
        I.u[o] = un;
        c2 = un;
      } break;
      case EQ_GS:
        I.u[o] = xf;
        break;
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
    if (c1 > 1e4 || c2 > 1e4) f |= ST_UNSTABLE;
    if (!isfinite(c1) || !isfinite(c2)) f |= ST_NONFINITE;
  }
  if (f) atomicOr(flags, f);
}

// ---------------------------------------------------------------------------
// The kernel
// ---------------------------------------------------------------------------

template <int K, typename TI>
__global__ void __launch_bounds__(kBlock) k_step(const Team* teams, const int* wg_team, const Inst* insts,
                                                const int* team_inst, int nsteps) {
  __shared__ Shared sh;
  const int tid_team = wg_team[blockIdx.x];
  if (tid_team < 0) return;
  const Team T = teams[tid_team];
  const int l = (int)blockIdx.x - T.first_wg;
  const int rank = xcd_rank(l, T.size);
  double* my_partial = T.partial + (size_t)rank * kMaxRed;

  for (int qi = T.inst_begin; qi < T.inst_end; qi++) {
    const Inst& I = insts[team_inst[qi]];
    const int n = I.n;
    if (n <= 0) continue;
    const int rows_per_wg = (n + T.size - 1) / T.size;
    const int row_lo = min(n, rank * rows_per_wg);
    const int row_hi = min(n, row_lo + rows_per_wg);
    const int jr = threadIdx.x % K;
    long long its_total = 0, sweeps_total = 0;
    int status = 0;

    for (int step = 0; step < nsteps; step++) {
      // ---- R: right-hand side and initial residual
      {
        double acc[3] = {0.0, 0.0, 0.0};
        phase_rhs<K, TI>(I, row_lo, row_hi, acc);
        block_partials<K, 3>(acc, sh, my_partial);
        if (!team_sync(T, 3 * K, sh)) return;
      }
      if (threadIdx.x < K) {
        const double bb = sh.red[threadIdx.x], rr = sh.red[K + threadIdx.x];
        sh.bb[threadIdx.x] = bb;
        sh.thr[threadIdx.x] = fmax(I.tol2 * bb, DBL_MIN);
        sh.active[threadIdx.x] = (bb != 0.0) && (rr >= sh.thr[threadIdx.x]);
        sh.gamma[threadIdx.x] = sh.red[2 * K + threadIdx.x];
        sh.ithr[threadIdx.x] = rr;  // residual^2 at sweep start
      }
      __syncthreads();
      int sweeps = 0, its = 0;
      const int max_iters = I.max_iters;
      while (true) {
        if (threadIdx.x == 0) {
          int any = 0;
          for (int j = 0; j < K; j++) any |= sh.active[j];
          sh.uni = any && sweeps < I.max_sweeps && its < max_iters;
        }
        __syncthreads();
        if (!sh.uni) break;
        if (threadIdx.x < K) {
          // inner threshold: FP32 sweeps reduce the residual by inner_tol;
          // an FP64 inner solve goes straight to the final tolerance.
          const double r0 = sh.ithr[threadIdx.x];
          sh.ithr[threadIdx.x] = (sizeof(TI) == 4) ? fmax(I.inner_tol2 * r0, sh.thr[threadIdx.x])
                                                   : sh.thr[threadIdx.x];
        }
        __syncthreads();
        int it_s = 0;
        TI* pa = (TI*)I.p0;
        TI* pb = (TI*)I.p1b;
        if (I.krylov == 0) {
          // ---------------- preconditioned CG
          bool first = true;
          while (true) {
            {
              double acc[1] = {0.0};
              phase_cg_a<K, TI>(I, row_lo, row_hi, pa, pb, first, sh, acc);
              block_partials<K, 1>(acc, sh, my_partial);
              if (!team_sync(T, K, sh)) return;
            }
            if (threadIdx.x < K) {
              const double pq = sh.red[threadIdx.x];
              sh.alpha[threadIdx.x] =
                  (sh.active[threadIdx.x] && pq != 0.0) ? sh.gamma[threadIdx.x] / pq : 0.0;
            }
            __syncthreads();
            {
              double acc[2] = {0.0, 0.0};
              phase_cg_b<K, TI>(I, row_lo, row_hi, pb, sh, acc);
              block_partials<K, 2>(acc, sh, my_partial);
              if (!team_sync(T, 2 * K, sh)) return;
            }
            it_s++;
            its++;
            if (threadIdx.x < K) {
              const double rz = sh.red[threadIdx.x];
              const double g = sh.gamma[threadIdx.x];
              sh.beta[threadIdx.x] = (g != 0.0) ? rz / g : 0.0;
              sh.gamma[threadIdx.x] = rz;
            }
            if (threadIdx.x == 0) {
              int alldone = 1;
              for (int j = 0; j < K; j++)
                alldone &= (!sh.active[j]) || (sh.red[K + j] < sh.ithr[j]) || !(sh.red[K + j] == sh.red[K + j]);
              sh.uni = alldone || it_s >= I.max_inner || its >= max_iters;
            }
            __syncthreads();
            if (sh.uni) break;
            first = false;
            TI* tmp = pa;
            pa = pb;
            pb = tmp;
          }
        } else {
          // ---------------- BiCGSTAB
          TI* va = (TI*)I.v0;
          TI* vb = (TI*)I.v1;
          bool first = true;
          while (true) {
            {
              double acc[1] = {0.0};
              phase_bi_p<K, TI>(I, row_lo, row_hi, pa, pb, va, vb, first, sh, acc);
              block_partials<K, 1>(acc, sh, my_partial);
              if (!team_sync(T, K, sh)) return;
            }
            if (threadIdx.x < K) {
              const double rv = sh.red[threadIdx.x];
              sh.alpha[threadIdx.x] =
                  (sh.active[threadIdx.x] && rv != 0.0) ? sh.gamma[threadIdx.x] / rv : 0.0;
            }
            __syncthreads();
            {
              double acc[3] = {0.0, 0.0, 0.0};
              phase_bi_s<K, TI>(I, row_lo, row_hi, vb, sh, acc);
              block_partials<K, 3>(acc, sh, my_partial);
              if (!team_sync(T, 3 * K, sh)) return;
            }
            if (threadIdx.x < K) {
              const double ts = sh.red[threadIdx.x], tt = sh.red[K + threadIdx.x], ss = sh.red[2 * K + threadIdx.x];
              const bool act = sh.active[threadIdx.x];
              sh.omega[threadIdx.x] = (act && tt > 0.0 && !(ss < sh.ithr[threadIdx.x])) ? ts / tt : 0.0;
            }
            __syncthreads();
            {
              double acc[2] = {0.0, 0.0};
              phase_bi_x<K, TI>(I, row_lo, row_hi, pb, sh, acc);
              block_partials<K, 2>(acc, sh, my_partial);
              if (!team_sync(T, 2 * K, sh)) return;
            }
            it_s++;
            its++;
            if (threadIdx.x == 0) {
              int alldone = 1, brk = 0;
              for (int j = 0; j < K; j++) {
                const double rho_new = sh.red[j], rr = sh.red[K + j];
                const bool act = sh.active[j];
                const bool done = !act || rr < sh.ithr[j] || !(rr == rr);
                alldone &= done;
                if (!done) {
                  const double om = sh.omega[j], rho = sh.gamma[j];
                  if (om == 0.0 || rho == 0.0 || rho_new == 0.0) brk = 1;  // breakdown: restart sweep
                  else sh.beta[j] = (rho_new / rho) * (sh.alpha[j] / om);
                }
                sh.gamma[j] = rho_new;
              }
              sh.uni = alldone || brk || it_s >= I.max_inner || its >= max_iters;
            }
            __syncthreads();
            if (sh.uni) break;
            first = false;
            TI* tmp = pa;
            pa = pb;
            pb = tmp;
            tmp = va;
            va = vb;
            vb = tmp;
          }
        }
        // ---- Res: FP64 residual of x + xi
        {
          double acc[2] = {0.0, 0.0};
          phase_res<K, TI>(I, row_lo, row_hi, acc);
          block_partials<K, 2>(acc, sh, my_partial);
          if (!team_sync(T, 2 * K, sh)) return;
        }
        sweeps++;
        if (threadIdx.x < K) {
          const double rr = sh.red[threadIdx.x];
          sh.active[threadIdx.x] = (sh.bb[threadIdx.x] != 0.0) && (rr >= sh.thr[threadIdx.x]) && (rr == rr);
          if (!(rr == rr)) sh.active[threadIdx.x] = 0;
          sh.gamma[threadIdx.x] = sh.red[K + threadIdx.x];
          sh.ithr[threadIdx.x] = rr;
        }
        __syncthreads();
      }
      // ---- U: state update
      {
        int notconv = 0;
        for (int j = 0; j < K; j++) notconv |= sh.active[j];
        if (notconv) status |= ST_NOT_CONVERGED;
        phase_update<K, TI>(I, row_lo, row_hi, sh, I.status);
        if (rank == 0 && threadIdx.x < K) {
          const double bb = sh.bb[threadIdx.x];
          I.last_res[threadIdx.x] = bb > 0.0 ? sqrt(sh.ithr[threadIdx.x] / bb) : 0.0;
        }
        its_total += its;
        sweeps_total += sweeps;
        // the next step (or the next instance of this team) gathers the new state
        if (!team_sync(T, 0, sh)) return;
      }
      (void)jr;
    }
    if (rank == 0 && threadIdx.x == 0) {
      I.counters[0] += its_total;
      I.counters[1] += sweeps_total;
      I.counters[2] += nsteps;
      if (status) atomicOr(I.status, status);
    }
  }
}

}  // namespace

// Launch one persistent cooperative grid for a prepared batch.
int launch_step(const StepLaunch& L, int nsteps, hipStream_t st) {
  void* args[] = {(void*)&L.teams, (void*)&L.wg_team, (void*)&L.insts, (void*)&L.team_inst, (void*)&nsteps};
  const void* fn = nullptr;
#define PICK(KK)                                                                        \
  case KK:                                                                              \
    fn = L.inner_fp32 ? (const void*)k_step<KK, float> : (const void*)k_step<KK, double>; \
    break;
  switch (L.k) {
    PICK(1)
    PICK(2)
    PICK(4)
    PICK(8)
    PICK(16)
    default:
      set_error("unsupported number of right-hand sides");
      return FEM_ERR_INVALID;
  }
#undef PICK
  hipError_t e = hipLaunchCooperativeKernel(fn, dim3(L.grid), dim3(kBlock), args, 0, st);
  if (e != hipSuccess) return hip_fail(e, "hipLaunchCooperativeKernel(k_step)");
  return 0;
}

int step_max_coresident(int k, int inner_fp32, int* blocks) {
  const void* fn = nullptr;
#define PICK(KK)                                                                        \
  case KK:                                                                              \
    fn = inner_fp32 ? (const void*)k_step<KK, float> : (const void*)k_step<KK, double>; \
    break;
  switch (k) {
    PICK(1)
    PICK(2)
    PICK(4)
    PICK(8)
    PICK(16)
    default:
      return FEM_ERR_INVALID;
  }
#undef PICK
  int per_cu = 0, dev = 0, cus = 0;
  FEM_HIP(hipGetDevice(&dev));
  FEM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  FEM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kBlock, 0));
  *blocks = per_cu * cus;
  return 0;
}

}  // namespace femhip
