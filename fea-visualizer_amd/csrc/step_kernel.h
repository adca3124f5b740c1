// step_kernel.h -- persistent, cooperative time-stepping kernel (included by
// one translation unit per exchange mode: step_plain.hip, step_coherent.hip,
// step_fenced.hip).
//
// One launch advances every instance of a batch by `nsteps` implicit time
// steps.  The grid is split into teams (contiguous workgroup ranges); a team
// owns a list of instances and walks them one after the other.  Inside a team
// each workgroup owns a contiguous block of CSR rows, and the team meets at a
// barrier between phases.
//
// Memory model.  On gfx950 a device-scope fence is a whole-L2 write-back plus
// invalidate (buffer_wbl2 sc1 / buffer_inv sc1) that the L2 serialises across
// workgroups, so a fenced barrier costs ~35 ns per participating workgroup.
// Two exchange modes serve teams of more than one workgroup:
//  - coherent (no fences): every value that one workgroup writes during the
//    launch and another reads later (state, solution, residual, Krylov
//    vectors, partial sums) is stored and loaded device-coherently (agent-
//    scope atomic accesses: global_load/store sc1).  A barrier is then: every
//    wave waits for its stores to complete (s_waitcnt 0 -- the release the
//    memory model would emit for such data, without the cache-wide
//    write-back), s_barrier, and the workgroup stores its partial dot
//    products coherently into its own slot.  Best when a workgroup's rows fit
//    its LDS row cache and the phases are latency-bound;
//  - fenced: ordinary cached accesses for all data; one wave writes the L2
//    back before the workgroup's arrival and invalidates once the barrier has
//    completed.  Best when phases stream megabytes between barriers.
// In both, each 64-bit partial carries the barrier's epoch tag in its lowest
// mantissa bit, so the partials are also the arrival signal: readers poll the
// values themselves, and arrivals never contend on one address.  Read-only
// data (pattern, matrices, preconditioner) uses ordinary cached loads.  A
// one-workgroup team needs none of this and runs the plain variant.
//
// Reductions: partials are four-way buffered by epoch in the slots and always
// summed in rank order, so results are deterministic and every workgroup
// takes the same decisions.
//
// Per step and instance (the reference's CPUSolver::advance_time,
// src/FEM/CPUSolver.cpp:64-147, re-expressed as fused phases):
//   R    right-hand side (FP64) incl. the Gray-Scott reaction, the initial
//        residual r = b - S x0 (FP64), r -> inner precision
//   I    CG start: w = S D^-1 r, <r, D^-1 r>, <D^-1 r, w>
//   C    one Chronopoulos-Gear PCG iteration per barrier: p, s = S p, xi, r
//        updated for own rows; w = S D^-1 r_new with the neighbours' r_new
//        recomputed from the double-buffered r, w, s of the previous
//        iteration; <r,u>, <u,w>, <r,r>
//        (classic two-phase CG and BiCGSTAB are also available)
//   Res  FP64 refinement residual r = b - S64 (x + xi), next sweep while
//        |r| >= tol |b|
//   U    state update (heat: u = x; wave: v = x, u += dt v; Gray-Scott: u,v;
//        Newmark: u, v, a) and the instability rule (CPUSolver.cpp:41-48)
// Vectors are [rows][k] with the right-hand side fastest, so the k lanes of a
// row gather one contiguous span per neighbour.
#pragma once

#include "femhip_internal.h"
#include "femhip_step.h"

#include <cfloat>

namespace femhip {

namespace {

// Every barrier wait is bounded by the team's wait_ticks (wall_clock64 ticks
// at 100 MHz; 4 s unless the caller set another limit): a longer wait aborts
// the launch.  It cannot happen in a correct run.

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
  int uni;                 // uniform control value broadcast
  int redo;                // refinement: repeat the residual pass
  int bkd;                 // BiCGSTAB: a scalar of this iteration broke down
  int reset[kMaxK];        // refinement: drop the correction xi of column j
};

// Pointers loaded from the instance table are generic; casting them to the
// global address space lets the compiler emit global_load/store instead of
// flat accesses.
#define GA __attribute__((address_space(1)))
template <typename T>
__device__ inline T GA* gp(T* p) { return (T GA*)p; }
template <typename T>
__device__ inline const T GA* gp(const T* p) { return (const T GA*)p; }

// Exchange modes of a launch (template parameter C):
//   kPlain     every team is one workgroup: nothing is shared across CUs;
//   kCoherent  data exchanged between the workgroups of a team is stored and
//              loaded device-coherently (below), barriers need no fence;
//   kFenced    data uses ordinary cached accesses; a barrier arrival first
//              writes the L2 back (release) and the workgroup invalidates its
//              caches once the barrier completes (acquire).  Costlier per
//              barrier, but every vector access stays cached: the mode for
//              teams whose phases stream megabytes between barriers.
constexpr int kPlain = 0, kCoherent = 1, kFenced = 2;

// Device-coherent (C == kCoherent) or plain access to data exchanged between
// the workgroups of a team.
template <int C, typename T>
__device__ inline T ld(const T* p) {
  if constexpr (C == kCoherent) return __hip_atomic_load(gp(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else return *gp(p);
}
template <int C, typename T>
__device__ inline void st(T* p, T v) {
  if constexpr (C == kCoherent) __hip_atomic_store(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else *gp(p) = v;
}


// Rank of this workgroup inside its team, ordered so that ranks on the same
// XCD (workgroups are dealt round-robin over the 8 XCDs) own adjacent row
// blocks and share halo rows in their XCD's L2.
__device__ inline int xcd_rank(int l, int g) {
  int x = l & 7, base = 0;
  for (int y = 0; y < x; y++) base += (g - y + 7) >> 3;
  return base + (l >> 3);
}

// Tag a partial sum for barrier epoch e: its lowest mantissa bit becomes
// (e >> 2) & 1.  The change is at most one unit in the last place and is the
// same on every workgroup, so reductions stay deterministic and every
// workgroup takes the same decisions.
__device__ inline double tag_value(double v, unsigned e) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  return __longlong_as_double((long long)((b & ~1ull) | (unsigned long long)((e >> 2) & 1u)));
}
__device__ inline bool has_tag(double v, unsigned e) {
  return ((unsigned long long)__double_as_longlong(v) & 1ull) == (unsigned long long)((e >> 2) & 1u);
}

// Values every barrier of a K-column instance writes per buffer.  A barrier
// reducing fewer values pads with tagged zeros: otherwise entries last
// written eight epochs earlier (same buffer, same tag) would look current.
template <int K>
constexpr int kPadVals = 4 * K;

// Per-lane partial sums -> this workgroup's share at q*K + j.  A one-workgroup
// team keeps it in LDS (sh.red).  Larger teams first wait until every wave's
// stores of the phase are complete, then store the share tagged and coherently
// into buffer e & 3 of their slot: the tagged values are the arrival signal.
template <int K, int NS, int C>
__device__ inline void block_partials(const double (&acc)[NS], Shared& sh, double* dst, unsigned e) {
  static_assert(NS * K <= kPadVals<K>, "reduction wider than the slot padding");
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
  for (int q = 0; q < NS; q++) {
    double v = acc[q];
#pragma unroll
    for (int m = kWave / 2; m >= K; m >>= 1) v += __shfl_xor(v, m, kWave);
    if (lane < K) sh.wred[wave][q * K + lane] = v;
  }
  if constexpr (C != kPlain) __builtin_amdgcn_s_waitcnt(0);  // the phase's stores have completed
  __syncthreads();
  if constexpr (C == kFenced) {
    // release: write the phase's data back from this XCD's L2 before arriving
    if (threadIdx.x < kWave) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
  }
  if constexpr (C != kPlain) {
    if (threadIdx.x < kPadVals<K>) {
      double s = 0.0;
      if (threadIdx.x < NS * K)
#pragma unroll
        for (int w = 0; w < kBlock / kWave; w++) s += sh.wred[w][threadIdx.x];
      st<kCoherent>(dst + threadIdx.x, tag_value(s, e));  // the team's slot in global memory
    }
  } else if (threadIdx.x < NS * K) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kBlock / kWave; w++) s += sh.wred[w][threadIdx.x];
    dst[threadIdx.x] = s;  // sh.red in LDS
  }
}

__device__ inline bool aborted(const Team& T) {
  return __hip_atomic_load(T.abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

__device__ void raise_abort(const Team& T) {
  if (threadIdx.x == 0) __hip_atomic_fetch_or(T.abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Re-load a tagged value until it carries epoch e's tag.  Sets *to on a
// timeout or when another workgroup aborted.
__device__ double wait_value(const Team& T, const double* p, unsigned e, int* to) {
  const long long t0 = wall_clock64();
  int spins = 0;
  while (true) {
    const double v = ld<kCoherent>(p);
    if (has_tag(v, e)) return v;
    if (T.wait_ticks <= 0) {  // zero limit: no waiting at all
      *to = 1;
      return 0.0;
    }
    __builtin_amdgcn_s_sleep(1);
    if ((++spins & 63) == 0 && (aborted(T) || wall_clock64() - t0 > T.wait_ticks)) {
      *to = 1;
      return 0.0;
    }
  }
}

// Wait for the tagged partials of every rank and sum them in rank order:
// thread (q, c) sums ranks c, c + chunks, ... with kR loads in flight, then
// the chunk sums are combined in LDS.  publish: also store the tagged sums
// for the other ranks (rank 0 of a large team).  Returns nonzero on abort.
__device__ int reduce_tagged(const Team& T, unsigned e, int nvals, int npad, Shared& sh, bool publish) {
  const int buf = e & 3;
  const int chunks = kBlock / nvals;
  const int q = threadIdx.x % nvals, c = threadIdx.x / nvals;
  double s = 0.0;
  int to = 0;
  if (c < chunks) {
    constexpr int kR = 8;
    for (int w0 = c; w0 < T.size && !to; w0 += kR * chunks) {
      double v[kR];
#pragma unroll
      for (int i = 0; i < kR; i++) {
        const int w = w0 + i * chunks;
        v[i] = ld<kCoherent>(&T.slots[w < T.size ? w : c].vals[buf][q]);
      }
#pragma unroll
      for (int i = 0; i < kR; i++) {
        const int w = w0 + i * chunks;
        if (w < T.size && !has_tag(v[i], e)) v[i] = wait_value(T, &T.slots[w].vals[buf][q], e, &to);
      }
#pragma unroll
      for (int i = 0; i < kR; i++)
        if (w0 + i * chunks < T.size) s += v[i];
    }
  }
  double* scratch = &sh.wred[0][0];  // kBlock doubles
  if (c < chunks) scratch[c * nvals + q] = s;
  if (__syncthreads_or(to)) return 1;
  if (threadIdx.x < nvals) {
    double t = 0.0;
    for (int cc = 0; cc < chunks; cc++) t += scratch[cc * nvals + threadIdx.x];
    if (publish) {
      // every rank must use bit-identical values: rank 0 uses the tagged sum too
      t = tag_value(t, e);
      st<kCoherent>(&T.scal[buf * kMaxRed + threadIdx.x], t);
    }
    sh.red[threadIdx.x] = t;
  } else if (publish && threadIdx.x < npad) {
    st<kCoherent>(&T.scal[buf * kMaxRed + threadIdx.x], tag_value(0.0, e));
  }
  return 0;
}

// Team barrier for epoch e; this workgroup's partials are already stored.
// Returns false when the launch must stop.
__device__ bool team_sync(const Team& T, int rank, unsigned e, int nvals, int npad, Shared& sh) {
  __syncthreads();
  if (T.size == 1) return true;
  if (T.allpoll || rank == 0) {
    if (reduce_tagged(T, e, nvals, npad, sh, !T.allpoll)) {
      raise_abort(T);
      return false;
    }
  } else {
    int to = 0;
    if (threadIdx.x < nvals) sh.red[threadIdx.x] = wait_value(T, &T.scal[(e & 3) * kMaxRed + threadIdx.x], e, &to);
    if (__syncthreads_or(to)) {
      raise_abort(T);
      return false;
    }
  }
  __syncthreads();
  return true;
}

// Reduction + barrier: acc -> team sums in sh.red[q*K + j].
template <int K, int NS, int C>
__device__ bool team_reduce(const double (&acc)[NS], const Team& T, int rank, unsigned& e, Shared& sh) {
  e++;
  // a one-workgroup team (also inside a multi-workgroup launch) reduces in LDS
  const bool lds = C == kPlain || T.size == 1;
  if (lds) block_partials<K, NS, kPlain>(acc, sh, sh.red, e);
  else block_partials<K, NS, C>(acc, sh, T.slots[rank].vals[e & 3], e);
  const bool ok = team_sync(T, rank, e, NS * K, kPadVals<K>, sh);
  if constexpr (C == kFenced) {
    if (ok && !lds) {
      // acquire: drop cached copies of data other workgroups wrote this phase
      if (threadIdx.x < kWave) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      __syncthreads();
    }
  }
  return ok;
}

// Plain barrier: K (ignored) values per workgroup are the arrival signal.
template <int K, int C>
__device__ bool team_barrier(const Team& T, int rank, unsigned& e, Shared& sh) {
  const double acc[1] = {0.0};
  return team_reduce<K, 1, C>(acc, T, rank, e, sh);
}

template <typename TI>
__device__ inline TI to_ti(double v) { return (TI)v; }

// Chronopoulos-Gear update of one entry, shared by the owner and by
// neighbours recomputing it (identical operations, identical rounding).
template <typename TI>
__device__ inline TI cg_s(TI w, TI s, TI be, bool first) { return first ? w : (w + be * s); }
template <typename TI>
__device__ inline TI cg_r(TI r, TI sv, TI al) { return r - al * sv; }

// ---------------------------------------------------------------------------
// Phases.  Lane layout: column jr = tid % K, rows tid / K + m * (kBlock / K).
//
// Row gathers are unrolled by kU with clamped (always in-bounds) indices, so a
// row's loads are all in flight before the first is consumed; the masked
// accumulation keeps the sequential summation order.
// ---------------------------------------------------------------------------

constexpr int kU = 8;

template <typename A>
__device__ inline void acc_add(A& acc, const A& t) { acc += t; }

struct D2 {
  double a, b;
  __device__ D2& operator+=(const D2& o) {
    a += o.a;
    b += o.b;
    return *this;
  }
};
struct D3 {
  double a, b, c;
  __device__ D3& operator+=(const D3& o) {
    a += o.a;
    b += o.b;
    c += o.c;
    return *this;
  }
};

// sum over s in [lo, hi) of f(s), in ascending s; U entries in flight
template <int U = kU, typename A, typename F>
__device__ inline A row_gather(int lo, int hi, A zero, F f) {
  A acc = zero;
  for (int s0 = lo; s0 < hi; s0 += U) {
    A t[U];
#pragma unroll
    for (int u = 0; u < U; u++) t[u] = f(min(s0 + u, hi - 1));
#pragma unroll
    for (int u = 0; u < U; u++)
      if (s0 + u < hi) acc_add(acc, t[u]);
  }
  return acc;
}

// Entries in flight for gathers that load five or more values per entry.
constexpr int kUWide = 4;

template <int K, typename TI, int C, int NA>
__device__ void phase_rhs(const Inst& I, int row_lo, int row_hi, TI* r, int step, double (&acc)[NA]) {
  static_assert(NA >= 3, "phase_rhs sums three values");
  constexpr int RPP = kBlock / K;
  const int jr = threadIdx.x % K;
  const int kv = I.kv, vj = (kv == 1) ? 0 : jr;
  TI* xi = (TI*)I.xi;
  const TI GA* dinv = gp((const TI*)I.dinvi);
  const double inv_dt = 1.0 / I.dt;
  const bool warm = I.warm != 0;
  // predictor order this step: 2 x_n - x_{n-1} once the previous solution is
  // known, 3 x_n - 3 x_{n-1} + x_{n-2} once two are, 4 x_n - 6 x_{n-1} +
  // 4 x_{n-2} - x_{n-3} once three are
  // (the guess gathered from the history levels here: quartic at most)
  const int lv = warm ? min(min(I.extrap, kPersistentExtrap), I.hist + step) : 0;
  const bool ex = lv > 0;
  const int GA* __restrict__ rp = gp(I.rp);
  const int GA* __restrict__ ci = gp(I.ci);
  const double GA* M64 = gp(I.M64);
  const double GA* K64 = gp(I.K64);
  const double GA* S64 = gp(I.S64);
  for (int row = row_lo + (int)threadIdx.x / K; row < row_hi; row += RPP) {
    const int lo = rp[row], hi = rp[row + 1];
    const size_t o = (size_t)row * K + jr;
    double b = 0.0, sx = 0.0, x0 = 0.0;
    switch (I.eq) {
      case EQ_HEAT:
      case EQ_ADVDIFF:
      case EQ_GS: {  // b = (M/dt) u                                  CPUSolver.cpp:78,96
        // Gray-Scott: b_u = (M/dt)u - uvv + F(1-u); b_v = (M/dt)v + uvv - (F+k)v   CPUSolver.cpp:134,137
        const D2 g = row_gather(lo, hi, D2{0.0, 0.0}, [&](int s) {
          const size_t oc = (size_t)ci[s] * K + jr;
          const double uc = ld<C>(I.u + oc);
          const double xc = lv == 4   ? 5.0 * (uc - ld<C>(I.h3 + oc)) + 10.0 * (ld<C>(I.v + oc) - ld<C>(I.a + oc)) +
                                            ld<C>(I.h4 + oc)
                            : lv == 3 ? 4.0 * (uc + ld<C>(I.v + oc)) - 6.0 * ld<C>(I.a + oc) - ld<C>(I.h3 + oc)
                            : lv == 2 ? 3.0 * (uc - ld<C>(I.a + oc)) + ld<C>(I.v + oc)
                            : ex      ? 2.0 * uc - ld<C>(I.a + oc)
                                      : uc;
          return D2{M64[s] * uc, warm ? S64[(size_t)s * kv + vj] * xc : 0.0};
        });
        b = g.a * inv_dt;
        if (I.eq == EQ_GS) {
          const double uu = ld<C>(I.u + (size_t)row * K), vv = ld<C>(I.u + (size_t)row * K + 1);
          const double uvv = uu * (vv * vv);
          const double f = (jr == 0) ? (-uvv + I.p3 * (1.0 - uu)) : (uvv - (I.p3 + I.p4) * vv);
          const double w = I.weighting ? gp(I.lumped)[row] : 1.0;
          b += w * f;
        }
        sx = g.b;
        const double uo = ld<C>(I.u + o);
        x0 = !warm     ? 0.0
             : lv == 4 ? 5.0 * (uo - ld<C>(I.h3 + o)) + 10.0 * (ld<C>(I.v + o) - ld<C>(I.a + o)) + ld<C>(I.h4 + o)
             : lv == 3 ? 4.0 * (uo + ld<C>(I.v + o)) - 6.0 * ld<C>(I.a + o) - ld<C>(I.h3 + o)
             : lv == 2 ? 3.0 * (uo - ld<C>(I.a + o)) + ld<C>(I.v + o)
             : ex      ? 2.0 * uo - ld<C>(I.a + o)
                       : uo;
      } break;
      case EQ_WAVE: {  // b = (M/dt) v - c^2 K u                    CPUSolver.cpp:114
        const D3 g = row_gather(lo, hi, D3{0.0, 0.0, 0.0}, [&](int s) {
          const size_t oc = (size_t)ci[s] * K + jr;
          const double vc = ld<C>(I.v + oc);
          const double xc = ex ? 2.0 * vc - ld<C>(I.a + oc) : vc;
          return D3{M64[s] * vc, K64[s] * ld<C>(I.u + oc), warm ? S64[(size_t)s * kv + vj] * xc : 0.0};
        });
        b = g.a * inv_dt - I.p1 * I.p1 * g.b;
        sx = g.c;
        const double vo = ld<C>(I.v + o);
        x0 = warm ? (ex ? 2.0 * vo - ld<C>(I.a + o) : vo) : 0.0;
      } break;
      case EQ_NEWMARK: {  // b = -c^2 K (u + dt v + (1/2 - beta) dt^2 a)
        // guess: the acceleration extrapolated (a_{n-1} in h3, a_{n-2} in h4)
        const double h = (0.5 - I.p2) * I.dt * I.dt;
        const D2 g = row_gather(lo, hi, D2{0.0, 0.0}, [&](int s) {
          const size_t oc = (size_t)ci[s] * K + jr;
          const double ac = ld<C>(I.a + oc);
          const double xc = lv == 2 ? 3.0 * (ac - ld<C>(I.h3 + oc)) + ld<C>(I.h4 + oc)
                            : ex    ? 2.0 * ac - ld<C>(I.h3 + oc)
                                    : ac;
          return D2{K64[s] * (ld<C>(I.u + oc) + I.dt * ld<C>(I.v + oc) + h * ac),
                    warm ? S64[(size_t)s * kv + vj] * xc : 0.0};
        });
        b = -I.p1 * I.p1 * g.a;
        sx = g.b;
        const double ao = ld<C>(I.a + o);
        x0 = !warm     ? 0.0
             : lv == 2 ? 3.0 * (ao - ld<C>(I.h3 + o)) + ld<C>(I.h4 + o)
             : ex      ? 2.0 * ao - ld<C>(I.h3 + o)
                       : ao;
      } break;
      default: {  // EQ_SOLVE: b supplied, x holds the guess
        b = gp(I.b)[o];
        if (warm)
          sx = row_gather(lo, hi, 0.0, [&](int s) {
            return S64[(size_t)s * kv + vj] * ld<C>(I.x + (size_t)ci[s] * K + jr);
          });
        x0 = warm ? ld<C>(I.x + o) : 0.0;
      } break;
    }
    const double r64 = b - sx;
    gp(I.b)[o] = b;
    st<C>(I.x + o, x0);
    const TI rr = to_ti<TI>(r64);
    st<C>(r + o, rr);
    st<C>(xi + o, (TI)0);
    acc[0] += b * b;
    acc[1] += r64 * r64;
    // BiCGSTAB starts from rho = <r, r>; CG from <r, D^-1 r>
    acc[2] += (I.krylov == 1) ? (double)rr * (double)rr : (double)rr * (double)(dinv[(size_t)row * kv + vj] * rr);
  }
}

// Predictor weights on (x_n, x_{n-1}, ..., x_{n-4}) for order lv: the
// alternating binomial coefficients of row lv + 1, 0 (1), 1 (2, -1),
// 2 (3, -3, 1), 3 (4, -6, 4, -1), 4 (5, -10, 10, -5, 1).
__device__ inline void pred_weights(int lv, double& w0, double& w1, double& w2, double& w3, double& w4) {
  w0 = lv + 1.0;
  w1 = lv == 4 ? -10.0 : lv == 3 ? -6.0 : lv == 2 ? -3.0 : lv == 1 ? -1.0 : 0.0;
  w2 = lv == 4 ? 10.0 : lv == 3 ? 4.0 : lv == 2 ? 1.0 : 0.0;
  w3 = lv == 4 ? -5.0 : lv == 3 ? -1.0 : 0.0;
  w4 = lv == 4 ? 1.0 : 0.0;
}

// The same weights for any order up to kMaxExtrap: w[j] = (-1)^j C(lv + 1, j + 1)
// on x_{n-j} (lv = 0: x_n), zero beyond the order.
__device__ inline void pred_weights_n(int lv, double (&w)[kMaxExtrap + 1]) {
  double c = 1.0;
  for (int j = 0; j <= kMaxExtrap; j++) {
    if (j <= lv) {
      c = c * (double)(lv + 1 - j) / (double)(j + 1);
      w[j] = (j & 1) ? -c : c;
    } else {
      w[j] = 0.0;
    }
  }
}

// LDS row cache of a workgroup's rows (filled by fill_cache below).
constexpr int kCacheE = 10;  // entries per cached row
constexpr int kCacheP = 2;   // cached rows per lane

template <typename TI>
struct RowCache {
  int col[kCacheP][kCacheE][kBlock];
  TI val[kCacheP][kCacheE][kBlock];
};

// One row of the FP64 residual pass of a fused step: S c, M c and S h, with
// c = x + xi the candidate solution (xi dropped for a column restarted from x)
// and h the next initial guess's history part.  HM: 0 no history part,
// 1 h gathered from hvec, 2 h formed from the history levels u, a, v, h3.
// cols (CC): the row's column indices in the LDS row cache, entry e at
// cols[e * kBlock], so the gathers need not wait for the column loads.
template <int HM, bool CC, int K, typename TI, int C>
__device__ __attribute__((always_inline)) inline D3 res_gather_fused(const Inst& I, int lo, int hi, int jr, int vj,
                                                                     bool drop, int lvn, double w1, double w2,
                                                                     double w3, double w4, const int* cols) {
  const int kv = I.kv;
  const TI* xi = (const TI*)I.xi;
  const int GA* __restrict__ ci = gp(I.ci);
  const double GA* S64 = gp(I.S64);
  const double GA* M64 = gp(I.M64);
  return row_gather<kUWide>(lo, hi, D3{0.0, 0.0, 0.0}, [&](int s) {
    const size_t oc = (size_t)(CC ? cols[(s - lo) * kBlock] : ci[s]) * K + jr;
    const double sv = S64[(size_t)s * kv + vj];
    const double xc = ld<C>(I.x + oc) + (drop ? 0.0 : (double)ld<C>(xi + oc));
    double hv = 0.0;
    if constexpr (HM == 1) hv = ld<C>(I.hvec + oc);
    if constexpr (HM == 2)
      hv = w1 * ld<C>(I.u + oc) + (lvn >= 2 ? w2 * ld<C>(I.a + oc) : 0.0) +
           (lvn >= 3 ? w3 * ld<C>(I.v + oc) : 0.0) + (lvn >= 4 ? w4 * ld<C>(I.h3 + oc) : 0.0);
    return D3{sv * xc, M64[s] * xc, sv * hv};
  });
}

// FP64 residual of x + xi into r (inner precision); acc: <r, r>, <r, D^-1 r>.
// hpre != nullptr: the next step's update is fused with its right-hand side
// and takes, should this pass turn out to be the last of the step, what it
// needs from here instead of gathering the candidate solution c = x + xi:
// the next right-hand side b = M c / dt (+ the Gray-Scott reaction at c)
// into xnext and the next initial residual b - (w0 S c + S h) into hpre
// (lvn: the next step's predictor order; hready: the update of this step
// formed h in hvec; RC: the workgroup's rows are in the LDS row cache).
template <int K, typename TI, int C>
__device__ __attribute__((always_inline)) inline void phase_res(const Inst& I, int row_lo, int row_hi, TI* r,
                                                                const Shared& sh, double (&acc)[2], int lvn = 0,
                                                                double* hpre = nullptr, bool hready = false,
                                                                const RowCache<TI>* RC = nullptr) {
  constexpr int RPP = kBlock / K;
  const int jr = threadIdx.x % K;
  const int kv = I.kv, vj = (kv == 1) ? 0 : jr;
  TI* xi = (TI*)I.xi;
  // a column whose sweep produced non-finite values restarts from x alone
  const bool drop = sh.reset[jr] != 0;
  // Eigen: rhs == 0 -> x = 0, so the next update sees a zero solution
  const bool zero = sh.bb[jr] == 0.0;
  const bool warm = I.warm != 0;
  const double inv_dt = 1.0 / I.dt;
  const TI GA* dinv = gp((const TI*)I.dinvi);
  const int GA* __restrict__ rp = gp(I.rp);
  const int GA* __restrict__ ci = gp(I.ci);
  const double GA* S64 = gp(I.S64);
  double w0, w1, w2, w3, w4;
  pred_weights(lvn, w0, w1, w2, w3, w4);
  const bool want_h = hpre && lvn > 0;
  int q = 0;  // pass over the workgroup's rows (row cache slot)
  for (int row = row_lo + (int)threadIdx.x / K; row < row_hi; row += RPP, q++) {
    const size_t o = (size_t)row * K + jr;
    const double bo = gp(I.b)[o];
    const TI dj = dinv[(size_t)row * kv + vj];  // loaded before the stores below (one completion counter)
    double sx;
    if (hpre) {
      const int lo = rp[row], hi = rp[row + 1];
      const int* cols = RC ? &RC->col[q][0][threadIdx.x] : nullptr;
      const D3 g =
          !want_h  ? res_gather_fused<0, false, K, TI, C>(I, lo, hi, jr, vj, drop, lvn, w1, w2, w3, w4, cols)
          : !hready ? res_gather_fused<2, false, K, TI, C>(I, lo, hi, jr, vj, drop, lvn, w1, w2, w3, w4, cols)
          : RC      ? res_gather_fused<1, true, K, TI, C>(I, lo, hi, jr, vj, drop, lvn, w1, w2, w3, w4, cols)
                    : res_gather_fused<1, false, K, TI, C>(I, lo, hi, jr, vj, drop, lvn, w1, w2, w3, w4, cols);
      sx = g.a;
      const double su = zero ? 0.0 : g.a;
      double bn = (zero ? 0.0 : g.b) * inv_dt;
      if (I.eq == EQ_GS) {  // CPUSolver.cpp:134,137 at the candidate
        auto cand = [&](int j) -> double {  // species j of this row
          const size_t oj = (size_t)row * K + j;
          return sh.bb[j] == 0.0 ? 0.0 : ld<C>(I.x + oj) + (sh.reset[j] ? 0.0 : (double)ld<C>(xi + oj));
        };
        double uu, vv;
        if constexpr (K == 2) {  // the row's other species is the pair lane
          const double xo = cand(jr);
          const double xp = __shfl_xor(xo, 1, kWave);
          uu = jr == 0 ? xo : xp;
          vv = jr == 0 ? xp : xo;
        } else {
          uu = cand(0);
          vv = cand(1);
        }
        const double uvv = uu * (vv * vv);
        const double fr = (jr == 0) ? (-uvv + I.p3 * (1.0 - uu)) : (uvv - (I.p3 + I.p4) * vv);
        bn += (I.weighting ? gp(I.lumped)[row] : 1.0) * fr;
      }
      gp(I.xnext)[o] = bn;
      st<C>(hpre + o, bn - (warm ? w0 * su + (want_h ? g.c : 0.0) : 0.0));
    } else {
      sx = row_gather(rp[row], rp[row + 1], 0.0, [&](int s) {
        const size_t oc = (size_t)ci[s] * K + jr;
        return S64[(size_t)s * kv + vj] * (ld<C>(I.x + oc) + (drop ? 0.0 : (double)ld<C>(xi + oc)));
      });
    }
    const double r64 = bo - sx;
    const TI rr = to_ti<TI>(r64);
    st<C>(r + o, rr);
    if (drop) st<C>(xi + o, (TI)0);
    acc[0] += r64 * r64;
    acc[1] += (I.krylov == 1) ? (double)rr * (double)rr : (double)rr * (double)(dj * rr);
  }
}

// Pipelined CG start: fold x += xi; w = S D^-1 r; <r, D^-1 r>, <D^-1 r, w>.
template <int K, typename TI, int C>
__device__ void phase_pcg_init(const Inst& I, int row_lo, int row_hi, const TI* rc, TI* wc, double (&acc)[2]) {
  constexpr int RPP = kBlock / K;
  const int jr = threadIdx.x % K;
  const int kv = I.kv, vj = (kv == 1) ? 0 : jr;
  TI* xi = (TI*)I.xi;
  const TI GA* S = gp((const TI*)I.Si);
  const TI GA* dinv = gp((const TI*)I.dinvi);
  const int GA* __restrict__ rp = gp(I.rp);
  const int GA* __restrict__ ci = gp(I.ci);
  for (int row = row_lo + (int)threadIdx.x / K; row < row_hi; row += RPP) {
    const size_t o = (size_t)row * K + jr;
    const double xo = ld<C>(I.x + o), xio = (double)ld<C>(xi + o);
    const TI ro = ld<C>(rc + o);
    const TI uo = dinv[(size_t)row * kv + vj] * ro;  // loaded before the stores below
    const TI wa = row_gather(rp[row], rp[row + 1], (TI)0, [&](int s) {
      const int c = ci[s];
      return S[(size_t)s * kv + vj] * (dinv[(size_t)c * kv + vj] * ld<C>(rc + (size_t)c * K + jr));
    });
    st<C>(I.x + o, xo + xio);
    st<C>(xi + o, (TI)0);
    st<C>(wc + o, wa);
    acc[0] += (double)ro * (double)uo;
    acc[1] += (double)uo * (double)wa;
  }
}

// One Chronopoulos-Gear iteration (see file header).
template <int K, typename TI, int C>
__device__ void phase_pcg_iter(const Inst& I, int row_lo, int row_hi, const TI* rc, const TI* wc, const TI* sc,
                               TI* rn, TI* wn, TI* sn, bool first, const Shared& sh, double (&acc)[3]) {
  constexpr int RPP = kBlock / K;
  const int jr = threadIdx.x % K;
  const int kv = I.kv, vj = (kv == 1) ? 0 : jr;
  TI* xi = (TI*)I.xi;
  TI GA* p = gp((TI*)I.p0);  // own rows only
  const TI GA* S = gp((const TI*)I.Si);
  const TI GA* dinv = gp((const TI*)I.dinvi);
  const TI al = to_ti<TI>(sh.alpha[jr]);
  const TI be = first ? (TI)0 : to_ti<TI>(sh.beta[jr]);
  const int GA* __restrict__ rp = gp(I.rp);
  const int GA* __restrict__ ci = gp(I.ci);
  for (int row = row_lo + (int)threadIdx.x / K; row < row_hi; row += RPP) {
    const size_t o = (size_t)row * K + jr;
    const TI ro = ld<C>(rc + o);
    const TI wo = ld<C>(wc + o);
    const TI so = first ? (TI)0 : ld<C>(sc + o);
    const TI po = first ? (TI)0 : p[o];
    const TI xio = ld<C>(xi + o);
    const TI wa = row_gather(rp[row], rp[row + 1], (TI)0, [&](int s) {
      const int c = ci[s];
      const size_t oc = (size_t)c * K + jr;
      const TI sv = cg_s<TI>(ld<C>(wc + oc), first ? (TI)0 : ld<C>(sc + oc), be, first);
      return S[(size_t)s * kv + vj] * (dinv[(size_t)c * kv + vj] * cg_r<TI>(ld<C>(rc + oc), sv, al));
    });
    const TI d = dinv[(size_t)row * kv + vj];
    const TI uo = d * ro;
    const TI pn = first ? uo : (uo + be * po);
    const TI sv = cg_s<TI>(wo, so, be, first);
    const TI rr = cg_r<TI>(ro, sv, al);
    p[o] = pn;
    st<C>(xi + o, xio + al * pn);
    st<C>(sn + o, sv);
    st<C>(rn + o, rr);
    st<C>(wn + o, wa);
    const TI un = d * rr;
    acc[0] += (double)rr * (double)un;
    acc[1] += (double)un * (double)wa;
    acc[2] += (double)rr * (double)rr;
  }
}

// Classic CG phase A: p = z + beta p_old (z = D^-1 r), q = S p, <p, q>.
// Neighbour p values are recomputed from r and p_old, so p needs no barrier of
// its own.  On the first iteration of a sweep the pending correction is folded.
template <int K, typename TI, int C>
__device__ void phase_cg_a(const Inst& I, int row_lo, int row_hi, const TI* r, const TI* pold, TI* pnew,
                           bool first, const Shared& sh, double (&acc)[1]) {
  constexpr int RPP = kBlock / K;
  const int jr = threadIdx.x % K;
  const int kv = I.kv, vj = (kv == 1) ? 0 : jr;
  TI GA* q = gp((TI*)I.q);  // own rows only
  TI* xi = (TI*)I.xi;
  const TI GA* S = gp((const TI*)I.Si);
  const TI GA* dinv = gp((const TI*)I.dinvi);
  const TI bj = first ? (TI)0 : to_ti<TI>(sh.beta[jr]);
  const int GA* __restrict__ rp = gp(I.rp);
  const int GA* __restrict__ ci = gp(I.ci);
  for (int row = row_lo + (int)threadIdx.x / K; row < row_hi; row += RPP) {
    const size_t o = (size_t)row * K + jr;
    if (first) {
      st<C>(I.x + o, ld<C>(I.x + o) + (double)ld<C>(xi + o));
      st<C>(xi + o, (TI)0);
    }
    const TI qa = row_gather(rp[row], rp[row + 1], (TI)0, [&](int s) {
      const int c = ci[s];
      const size_t oc = (size_t)c * K + jr;
      const TI zc = dinv[(size_t)c * kv + vj] * ld<C>(r + oc);
      return S[(size_t)s * kv + vj] * (first ? zc : (zc + bj * ld<C>(pold + oc)));
    });
    const TI zo = dinv[(size_t)row * kv + vj] * ld<C>(r + o);
    const TI pn = first ? zo : (zo + bj * ld<C>(pold + o));
    st<C>(pnew + o, pn);
    q[o] = qa;
    acc[0] += (double)pn * (double)qa;
  }
}

// Classic CG phase B: xi += alpha p, r -= alpha q, <r, D^-1 r>, <r, r>.
template <int K, typename TI, int C>
__device__ void phase_cg_b(const Inst& I, int row_lo, int row_hi, TI* r, const TI* p, const Shared& sh,
                           double (&acc)[2]) {
  constexpr int RPP = kBlock / K;
  const int jr = threadIdx.x % K;
  const int kv = I.kv, vj = (kv == 1) ? 0 : jr;
  TI* xi = (TI*)I.xi;
  const TI GA* q = gp((const TI*)I.q);
  const TI GA* dinv = gp((const TI*)I.dinvi);
  const TI al = to_ti<TI>(sh.alpha[jr]);
  for (int row = row_lo + (int)threadIdx.x / K; row < row_hi; row += RPP) {
    const size_t o = (size_t)row * K + jr;
    st<C>(xi + o, ld<C>(xi + o) + al * ld<C>(p + o));
    const TI rr = ld<C>(r + o) - al * q[o];
    st<C>(r + o, rr);
    acc[0] += (double)rr * (double)(dinv[(size_t)row * kv + vj] * rr);
    acc[1] += (double)rr * (double)rr;
  }
}

// BiCGSTAB (right Jacobi preconditioning) phase P:
// p = r + beta (p_old - omega v_old); v = S D^-1 p; <rh, v>.
template <int K, typename TI, int C>
__device__ void phase_bi_p(const Inst& I, int row_lo, int row_hi, const TI* r, const TI* pold, TI* pnew,
                           const TI* vold, TI* vnew, bool first, const Shared& sh, double (&acc)[1]) {
  constexpr int RPP = kBlock / K;
  const int jr = threadIdx.x % K;
  const int kv = I.kv, vj = (kv == 1) ? 0 : jr;
  TI GA* rh = gp((TI*)I.rh);  // own rows only
  TI* xi = (TI*)I.xi;
  const TI GA* S = gp((const TI*)I.Si);
  const TI GA* dinv = gp((const TI*)I.dinvi);
  const TI be = first ? (TI)0 : to_ti<TI>(sh.beta[jr]);
  const TI om = first ? (TI)0 : to_ti<TI>(sh.omega[jr]);
  const int GA* __restrict__ rp = gp(I.rp);
  const int GA* __restrict__ ci = gp(I.ci);
  for (int row = row_lo + (int)threadIdx.x / K; row < row_hi; row += RPP) {
    const size_t o = (size_t)row * K + jr;
    const TI ro = ld<C>(r + o);
    if (first) {
      st<C>(I.x + o, ld<C>(I.x + o) + (double)ld<C>(xi + o));
      st<C>(xi + o, (TI)0);
      rh[o] = ro;
    }
    const TI va = row_gather(rp[row], rp[row + 1], (TI)0, [&](int s) {
      const int c = ci[s];
      const size_t oc = (size_t)c * K + jr;
      const TI rcv = ld<C>(r + oc);
      const TI pc = first ? rcv : (rcv + be * (ld<C>(pold + oc) - om * ld<C>(vold + oc)));
      return S[(size_t)s * kv + vj] * (dinv[(size_t)c * kv + vj] * pc);
    });
    const TI pn = first ? ro : (ro + be * (ld<C>(pold + o) - om * ld<C>(vold + o)));
    st<C>(pnew + o, pn);
    st<C>(vnew + o, va);
    acc[0] += (double)(first ? ro : rh[o]) * (double)va;
  }
}

// BiCGSTAB phase S: s = r - alpha v; t = S D^-1 s; <t,s>, <t,t>, <s,s>.
template <int K, typename TI, int C>
__device__ void phase_bi_s(const Inst& I, int row_lo, int row_hi, const TI* r, const TI* vnew, const Shared& sh,
                           double (&acc)[3]) {
  constexpr int RPP = kBlock / K;
  const int jr = threadIdx.x % K;
  const int kv = I.kv, vj = (kv == 1) ? 0 : jr;
  TI GA* sv = gp((TI*)I.s);  // own rows only
  TI GA* tv = gp((TI*)I.t);  // own rows only
  const TI GA* S = gp((const TI*)I.Si);
  const TI GA* dinv = gp((const TI*)I.dinvi);
  const TI al = to_ti<TI>(sh.alpha[jr]);
  const int GA* __restrict__ rp = gp(I.rp);
  const int GA* __restrict__ ci = gp(I.ci);
  for (int row = row_lo + (int)threadIdx.x / K; row < row_hi; row += RPP) {
    const size_t o = (size_t)row * K + jr;
    const TI so = ld<C>(r + o) - al * ld<C>(vnew + o);
    const TI ta = row_gather(rp[row], rp[row + 1], (TI)0, [&](int s) {
      const int c = ci[s];
      const size_t oc = (size_t)c * K + jr;
      return S[(size_t)s * kv + vj] * (dinv[(size_t)c * kv + vj] * (ld<C>(r + oc) - al * ld<C>(vnew + oc)));
    });
    sv[o] = so;
    tv[o] = ta;
    acc[0] += (double)ta * (double)so;
    acc[1] += (double)ta * (double)ta;
    acc[2] += (double)so * (double)so;
  }
}

// BiCGSTAB phase X: xi += D^-1 (alpha p + omega s); r = s - omega t; <rh,r>, <r,r>.
template <int K, typename TI, int C>
__device__ void phase_bi_x(const Inst& I, int row_lo, int row_hi, TI* r, const TI* p, const Shared& sh,
                           double (&acc)[2]) {
  constexpr int RPP = kBlock / K;
  const int jr = threadIdx.x % K;
  const int kv = I.kv, vj = (kv == 1) ? 0 : jr;
  TI* xi = (TI*)I.xi;
  const TI GA* rh = gp((const TI*)I.rh);
  const TI GA* sv = gp((const TI*)I.s);
  const TI GA* tv = gp((const TI*)I.t);
  const TI GA* dinv = gp((const TI*)I.dinvi);
  const TI al = to_ti<TI>(sh.alpha[jr]);
  const TI om = to_ti<TI>(sh.omega[jr]);
  for (int row = row_lo + (int)threadIdx.x / K; row < row_hi; row += RPP) {
    const size_t o = (size_t)row * K + jr;
    const TI d = dinv[(size_t)row * kv + vj];
    st<C>(xi + o, ld<C>(xi + o) + d * (al * ld<C>(p + o) + om * sv[o]));
    const TI rr = sv[o] - om * tv[o];
    st<C>(r + o, rr);
    acc[0] += (double)rh[o] * (double)rr;
    acc[1] += (double)rr * (double)rr;
  }
}

// ---------------------------------------------------------------------------
// LDS row cache.  Rows are owned by the same lane for the whole launch, and the
// inner-precision matrix and preconditioner do not change during it, so each
// lane's rows are staged once per instance as (column, S_ij * D^-1_jj) pairs.
// The Krylov gathers then have one dependent memory level (the vector loads)
// instead of three (row pointer -> column index -> vector).  Unused entries
// point at the row itself with weight 0, so the unrolled gather is
// branch-free and every address is valid (RowCache, defined with the FP64
// residual pass, which also takes its column indices).
// ---------------------------------------------------------------------------

// Fill the cache for this workgroup's rows.  Whether the team caches is
// decided from team-wide quantities -- the rows every workgroup is dealt and
// the instance's longest row -- so all workgroups of a team agree: with the
// cache the pipelined CG starts inside the update phase, one barrier fewer,
// and a team whose workgroups disagreed would wait at different barriers.
template <int K, typename TI>
__device__ bool fill_cache(const Inst& I, int row_lo, int row_hi, int rows_per_wg, RowCache<TI>& RC) {
  constexpr int RPP = kBlock / K;
  const int jr = threadIdx.x % K;
  const int kv = I.kv, vj = (kv == 1) ? 0 : jr;
  const TI GA* S = gp((const TI*)I.Si);
  const TI GA* dinv = gp((const TI*)I.dinvi);
  const int GA* __restrict__ rp = gp(I.rp);
  const int GA* __restrict__ ci = gp(I.ci);
  const bool ok = (rows_per_wg + RPP - 1) / RPP <= kCacheP && I.max_row <= kCacheE;
  if (ok) {
    for (int q = 0; q < kCacheP; q++) {
      const int row = row_lo + (int)threadIdx.x / K + q * RPP;
      int lo = 0, hi = 0;
      if (row < row_hi) {
        lo = rp[row];
        hi = rp[row + 1];
      }
      for (int e = 0; e < kCacheE; e++) {
        const int s = lo + e;
        const bool v = s < hi;
        const int c = v ? ci[s] : (row < row_hi ? row : 0);
        RC.col[q][e][threadIdx.x] = c;
        RC.val[q][e][threadIdx.x] = v ? S[(size_t)s * kv + vj] * dinv[(size_t)c * kv + vj] : (TI)0;
      }
    }
  }
  __syncthreads();
  return ok;
}

// Pipelined CG start from the cache (same arithmetic as phase_pcg_init).
template <int K, typename TI, int C>
__device__ void phase_pcg_init_c(const Inst& I, int row_lo, int row_hi, const TI* rc, TI* wc,
                                 const RowCache<TI>& RC, double (&acc)[2]) {
  constexpr int RPP = kBlock / K;
  const int jr = threadIdx.x % K;
  const int kv = I.kv, vj = (kv == 1) ? 0 : jr;
  TI* xi = (TI*)I.xi;
  const TI GA* dinv = gp((const TI*)I.dinvi);
  for (int q = 0; q < kCacheP; q++) {
    const int row = row_lo + (int)threadIdx.x / K + q * RPP;
    if (row >= row_hi) break;
    const size_t o = (size_t)row * K + jr;
    const double xo = ld<C>(I.x + o), xio = (double)ld<C>(xi + o);
    const TI ro = ld<C>(rc + o);
    const TI uo = dinv[(size_t)row * kv + vj] * ro;  // loaded before the stores below
    TI g[kCacheE];
#pragma unroll
    for (int e = 0; e < kCacheE; e++) g[e] = ld<C>(rc + (size_t)RC.col[q][e][threadIdx.x] * K + jr);
    TI wa = (TI)0;
#pragma unroll
    for (int e = 0; e < kCacheE; e++) wa += RC.val[q][e][threadIdx.x] * g[e];
    st<C>(I.x + o, xo + xio);
    st<C>(xi + o, (TI)0);
    st<C>(wc + o, wa);
    acc[0] += (double)ro * (double)uo;
    acc[1] += (double)uo * (double)wa;
  }
}

// One Chronopoulos-Gear iteration from the cache (same arithmetic as phase_pcg_iter).
template <int K, typename TI, int C>
__device__ void phase_pcg_iter_c(const Inst& I, int row_lo, int row_hi, const TI* rc, const TI* wc, const TI* sc,
                                 TI* rn, TI* wn, TI* sn, bool first, const Shared& sh, const RowCache<TI>& RC,
                                 double (&acc)[3]) {
  constexpr int RPP = kBlock / K;
  const int jr = threadIdx.x % K;
  const int kv = I.kv, vj = (kv == 1) ? 0 : jr;
  TI* xi = (TI*)I.xi;
  TI GA* p = gp((TI*)I.p0);  // own rows only
  const TI GA* dinv = gp((const TI*)I.dinvi);
  const TI al = to_ti<TI>(sh.alpha[jr]);
  const TI be = first ? (TI)0 : to_ti<TI>(sh.beta[jr]);
  for (int q = 0; q < kCacheP; q++) {
    const int row = row_lo + (int)threadIdx.x / K + q * RPP;
    if (row >= row_hi) break;
    const size_t o = (size_t)row * K + jr;
    const TI ro = ld<C>(rc + o);
    const TI wo = ld<C>(wc + o);
    const TI so = first ? (TI)0 : ld<C>(sc + o);
    const TI po = first ? (TI)0 : p[o];
    const TI xio = ld<C>(xi + o);
    TI gr[kCacheE], gw[kCacheE], gs[kCacheE];
#pragma unroll
    for (int e = 0; e < kCacheE; e++) {
      const size_t oc = (size_t)RC.col[q][e][threadIdx.x] * K + jr;
      gr[e] = ld<C>(rc + oc);
      gw[e] = ld<C>(wc + oc);
      gs[e] = first ? (TI)0 : ld<C>(sc + oc);
    }
    TI wa = (TI)0;
#pragma unroll
    for (int e = 0; e < kCacheE; e++)
      wa += RC.val[q][e][threadIdx.x] * cg_r<TI>(gr[e], cg_s<TI>(gw[e], gs[e], be, first), al);
    const TI d = dinv[(size_t)row * kv + vj];
    const TI uo = d * ro;
    const TI pn = first ? uo : (uo + be * po);
    const TI sv = cg_s<TI>(wo, so, be, first);
    const TI rr = cg_r<TI>(ro, sv, al);
    p[o] = pn;
    st<C>(xi + o, xio + al * pn);
    st<C>(sn + o, sv);
    st<C>(rn + o, rr);
    st<C>(wn + o, wa);
    const TI un = d * rr;
    acc[0] += (double)rr * (double)un;
    acc[1] += (double)un * (double)wa;
    acc[2] += (double)rr * (double)rr;
  }
}

// BiCGSTAB phase P from the cache.
template <int K, typename TI, int C>
__device__ void phase_bi_p_c(const Inst& I, int row_lo, int row_hi, const TI* r, const TI* pold, TI* pnew,
                             const TI* vold, TI* vnew, bool first, const Shared& sh, const RowCache<TI>& RC,
                             double (&acc)[1]) {
  constexpr int RPP = kBlock / K;
  const int jr = threadIdx.x % K;
  TI GA* rh = gp((TI*)I.rh);  // own rows only
  TI* xi = (TI*)I.xi;
  const TI be = first ? (TI)0 : to_ti<TI>(sh.beta[jr]);
  const TI om = first ? (TI)0 : to_ti<TI>(sh.omega[jr]);
  for (int q = 0; q < kCacheP; q++) {
    const int row = row_lo + (int)threadIdx.x / K + q * RPP;
    if (row >= row_hi) break;
    const size_t o = (size_t)row * K + jr;
    const TI ro = ld<C>(r + o);
    if (first) {
      st<C>(I.x + o, ld<C>(I.x + o) + (double)ld<C>(xi + o));
      st<C>(xi + o, (TI)0);
      rh[o] = ro;
    }
    TI gr[kCacheE], gp_[kCacheE], gv[kCacheE];
#pragma unroll
    for (int e = 0; e < kCacheE; e++) {
      const size_t oc = (size_t)RC.col[q][e][threadIdx.x] * K + jr;
      gr[e] = ld<C>(r + oc);
      gp_[e] = first ? (TI)0 : ld<C>(pold + oc);
      gv[e] = first ? (TI)0 : ld<C>(vold + oc);
    }
    TI va = (TI)0;
#pragma unroll
    for (int e = 0; e < kCacheE; e++)
      va += RC.val[q][e][threadIdx.x] * (first ? gr[e] : (gr[e] + be * (gp_[e] - om * gv[e])));
    const TI pn = first ? ro : (ro + be * (ld<C>(pold + o) - om * ld<C>(vold + o)));
    st<C>(pnew + o, pn);
    st<C>(vnew + o, va);
    acc[0] += (double)(first ? ro : rh[o]) * (double)va;
  }
}

// BiCGSTAB phase S from the cache.
template <int K, typename TI, int C>
__device__ void phase_bi_s_c(const Inst& I, int row_lo, int row_hi, const TI* r, const TI* vnew, const Shared& sh,
                             const RowCache<TI>& RC, double (&acc)[3]) {
  constexpr int RPP = kBlock / K;
  const int jr = threadIdx.x % K;
  TI GA* sv = gp((TI*)I.s);  // own rows only
  TI GA* tv = gp((TI*)I.t);  // own rows only
  const TI al = to_ti<TI>(sh.alpha[jr]);
  for (int q = 0; q < kCacheP; q++) {
    const int row = row_lo + (int)threadIdx.x / K + q * RPP;
    if (row >= row_hi) break;
    const size_t o = (size_t)row * K + jr;
    const TI so = ld<C>(r + o) - al * ld<C>(vnew + o);
    TI gr[kCacheE], gv[kCacheE];
#pragma unroll
    for (int e = 0; e < kCacheE; e++) {
      const size_t oc = (size_t)RC.col[q][e][threadIdx.x] * K + jr;
      gr[e] = ld<C>(r + oc);
      gv[e] = ld<C>(vnew + oc);
    }
    TI ta = (TI)0;
#pragma unroll
    for (int e = 0; e < kCacheE; e++) ta += RC.val[q][e][threadIdx.x] * (gr[e] - al * gv[e]);
    sv[o] = so;
    tv[o] = ta;
    acc[0] += (double)ta * (double)so;
    acc[1] += (double)ta * (double)ta;
    acc[2] += (double)so * (double)so;
  }
}

// State update and the instability rule (CPUSolver.cpp:41-48, widened to
// flag NaN/Inf separately).
// With lvn >= 0 (heat, advection-diffusion, Gray-Scott), also the next
// step's initial guess of predictor order lvn into xnext, from this row's own
// history: the next right-hand side then gathers one vector for S x0.
template <int K, typename TI, int C>
__device__ void phase_update(const Inst& I, int row_lo, int row_hi, const Shared& sh, int* flags, int step,
                             int lvn = -1) {
  constexpr int RPP = kBlock / K;
  const int jr = threadIdx.x % K;
  TI* xi = (TI*)I.xi;
  int f = 0;
  double* traj = nullptr;  // trajectory record of this step, if any
  if (I.traj && (step + 1) % I.traj_every == 0)
    traj = I.traj + (size_t)((step + 1) / I.traj_every - 1) * I.n * K;
  const bool first_order = I.eq == EQ_HEAT || I.eq == EQ_ADVDIFF || I.eq == EQ_GS;
  for (int row = row_lo + (int)threadIdx.x / K; row < row_hi; row += RPP) {
    const size_t o = (size_t)row * K + jr;
    // every load of the row first: loads and stores share one completion
    // counter, so a load issued after a store also waits for the store
    double xf = ld<C>(I.x + o) + (double)ld<C>(xi + o);
    const double uo = ld<C>(I.u + o);
    const double ao = (first_order ? I.extrap >= 2 : I.eq == EQ_NEWMARK) ? ld<C>(I.a + o) : 0.0;
    const double vo = (first_order ? I.extrap >= 3 : (I.eq == EQ_WAVE || I.eq == EQ_NEWMARK)) ? ld<C>(I.v + o) : 0.0;
    const double ho = (first_order ? I.extrap >= 4 : (I.eq == EQ_NEWMARK && I.extrap >= 2)) ? ld<C>(I.h3 + o) : 0.0;
    if (sh.bb[jr] == 0.0) xf = 0.0;  // Eigen: rhs == 0 -> x = 0
    st<C>(I.x + o, xf);
    st<C>(xi + o, (TI)0);
    double c1 = xf, c2 = 0.0, up = xf;  // up: new primary variable
    switch (I.eq) {
      case EQ_HEAT:
      case EQ_ADVDIFF:
      case EQ_GS:
        if (I.extrap) {  // history for the predictor: x_{n-4} <- x_{n-3} <- x_{n-2} <- x_{n-1} <- x_n
          if (I.extrap >= 4) st<C>(I.h4 + o, ho);
          if (I.extrap >= 3) st<C>(I.h3 + o, vo);
          if (I.extrap >= 2) st<C>(I.v + o, ao);
          st<C>(I.a + o, uo);
          if (lvn >= 0 && I.xnext) {
            double w0, w1, w2, w3, w4;
            pred_weights(lvn, w0, w1, w2, w3, w4);
            gp(I.xnext)[o] = w0 * xf + w1 * uo + w2 * ao + w3 * vo + w4 * ho;
          }
        }
        st<C>(I.u + o, xf);
        break;
      case EQ_WAVE: {
        if (I.extrap) st<C>(I.a + o, vo);
        st<C>(I.v + o, xf);
        const double un = uo + xf * I.dt;  // CPUSolver.cpp:118
        st<C>(I.u + o, un);
        c2 = un;
        up = un;
      } break;
      case EQ_NEWMARK: {
        const double dt = I.dt, be = I.p2, ga = I.p3;
        const double ut = uo + dt * vo + (0.5 - be) * dt * dt * ao;
        const double un = ut + be * dt * dt * xf;
        const double vn = vo + dt * ((1.0 - ga) * ao + ga * xf);
        st<C>(I.u + o, un);
        st<C>(I.v + o, vn);
        if (I.extrap >= 2) st<C>(I.h4 + o, ho);  // acceleration history for the predictor
        if (I.extrap >= 1) st<C>(I.h3 + o, ao);
        st<C>(I.a + o, xf);
        c1 = un;
        c2 = vn;
        up = un;
      } break;
      default:
        break;
    }
    if (traj) gp(traj)[o] = up;
    if (c1 > 1e4 || c2 > 1e4) f |= ST_UNSTABLE;
    if (!isfinite(c1) || !isfinite(c2)) f |= ST_NONFINITE;
  }
  if (f) atomicOr(flags, f);
}

// The state update of step `step` and the right-hand side of step + 1 as one
// phase (heat, advection-diffusion, Gray-Scott: the new state is the
// solution).  Neighbours are gathered only through x + xi, which no
// workgroup writes in this phase; everything written belongs to the
// workgroup's own rows: the state and predictor history, b, r, and the next
// initial guess in xnext, which phase_install moves into x (xi = 0) after the
// barrier.  The history part of S x0 comes from hpre (phase_res or
// phase_hpre of this step); lvn is the next step's predictor order.  pre:
// the last residual pass of this step already formed b (in xnext) and the
// initial residual r (in hpre), so no neighbour state is gathered.  lvh > 0:
// also the history part of the initial guess after that into hvec.
// RCi (with pre; pipelined CG on cached rows): also the start of that CG,
// w = S D^-1 r into wq from the neighbours' r in hpre, so the sweep needs no
// start phase of its own; acc[3] then sums <D^-1 r, w>.
template <int K, typename TI, int C>
__device__ void phase_update_rhs(const Inst& I, int row_lo, int row_hi, TI* r, const Shared& sh, int* flags, int step,
                                 int lvn, int lvh, bool pre, const RowCache<TI>* RCi, TI* wq, double (&acc)[4]) {
  constexpr int RPP = kBlock / K;
  const int jr = threadIdx.x % K;
  const int kv = I.kv, vj = (kv == 1) ? 0 : jr;
  const TI* xi = (const TI*)I.xi;
  const TI GA* dinv = gp((const TI*)I.dinvi);
  const double inv_dt = 1.0 / I.dt;
  const bool warm = I.warm != 0;
  double w0, w1, w2, w3, w4;
  pred_weights(lvn, w0, w1, w2, w3, w4);
  double h0, h1, h2, h3, h4;  // lvh: the step after this one (its history combination goes to hvec)
  pred_weights(lvh, h0, h1, h2, h3, h4);
  const int GA* __restrict__ rp = gp(I.rp);
  const int GA* __restrict__ ci = gp(I.ci);
  const double GA* M64 = gp(I.M64);
  const double GA* S64 = gp(I.S64);
  int f = 0;
  double* traj = nullptr;  // trajectory record of this step, if any
  if (I.traj && (step + 1) % I.traj_every == 0)
    traj = I.traj + (size_t)((step + 1) / I.traj_every - 1) * I.n * K;
  // this step's solution of column j at offset idx (Eigen: rhs == 0 -> x = 0)
  auto xn_at = [&](size_t idx, int j) -> double {
    return sh.bb[j] == 0.0 ? 0.0 : ld<C>(I.x + idx) + (double)ld<C>(xi + idx);
  };
  int q = 0;  // pass over the workgroup's rows (row cache slot)
  for (int row = row_lo + (int)threadIdx.x / K; row < row_hi; row += RPP, q++) {
    const size_t o = (size_t)row * K + jr;
    // every load of the row first: loads and stores share one completion
    // counter, so a load issued after a store also waits for the store
    const double xn = xn_at(o, jr);
    const double uo = ld<C>(I.u + o);
    const double ao = I.extrap >= 2 ? ld<C>(I.a + o) : 0.0;
    const double vo = I.extrap >= 3 ? ld<C>(I.v + o) : 0.0;
    const double ho = I.extrap >= 4 ? ld<C>(I.h3 + o) : 0.0;
    // ---- next right-hand side: b = (M/dt) x_n (+ Gray-Scott reaction), r = b - S x0
    double b, r64;
    TI wa = (TI)0;  // S D^-1 r (with RCi)
    if (pre) {      // both formed by the last residual pass of this step
      b = gp(I.xnext)[o];
      r64 = ld<C>(I.hpre + o);
      if (RCi) {
        TI g[kCacheE];
#pragma unroll
        for (int e = 0; e < kCacheE; e++)
          g[e] = to_ti<TI>(ld<C>(I.hpre + (size_t)RCi->col[q][e][threadIdx.x] * K + jr));
#pragma unroll
        for (int e = 0; e < kCacheE; e++) wa += RCi->val[q][e][threadIdx.x] * g[e];
      }
    } else {
      const D2 g = row_gather(rp[row], rp[row + 1], D2{0.0, 0.0}, [&](int s) {
        const double xc = xn_at((size_t)ci[s] * K + jr, jr);
        return D2{M64[s] * xc, warm ? S64[(size_t)s * kv + vj] * xc : 0.0};
      });
      b = g.a * inv_dt;
      if (I.eq == EQ_GS) {  // CPUSolver.cpp:134,137 with the new state
        const double uu = xn_at((size_t)row * K, 0), vv = xn_at((size_t)row * K + 1, 1);
        const double uvv = uu * (vv * vv);
        const double fr = (jr == 0) ? (-uvv + I.p3 * (1.0 - uu)) : (uvv - (I.p3 + I.p4) * vv);
        const double w = I.weighting ? gp(I.lumped)[row] : 1.0;
        b += w * fr;
      }
      r64 = b - (warm ? w0 * g.b + (lvn > 0 ? gp(I.hpre)[o] : 0.0) : 0.0);
    }
    const TI dj = dinv[(size_t)row * kv + vj];
    const double x0 = warm ? w0 * xn + w1 * uo + w2 * ao + w3 * vo + w4 * ho : 0.0;
    const TI rr = to_ti<TI>(r64);
    // ---- update: state and history x_{n-3} <- x_{n-2} <- x_{n-1} <- x_n (own row)
    if (I.extrap >= 4) st<C>(I.h4 + o, ho);
    if (I.extrap >= 3) st<C>(I.h3 + o, vo);
    if (I.extrap >= 2) st<C>(I.v + o, ao);
    if (I.extrap >= 1) st<C>(I.a + o, uo);
    st<C>(I.u + o, xn);
    // the shifted history is x_n, x_{n-1}, x_{n-2}, x_{n-3}
    if (lvh > 0) st<C>(I.hvec + o, h1 * xn + h2 * uo + h3 * ao + h4 * vo);
    if (traj) gp(traj)[o] = xn;
    if (xn > 1e4) f |= ST_UNSTABLE;
    if (!isfinite(xn)) f |= ST_NONFINITE;
    gp(I.b)[o] = b;
    gp(I.xnext)[o] = x0;
    st<C>(r + o, rr);
    if (RCi) st<C>(wq + o, wa);
    acc[0] += b * b;
    acc[1] += r64 * r64;
    const TI uz = dj * rr;
    acc[2] += (I.krylov == 1) ? (double)rr * (double)rr : (double)rr * (double)uz;
    if (RCi) acc[3] += (double)uz * (double)wa;
  }
  if (f) atomicOr(flags, f);
}

// Own rows: x = xnext, xi = 0 (the initial guess phase_update_rhs prepared).
template <int K, typename TI, int C>
__device__ void phase_install(const Inst& I, int row_lo, int row_hi) {
  constexpr int RPP = kBlock / K;
  const int jr = threadIdx.x % K;
  TI* xi = (TI*)I.xi;
  for (int row = row_lo + (int)threadIdx.x / K; row < row_hi; row += RPP) {
    const size_t o = (size_t)row * K + jr;
    st<C>(I.x + o, gp(I.xnext)[o]);
    st<C>(xi + o, (TI)0);
  }
}

// The history part of the next step's initial residual (as phase_res does)
// for a step that ran no refinement sweep.
template <int K, typename TI, int C>
__device__ void phase_hpre(const Inst& I, int row_lo, int row_hi, int lvn, bool hready) {
  constexpr int RPP = kBlock / K;
  const int jr = threadIdx.x % K;
  const int kv = I.kv, vj = (kv == 1) ? 0 : jr;
  double w0, w1, w2, w3, w4;
  pred_weights(lvn, w0, w1, w2, w3, w4);
  const int GA* __restrict__ rp = gp(I.rp);
  const int GA* __restrict__ ci = gp(I.ci);
  const double GA* S64 = gp(I.S64);
  for (int row = row_lo + (int)threadIdx.x / K; row < row_hi; row += RPP) {
    const size_t o = (size_t)row * K + jr;
    if (hready) {
      gp(I.hpre)[o] = row_gather(rp[row], rp[row + 1], 0.0, [&](int s) {
        return S64[(size_t)s * kv + vj] * ld<C>(I.hvec + (size_t)ci[s] * K + jr);
      });
      continue;
    }
    gp(I.hpre)[o] = row_gather<kUWide>(rp[row], rp[row + 1], 0.0, [&](int s) {
      const size_t oc = (size_t)ci[s] * K + jr;
      const double hv = w1 * ld<C>(I.u + oc) + (lvn >= 2 ? w2 * ld<C>(I.a + oc) : 0.0) +
                        (lvn >= 3 ? w3 * ld<C>(I.v + oc) : 0.0) + (lvn >= 4 ? w4 * ld<C>(I.h3 + oc) : 0.0);
      return S64[(size_t)s * kv + vj] * hv;
    });
  }
}

// ---------------------------------------------------------------------------
// Krylov drivers (uniform control flow across the team)
// ---------------------------------------------------------------------------

// Pipelined (Chronopoulos-Gear) PCG on the current sweep; r starts in I.r.
// started: the update phase already formed w = S D^-1 r and the start
// scalars (x holds the initial guess, xi = 0).
template <int K, typename TI, int C>
__device__ bool krylov_pcg(const Inst& I, const Team& T, int rank, unsigned& e, Shared& sh, int row_lo, int row_hi,
                           int& its, const RowCache<TI>& RC, bool cached, bool started) {
  TI* rb[2] = {(TI*)I.r, (TI*)I.z};
  TI* wb[2] = {(TI*)I.q, (TI*)I.rh};
  TI* sb[2] = {(TI*)I.s, (TI*)I.t};
  int par = 0;
  if (!started) {
    {
      double acc[2] = {0.0, 0.0};
      if (cached) phase_pcg_init_c<K, TI, C>(I, row_lo, row_hi, rb[0], wb[0], RC, acc);
      else phase_pcg_init<K, TI, C>(I, row_lo, row_hi, rb[0], wb[0], acc);
      if (!team_reduce<K, 2, C>(acc, T, rank, e, sh)) return false;
    }
    if (threadIdx.x < K) {
      const int j = threadIdx.x;
      const double g = sh.red[j], d = sh.red[K + j];
      sh.gamma[j] = g;
      sh.alpha[j] = (sh.active[j] && d != 0.0) ? g / d : 0.0;
      sh.beta[j] = 0.0;
    }
    __syncthreads();
  }
  int it_s = 0;
  bool first = true;
  while (true) {
    {
      double acc[3] = {0.0, 0.0, 0.0};
      if (cached)
        phase_pcg_iter_c<K, TI, C>(I, row_lo, row_hi, rb[par], wb[par], sb[par], rb[par ^ 1], wb[par ^ 1],
                                   sb[par ^ 1], first, sh, RC, acc);
      else
        phase_pcg_iter<K, TI, C>(I, row_lo, row_hi, rb[par], wb[par], sb[par], rb[par ^ 1], wb[par ^ 1], sb[par ^ 1],
                                 first, sh, acc);
      if (!team_reduce<K, 3, C>(acc, T, rank, e, sh)) return false;
    }
    par ^= 1;
    first = false;
    it_s++;
    its++;
    // the recurrences of the K columns and the stop vote in wave 0 (K <= 64
    // columns, one lane each), then one workgroup barrier
    if (threadIdx.x < kWave) {
      bool done = true;
      if (threadIdx.x < K) {
        const int j = threadIdx.x;
        const double g = sh.red[j], d = sh.red[K + j], rr = sh.red[2 * K + j];
        const double gold = sh.gamma[j], aold = sh.alpha[j];
        done = !sh.active[j] || rr < sh.ithr[j] || !(rr == rr);
        double be = 0.0, al = 0.0;
        if (!done && gold != 0.0 && aold != 0.0) {
          be = g / gold;
          const double den = d - be * g / aold;
          al = (den != 0.0) ? g / den : 0.0;
        }
        sh.beta[j] = be;
        sh.alpha[j] = al;
        sh.gamma[j] = g;
      }
      const bool alldone = __all(done);
      if (threadIdx.x == 0) sh.uni = alldone || it_s >= I.max_inner || its >= I.max_iters;
    }
    __syncthreads();
    if (sh.uni) break;
  }
  return true;
}

template <int K, typename TI, int C>
__device__ bool krylov_cg_classic(const Inst& I, const Team& T, int rank, unsigned& e, Shared& sh, int row_lo,
                                  int row_hi, int& its) {
  TI* r = (TI*)I.r;
  TI* pa = (TI*)I.p0;
  TI* pb = (TI*)I.p1b;
  int it_s = 0;
  bool first = true;
  while (true) {
    {
      double acc[1] = {0.0};
      phase_cg_a<K, TI, C>(I, row_lo, row_hi, r, pa, pb, first, sh, acc);
      if (!team_reduce<K, 1, C>(acc, T, rank, e, sh)) return false;
    }
    if (threadIdx.x < K) {
      const double pq = sh.red[threadIdx.x];
      sh.alpha[threadIdx.x] = (sh.active[threadIdx.x] && pq != 0.0) ? sh.gamma[threadIdx.x] / pq : 0.0;
    }
    __syncthreads();
    {
      double acc[2] = {0.0, 0.0};
      phase_cg_b<K, TI, C>(I, row_lo, row_hi, r, pb, sh, acc);
      if (!team_reduce<K, 2, C>(acc, T, rank, e, sh)) return false;
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
      sh.uni = alldone || it_s >= I.max_inner || its >= I.max_iters;
    }
    __syncthreads();
    if (sh.uni) break;
    first = false;
    TI* tmp = pa;
    pa = pb;
    pb = tmp;
  }
  return true;
}

template <int K, typename TI, int C>
__device__ bool krylov_bicgstab(const Inst& I, const Team& T, int rank, unsigned& e, Shared& sh, int row_lo,
                                int row_hi, int& its, const RowCache<TI>& RC, bool cached) {
  TI* r = (TI*)I.r;
  TI* pa = (TI*)I.p0;
  TI* pb = (TI*)I.p1b;
  TI* va = (TI*)I.v0;
  TI* vb = (TI*)I.v1;
  int it_s = 0;
  bool first = true;
  while (true) {
    if (threadIdx.x == 0) sh.bkd = 0;
    {
      double acc[1] = {0.0};
      if (cached) phase_bi_p_c<K, TI, C>(I, row_lo, row_hi, r, pa, pb, va, vb, first, sh, RC, acc);
      else phase_bi_p<K, TI, C>(I, row_lo, row_hi, r, pa, pb, va, vb, first, sh, acc);
      if (!team_reduce<K, 1, C>(acc, T, rank, e, sh)) return false;
    }
    if (threadIdx.x < K) {
      const double rv = sh.red[threadIdx.x];
      const double a = (sh.active[threadIdx.x] && rv != 0.0) ? sh.gamma[threadIdx.x] / rv : 0.0;
      // near-breakdown: a step the inner precision cannot represent restarts the sweep
      const bool fine = fabs(a) < 1e30;
      sh.alpha[threadIdx.x] = fine ? a : 0.0;
      if (!fine) sh.bkd = 1;
    }
    __syncthreads();
    {
      double acc[3] = {0.0, 0.0, 0.0};
      if (cached) phase_bi_s_c<K, TI, C>(I, row_lo, row_hi, r, vb, sh, RC, acc);
      else phase_bi_s<K, TI, C>(I, row_lo, row_hi, r, vb, sh, acc);
      if (!team_reduce<K, 3, C>(acc, T, rank, e, sh)) return false;
    }
    if (threadIdx.x < K) {
      const double ts = sh.red[threadIdx.x], tt = sh.red[K + threadIdx.x], ss = sh.red[2 * K + threadIdx.x];
      const bool act = sh.active[threadIdx.x];
      const double om = (act && tt > 0.0 && !(ss < sh.ithr[threadIdx.x])) ? ts / tt : 0.0;
      const bool fine = fabs(om) < 1e30;
      sh.omega[threadIdx.x] = fine ? om : 0.0;
      if (!fine) sh.bkd = 1;
    }
    __syncthreads();
    {
      double acc[2] = {0.0, 0.0};
      phase_bi_x<K, TI, C>(I, row_lo, row_hi, r, pb, sh, acc);
      if (!team_reduce<K, 2, C>(acc, T, rank, e, sh)) return false;
    }
    it_s++;
    its++;
    if (threadIdx.x == 0) {
      int alldone = 1, brk = sh.bkd;
      for (int j = 0; j < K; j++) {
        const double rho_new = sh.red[j], rr = sh.red[K + j];
        const bool act = sh.active[j];
        const bool done = !act || rr < sh.ithr[j] || !(rr == rr);
        alldone &= done;
        if (!done) {
          const double om = sh.omega[j], rho = sh.gamma[j];
          const double be = (om == 0.0 || rho == 0.0) ? 0.0 : (rho_new / rho) * (sh.alpha[j] / om);
          if (be == 0.0 || !(fabs(be) < 1e30)) brk = 1;  // breakdown: restart the sweep
          else sh.beta[j] = be;
        }
        sh.gamma[j] = rho_new;
      }
      sh.uni = alldone || brk || it_s >= I.max_inner || its >= I.max_iters;
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
  return true;
}

// All steps of one instance by one team.  KR: 0 pipelined CG, 1 BiCGSTAB,
// 2 classic CG (one Krylov method per kernel keeps register pressure down).
// Returns false when the launch aborts.
template <int K, typename TI, int KR, int C>
__device__ bool run_instance(const Inst& I, const Team& T, int rank, unsigned& e, Shared& sh, RowCache<TI>& RC,
                             int nsteps) {
  const int n = I.n;
  const int rows_per_wg = (n + T.size - 1) / T.size;
  const int row_lo = min(n, rank * rows_per_wg);
  const int row_hi = min(n, row_lo + rows_per_wg);
  bool cached = false;
  if constexpr (KR != 2) cached = fill_cache<K, TI>(I, row_lo, row_hi, rows_per_wg, RC);
  long long its_total = 0, sweeps_total = 0;
  int status = 0;
  TI* r0 = (TI*)I.r;
  // first-order equations: the update of one step and the right-hand side of
  // the next share a phase (phase_update_rhs); have_rhs: this step's
  // right-hand side is already formed and reduced (sh.red)
  // (each barrier has one call site: inlined copies of the wait loops spill)
  const bool fuse_eq = (I.eq == EQ_HEAT || I.eq == EQ_ADVDIFF || I.eq == EQ_GS) && I.xnext && I.hpre;
  bool have_rhs = false;  // the previous step's update is pending: it runs with this step's right-hand side
  bool rhs_pre = false;   // ... and the previous step's last residual pass formed its products
  for (int step = 0; step < nsteps; step++) {
    const bool fuse_next = fuse_eq && step + 1 < nsteps;
    const int lvn = (fuse_next && I.warm) ? min(I.extrap, I.hist + step + 1) : 0;  // next step's predictor order
    // ---- R: right-hand side and initial residual (after the previous step's
    // update, in the same phase, when that was deferred)
    // the pipelined CG of the first sweep starts in the update phase (from
    // the initial residuals the last residual pass stored)
    const bool started = KR == 0 && have_rhs && rhs_pre && cached;
    {
      double acc[4] = {0.0, 0.0, 0.0, 0.0};
      if (have_rhs)
        phase_update_rhs<K, TI, C>(I, row_lo, row_hi, r0, sh, I.status, step - 1,
                                   I.warm ? min(I.extrap, I.hist + step) : 0, I.hvec ? lvn : 0, rhs_pre,
                                   started ? &RC : nullptr, (TI*)I.q, acc);
      else
        phase_rhs<K, TI, C>(I, row_lo, row_hi, r0, step, acc);
      if (!team_reduce<K, 4, C>(acc, T, rank, e, sh)) return false;
    }
    if (threadIdx.x < K) {
      const int j = threadIdx.x;
      const double bb = sh.red[j], rr = sh.red[K + j], g = sh.red[2 * K + j];
      sh.bb[j] = bb;
      sh.thr[j] = fmax(I.tol2 * bb, DBL_MIN);
      const bool act = (bb != 0.0) && (rr >= sh.thr[j]);
      sh.active[j] = act;
      sh.gamma[j] = g;
      sh.ithr[j] = rr;  // residual^2 at sweep start
      sh.reset[j] = 0;
      if (started) {  // as krylov_pcg after its start phase
        const double d = sh.red[3 * K + j];
        sh.alpha[j] = (act && d != 0.0) ? g / d : 0.0;
        sh.beta[j] = 0.0;
      }
    }
    __syncthreads();
    // own rows only: the phases up to the next barrier read no other row's x
    if (have_rhs) phase_install<K, TI, C>(I, row_lo, row_hi);
    const bool hready = have_rhs && I.hvec;  // the update above formed hvec
    int sweeps = 0, its = 0;
    // Whether a sweep follows, and its inner threshold, are decided in wave 0
    // (one lane per column) together with the state they depend on: here for
    // the first sweep, after the last residual pass of a sweep for the next.
    auto next_sweep = [&](int sweeps_done) {
      const bool act = threadIdx.x < K && sh.active[threadIdx.x];
      const bool go = __any(act) && sweeps_done < I.max_sweeps && its < I.max_iters;
      if (go && threadIdx.x < K) {
        // FP32 sweeps reduce the residual by inner_tol; an FP64 inner
        // solve goes straight to the final tolerance.
        const double rs = sh.ithr[threadIdx.x];
        sh.ithr[threadIdx.x] = (sizeof(TI) == 4) ? fmax(I.inner_tol2 * rs, sh.thr[threadIdx.x])
                                                 : sh.thr[threadIdx.x];
      }
      if (threadIdx.x == 0) sh.uni = go;
    };
    if (threadIdx.x < kWave) next_sweep(0);
    __syncthreads();
    while (sh.uni) {
      bool ok;
      if constexpr (KR == 1) ok = krylov_bicgstab<K, TI, C>(I, T, rank, e, sh, row_lo, row_hi, its, RC, cached);
      else if constexpr (KR == 2) ok = krylov_cg_classic<K, TI, C>(I, T, rank, e, sh, row_lo, row_hi, its);
      else ok = krylov_pcg<K, TI, C>(I, T, rank, e, sh, row_lo, row_hi, its, RC, cached, started && sweeps == 0);
      if (!ok) return false;
      // ---- Res: FP64 residual of x + xi.  A column whose residual is not
      // finite (the sweep broke down in the inner precision) drops its
      // correction and takes the residual of x alone.
      for (int pass = 0; pass < 2; pass++) {
        double acc[2] = {0.0, 0.0};
        phase_res<K, TI, C>(I, row_lo, row_hi, r0, sh, acc, lvn, fuse_next ? I.hpre : nullptr, hready,
                            cached ? &RC : nullptr);
        if (!team_reduce<K, 2, C>(acc, T, rank, e, sh)) return false;
        if (threadIdx.x < kWave) {
          const int j = threadIdx.x;
          bool redo_j = false;
          if (j < K) {
            const bool bad = !(sh.red[j] <= DBL_MAX);  // NaN or infinite
            redo_j = bad && !sh.reset[j];
            sh.reset[j] = bad;
          }
          const bool redo = __any(redo_j) && pass == 0;
          if (!redo) {  // the sweep's last pass: its state, then the next sweep's decision
            if (j < K) {
              sh.reset[j] = 0;
              const double rr = sh.red[j];
              sh.active[j] = (sh.bb[j] != 0.0) && (rr >= sh.thr[j]) && (rr == rr);
              sh.gamma[j] = sh.red[K + j];
              sh.ithr[j] = rr;
            }
            next_sweep(sweeps + 1);
          }
          if (j == 0) sh.redo = redo;
        }
        __syncthreads();
        if (!sh.redo) break;
      }
      sweeps++;
    }
    // ---- U: state update
    int notconv = 0;
    for (int j = 0; j < K; j++) notconv |= sh.active[j];
    if (notconv) status |= ST_NOT_CONVERGED;
    if (rank == 0 && threadIdx.x < K) {
      const double bb = sh.bb[threadIdx.x];
      I.last_res[threadIdx.x] = bb > 0.0 ? sqrt(sh.ithr[threadIdx.x] / bb) : 0.0;
    }
    its_total += its;
    sweeps_total += sweeps;
    bool barrier;
    if (fuse_next) {
      // ---- U deferred: it runs with the next step's right-hand side.  With
      // no FP64 residual this step, form the history term here and share the
      // installed x before neighbours gather it.
      if (sweeps == 0 && lvn > 0) phase_hpre<K, TI, C>(I, row_lo, row_hi, lvn, hready);
      barrier = sweeps == 0 && have_rhs;
      have_rhs = true;
      rhs_pre = sweeps > 0;
    } else {
      phase_update<K, TI, C>(I, row_lo, row_hi, sh, I.status, step);
      barrier = true;  // the next step (or the next instance of this team) gathers the new state
      have_rhs = false;
    }
    if (barrier && !team_barrier<K, C>(T, rank, e, sh)) return false;
  }
  if (rank == 0 && threadIdx.x == 0) {
    I.counters[0] += its_total;
    I.counters[1] += sweeps_total;
    I.counters[2] += nsteps;
    if (status) atomicOr(I.status, status);
    if (I.hist_dev) *I.hist_dev = min(4, I.hist_raw + nsteps);
  }
  return true;
}

// ---------------------------------------------------------------------------
// The kernel
// ---------------------------------------------------------------------------

template <int K, typename TI, int KR, int C>
__global__ void __launch_bounds__(kBlock) k_step(const Team* __restrict__ teams, const int* __restrict__ wg_team,
                                                const Inst* __restrict__ insts, const int* __restrict__ team_inst,
                                                int nsteps) {
  __shared__ Shared sh;
  __shared__ RowCache<TI> RC;
  // The instance record lives in LDS while it runs: its fields are read in
  // every phase, and a global load of one would wait for every store the
  // workgroup still has in flight (loads and stores share one counter).
  __shared__ Inst sI;
  static_assert(sizeof(Inst) % sizeof(int) == 0, "Inst copied as ints");
  const int tid_team = wg_team[blockIdx.x];
  if (tid_team < 0) return;
  const Team T = teams[tid_team];
  const int l = (int)blockIdx.x - T.first_wg;
  const int rank = xcd_rank(l, T.size);
  unsigned e = 0;  // barrier epoch, identical across the team
  for (int qi = T.inst_begin; qi < T.inst_end; qi++) {
    __syncthreads();  // the previous instance's readers are done
    const int* src = (const int*)(insts + team_inst[qi]);
    for (int w = threadIdx.x; w < (int)(sizeof(Inst) / sizeof(int)); w += kBlock) ((int*)&sI)[w] = src[w];
    __syncthreads();
    // predictor history held in device memory: every workgroup of the team
    // reads it here, before the instance's first barrier; rank 0 stores the
    // new count after the last one (run_instance)
    if (threadIdx.x == 0 && sI.hist_dev) {
      const int held = *sI.hist_dev;
      sI.hist_raw = held;
      sI.hist = min(held, sI.extrap);
    }
    __syncthreads();
    if (sI.n <= 0) continue;
    if (!run_instance<K, TI, KR, C>(sI, T, rank, e, sh, RC, nsteps)) return;
  }
}

// Instantiation for (columns, inner precision, Krylov method) in exchange mode C.
template <int C, int K, typename TI>
const void* pick_kr(int kr) {
  switch (kr) {
    case 0: return (const void*)k_step<K, TI, 0, C>;
    case 1: return (const void*)k_step<K, TI, 1, C>;
    case 2: return (const void*)k_step<K, TI, 2, C>;
    default: return nullptr;
  }
}

template <int C>
const void* pick_for_mode(int k, int fp32, int kr) {
  switch (k) {
    case 1: return fp32 ? pick_kr<C, 1, float>(kr) : pick_kr<C, 1, double>(kr);
    case 2: return fp32 ? pick_kr<C, 2, float>(kr) : pick_kr<C, 2, double>(kr);
    case 4: return fp32 ? pick_kr<C, 4, float>(kr) : pick_kr<C, 4, double>(kr);
    case 16: return fp32 ? pick_kr<C, 16, float>(kr) : pick_kr<C, 16, double>(kr);
    default: return nullptr;
  }
}

}  // namespace
}  // namespace femhip
