// stream.hip -- the streaming engine: the stepping kernel's phases as
// separate full-occupancy launches over every row block of every instance.
//
// The persistent kernel (step_kernel.h) keeps one workgroup pair per CU alive
// for the whole run; that suits latency-bound instances whose rows fit the
// LDS row cache, but large instances then walk hundreds of rows per lane per
// phase with only 8 waves per CU in flight.  This engine runs each phase
// (right-hand side, Krylov phases, FP64 residual, state update) as its own
// kernel over a table of row blocks of all instances: kernel boundaries are
// the barriers, every kernel carries only its own phase's registers, and the
// GPU holds several times more rows in flight.  Each block writes its partial
// dot products; a reduction kernel per phase sums them in block order
// (deterministic) and applies the same scalar recurrences as the persistent
// drivers to a per-instance scalar record in device memory, which the next
// phase reads.  All instances iterate in lockstep; instances whose sweep or
// solve has finished skip the launches.  The host reads one "all finished"
// flag per iteration to stop the loop.
//
// The phase bodies are the persistent kernel's own (plain exchange mode), so
// both engines compute the same arithmetic.
#include "step_kernel.h"

#include <algorithm>

namespace femhip {

namespace {

// Per-instance scalars of the running solve (the persistent kernel's Shared).
struct SState {
  double gamma[kMaxK], alpha[kMaxK], beta[kMaxK], omega[kMaxK];
  double thr[kMaxK], ithr[kMaxK], bb[kMaxK];
  int active[kMaxK], reset[kMaxK];
  int sweep_done, solve_done, redo, bkd;
  int it_s, its, sweeps, status;
  long long its_total, sweeps_total;
};

enum Phase : int {
  PH_RHS = 0, PH_RES, PH_PCG_INIT, PH_PCG_ITER, PH_CG_A, PH_CG_B, PH_BI_P, PH_BI_S, PH_BI_X, PH_UPD
};

template <int PH>
constexpr int nsums() {
  return PH == PH_RHS || PH == PH_PCG_ITER || PH == PH_BI_S ? 3
         : PH == PH_RES || PH == PH_PCG_INIT || PH == PH_CG_B || PH == PH_BI_X ? 2
         : PH == PH_UPD ? 0 : 1;
}

struct SArgs {
  const Inst* insts;
  SState* ss;
  const int* blk_inst;  // [nblocks]
  const int* blk_row;   // [nblocks] first row
  int rows;             // rows per block
  int par;              // buffer parity of the current iteration
  int first;            // first iteration of the sweep
  int step;             // step index inside the launch sequence (trajectory records)
  int pass;             // FP64 residual: 0 normal, 1 redo with the correction dropped
  double* partials;     // [nblocks][3 K]
};

template <int K, typename TI, int PH>
__global__ void __launch_bounds__(kBlock) k_sphase(SArgs a) {
  __shared__ Shared sh;
  const int b = blockIdx.x;
  const int i = a.blk_inst[b];
  const Inst& I = a.insts[i];
  const SState& S = a.ss[i];
  // skip rules (uniform per block)
  if constexpr (PH == PH_RES) {
    if (a.pass ? !S.redo : S.solve_done) return;
  } else if constexpr (PH != PH_RHS && PH != PH_UPD) {
    if (S.sweep_done) return;
  }
  if (threadIdx.x < K) {
    const int j = threadIdx.x;
    sh.gamma[j] = S.gamma[j];
    sh.alpha[j] = S.alpha[j];
    sh.beta[j] = S.beta[j];
    sh.omega[j] = S.omega[j];
    sh.bb[j] = S.bb[j];
    sh.reset[j] = (PH == PH_RES && a.pass) ? S.reset[j] : 0;
  }
  __syncthreads();
  const int row_lo = a.blk_row[b];
  const int row_hi = min(I.n, row_lo + a.rows);
  constexpr int NS = nsums<PH>();
  double acc[NS > 0 ? NS : 1];
  for (int q = 0; q < (NS > 0 ? NS : 1); q++) acc[q] = 0.0;
  const bool first = a.first != 0;
  const int par = a.par;
  TI* r0 = (TI*)I.r;
  if constexpr (PH == PH_RHS) {
    phase_rhs<K, TI, kPlain>(I, row_lo, row_hi, r0, a.step, acc);
  } else if constexpr (PH == PH_RES) {
    phase_res<K, TI, kPlain>(I, row_lo, row_hi, r0, sh, acc);
  } else if constexpr (PH == PH_PCG_INIT) {
    phase_pcg_init<K, TI, kPlain>(I, row_lo, row_hi, (TI*)I.r, (TI*)I.q, acc);
  } else if constexpr (PH == PH_PCG_ITER) {
    TI* rb[2] = {(TI*)I.r, (TI*)I.z};
    TI* wb[2] = {(TI*)I.q, (TI*)I.rh};
    TI* sb[2] = {(TI*)I.s, (TI*)I.t};
    phase_pcg_iter<K, TI, kPlain>(I, row_lo, row_hi, rb[par], wb[par], sb[par], rb[par ^ 1], wb[par ^ 1],
                                  sb[par ^ 1], first, sh, acc);
  } else if constexpr (PH == PH_CG_A) {
    TI* pb[2] = {(TI*)I.p0, (TI*)I.p1b};
    phase_cg_a<K, TI, kPlain>(I, row_lo, row_hi, (TI*)I.r, pb[par], pb[par ^ 1], first, sh, acc);
  } else if constexpr (PH == PH_CG_B) {
    TI* pb[2] = {(TI*)I.p0, (TI*)I.p1b};
    phase_cg_b<K, TI, kPlain>(I, row_lo, row_hi, (TI*)I.r, pb[par ^ 1], sh, acc);
  } else if constexpr (PH == PH_BI_P) {
    TI* pb[2] = {(TI*)I.p0, (TI*)I.p1b};
    TI* vb[2] = {(TI*)I.v0, (TI*)I.v1};
    phase_bi_p<K, TI, kPlain>(I, row_lo, row_hi, (TI*)I.r, pb[par], pb[par ^ 1], vb[par], vb[par ^ 1], first, sh, acc);
  } else if constexpr (PH == PH_BI_S) {
    TI* vb[2] = {(TI*)I.v0, (TI*)I.v1};
    phase_bi_s<K, TI, kPlain>(I, row_lo, row_hi, (TI*)I.r, vb[par ^ 1], sh, acc);
  } else if constexpr (PH == PH_BI_X) {
    TI* pb[2] = {(TI*)I.p0, (TI*)I.p1b};
    phase_bi_x<K, TI, kPlain>(I, row_lo, row_hi, (TI*)I.r, pb[par ^ 1], sh, acc);
  } else {
    phase_update<K, TI, kPlain>(I, row_lo, row_hi, sh, I.status, a.step);
  }
  if constexpr (NS > 0) {
    block_partials<K, NS, kPlain>(acc, sh, sh.red, 0);
    __syncthreads();
    if (threadIdx.x < NS * K) a.partials[(size_t)b * 3 * K + threadIdx.x] = sh.red[threadIdx.x];
  }
}

// Sum of the instance's block partials, in block order: thread (q, c) sums
// blocks c, c + chunks, ... of value q; the chunk sums are combined in order.
template <int K, int NS>
__device__ void sum_blocks(const SArgs& a, const int* inst_blk, int i, double* red, double* scratch) {
  constexpr int nv = NS * K;
  constexpr int chunks = kBlock / nv;
  const int q = threadIdx.x % nv, c = threadIdx.x / nv;
  const int b0 = inst_blk[i], b1 = inst_blk[i + 1];
  double s = 0.0;
  if (c < chunks) {
    constexpr int kR = 8;
    for (int w0 = b0 + c; w0 < b1; w0 += kR * chunks) {
      double v[kR];
#pragma unroll
      for (int u = 0; u < kR; u++) {
        const int w = w0 + u * chunks;
        v[u] = (w < b1) ? a.partials[(size_t)w * 3 * K + q] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < kR; u++) s += v[u];
    }
    scratch[c * nv + q] = s;
  }
  __syncthreads();
  if (threadIdx.x < nv) {
    double t = 0.0;
    for (int cc = 0; cc < chunks; cc++) t += scratch[cc * nv + threadIdx.x];
    red[threadIdx.x] = t;
  }
  __syncthreads();
}

// One block per instance: reduce the phase's partials and advance the scalar
// recurrences exactly as the persistent drivers do after their barriers.
template <int K, int PH>
__global__ void __launch_bounds__(kBlock) k_sreduce(SArgs a, const int* inst_blk, int ninst, int fp32) {
  __shared__ double scratch[kBlock];
  __shared__ double red[3 * kMaxK];
  const int i = blockIdx.x;
  if (i >= ninst) return;
  const Inst& I = a.insts[i];
  SState& S = a.ss[i];
  if constexpr (PH == PH_RES) {
    if (a.pass ? !S.redo : S.solve_done) return;
  } else if constexpr (PH != PH_RHS) {
    if (S.sweep_done) return;
  }
  sum_blocks<K, nsums<PH>()>(a, inst_blk, i, red, scratch);
  if (threadIdx.x != 0) return;
  if constexpr (PH == PH_RHS) {
    int any = 0;
    for (int j = 0; j < K; j++) {
      const double bb = red[j], rr = red[K + j];
      S.bb[j] = bb;
      S.thr[j] = fmax(I.tol2 * bb, DBL_MIN);
      S.active[j] = (bb != 0.0) && (rr >= S.thr[j]);
      S.gamma[j] = red[2 * K + j];
      S.ithr[j] = rr;
      S.reset[j] = 0;
      any |= S.active[j];
    }
    S.solve_done = !any;
    S.sweep_done = 1;
    S.redo = 0;
    S.sweeps = 0;
    S.its = 0;
  } else if constexpr (PH == PH_RES) {
    int redo = 0;
    if (!a.pass) {
      for (int j = 0; j < K; j++) {
        const bool bad = !(red[j] <= DBL_MAX);  // NaN or infinite: drop the sweep's correction
        redo |= bad;
        S.reset[j] = bad;
      }
    }
    S.redo = redo;
    if (redo) return;  // finalised by the second pass
    int any = 0;
    for (int j = 0; j < K; j++) {
      const double rr = red[j];
      S.active[j] = (S.bb[j] != 0.0) && (rr >= S.thr[j]) && (rr == rr);
      S.gamma[j] = red[K + j];
      S.ithr[j] = rr;
      S.reset[j] = 0;
      any |= S.active[j];
    }
    S.sweeps++;
    S.solve_done = !any;
  } else if constexpr (PH == PH_PCG_INIT) {
    for (int j = 0; j < K; j++) {
      const double g = red[j], d = red[K + j];
      S.gamma[j] = g;
      S.alpha[j] = (S.active[j] && d != 0.0) ? g / d : 0.0;
      S.beta[j] = 0.0;
    }
  } else if constexpr (PH == PH_PCG_ITER) {
    S.it_s++;
    S.its++;
    int alldone = 1;
    for (int j = 0; j < K; j++) {
      const double g = red[j], d = red[K + j], rr = red[2 * K + j];
      const double gold = S.gamma[j], aold = S.alpha[j];
      const bool done = !S.active[j] || rr < S.ithr[j] || !(rr == rr);
      double be = 0.0, al = 0.0;
      if (!done && gold != 0.0 && aold != 0.0) {
        be = g / gold;
        const double den = d - be * g / aold;
        al = (den != 0.0) ? g / den : 0.0;
      }
      S.beta[j] = be;
      S.alpha[j] = al;
      S.gamma[j] = g;
      alldone &= done;
    }
    S.sweep_done = alldone || S.it_s >= I.max_inner || S.its >= I.max_iters;
  } else if constexpr (PH == PH_CG_A) {
    for (int j = 0; j < K; j++) {
      const double pq = red[j];
      S.alpha[j] = (S.active[j] && pq != 0.0) ? S.gamma[j] / pq : 0.0;
    }
  } else if constexpr (PH == PH_CG_B) {
    S.it_s++;
    S.its++;
    int alldone = 1;
    for (int j = 0; j < K; j++) {
      const double rz = red[j], rr = red[K + j];
      const double g = S.gamma[j];
      S.beta[j] = (g != 0.0) ? rz / g : 0.0;
      S.gamma[j] = rz;
      alldone &= (!S.active[j]) || (rr < S.ithr[j]) || !(rr == rr);
    }
    S.sweep_done = alldone || S.it_s >= I.max_inner || S.its >= I.max_iters;
  } else if constexpr (PH == PH_BI_P) {
    S.bkd = 0;
    for (int j = 0; j < K; j++) {
      const double rv = red[j];
      const double al = (S.active[j] && rv != 0.0) ? S.gamma[j] / rv : 0.0;
      const bool fine = fabs(al) < 1e30;
      S.alpha[j] = fine ? al : 0.0;
      if (!fine) S.bkd = 1;
    }
  } else if constexpr (PH == PH_BI_S) {
    for (int j = 0; j < K; j++) {
      const double ts = red[j], tt = red[K + j], ssv = red[2 * K + j];
      const double om = (S.active[j] && tt > 0.0 && !(ssv < S.ithr[j])) ? ts / tt : 0.0;
      const bool fine = fabs(om) < 1e30;
      S.omega[j] = fine ? om : 0.0;
      if (!fine) S.bkd = 1;
    }
  } else if constexpr (PH == PH_BI_X) {
    S.it_s++;
    S.its++;
    int alldone = 1, brk = S.bkd;
    for (int j = 0; j < K; j++) {
      const double rho_new = red[j], rr = red[K + j];
      const bool done = !S.active[j] || rr < S.ithr[j] || !(rr == rr);
      alldone &= done;
      if (!done) {
        const double om = S.omega[j], rho = S.gamma[j];
        const double be = (om == 0.0 || rho == 0.0) ? 0.0 : (rho_new / rho) * (S.alpha[j] / om);
        if (be == 0.0 || !(fabs(be) < 1e30)) brk = 1;
        else S.beta[j] = be;
      }
      S.gamma[j] = rho_new;
    }
    S.sweep_done = alldone || brk || S.it_s >= I.max_inner || S.its >= I.max_iters;
  }
  (void)fp32;
}

// Start of a refinement sweep: inner threshold, iteration counter, and the
// decision whether the instance sweeps at all.
__global__ void k_sweep_begin(SState* ss, const Inst* insts, int ninst, int k, int fp32) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ninst) return;
  SState& S = ss[i];
  const Inst& I = insts[i];
  S.it_s = 0;
  if (S.solve_done || S.sweeps >= I.max_sweeps || S.its >= I.max_iters) {
    S.sweep_done = 1;
    S.solve_done = 1;
    return;
  }
  for (int j = 0; j < k; j++) S.ithr[j] = fp32 ? fmax(I.inner_tol2 * S.ithr[j], S.thr[j]) : S.thr[j];
  S.sweep_done = 0;
}

// End of a step: counters, convergence status, last residual.
__global__ void k_step_end(SState* ss, const Inst* insts, int ninst, int k) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ninst) return;
  SState& S = ss[i];
  const Inst& I = insts[i];
  int notconv = 0;
  for (int j = 0; j < k; j++) {
    notconv |= S.active[j];
    I.last_res[j] = S.bb[j] > 0.0 ? sqrt(S.ithr[j] / S.bb[j]) : 0.0;
  }
  if (notconv) atomicOr(I.status, ST_NOT_CONVERGED);
  I.counters[0] += S.its;
  I.counters[1] += S.sweeps;
  I.counters[2] += 1;
}

__global__ void k_sinit(SState* ss, const Inst* insts, int ninst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ninst) return;
  SState z;
  memset(&z, 0, sizeof(z));
  z.sweep_done = 1;
  z.solve_done = 1;
  ss[i] = z;
  insts[i].counters[0] = 0;
  insts[i].counters[1] = 0;
  insts[i].counters[2] = 0;
  *insts[i].status = 0;
}

// *out = 1 when every instance has `field` set (0: sweep_done, 1: solve_done).
__global__ void k_all(const SState* ss, int ninst, int field, int* out) {
  __shared__ int any;
  if (threadIdx.x == 0) any = 0;
  __syncthreads();
  int mine = 0;
  for (int i = threadIdx.x; i < ninst; i += blockDim.x) mine |= !(field ? ss[i].solve_done : ss[i].sweep_done);
  if (mine) atomicOr(&any, 1);
  __syncthreads();
  if (threadIdx.x == 0) *out = !any;
}

template <int K, typename TI>
const void* phase_fn(int ph) {
  switch (ph) {
    case PH_RHS: return (const void*)k_sphase<K, TI, PH_RHS>;
    case PH_RES: return (const void*)k_sphase<K, TI, PH_RES>;
    case PH_PCG_INIT: return (const void*)k_sphase<K, TI, PH_PCG_INIT>;
    case PH_PCG_ITER: return (const void*)k_sphase<K, TI, PH_PCG_ITER>;
    case PH_CG_A: return (const void*)k_sphase<K, TI, PH_CG_A>;
    case PH_CG_B: return (const void*)k_sphase<K, TI, PH_CG_B>;
    case PH_BI_P: return (const void*)k_sphase<K, TI, PH_BI_P>;
    case PH_BI_S: return (const void*)k_sphase<K, TI, PH_BI_S>;
    case PH_BI_X: return (const void*)k_sphase<K, TI, PH_BI_X>;
    default: return (const void*)k_sphase<K, TI, PH_UPD>;
  }
}

template <int K>
const void* reduce_fn(int ph) {
  switch (ph) {
    case PH_RHS: return (const void*)k_sreduce<K, PH_RHS>;
    case PH_RES: return (const void*)k_sreduce<K, PH_RES>;
    case PH_PCG_INIT: return (const void*)k_sreduce<K, PH_PCG_INIT>;
    case PH_PCG_ITER: return (const void*)k_sreduce<K, PH_PCG_ITER>;
    case PH_CG_A: return (const void*)k_sreduce<K, PH_CG_A>;
    case PH_CG_B: return (const void*)k_sreduce<K, PH_CG_B>;
    case PH_BI_P: return (const void*)k_sreduce<K, PH_BI_P>;
    case PH_BI_S: return (const void*)k_sreduce<K, PH_BI_S>;
    case PH_BI_X: return (const void*)k_sreduce<K, PH_BI_X>;
    default: return nullptr;
  }
}

const void* pick_phase(int k, int fp32, int ph) {
  switch (k) {
    case 1: return fp32 ? phase_fn<1, float>(ph) : phase_fn<1, double>(ph);
    case 2: return fp32 ? phase_fn<2, float>(ph) : phase_fn<2, double>(ph);
    case 4: return fp32 ? phase_fn<4, float>(ph) : phase_fn<4, double>(ph);
    case 16: return fp32 ? phase_fn<16, float>(ph) : phase_fn<16, double>(ph);
    default: return nullptr;
  }
}

const void* pick_reduce(int k, int ph) {
  switch (k) {
    case 1: return reduce_fn<1>(ph);
    case 2: return reduce_fn<2>(ph);
    case 4: return reduce_fn<4>(ph);
    case 16: return reduce_fn<16>(ph);
    default: return nullptr;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

struct StreamPlan {
  int ninst = 0, nblocks = 0, k = 1, fp32 = 1, kr = 0, rows = 0;
  SState* d_ss = nullptr;
  int* d_blk_inst = nullptr;
  int* d_blk_row = nullptr;
  int* d_inst_blk = nullptr;
  double* d_partials = nullptr;
  int* d_flag = nullptr;
  int* h_flag = nullptr;  // pinned
};

void stream_plan_free(StreamPlan* p) {
  if (!p) return;
  void* ptrs[] = {p->d_ss, p->d_blk_inst, p->d_blk_row, p->d_inst_blk, p->d_partials, p->d_flag};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  if (p->h_flag) (void)hipHostFree(p->h_flag);
  delete p;
}

// Row-block table over all instances; `rows` per block (a multiple of 256/k).
int stream_plan_build(const std::vector<InstHost>& insts, const Inst* d_insts, int k, int fp32, int kr, int rows,
                      hipStream_t st, StreamPlan** out) {
  (void)d_insts;
  StreamPlan* p = new StreamPlan();
  p->ninst = (int)insts.size();
  p->k = k;
  p->fp32 = fp32;
  p->kr = kr;
  p->rows = rows;
  std::vector<int> bi, br, ib(p->ninst + 1, 0);
  for (int i = 0; i < p->ninst; i++) {
    ib[i] = (int)bi.size();
    const int n = insts[i].ctx->n;
    for (int r = 0; r < n; r += rows) {
      bi.push_back(i);
      br.push_back(r);
    }
  }
  ib[p->ninst] = (int)bi.size();
  p->nblocks = (int)bi.size();
  int rc = 0;
  p->d_ss = dalloc<SState>(p->ninst, &rc);
  p->d_blk_inst = dalloc<int>(std::max(1, p->nblocks), &rc);
  p->d_blk_row = dalloc<int>(std::max(1, p->nblocks), &rc);
  p->d_inst_blk = dalloc<int>(p->ninst + 1, &rc);
  p->d_partials = dalloc<double>((size_t)std::max(1, p->nblocks) * 3 * k, &rc);
  p->d_flag = dalloc<int>(1, &rc);
  if (!rc) {
    hipError_t e = hipHostMalloc((void**)&p->h_flag, sizeof(int), hipHostMallocDefault);
    if (e != hipSuccess) rc = hip_fail(e, "hipHostMalloc");
  }
  if (rc) {
    stream_plan_free(p);
    return rc;
  }
  if (p->nblocks > 0) {
    FEM_HIP(hipMemcpyAsync(p->d_blk_inst, bi.data(), sizeof(int) * bi.size(), hipMemcpyHostToDevice, st));
    FEM_HIP(hipMemcpyAsync(p->d_blk_row, br.data(), sizeof(int) * br.size(), hipMemcpyHostToDevice, st));
  }
  FEM_HIP(hipMemcpyAsync(p->d_inst_blk, ib.data(), sizeof(int) * ib.size(), hipMemcpyHostToDevice, st));
  FEM_HIP(hipStreamSynchronize(st));
  *out = p;
  return 0;
}

namespace {

int launch_phase(const StreamPlan* p, const SArgs& a, int ph, hipStream_t st) {
  const void* fn = pick_phase(p->k, p->fp32, ph);
  if (!fn) {
    set_error("stream engine: unsupported column count");
    return FEM_ERR_INVALID;
  }
  if (p->nblocks == 0) return 0;
  SArgs aa = a;
  void* args[] = {(void*)&aa};
  FEM_HIP(hipLaunchKernel(fn, dim3(p->nblocks), dim3(kBlock), args, 0, st));
  if (ph == PH_UPD) return 0;
  const void* rf = pick_reduce(p->k, ph);
  const int* ib = p->d_inst_blk;
  int ninst = p->ninst, fp32 = p->fp32;
  void* rargs[] = {(void*)&aa, (void*)&ib, (void*)&ninst, (void*)&fp32};
  FEM_HIP(hipLaunchKernel(rf, dim3(p->ninst), dim3(kBlock), rargs, 0, st));
  return 0;
}

// Host read of "every instance has finished its sweep (0) / solve (1)".
int all_done(const StreamPlan* p, int field, hipStream_t st, bool* done) {
  hipLaunchKernelGGL(k_all, dim3(1), dim3(kBlock), 0, st, p->d_ss, p->ninst, field, p->d_flag);
  FEM_HIP(hipGetLastError());
  FEM_HIP(hipMemcpyAsync(p->h_flag, p->d_flag, sizeof(int), hipMemcpyDeviceToHost, st));
  FEM_HIP(hipStreamSynchronize(st));
  *done = *p->h_flag != 0;
  return 0;
}

}  // namespace

// Advance every instance by nsteps with the streaming engine.  The instance
// descriptors are already on the device (d_insts).
int stream_run(const StreamPlan* p, const Inst* d_insts, int max_sweeps, int max_inner, int nsteps,
               hipStream_t st) {
  SArgs a{};
  a.insts = d_insts;
  a.ss = p->d_ss;
  a.blk_inst = p->d_blk_inst;
  a.blk_row = p->d_blk_row;
  a.rows = p->rows;
  a.partials = p->d_partials;
  hipLaunchKernelGGL(k_sinit, dim3(div_up(p->ninst, kBlock)), dim3(kBlock), 0, st, p->d_ss, d_insts, p->ninst);
  FEM_HIP(hipGetLastError());
  int rc = 0;
#define STREAM_PH(ph)                                   \
  do {                                                  \
    if ((rc = launch_phase(p, a, (ph), st))) return rc; \
  } while (0)
  for (int step = 0; step < nsteps; step++) {
    a.step = step;
    a.par = 0;
    a.first = 1;
    a.pass = 0;
    STREAM_PH(PH_RHS);
    for (int sw = 0; sw < max_sweeps; sw++) {
      hipLaunchKernelGGL(k_sweep_begin, dim3(div_up(p->ninst, kBlock)), dim3(kBlock), 0, st, p->d_ss, d_insts,
                         p->ninst, p->k, p->fp32);
      FEM_HIP(hipGetLastError());
      bool done = false;
      if ((rc = all_done(p, 1, st, &done))) return rc;
      if (done) break;
      int par = 0;
      if (p->kr == 0) {  // pipelined CG
        a.par = 0;
        a.first = 1;
        STREAM_PH(PH_PCG_INIT);
        for (int it = 0; it < max_inner; it++) {
          a.par = par;
          a.first = it == 0;
          STREAM_PH(PH_PCG_ITER);
          par ^= 1;
          if ((rc = all_done(p, 0, st, &done))) return rc;
          if (done) break;
        }
      } else {  // classic CG (2) or BiCGSTAB (1): p / v buffers swap every iteration
        for (int it = 0; it < max_inner; it++) {
          a.par = par;
          a.first = it == 0;
          if (p->kr == 2) {
            STREAM_PH(PH_CG_A);
            STREAM_PH(PH_CG_B);
          } else {
            STREAM_PH(PH_BI_P);
            STREAM_PH(PH_BI_S);
            STREAM_PH(PH_BI_X);
          }
          par ^= 1;
          if ((rc = all_done(p, 0, st, &done))) return rc;
          if (done) break;
        }
      }
      a.pass = 0;
      STREAM_PH(PH_RES);
      a.pass = 1;
      STREAM_PH(PH_RES);
      a.pass = 0;
    }
    STREAM_PH(PH_UPD);
    hipLaunchKernelGGL(k_step_end, dim3(div_up(p->ninst, kBlock)), dim3(kBlock), 0, st, p->d_ss, d_insts, p->ninst,
                       p->k);
    FEM_HIP(hipGetLastError());
  }
#undef STREAM_PH
  return 0;
}

}  // namespace femhip
