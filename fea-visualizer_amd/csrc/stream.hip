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
// solve has finished skip the launches.  The host reads the counts of the
// instances still iterating every few iterations, one read in flight behind
// the queued launches, to stop the loop.
//
// For one or two right-hand sides the phase bodies are the persistent
// kernel's own (plain exchange mode); for four or more, each lane owns 16
// bytes of adjacent columns (stream_vec.h).  Both compute the persistent
// kernel's per-element arithmetic.
#include "stream_vec.h"

#include <algorithm>
#include <chrono>

namespace femhip {

namespace {

// Per-instance scalars of the running solve (the persistent kernel's Shared).
struct SState {
  double gamma[kMaxK], alpha[kMaxK], beta[kMaxK], omega[kMaxK];
  double thr[kMaxK], ithr[kMaxK], bb[kMaxK];
  int active[kMaxK], reset[kMaxK];
  int bbad[kMaxK];  // BiCGSTAB, multi-column layout: beta of column j broke down (phase S)
  int sweep_done, solve_done, redo, bkd;
  int it_s, its, sweeps, status;
  long long its_total, sweeps_total;
};

enum Phase : int {
  PH_RHS = 0, PH_RES, PH_PCG_INIT, PH_PCG_ITER, PH_CG_A, PH_CG_B, PH_BI_P, PH_BI_S, PH_BI_X, PH_UPD
};

// Dot products a phase sums per column.  BiCGSTAB on the multi-column layout
// (stream_vec.h) sums rho's two parts in phase S and only <r,r> in phase X.
template <int K, int PH>
constexpr int nsums() {
  return PH == PH_BI_S ? (vec_layout<K>() ? 5 : 3)
         : PH == PH_BI_X ? (vec_layout<K>() ? 1 : 2)
         : PH == PH_RHS || PH == PH_PCG_ITER ? 3
         : PH == PH_RES || PH == PH_PCG_INIT || PH == PH_CG_B ? 2
         : PH == PH_UPD ? 0 : 1;
}

// Partial sums per row block and column: the most any phase sums.
constexpr int kPartW = 5;

struct SArgs {
  const Inst* insts;
  SState* ss;
  const int* blk_inst;  // [nblocks]
  const int* blk_row;   // [nblocks] first row
  int par;              // buffer parity of the current iteration
  int first;            // first iteration of the sweep
  int step;             // step index inside the launch sequence (trajectory records)
  int pass;             // FP64 residual: 0 normal, 1 redo with the correction dropped
  int more;             // update: another step follows in this launch (it reads xnext)
  double* partials;     // [nblocks][kPartW K]
  const int* act_blk;   // launched block -> row block (null: every row block, in order)
  const int* act_inst;  // reduction block -> instance (null: every instance)
  const int* act_cnt;   // {instances, blocks} of the lists (device); the grid may be larger
  const int* blk_end;   // [nblocks] end row
  int nall;             // row blocks of the plan (launches that cover all of them)
  int stage_rows;       // LDS stage capacity: rows and nonzeros per row block
  int stage_nnz;
  unsigned long long* post;  // one instance: pinned word for the host's read of sweep_done, written by the
  unsigned post_seq;         //   reduction that ends the iteration (null: no read), tagged with post_seq
  int rhs_stage;             // right-hand side: the launch carries the FP64 stage
  int res_stage;             // FP64 residual: likewise
};

// Nonzeros a row block may hold: blocks are cut so that their pattern fits
// the LDS stage; a single longer row gets a block of its own and the stage
// grows to it (up to the workgroup's 64 KiB of LDS).
constexpr int kStageCap = 2048;
constexpr size_t kMaxLds = 65536 - sizeof(Shared);

template <int PH>
constexpr bool staged_phase() {
  return PH == PH_PCG_INIT || PH == PH_PCG_ITER || PH == PH_CG_A || PH == PH_BI_P || PH == PH_BI_S;
}


// One column per lane (the persistent kernel's phase bodies).
template <int K, typename TI, int PH>
__device__ void run_scalar(const Inst& I, const SArgs& a, Shared& sh, int b, int row_lo, int row_hi) {
  constexpr int NS = nsums<K, PH>();
  const bool first = a.first != 0;
  const int par = a.par;
  double acc[NS > 0 ? NS : 1];
  for (int q = 0; q < (NS > 0 ? NS : 1); q++) acc[q] = 0.0;
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
    phase_bi_p<K, TI, kPlain>(I, row_lo, row_hi, (TI*)I.r, pb[par], pb[par ^ 1], vb[par], vb[par ^ 1], first, sh,
                              acc);
  } else if constexpr (PH == PH_BI_S) {
    TI* vb[2] = {(TI*)I.v0, (TI*)I.v1};
    phase_bi_s<K, TI, kPlain>(I, row_lo, row_hi, (TI*)I.r, vb[par ^ 1], sh, acc);
  } else if constexpr (PH == PH_BI_X) {
    TI* pb[2] = {(TI*)I.p0, (TI*)I.p1b};
    phase_bi_x<K, TI, kPlain>(I, row_lo, row_hi, (TI*)I.r, pb[par ^ 1], sh, acc);
  }
  if constexpr (NS > 0) {
    block_partials<K, NS, kPlain>(acc, sh, sh.red, 0);
    __syncthreads();
    if (threadIdx.x < NS * K) a.partials[(size_t)b * kPartW * K + threadIdx.x] = sh.red[threadIdx.x];
  }
}

// State update of heat, advection-diffusion and Gray-Scott (phase_update's
// arithmetic), with the predictor history turned by pointer instead of by
// copies: a row moves u_n into the oldest history slot, which k_step_end then
// makes the newest (a), and stores the new solution into u.  x + xi is not
// stored back, since the next right-hand side phase rewrites x and xi.
// guess: also the next step's initial guess of order lvn into x, which the
// next right-hand side phase gathers (vphase_rhs).
template <int K, typename TI>
__device__ void stream_update_first(const Inst& I, int row_lo, int row_hi, const Shared& sh, int* flags, int step,
                                    int lvn, bool guess) {
  constexpr int RPP = kBlock / K;
  const int jr = threadIdx.x % K;
  const TI GA* xi = gp((const TI*)I.xi);
  const int e = I.extrap;
  double GA* oldest = gp(e == 1 ? I.a : e == 2 ? I.v : e == 3 ? I.h3 : e == 4 ? I.h4 : e == 5 ? I.h5 : I.h6);
  double w[kMaxExtrap + 1] = {};
  if (guess) pred_weights_n(lvn, w);
  int f = 0;
  double* traj = nullptr;  // trajectory record of this step, if any
  if (I.traj && (step + 1) % I.traj_every == 0)
    traj = I.traj + (size_t)((step + 1) / I.traj_every - 1) * I.n * K;
  for (int row = row_lo + (int)threadIdx.x / K; row < row_hi; row += RPP) {
    const size_t o = (size_t)row * K + jr;
    double xf = gp(I.x)[o] + (double)xi[o];
    const double uo = gp(I.u)[o];
    const double ao = e >= 2 ? gp(I.a)[o] : 0.0;
    const double vo = e >= 3 ? gp(I.v)[o] : 0.0;
    const double ho = e >= 4 ? gp(I.h3)[o] : 0.0;
    const double h4o = e >= 5 ? gp(I.h4)[o] : 0.0;
    const double h5o = e >= 6 ? gp(I.h5)[o] : 0.0;
    if (sh.bb[jr] == 0.0) xf = 0.0;  // Eigen: rhs == 0 -> x = 0
    if (guess) gp(I.x)[o] = w[0] * xf + w[1] * uo + w[2] * ao + w[3] * vo + w[4] * ho + w[5] * h4o + w[6] * h5o;
    if (e >= 1) oldest[o] = uo;
    gp(I.u)[o] = xf;
    if (traj) gp(traj)[o] = xf;
    if (xf > 1e4) f |= ST_UNSTABLE;
    if (!isfinite(xf)) f |= ST_NONFINITE;
  }
  if (f) atomicOr(flags, f);
}

// Newmark update (phase_update's arithmetic, step_kernel.h).  more: another
// step follows in this launch, so also its displacement predictor
// u + dt v + (1/2 - beta) dt^2 a into hvec and, lvn >= 0, its guess of the
// acceleration (order lvn) into x: the next right-hand side (vphase_rhs)
// then gathers two vectors instead of five.
template <int K, typename TI>
__device__ void stream_update_newmark(const Inst& I, int row_lo, int row_hi, const Shared& sh, int* flags, int step,
                                      int lvn, bool more) {
  constexpr int RPP = kBlock / K;
  const int jr = threadIdx.x % K;
  TI GA* xi = gp((TI*)I.xi);
  const double dt = I.dt, be = I.p2, ga = I.p3;
  const double hh = (0.5 - be) * dt * dt;
  int f = 0;
  double* traj = nullptr;  // trajectory record of this step, if any
  if (I.traj && (step + 1) % I.traj_every == 0)
    traj = I.traj + (size_t)((step + 1) / I.traj_every - 1) * I.n * K;
  for (int row = row_lo + (int)threadIdx.x / K; row < row_hi; row += RPP) {
    const size_t o = (size_t)row * K + jr;
    double xf = gp(I.x)[o] + (double)xi[o];
    const double uo = gp(I.u)[o], ao = gp(I.a)[o], vo = gp(I.v)[o];
    const double ho = I.extrap >= 2 ? gp(I.h3)[o] : 0.0;
    if (sh.bb[jr] == 0.0) xf = 0.0;  // Eigen: rhs == 0 -> x = 0
    const double ut = uo + dt * vo + hh * ao;
    const double un = ut + be * dt * dt * xf;
    const double vn = vo + dt * ((1.0 - ga) * ao + ga * xf);
    gp(I.u)[o] = un;
    gp(I.v)[o] = vn;
    if (I.extrap >= 2) gp(I.h4)[o] = ho;  // acceleration history for the predictor
    if (I.extrap >= 1) gp(I.h3)[o] = ao;
    gp(I.a)[o] = xf;
    if (more) {
      gp(I.hvec)[o] = un + dt * vn + hh * xf;
      if (lvn >= 0) gp(I.x)[o] = lvn == 2 ? 3.0 * (xf - ao) + ho : lvn == 1 ? 2.0 * xf - ao : xf;
    } else {
      gp(I.x)[o] = xf;
      xi[o] = (TI)0;
    }
    if (traj) gp(traj)[o] = un;
    if (un > 1e4 || vn > 1e4) f |= ST_UNSTABLE;
    if (!isfinite(un) || !isfinite(vn)) f |= ST_NONFINITE;
  }
  if (f) atomicOr(flags, f);
}

// V adjacent columns per lane (stream_vec.h); g: the staged pattern of the
// gathering Krylov phases.
template <int K, int V, typename TI, int PH>
__device__ void run_vector(const Inst& I, const SArgs& a, Shared& sh, int b, int row_lo, int row_hi,
                           const Stage<TI>& g, int rcase = 0, const RhsStage& rs = RhsStage{}) {
  constexpr int NS = nsums<K, PH>();
  const bool first = a.first != 0;
  const int par = a.par;
  vec<double, V> acc[NS];
  for (int q = 0; q < NS; q++) acc[q] = splat<V>(0.0);
  if constexpr (PH == PH_RHS) {
    if (rcase)
      vphase_rhs_staged<K, V, TI>(I, row_lo, row_hi, rcase, rs, acc);
    else
      vphase_rhs<K, V, TI>(I, row_lo, row_hi, a.step, acc);
  } else if constexpr (PH == PH_RES) {
    if (rcase)
      vphase_res<K, V, TI, true>(I, row_lo, row_hi, sh, acc, rs);
    else
      vphase_res<K, V, TI>(I, row_lo, row_hi, sh, acc);
  } else if constexpr (PH == PH_PCG_INIT) {
    vphase_pcg_init<K, V, TI>(I, row_lo, row_hi, g, (TI*)I.r, (TI*)I.q, acc);
  } else if constexpr (PH == PH_PCG_ITER) {
    TI* rb[2] = {(TI*)I.r, (TI*)I.z};
    TI* wb[2] = {(TI*)I.q, (TI*)I.rh};
    TI* sb[2] = {(TI*)I.s, (TI*)I.t};
    vphase_pcg_iter<K, V, TI>(I, row_lo, row_hi, g, rb[par], wb[par], sb[par], rb[par ^ 1], wb[par ^ 1],
                              sb[par ^ 1], first, sh, acc);
  } else if constexpr (PH == PH_CG_A) {
    TI* pb[2] = {(TI*)I.p0, (TI*)I.p1b};
    vphase_cg_a<K, V, TI>(I, row_lo, row_hi, g, (TI*)I.r, pb[par], pb[par ^ 1], first, sh, acc);
  } else if constexpr (PH == PH_CG_B) {
    TI* pb[2] = {(TI*)I.p0, (TI*)I.p1b};
    vphase_cg_b<K, V, TI>(I, row_lo, row_hi, (TI*)I.r, pb[par ^ 1], sh, acc);
  } else if constexpr (PH == PH_BI_P) {  // BiCGSTAB: one p and one v buffer (stream_vec.h)
    vphase_bi_p<K, V, TI>(I, row_lo, row_hi, g, (TI*)I.r, (TI*)I.p0, (TI*)I.v0, first, acc);
  } else if constexpr (PH == PH_BI_S) {
    vphase_bi_s<K, V, TI>(I, row_lo, row_hi, g, (TI*)I.r, (const TI*)I.v0, sh, acc);
  } else {
    vphase_bi_x<K, V, TI>(I, row_lo, row_hi, (TI*)I.r, (TI*)I.p0, (const TI*)I.v0, sh, acc);
  }
  // block sums in rounds of at most kMaxSlots values per column
  constexpr int N0 = NS < kMaxSlots ? NS : kMaxSlots;
  vblock_partials<K, V, N0>(acc, sh);
  __syncthreads();
  if (threadIdx.x < N0 * K) a.partials[(size_t)b * kPartW * K + threadIdx.x] = sh.red[threadIdx.x];
  if constexpr (NS > N0) {
    static_assert(NS - N0 <= kMaxSlots, "two rounds of block sums");
    __syncthreads();
    vblock_partials<K, V, NS - N0>(acc + N0, sh);
    __syncthreads();
    if (threadIdx.x < (NS - N0) * K) a.partials[(size_t)b * kPartW * K + N0 * K + threadIdx.x] = sh.red[threadIdx.x];
  }
}

// One launch covers nb entries of the block list (every row block when the
// launch covers all of them).  Workgroup blockIdx sits on XCD x = blockIdx
// mod 8, which owns one contiguous chunk of the list (halo rows then hit its
// L2); the XCD's workgroups take the chunk's entries in turn.  With one
// workgroup per entry (the host knows the list's size) each takes one entry;
// WALK: a fixed smaller grid (graph replays, where the size is known only on
// the device) walks several (a separate variant: the loop costs registers).
template <int K, typename TI, int PH, bool WALK = false>
__global__ void __launch_bounds__(kBlock) k_sphase(SArgs a) {
  __shared__ Shared sh;
  const int nb = a.act_cnt ? a.act_cnt[1] : a.nall;
  const int x = (int)blockIdx.x & 7, q = nb >> 3, r = nb & 7;
  const int cx = q + (x < r);                      // entries of XCD x
  const int c0 = x * q + min(x, r);                // its first entry
  const int gx = ((int)gridDim.x - x + 7) >> 3;    // workgroups on XCD x
  for (int jx = (int)blockIdx.x >> 3; jx < cx; jx += gx) {
    const int c = c0 + jx;
    const int b = a.act_blk ? a.act_blk[c] : c;
    const int i = a.blk_inst[b];
    const Inst& I = a.insts[i];
    const SState& S = a.ss[i];
    // skip rules (uniform per block)
    bool skip = false;
    if constexpr (PH == PH_RES) skip = a.pass ? !S.redo : S.solve_done;
    else if constexpr (PH != PH_RHS && PH != PH_UPD) skip = S.sweep_done != 0;
    if (skip) {
      if constexpr (WALK) continue;
      else return;
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
    const int row_lo = a.blk_row[b];
    const int row_hi = a.blk_end[b];
    constexpr int V = vec_cols<K, TI>();
    if constexpr (PH == PH_UPD) {
      __syncthreads();
      const int lvn = (a.more && I.warm) ? min(I.extrap, I.hist + a.step + 1) : -1;
      if (first_order_eq(I.eq))  // the guess goes to x where vphase_rhs gathers it
        stream_update_first<K, TI>(I, row_lo, row_hi, sh, I.status, a.step, lvn,
                                   vec_layout<K>() && lvn > 0 && (I.eq == EQ_HEAT || I.eq == EQ_ADVDIFF));
      else if (vec_layout<K>() && I.eq == EQ_NEWMARK && I.hvec)
        stream_update_newmark<K, TI>(I, row_lo, row_hi, sh, I.status, a.step, lvn, a.more != 0);
      else
        phase_update<K, TI, kPlain>(I, row_lo, row_hi, sh, I.status, a.step, lvn);
    } else if constexpr (vec_layout<K>() && staged_phase<PH>()) {
      // stage the block's pattern: row pointers, columns, S D^-1 (phase A of classic CG: S)
      extern __shared__ int dyn[];
      const int GA* rp = gp(I.rp);
      const int s_lo = rp[row_lo], s_hi = rp[row_hi];
      const int nrow = row_hi - row_lo, nnz = s_hi - s_lo;  // within the stage (plan)
      int LA* lrp = (int LA*)dyn;
      int LA* lci = lrp + ((a.stage_rows + 4) & ~3);
      TI LA* lw = (TI LA*)(lci + ((a.stage_nnz + 3) & ~3));
      const int GA* ci = gp(I.ci);
      const TI GA* w = gp((const TI*)(PH == PH_CG_A ? I.Si : I.SDi));
      for (int t = threadIdx.x; t <= nrow; t += kBlock) lrp[t] = rp[row_lo + t] - s_lo;
      for (int t = threadIdx.x; t < nnz; t += kBlock) {
        lci[t] = ci[s_lo + t];
        lw[t] = w[s_lo + t];
      }
      __syncthreads();
      run_vector<K, V, TI, PH>(I, a, sh, b, row_lo, row_hi, Stage<TI>{lrp, lci, lw, row_lo});
    } else if constexpr (vec_layout<K>() && (PH == PH_RHS || PH == PH_RES)) {
      // right-hand side (steady cases) and FP64 residual: stage the block's
      // pattern and its FP64 value rows (K or M, and S; the residual: S),
      // which one lane per row would read with many distinct lines per load
      const int rcase = PH == PH_RES ? (a.res_stage ? 3 : 0) : a.rhs_stage ? rhs_case(I, a.step) : 0;
      if (rcase) {
        extern __shared__ __attribute__((aligned(16))) int dyn[];
        const int GA* rp = gp(I.rp);
        const int s_lo = rp[row_lo], s_hi = rp[row_hi];
        const int nrow = row_hi - row_lo, nnz = s_hi - s_lo;  // within the stage (plan)
        int LA* lrp = (int LA*)dyn;
        int LA* lci = lrp + ((a.stage_rows + 4) & ~3);
        double LA* la = (double LA*)(lci + ((a.stage_nnz + 3) & ~3));
        double LA* ls = la + a.stage_nnz;
        const int GA* ci = gp(I.ci);
        const double GA* wa = gp(rcase == 1 ? I.K64 : rcase == 2 ? I.M64 : I.S64);
        const double GA* ws = gp(I.S64);
        for (int t = threadIdx.x; t <= nrow; t += kBlock) lrp[t] = rp[row_lo + t] - s_lo;
        for (int t = threadIdx.x; t < nnz; t += kBlock) {
          lci[t] = ci[s_lo + t];
          la[t] = wa[s_lo + t];
          if (rcase != 3) ls[t] = ws[s_lo + t];
        }
        __syncthreads();
        run_vector<K, V, TI, PH>(I, a, sh, b, row_lo, row_hi, Stage<TI>{}, rcase, RhsStage{lrp, lci, la, ls});
      } else {
        __syncthreads();
        run_vector<K, V, TI, PH>(I, a, sh, b, row_lo, row_hi, Stage<TI>{});
      }
    } else if constexpr (vec_layout<K>()) {
      __syncthreads();
      run_vector<K, V, TI, PH>(I, a, sh, b, row_lo, row_hi, Stage<TI>{});
    } else {
      __syncthreads();
      run_scalar<K, TI, PH>(I, a, sh, b, row_lo, row_hi);
    }
    if constexpr (!WALK) return;
    __syncthreads();  // the next entry reuses the scalars and the stage
  }
}

// Threads of a reduction workgroup: the largest instances have thousands of
// row blocks, and the sum is bound by how many partials are in flight.
constexpr int kRedBlock = 1024;

// Sum of the instance's block partials in a fixed order: thread (q, c) sums
// blocks c, c + chunks, ... of value q, kR loads in flight at a time; a fixed
// pairwise tree combines the chunk sums.
template <int K, int NS>
__device__ void sum_blocks(const SArgs& a, const int* inst_blk, int i, double* red, double* scratch) {
  constexpr int nv = NS * K;
  constexpr int chunks = kRedBlock / nv;
  const int q = threadIdx.x % nv, c = threadIdx.x / nv;
  const int b0 = inst_blk[i], b1 = inst_blk[i + 1];
  double s = 0.0;
  if (c < chunks) {
    constexpr int kR = 16;
    for (int w0 = b0 + c; w0 < b1; w0 += kR * chunks) {
      double v[kR];
#pragma unroll
      for (int u = 0; u < kR; u++) {
        const int w = w0 + u * chunks;
        v[u] = (w < b1) ? a.partials[(size_t)w * kPartW * K + q] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < kR; u++) s += v[u];
    }
    scratch[c * nv + q] = s;
  }
  __syncthreads();
  // pairwise tree over the chunks in a fixed order (deterministic): a chain
  // of up to kRedBlock dependent adds would dominate the reduction
  int span = 1;
  while (span < chunks) span <<= 1;
  for (int h = span >> 1; h > 0; h >>= 1) {
    if (c < h && c + h < chunks) scratch[c * nv + q] += scratch[(c + h) * nv + q];
    __syncthreads();
  }
  if (threadIdx.x < nv) red[threadIdx.x] = scratch[threadIdx.x];
  __syncthreads();
}

// Phases whose reduction ends a Krylov iteration (and sets sweep_done).
template <int PH>
constexpr bool ends_iteration() {
  return PH == PH_PCG_ITER || PH == PH_CG_B || PH == PH_BI_X;
}

// The host's read of one instance's sweep_done: one 8-byte word with the
// read's sequence number in the high half (the layout k_post writes).
__device__ inline void post_flag_word(const SArgs& a, int v) {
  __hip_atomic_store(a.post, ((unsigned long long)a.post_seq << 32) | (unsigned)v, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_SYSTEM);
}

// One block per instance: reduce the phase's partials and advance the scalar
// recurrences exactly as the persistent drivers do after their barriers.
template <int K, int PH>
__global__ void __launch_bounds__(kRedBlock) k_sreduce(SArgs a, const int* inst_blk, int ninst, int fp32) {
  __shared__ double scratch[kRedBlock];
  __shared__ double red[kPartW * kMaxK];
  if ((int)blockIdx.x >= ninst || (a.act_cnt && (int)blockIdx.x >= a.act_cnt[0])) return;
  const int i = a.act_inst ? a.act_inst[blockIdx.x] : blockIdx.x;
  const Inst& I = a.insts[i];
  SState& S = a.ss[i];
  if constexpr (PH == PH_RES) {
    if (a.pass ? !S.redo : S.solve_done) return;
  } else if constexpr (PH != PH_RHS) {
    if (S.sweep_done) {
      if constexpr (ends_iteration<PH>())
        if (a.post && threadIdx.x == 0) post_flag_word(a, 1);
      return;
    }
  }
  // pipelined CG, few columns: thread 0 loads the recurrence's scalars before
  // the sums, so the loads overlap them (no dependent round trip at the end)
  constexpr bool pre = PH == PH_PCG_ITER && K <= 4;
  double p_gamma[pre ? K : 1], p_alpha[pre ? K : 1], p_ithr[pre ? K : 1];
  int p_active[pre ? K : 1], p_it_s = 0, p_its = 0, p_max_inner = 0, p_max_iters = 0;
  if constexpr (pre) {
    if (threadIdx.x == 0) {
      for (int j = 0; j < K; j++) {
        p_gamma[j] = S.gamma[j];
        p_alpha[j] = S.alpha[j];
        p_ithr[j] = S.ithr[j];
        p_active[j] = S.active[j];
      }
      p_it_s = S.it_s;
      p_its = S.its;
      p_max_inner = I.max_inner;
      p_max_iters = I.max_iters;
    }
  }
  sum_blocks<K, nsums<K, PH>()>(a, inst_blk, i, red, scratch);
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
    const int it_s = (pre ? p_it_s : S.it_s) + 1, its = (pre ? p_its : S.its) + 1;
    S.it_s = it_s;
    S.its = its;
    int alldone = 1;
    for (int j = 0; j < K; j++) {
      const double g = red[j], d = red[K + j], rr = red[2 * K + j];
      const double gold = pre ? p_gamma[j] : S.gamma[j], aold = pre ? p_alpha[j] : S.alpha[j];
      const bool done = !(pre ? p_active[j] : S.active[j]) || rr < (pre ? p_ithr[j] : S.ithr[j]) || !(rr == rr);
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
    S.sweep_done = alldone || it_s >= (pre ? p_max_inner : I.max_inner) || its >= (pre ? p_max_iters : I.max_iters);
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
      if constexpr (vec_layout<K>()) {
        // next rho = <rh, s - omega t>, and beta for the x phase's new p;
        // a breakdown stops the sweep only if the column goes on (phase X)
        const double w = S.omega[j], rho = S.gamma[j];
        const double rho_new = red[3 * K + j] - w * red[4 * K + j];
        const double be = (w == 0.0 || rho == 0.0) ? 0.0 : (rho_new / rho) * (S.alpha[j] / w);
        const bool bad = be == 0.0 || !(fabs(be) < 1e30);
        S.bbad[j] = bad;
        S.beta[j] = bad ? 0.0 : be;
        S.gamma[j] = rho_new;
      }
    }
  } else if constexpr (PH == PH_BI_X && vec_layout<K>()) {
    S.it_s++;
    S.its++;
    int alldone = 1, brk = S.bkd;
    for (int j = 0; j < K; j++) {
      const double rr = red[j];
      const bool done = !S.active[j] || rr < S.ithr[j] || !(rr == rr);
      alldone &= done;
      if (!done && S.bbad[j]) brk = 1;
    }
    S.sweep_done = alldone || brk || S.it_s >= I.max_inner || S.its >= I.max_iters;
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
  if constexpr (ends_iteration<PH>())
    if (a.post) post_flag_word(a, S.sweep_done);
  (void)fp32;
}

// Start of a refinement sweep: inner threshold, iteration counter, and the
// decision whether the instance sweeps at all.
// post (one instance, else null): the host's read of the sweep_done flag,
// written here as k_post would write it.
__global__ void k_sweep_begin(SState* ss, const Inst* insts, int ninst, int k, int fp32, unsigned long long* post,
                              unsigned seq) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ninst) return;
  SState& S = ss[i];
  const Inst& I = insts[i];
  S.it_s = 0;
  int done = 1;
  if (S.solve_done || S.sweeps >= I.max_sweeps || S.its >= I.max_iters) {
    S.solve_done = 1;
  } else {
    for (int j = 0; j < k; j++) S.ithr[j] = fp32 ? fmax(I.inner_tol2 * S.ithr[j], S.thr[j]) : S.thr[j];
    done = 0;
  }
  S.sweep_done = done;
  if (post && i == 0)
    __hip_atomic_store(post, ((unsigned long long)seq << 32) | (unsigned)done, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_SYSTEM);
}

// End of a step: counters, convergence status, last residual, and the turn
// of the predictor history (stream_update_first; the host turns its copies
// of the pointers by the same count after the run).
__global__ void k_step_end(SState* ss, Inst* insts, int ninst, int k) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ninst) return;
  SState& S = ss[i];
  Inst& I = insts[i];
  if (first_order_eq(I.eq)) turn_history(&I.a, &I.v, &I.h3, &I.h4, &I.h5, &I.h6, I.extrap);
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

// Predictor history levels of a run: the held count is read once into the
// run's descriptor copy (every phase launch reads that), and the new count is
// stored after the run.
__global__ void k_hist_load(Inst* insts, int ninst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ninst || !insts[i].hist_dev) return;
  const int held = *insts[i].hist_dev;
  insts[i].hist_raw = held;
  insts[i].hist = min(held, insts[i].extrap);
}
__global__ void k_hist_store(const Inst* insts, int ninst, int nsteps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ninst || !insts[i].hist_dev) return;
  *insts[i].hist_dev = min(kMaxExtrap, insts[i].hist_raw + nsteps);
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

// Instances that take part in the next launches: pred 0 the sweep is running,
// 1 the solve is running, 2 the residual pass must be redone.  act_base[i] is
// the instance's first position in the active block list (-1: inactive),
// act_inst lists the active instances in order, cnt = {instances, blocks}.
// One workgroup; instances in chunks of kBlock with a running offset.
__global__ void __launch_bounds__(kBlock) k_compact(const SState* ss, const int* inst_blk, int ninst, int pred,
                                                    int* act_inst, int* act_base, int* cnt) {
  __shared__ int sb[kBlock], si[kBlock];
  __shared__ int tot_b, tot_i;
  const int t = threadIdx.x;
  if (t == 0) tot_b = tot_i = 0;
  __syncthreads();
  for (int i0 = 0; i0 < ninst; i0 += kBlock) {
    const int i = i0 + t;
    int on = 0, nb = 0;
    if (i < ninst) {
      const SState& S = ss[i];
      on = pred == 0 ? !S.sweep_done : pred == 1 ? !S.solve_done : S.redo != 0;
      nb = on ? inst_blk[i + 1] - inst_blk[i] : 0;
    }
    sb[t] = nb;
    si[t] = on;
    __syncthreads();
    for (int d = 1; d < kBlock; d <<= 1) {  // inclusive scans
      const int vb = t >= d ? sb[t - d] : 0, vi = t >= d ? si[t - d] : 0;
      __syncthreads();
      sb[t] += vb;
      si[t] += vi;
      __syncthreads();
    }
    if (i < ninst) {
      act_base[i] = on ? tot_b + sb[t] - nb : -1;
      if (on) act_inst[tot_i + si[t] - 1] = i;
    }
    __syncthreads();
    if (t == kBlock - 1) {
      tot_b += sb[t];
      tot_i += si[t];
    }
    __syncthreads();
  }
  if (t == 0) {
    cnt[0] = tot_i;
    cnt[1] = tot_b;
  }
}

// Active block list: every row block of an active instance, in row-block order.
__global__ void k_expand(const int* blk_inst, const int* inst_blk, const int* act_base, int nblocks, int* act_blk) {
  const int gb = blockIdx.x * blockDim.x + threadIdx.x;
  if (gb >= nblocks) return;
  const int i = blk_inst[gb];
  const int base = act_base[i];
  if (base >= 0) act_blk[base + gb - inst_blk[i]] = gb;
}

template <int K, typename TI>
const void* phase_fn(int ph, bool walk) {
  if (walk) {  // the phases a budgeted run launches over its fixed grid
    switch (ph) {
      case PH_RES: return (const void*)k_sphase<K, TI, PH_RES, true>;
      case PH_PCG_INIT: return (const void*)k_sphase<K, TI, PH_PCG_INIT, true>;
      case PH_PCG_ITER: return (const void*)k_sphase<K, TI, PH_PCG_ITER, true>;
      case PH_CG_A: return (const void*)k_sphase<K, TI, PH_CG_A, true>;
      case PH_CG_B: return (const void*)k_sphase<K, TI, PH_CG_B, true>;
      case PH_BI_P: return (const void*)k_sphase<K, TI, PH_BI_P, true>;
      case PH_BI_S: return (const void*)k_sphase<K, TI, PH_BI_S, true>;
      case PH_BI_X: return (const void*)k_sphase<K, TI, PH_BI_X, true>;
      default: break;
    }
  }
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

const void* pick_phase(int k, int fp32, int ph, bool walk) {
  switch (k) {
    case 1: return fp32 ? phase_fn<1, float>(ph, walk) : phase_fn<1, double>(ph, walk);
    case 2: return fp32 ? phase_fn<2, float>(ph, walk) : phase_fn<2, double>(ph, walk);
    case 4: return fp32 ? phase_fn<4, float>(ph, walk) : phase_fn<4, double>(ph, walk);
    case 16: return fp32 ? phase_fn<16, float>(ph, walk) : phase_fn<16, double>(ph, walk);
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

// Publish one count read to pinned host memory: each value (a flag, or the
// lists' two counts) in a 64-bit word of its own, tagged with the read's
// sequence number in the high half, which the host polls.  A word is written
// by one aligned 8-byte store, so a reader sees a value and its tag together
// and no release fence is needed (one at system scope writes back the L2).
// A device-side write needs no copy command and no event.
__global__ void k_post(const int* src, int nvals, unsigned long long* dst, unsigned seq) {
  if (threadIdx.x != 0) return;
  const unsigned long long tag = (unsigned long long)seq << 32;
  for (int q = 0; q < nvals; q++)
    __hip_atomic_store(dst + q, tag | (unsigned)src[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

bool stream_vec_layout(int k) { return k == 1 || k >= 4; }

// Pinned slots for the host's reads of the list counts: a read is queued in
// one slot while the previous one is still being waited for.  Each slot holds
// two tagged words (instances, blocks; a flag uses the first).
constexpr int kCountSlots = 2, kPostWords = 2;

struct StreamPlan {
  int ninst = 0, nblocks = 0, k = 1, fp32 = 1, kr = 0, rows = 0;
  int stage_rows = 0, stage_nnz = 0;  // largest row block: rows, nonzeros (up to kStageCap)
  size_t lds = 0;                     // dynamic LDS of the staged phases
  size_t lds_rhs = 0;                 // and of the right-hand side's FP64 stage (0: not staged)
  size_t lds_res = 0;                 // and of the FP64 residual's (one value row)
  SState* d_ss = nullptr;
  int* d_blk_inst = nullptr;
  int* d_blk_row = nullptr;
  int* d_blk_end = nullptr;
  int* d_inst_blk = nullptr;
  double* d_partials = nullptr;
  int* d_act_inst = nullptr;  // [ninst]
  int* d_act_base = nullptr;  // [ninst]
  int* d_act_blk = nullptr;   // [nblocks]
  int* d_cnt = nullptr;       // {active instances, active blocks}
  unsigned long long* h_post = nullptr;  // coherent pinned memory: kCountSlots x kPostWords, written by k_post
  unsigned long long* d_post = nullptr;  // its device address
  mutable unsigned post_seq = 0;                // sequence number of the last queued read
  mutable unsigned slot_seq[kCountSlots] = {};  // the read each slot waits for
  mutable int slot_vals[kCountSlots] = {};      // and how many words it writes
};

void stream_plan_free(StreamPlan* p) {
  if (!p) return;
  void* ptrs[] = {p->d_ss,       p->d_blk_inst, p->d_blk_row,  p->d_blk_end, p->d_inst_blk,
                  p->d_partials, p->d_act_inst, p->d_act_base, p->d_act_blk,  p->d_cnt};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  if (p->h_post) (void)hipHostFree(p->h_post);
  delete p;
}

// Row-block table over all instances: up to `rows` rows per block, cut
// shorter where the block's pattern would exceed kStageCap nonzeros.
int stream_plan_build(const std::vector<InstHost>& insts, const Inst* d_insts, int k, int fp32, int kr, int rows,
                      hipStream_t st, StreamPlan** out) {
  (void)d_insts;
  if (rows < 1) {
    set_error("stream engine: rows per block must be positive");
    return FEM_ERR_INVALID;
  }
  StreamPlan* p = new StreamPlan();
  p->ninst = (int)insts.size();
  p->k = k;
  p->fp32 = fp32;
  p->kr = kr;
  p->rows = rows;
  std::vector<int> bi, br, be, ib(p->ninst + 1, 0);
  double s_bytes = 0.0;  // FP64 S rows of all instances
  for (int i = 0; i < p->ninst; i++) {
    ib[i] = (int)bi.size();
    const fem_ctx* c = insts[i].ctx;
    const int n = c->n;
    const std::vector<int>& rp = c->h_rp;
    if ((int)rp.size() != n + 1) {
      delete p;
      set_error("stream engine: context pattern not assembled");
      return FEM_ERR_STATE;
    }
    for (int r = 0; r < n;) {
      int e = std::min(n, r + rows);
      while (e > r + 1 && rp[e] - rp[r] > kStageCap) e--;
      bi.push_back(i);
      br.push_back(r);
      be.push_back(e);
      p->stage_rows = std::max(p->stage_rows, e - r);
      p->stage_nnz = std::max(p->stage_nnz, rp[e] - rp[r]);
      r = e;
    }
    s_bytes += 8.0 * rp[n];
  }
  ib[p->ninst] = (int)bi.size();
  p->nblocks = (int)bi.size();
  const size_t tb = fp32 ? 4 : 8;
  p->lds = 4 * (size_t)((p->stage_rows + 4) & ~3) + 4 * (size_t)((p->stage_nnz + 3) & ~3) + tb * p->stage_nnz;
  if (stream_vec_layout(k) && p->lds > kMaxLds) {
    delete p;
    set_error("stream engine: a matrix row has too many nonzeros for the LDS stage; use the persistent engine");
    return FEM_ERR_INVALID;
  }
  // the right-hand side's stage: pattern and two FP64 value rows, where it fits
  const size_t lds_rhs =
      4 * (size_t)((p->stage_rows + 4) & ~3) + 4 * (size_t)((p->stage_nnz + 3) & ~3) + 16 * (size_t)p->stage_nnz;
  p->lds_rhs = stream_vec_layout(k) && lds_rhs <= kMaxLds ? lds_rhs : 0;
  // the residual's S rows: staged when they exceed half of the 256 MB
  // Infinity Cache (they then come from HBM); below that they stay cached
  // between sweeps and one lane per row reads them faster than a staged copy
  p->lds_res = p->lds_rhs && s_bytes > 128.0 * (1 << 20) ? lds_rhs - 8 * (size_t)p->stage_nnz : 0;
  int rc = 0;
  p->d_ss = dalloc<SState>(p->ninst, &rc);
  p->d_blk_inst = dalloc<int>(std::max(1, p->nblocks), &rc);
  p->d_blk_row = dalloc<int>(std::max(1, p->nblocks), &rc);
  p->d_blk_end = dalloc<int>(std::max(1, p->nblocks), &rc);
  p->d_inst_blk = dalloc<int>(p->ninst + 1, &rc);
  p->d_partials = dalloc<double>((size_t)std::max(1, p->nblocks) * kPartW * k, &rc);
  p->d_act_inst = dalloc<int>(p->ninst, &rc);
  p->d_act_base = dalloc<int>(p->ninst, &rc);
  p->d_act_blk = dalloc<int>(std::max(1, p->nblocks), &rc);
  p->d_cnt = dalloc<int>(2, &rc);
  if (!rc) {
    hipError_t e = hipHostMalloc((void**)&p->h_post, kPostWords * kCountSlots * sizeof(unsigned long long),
                                 hipHostMallocCoherent | hipHostMallocMapped);
    if (e != hipSuccess) rc = hip_fail(e, "hipHostMalloc");
    if (!rc) {
      for (int q = 0; q < kPostWords * kCountSlots; q++) p->h_post[q] = 0;
      e = hipHostGetDevicePointer((void**)&p->d_post, p->h_post, 0);
      if (e != hipSuccess) rc = hip_fail(e, "hipHostGetDevicePointer");
    }
  }
  if (rc) {
    stream_plan_free(p);
    return rc;
  }
  if (p->nblocks > 0) {
    FEM_HIP(hipMemcpyAsync(p->d_blk_inst, bi.data(), sizeof(int) * bi.size(), hipMemcpyHostToDevice, st));
    FEM_HIP(hipMemcpyAsync(p->d_blk_row, br.data(), sizeof(int) * br.size(), hipMemcpyHostToDevice, st));
    FEM_HIP(hipMemcpyAsync(p->d_blk_end, be.data(), sizeof(int) * be.size(), hipMemcpyHostToDevice, st));
  }
  FEM_HIP(hipMemcpyAsync(p->d_inst_blk, ib.data(), sizeof(int) * ib.size(), hipMemcpyHostToDevice, st));
  FEM_HIP(hipStreamSynchronize(st));
  *out = p;
  return 0;
}

namespace {

// The blocks and instances one launch covers: every row block and instance
// (all = true), or the lists of the last compaction, launched over nab blocks
// and nact reduction workgroups (the last counts the host read: a superset of
// the lists, whose own counts the kernels read on the device).
struct Cover {
  bool all;
  int nact, nab;
  bool walk = false;  // a fixed grid walks the lists (budgeted runs)
};

// Host checks of "anything left" in the Krylov loop: every kCheckEvery-th
// iteration, and every iteration from one before the count the same sweep
// ran in the previous step (consecutive steps need similar counts).
// Instances only ever leave the lists, so between checks the launches cover
// a superset and finished blocks exit at once.
constexpr int kCheckEvery = 4;

int launch_phase(const StreamPlan* p, const SArgs& a, int ph, const Cover& cv, hipStream_t st) {
  const void* fn = pick_phase(p->k, p->fp32, ph, !cv.all && cv.walk);
  if (!fn) {
    set_error("stream engine: unsupported column count");
    return FEM_ERR_INVALID;
  }
  const int nab = cv.all ? p->nblocks : cv.nab;
  const int nact = cv.all ? p->ninst : cv.nact;
  if (nab == 0) return 0;
  SArgs aa = a;
  // one instance: its list is all of its blocks or none, in order, so the
  // kernels take the identity with the device's counts (k_expand is skipped)
  const bool lists = !cv.all && p->ninst > 1;
  aa.act_blk = lists ? p->d_act_blk : nullptr;
  aa.act_inst = lists ? p->d_act_inst : nullptr;
  aa.act_cnt = cv.all ? nullptr : p->d_cnt;
  void* args[] = {(void*)&aa};
  const bool staged = stream_vec_layout(p->k) && (ph == PH_PCG_INIT || ph == PH_PCG_ITER || ph == PH_CG_A ||
                                                  ph == PH_BI_P || ph == PH_BI_S);
  const size_t lds = staged ? p->lds : ph == PH_RHS ? p->lds_rhs : ph == PH_RES ? p->lds_res : 0;
  FEM_HIP(hipLaunchKernel(fn, dim3(nab), dim3(kBlock), args, lds, st));
  if (ph == PH_UPD) return 0;
  const void* rf = pick_reduce(p->k, ph);
  const int* ib = p->d_inst_blk;
  int fp32 = p->fp32;
  void* rargs[] = {(void*)&aa, (void*)&ib, (void*)&nact, (void*)&fp32};
  FEM_HIP(hipLaunchKernel(rf, dim3(nact), dim3(kRedBlock), rargs, 0, st));
  return 0;
}

// Take pinned slot `slot` for a read of nvals words under a new sequence
// number, which wait_counts polls for.
unsigned claim_slot(const StreamPlan* p, int slot, int nvals) {
  if (++p->post_seq == 0) ++p->post_seq;  // 0 is the slots' initial tag
  p->slot_seq[slot] = p->post_seq;
  p->slot_vals[slot] = nvals;
  return p->post_seq;
}

// Queue a read of nvals ints at src into pinned slot `slot` (k_post).
int post_read(const StreamPlan* p, const int* src, int nvals, int slot, hipStream_t st) {
  const unsigned seq = claim_slot(p, slot, nvals);
  hipLaunchKernelGGL(k_post, dim3(1), dim3(kWave), 0, st, src, nvals, p->d_post + kPostWords * slot, seq);
  FEM_HIP(hipGetLastError());
  return 0;
}

// Compact the instances and row blocks that satisfy `pred` into the device
// lists (k_compact, k_expand).  post >= 0: also queue the read of the lists'
// counts into that pinned slot (read back by wait_counts).  The launch sizes
// in *cv are left as they are: a superset.
int compact(const StreamPlan* p, int pred, hipStream_t st, Cover* cv, int post = -1) {
  hipLaunchKernelGGL(k_compact, dim3(1), dim3(kBlock), 0, st, p->d_ss, p->d_inst_blk, p->ninst, pred, p->d_act_inst,
                     p->d_act_base, p->d_cnt);
  FEM_HIP(hipGetLastError());
  if (post >= 0) {
    const int rc = post_read(p, p->d_cnt, 2, post, st);
    if (rc) return rc;
  }
  if (p->nblocks > 0 && p->ninst > 1) {  // queued before the host waits, so it overlaps the read-back
    hipLaunchKernelGGL(k_expand, dim3(div_up(p->nblocks, kBlock)), dim3(kBlock), 0, st, p->d_blk_inst, p->d_inst_blk,
                       p->d_act_base, p->nblocks, p->d_act_blk);
    FEM_HIP(hipGetLastError());
  }
  cv->all = false;
  return 0;
}

// One instance: queue the read of its sweep_done flag into `slot` instead of
// a compaction (the kernels skip its blocks by the flag itself).
int post_flag(const StreamPlan* p, int slot, hipStream_t st) {
  return post_read(p, (const int*)((const char*)p->d_ss + offsetof(SState, sweep_done)), 1, slot, st);
}

// Wait for the counts queued in `slot` (flag: a post_flag read, as counts):
// nab == 0 means no block is left.  The host polls the words' tags; past 1 ms
// it also asks whether the stream has drained, so that a read that can no
// longer arrive (an earlier launch failed) ends the wait with an error.
int wait_counts(const StreamPlan* p, int slot, hipStream_t st, Cover* cv, bool flag = false) {
  const unsigned want = p->slot_seq[slot];
  const int nw = p->slot_vals[slot];
  const unsigned long long* h = p->h_post + kPostWords * slot;
  unsigned long long w[kPostWords] = {};
  auto arrived = [&]() {
    for (int q = 0; q < nw; q++) {
      w[q] = __atomic_load_n(h + q, __ATOMIC_ACQUIRE);
      if ((unsigned)(w[q] >> 32) != want) return false;
    }
    return true;
  };
  if (!arrived()) {
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    auto next_query = t0 + std::chrono::milliseconds(1);
    for (unsigned spin = 1; !arrived(); spin++) {
      if (spin % 256) continue;
      const auto now = clk::now();
      if (now < next_query) continue;
      next_query = now + std::chrono::milliseconds(1);
      const hipError_t e = hipStreamQuery(st);
      if (e == hipErrorNotReady) continue;
      if (e != hipSuccess) return hip_fail(e, "hipStreamQuery");
      if (arrived()) break;
      set_error("stream engine: a count read did not arrive although the stream drained");
      return FEM_ERR_HIP;
    }
  }
  const int v0 = (int)(unsigned)w[0];
  const int v1 = nw > 1 ? (int)(unsigned)w[1] : 0;
  cv->all = false;
  if (flag) {
    const bool done = v0 != 0;
    cv->nact = done ? 0 : 1;
    cv->nab = done ? 0 : p->nblocks;
    return 0;
  }
  cv->nact = v0;
  cv->nab = v1;
  return 0;
}

}  // namespace

// Advance every instance by nsteps with the streaming engine.  The instance
// descriptors are already on the device (d_insts); *turned counts the steps
// whose end (and turn of the predictor history) was queued.  ctl: eager or
// budgeted (graph-recorded) control, and the eager run's observations.
int stream_run(const StreamPlan* p, const Inst* d_insts, int max_sweeps, int max_inner, int nsteps,
               hipStream_t st, int* turned, StreamCtl* ctl) {
  *turned = 0;
  const bool budget = ctl->budget;
  if (budget && (ctl->sweeps < 1 || ctl->sweeps > kBudgetSweeps || ctl->grid < 8)) {
    set_error("stream engine: bad iteration budget");
    return FEM_ERR_INVALID;
  }
  SArgs a{};
  a.insts = d_insts;
  a.ss = p->d_ss;
  a.blk_inst = p->d_blk_inst;
  a.blk_row = p->d_blk_row;
  a.blk_end = p->d_blk_end;
  a.nall = p->nblocks;
  a.stage_rows = p->stage_rows;
  a.stage_nnz = p->stage_nnz;
  a.rhs_stage = p->lds_rhs > 0;
  a.res_stage = p->lds_res > 0;
  a.partials = p->d_partials;
  hipLaunchKernelGGL(k_hist_load, dim3(div_up(p->ninst, kBlock)), dim3(kBlock), 0, st, const_cast<Inst*>(d_insts),
                     p->ninst);
  hipLaunchKernelGGL(k_sinit, dim3(div_up(p->ninst, kBlock)), dim3(kBlock), 0, st, p->d_ss, d_insts, p->ninst);
  FEM_HIP(hipGetLastError());
  int rc = 0;
  const Cover full{true, 0, 0};
  // budgeted: the lists' sizes stay on the device; a fixed grid walks them
  const Cover fixed{false, p->ninst, std::min(ctl->grid, std::max(p->nblocks, 8)), true};
  Cover cv{};
  const bool multi = p->ninst > 1;  // one instance: its list is all of its blocks or none
  const bool single = !multi;
  std::vector<int> prev_its(std::max(0, max_sweeps), 0);  // iterations each sweep ran in the last step
  if (!budget) {
    ctl->obs_sweeps = 0;
    for (int& v : ctl->obs_iters) v = 0;
  }
  const int nsweeps = budget ? ctl->sweeps : max_sweeps;
  // Eager runs read the lists' counts on the host, at most one read in
  // flight: each read is queued in a pinned slot and waited for only after
  // the next launches are queued, so the device does not drain while the
  // host waits.  The launches in between cover a superset (instances only
  // leave the lists) and blocks past the device's counts exit at once.
  int slot = 0;
  const Cover every{false, p->ninst, p->nblocks};
#define STREAM_PH(ph, cover)                                     \
  do {                                                           \
    if ((rc = launch_phase(p, a, (ph), (cover), st))) return rc; \
  } while (0)
  for (int step = 0; step < nsteps; step++) {
    a.step = step;
    a.par = 0;
    a.first = 1;
    a.pass = 0;
    STREAM_PH(PH_RHS, full);
    Cover prev_sweep = every;  // a superset of the next sweep's instances
    for (int sw = 0; sw < nsweeps; sw++) {
      // one instance, eager: k_sweep_begin writes the read of its flag
      int sweep_slot = -1;
      unsigned long long* sb_post = nullptr;
      unsigned sb_seq = 0;
      if (single && !budget) {
        slot ^= 1;
        sweep_slot = slot;
        sb_seq = claim_slot(p, slot, 1);
        sb_post = p->d_post + kPostWords * slot;
      }
      hipLaunchKernelGGL(k_sweep_begin, dim3(div_up(p->ninst, kBlock)), dim3(kBlock), 0, st, p->d_ss, d_insts,
                         p->ninst, p->k, p->fp32, sb_post, sb_seq);
      FEM_HIP(hipGetLastError());
      // instances that sweep; eager: their counts arrive after the sweep's
      // first launches, which run over the previous sweep's cover meanwhile
      // (one instance: every launch covers all of its blocks, which skip by
      // its flags, and the host reads its sweep_done flag)
      if (single) {
        cv = full;
      } else if (budget) {
        if ((rc = compact(p, 0, st, &cv))) return rc;
        cv = fixed;
      } else {
        slot ^= 1;
        sweep_slot = slot;
        if ((rc = compact(p, 0, st, &cv, sweep_slot))) return rc;
        cv = prev_sweep;
      }
      Cover sweep = cv;  // covers every instance of this sweep
      bool none = false;
      auto settle = [&]() -> int {
        if (sweep_slot < 0) return 0;
        Cover got{};
        const int r = wait_counts(p, sweep_slot, st, &got, single);
        sweep_slot = -1;
        none = got.nab == 0;
        if (!single) sweep = cv = got;
        return r;
      };
      // one iteration queued: refresh the lists (several instances), take
      // the counts read after the previous iteration, and every
      // kCheckEvery-th time queue another read (budgeted: refresh only)
      int since = 0, it_run = 0, pend = -1, pend_it = 0, need = 0;
      const int expect = sw < (int)prev_its.size() ? prev_its[sw] : 0;
      // before an iteration's launches: whether it ends with a read (the slot,
      // or -1); one instance with row blocks: the iteration's last reduction
      // writes the flag (folded), otherwise k_post after the launches
      int post = -1;
      bool folded = false;
      auto plan_read = [&]() {
        post = -1;
        folded = false;
        a.post = nullptr;
        if (budget) return;
        // one instance with the previous step's count: checks before it rarely
        // end the sweep, so they are sparser
        const int every = single && expect > 0 ? 2 * kCheckEvery : kCheckEvery;
        if (++since >= every || (expect > 0 && it_run + 1 >= expect - 1)) {
          since = 0;
          slot ^= 1;
          post = slot;
          if (single && p->nblocks > 0) {
            a.post_seq = claim_slot(p, slot, 1);
            a.post = p->d_post + kPostWords * slot;
            folded = true;
          }
        }
      };
      auto next = [&](bool* stop) -> int {
        ++it_run;
        *stop = false;
        a.post = nullptr;
        if (budget) {
          if (single) return 0;
          const int r = compact(p, 0, st, &cv);
          cv = fixed;
          return r;
        }
        int r = 0;
        if (multi) {
          if ((r = compact(p, 0, st, &cv, post))) return r;
        } else if (post >= 0 && !folded && (r = post_flag(p, post, st))) {
          return r;
        }
        if (pend >= 0) {
          Cover got{};
          if ((r = wait_counts(p, pend, st, &got, !multi))) return r;
          pend = -1;
          if (got.nab == 0) {  // done after iteration pend_it: this one is empty
            *stop = true;
            need = pend_it;
            if (sw < (int)prev_its.size()) prev_its[sw] = pend_it;
            return 0;
          }
          if (multi) {
            cv.nact = got.nact;
            cv.nab = got.nab;
          }
        }
        if (post >= 0) {
          pend = post;
          pend_it = it_run;
        }
        return 0;
      };
      const int nits = budget ? (sw < kBudgetSweeps ? ctl->iters[sw] : 0) : max_inner;
      bool stop = false;
      int par = 0;
      if (p->kr == 0) {  // pipelined CG
        a.par = 0;
        a.first = 1;
        STREAM_PH(PH_PCG_INIT, cv);
        if ((rc = settle())) return rc;
        for (int it = 0; it < nits && !stop && !none; it++) {
          a.par = par;
          a.first = it == 0;
          plan_read();
          STREAM_PH(PH_PCG_ITER, cv);
          par ^= 1;
          if ((rc = next(&stop))) return rc;
        }
      } else {  // classic CG (2) or BiCGSTAB (1): classic CG's p buffers swap every iteration
        for (int it = 0; it < nits && !stop && !none; it++) {
          a.par = par;
          a.first = it == 0;
          plan_read();
          if (p->kr == 2) {
            STREAM_PH(PH_CG_A, cv);
            STREAM_PH(PH_CG_B, cv);
          } else {
            STREAM_PH(PH_BI_P, cv);
            STREAM_PH(PH_BI_S, cv);
            STREAM_PH(PH_BI_X, cv);
          }
          par ^= 1;
          if (it == 0) {
            if ((rc = settle())) return rc;
            if (none) {
              a.post = nullptr;
              break;
            }
          }
          if ((rc = next(&stop))) return rc;
        }
        if ((rc = settle())) return rc;  // no iteration ran (nits == 0)
      }
      if (none) break;  // no instance swept: its launches were empty
      if (!stop) need = it_run;
      if (!budget && sw < kBudgetSweeps) {
        ctl->obs_sweeps = std::max(ctl->obs_sweeps, sw + 1);
        ctl->obs_iters[sw] = std::max(ctl->obs_iters[sw], need);
      }
      // FP64 residual over the sweep's instances (device lists, no host wait),
      // then the redo pass for columns that produced non-finite values
      cv = sweep;
      a.pass = 0;
      if (!single && (rc = compact(p, 1, st, &cv))) return rc;
      if (budget && !single) cv = fixed;
      STREAM_PH(PH_RES, cv);
      if (!single && (rc = compact(p, 2, st, &cv))) return rc;
      if (budget && !single) cv = fixed;
      a.pass = 1;
      STREAM_PH(PH_RES, cv);
      a.pass = 0;
      prev_sweep = sweep;
    }
    a.more = step + 1 < nsteps;
    STREAM_PH(PH_UPD, full);
    hipLaunchKernelGGL(k_step_end, dim3(div_up(p->ninst, kBlock)), dim3(kBlock), 0, st, p->d_ss,
                       const_cast<Inst*>(d_insts), p->ninst, p->k);
    FEM_HIP(hipGetLastError());
    ++*turned;
  }
#undef STREAM_PH
  hipLaunchKernelGGL(k_hist_store, dim3(div_up(p->ninst, kBlock)), dim3(kBlock), 0, st, d_insts, p->ninst, nsteps);
  FEM_HIP(hipGetLastError());
  return 0;
}

}  // namespace femhip
