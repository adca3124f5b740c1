// noalloc.hip -- fem_ctx_solve_noalloc: one linear solve, as fem_ctx_solve,
// that allocates nothing.
//
// Everything the solve needs beyond the caller's inputs and outputs (the
// formed system matrix and its Jacobi diagonal, the Krylov vectors, the
// right-hand side and solution on the device, the barrier slots of the
// stepping kernel, its launch descriptors and outcome record) lives in one
// file-scope device array, `scratch`, laid out afresh by every call.  Each
// GPU the library runs on holds its own copy of the array until the process
// ends.  Host code never passes the array's name to a kernel: it takes the
// array's device address from hipGetSymbolAddress once per device, with that
// device current, and uses that address everywhere.
//
// Every call on one GPU shares that GPU's array; calls that overlap on one GPU
// (several host threads, or streams the caller synchronises differently) race
// on it.  Each call clears what must start at zero.
#include "femhip_step.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>
#include <string>

namespace femhip {

// The scratch array (size: kScratchBytes below, 32 MiB).
static __device__ __align__(256) unsigned char scratch[size_t(32) << 20];

namespace {

// Largest footprint a call may lay out: 32 MiB per GPU.  One right-hand side
// of an FP32 + refinement solve uses about 12 nnz + 80 N bytes plus 2 KiB
// per workgroup of barrier slots, so meshes up to roughly 180,000 unknowns
// fit at one column (about 25,000 at sixteen).
constexpr size_t kScratchBytes = size_t(32) << 20;

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices];  // device address of `scratch` on each device, once looked up

// Device address of this device's `scratch`, looked up on the first call.
int scratch_ptr(unsigned char** out) {
  int dev = 0;
  FEM_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("fem_ctx_solve_noalloc: device index beyond the library's table");
    return FEM_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) {
      set_error("hipGetSymbolAddress(scratch) returned null");
      return FEM_ERR_HIP;
    }
    g_ptr[dev] = p;
  }
  *out = (unsigned char*)g_ptr[dev];
  return 0;
}

// Bump layout over the scratch array: offsets only, so the footprint is known
// before any device work.
struct Layout {
  size_t used = 0;
  size_t take(size_t bytes) {
    const size_t at = (used + 255) & ~size_t(255);
    used = at + std::max<size_t>(bytes, 1);
    return at;
  }
};

// Writes the launch descriptors of the one-team plan and zeroes the buffer
// the update phase reads as the previous state (a fill kernel: no memset, so
// the sequence stays capturable).
__global__ void k_noalloc_setup(Team team, Inst inst, int grid, Team* d_team, int* d_wg_team, int* d_team_inst,
                                Inst* d_inst, double* zero, long long nzero) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    *d_team = team;
    *d_inst = inst;
    *d_team_inst = 0;
  }
  if (i < grid) d_wg_team[i] = 0;
  for (long long q = i; q < nzero; q += (long long)gridDim.x * blockDim.x) zero[q] = 0.0;
}

}  // namespace

int linear_solve_noalloc(fem_ctx* c, const double coef[3], const fem_solver_opts* opts, const double* bh, double* xh,
                         fem_step_stats* stats) {
  fem_solver_opts o;
  fem_opts_default(&o);
  if (opts) o = *opts;
  const int k = o.nrhs <= 0 ? 1 : o.nrhs;
  if (k != 1 && k != 2 && k != 4 && k != 16) {
    set_error("nrhs must be 1, 2, 4 or 16");
    return FEM_ERR_INVALID;
  }
  if (o.precision != FEM_PREC_FP64 && o.precision != FEM_PREC_FP32_IR) {
    set_error("unknown precision");
    return FEM_ERR_INVALID;
  }
  if (o.engine != FEM_ENGINE_PERSISTENT) {
    set_error("fem_ctx_solve_noalloc runs the persistent engine only");
    return FEM_ERR_INVALID;
  }
  if (stats) std::memset(stats, 0, sizeof(*stats));
  const long long n = c->n, nnz = c->nnz;
  if (n == 0) return FEM_OK;
  const int fp32 = o.precision == FEM_PREC_FP32_IR;
  const size_t tb = fp32 ? 4 : 8;
  int kr = o.krylov;
  if (kr == FEM_KRYLOV_AUTO) kr = coef[2] != 0.0 ? FEM_KRYLOV_BICGSTAB : FEM_KRYLOV_CG;
  const int kri = (kr == FEM_KRYLOV_BICGSTAB) ? 1 : (kr == FEM_KRYLOV_CG_CLASSIC) ? 2 : 0;

  // plan: one team, sized and moded as a one-instance batch plan
  int maxv[3] = {0, 0, 0};
  for (int m = 0; m < 3; m++) {
    const int rc = step_max_coresident(k, fp32, kri, m, &maxv[m]);
    if (rc) return rc;
  }
  const int forced = (o.exchange == FEM_EXCHANGE_COHERENT || o.exchange == FEM_EXCHANGE_FENCED) ? o.exchange : 0;
  int grid = single_team_size(n, k, o.team_size, maxv[kExchangeCoherent]);
  int mode = exchange_mode(grid > 1, team_streams(n, k, grid), forced, 0);
  if (grid > maxv[mode]) {
    grid = single_team_size(n, k, o.team_size, std::min(maxv[0], std::min(maxv[1], maxv[2])));
    mode = exchange_mode(grid > 1, team_streams(n, k, grid), forced, 0);
  }
  if (grid <= 0 || grid > maxv[mode]) {
    set_error("persistent kernel cannot be resident on this device");
    return FEM_ERR_HIP;
  }

  // footprint
  Layout L;
  const size_t o_slots = L.take(sizeof(Slot) * (size_t)grid);
  const size_t o_scal = L.take(sizeof(double) * kSlotBufs * kMaxRed);
  const size_t o_team = L.take(sizeof(Team));
  const size_t o_wgt = L.take(sizeof(int) * (size_t)grid);
  const size_t o_ti = L.take(sizeof(int));
  const size_t o_inst = L.take(sizeof(Inst));
  const size_t o_info = L.take(kInfoBytes + sizeof(int));
  const size_t o_S64 = L.take(sizeof(double) * (size_t)nnz);
  const size_t o_Si = L.take(tb * (size_t)nnz);
  const size_t o_d64 = L.take(sizeof(double) * (size_t)n);
  const size_t o_di = L.take(tb * (size_t)n);
  const size_t o_u = L.take(sizeof(double) * (size_t)n * k);
  const size_t o_x = L.take(sizeof(double) * (size_t)n * k);
  const size_t o_b = L.take(sizeof(double) * (size_t)n * k);
  const size_t o_vec = L.take(tb * (size_t)n * k * 11);
  static_assert(sizeof(scratch) == kScratchBytes, "scratch size");
  if (L.used > kScratchBytes) {
    set_error("fem_ctx_solve_noalloc: this system needs " + std::to_string(L.used) + " bytes of scratch, more than the " +
              std::to_string(kScratchBytes) + " available (use fem_ctx_solve)");
    return FEM_ERR_INVALID;
  }
  unsigned char* base = nullptr;
  int rc = scratch_ptr(&base);
  if (rc) return rc;

  Slot* slots = (Slot*)(base + o_slots);
  double* scal = (double*)(base + o_scal);
  Team* d_team = (Team*)(base + o_team);
  int* d_wg_team = (int*)(base + o_wgt);
  int* d_team_inst = (int*)(base + o_ti);
  Inst* d_inst = (Inst*)(base + o_inst);
  char* info = (char*)(base + o_info);
  int* abort_flag = (int*)(info + kInfoBytes);

  // instance descriptor of one solve (the EQ_SOLVE branch of the stepping
  // kernel: b given, x the guess); the update phase reads u, v, a only as the
  // previous state of time-stepping equations, here the zeroed u
  Inst I;
  std::memset(&I, 0, sizeof(I));
  I.n = (int)n;
  I.nnz = (int)nnz;
  I.k = k;
  I.kv = 1;
  I.eq = EQ_SOLVE;
  I.krylov = kri;
  I.max_row = c->max_row;
  I.warm = o.warm_start;
  I.max_inner = o.max_inner > 0 ? o.max_inner : 200;
  I.max_sweeps = o.max_sweeps > 0 ? o.max_sweeps : 30;
  I.max_iters = o.max_iters > 0 ? o.max_iters : (int)std::max(1LL, 2 * n);
  const double tol = o.tol > 0 ? o.tol : 1e-10;
  const double itol = o.inner_tol > 0 ? o.inner_tol : 1e-4;
  I.tol2 = tol * tol;
  I.inner_tol2 = itol * itol;
  I.dt = 1.0;
  I.p1 = coef[0];
  I.p2 = coef[1];
  I.p3 = coef[2];
  I.rp = c->d_rp;
  I.ci = c->d_ci;
  I.M64 = c->d_M;
  I.K64 = c->d_K;
  I.S64 = (const double*)(base + o_S64);
  I.Si = base + o_Si;
  I.dinv64 = (const double*)(base + o_d64);
  I.dinvi = base + o_di;
  I.lumped = c->d_lumped;
  I.u = (double*)(base + o_u);
  I.v = I.u;
  I.a = I.u;
  I.x = (double*)(base + o_x);
  I.b = (double*)(base + o_b);
  void** slot_ptrs[] = {&I.xi, &I.r, &I.z, &I.p0, &I.p1b, &I.q, &I.rh, &I.v0, &I.v1, &I.s, &I.t};
  for (int q = 0; q < 11; q++) *slot_ptrs[q] = base + o_vec + (size_t)q * tb * n * k;
  I.counters = (long long*)info;
  I.status = (int*)(info + kInfoStatus);
  I.last_res = (double*)(info + kInfoLastRes);
  I.traj_every = 1;

  Team T;
  std::memset(&T, 0, sizeof(T));
  T.first_wg = 0;
  T.size = grid;
  T.inst_begin = 0;
  T.inst_end = 1;
  T.allpoll = ((long long)grid * 3 * k <= 2048) ? 1 : 0;
  T.slots = slots;
  T.scal = scal;
  T.abort_flag = abort_flag;
  T.wait_ticks = kDefaultWaitTicks;

  hipStream_t st = c->stream;
  const auto t0 = std::chrono::steady_clock::now();
  const size_t vbytes = sizeof(double) * (size_t)n * k;
  FEM_HIP(hipMemcpyAsync(I.b, bh, vbytes, hipMemcpyHostToDevice, st));
  if (o.warm_start) FEM_HIP(hipMemcpyAsync(I.x, xh, vbytes, hipMemcpyHostToDevice, st));
  const long long nzero = n * k;
  hipLaunchKernelGGL(k_noalloc_setup, dim3(div_up(std::max<long long>(grid, std::min<long long>(nzero, 1 << 20)),
                                                  kBlock)),
                     dim3(kBlock), 0, st, T, I, grid, d_team, d_wg_team, d_team_inst, d_inst, I.u, nzero);
  FEM_HIP(hipGetLastError());
  rc = form_system(c, coef, fp32, (double*)I.S64, (void*)I.Si, (double*)I.dinv64, (void*)I.dinvi, st);
  if (rc) return rc;
  rc = launch_slot_reset(slots, grid, scal, 1, abort_flag, d_inst, 1, st);
  if (rc) return rc;
  StepLaunch SL;
  SL.teams = d_team;
  SL.wg_team = d_wg_team;
  SL.insts = d_inst;
  SL.team_inst = d_team_inst;
  SL.grid = grid;
  SL.k = k;
  SL.inner_fp32 = fp32;
  SL.kr = kri;
  SL.coh = mode;
  rc = launch_step(SL, 1, st);
  if (rc) return rc;
  alignas(8) char rec[kInfoBytes + sizeof(int)];
  FEM_HIP(hipMemcpyAsync(rec, info, sizeof(rec), hipMemcpyDeviceToHost, st));
  FEM_HIP(hipMemcpyAsync(xh, I.x, vbytes, hipMemcpyDeviceToHost, st));
  FEM_HIP(hipStreamSynchronize(st));
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();

  const long long* cnt = (const long long*)rec;
  const int status = *(const int*)(rec + kInfoStatus);
  const double* lr = (const double*)(rec + kInfoLastRes);
  double res = 0.0;
  for (int j = 0; j < k; j++) res = std::max(res, lr[j]);
  if (stats) {
    stats->iterations = cnt[0];
    stats->sweeps = cnt[1];
    stats->steps = cnt[2];
    stats->residual = res;
    stats->ms = ms;
    stats->status = status;
  }
  if (*(const int*)(rec + kInfoBytes) || (status & ST_BARRIER_TIMEOUT)) {
    set_error("device barrier wait timed out; results of this call are invalid");
    return FEM_ERR_TIMEOUT;
  }
  if (status & (ST_UNSTABLE | ST_NONFINITE)) return FEM_DIVERGED;
  if (status & ST_NOT_CONVERGED) return FEM_NOT_CONVERGED;
  return FEM_OK;
}

}  // namespace femhip
