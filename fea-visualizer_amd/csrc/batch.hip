// batch.hip -- host side of the solver: per-instance buffers, system-matrix
// forming, state transfer, the team plan (which workgroups advance which
// instance) and the launch of the persistent stepping kernel.
//
// A single solver is a batch of one instance.  The system matrices of the
// reference (CPUSolver.cpp:77,95,113,133,136) are formed once per parameter
// change instead of once per step: they only depend on the parameters, so the
// result is the same and the per-step cost is the solve alone.
#include "femhip_step.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

namespace femhip {

namespace {

struct Coef {
  double cM[kMaxK], cK[kMaxK], cA[kMaxK];
};

// tperm (optional) forms the transpose: entry (r, c) takes the values of (c, r).
template <typename TI>
__global__ void k_form(int n, const int* rp, const int* diag, const int* tperm, const double* M, const double* Kk,
                       const double* A, int kv, Coef cf, double* S64, TI* Si, double* dinv64, TI* dinvi) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  for (int s = rp[r]; s < rp[r + 1]; s++) {
    const int t = tperm ? tperm[s] : s;
    const double m = M[t], kk = Kk[t], a = A[t];
    for (int j = 0; j < kv; j++) {
      double val = 0.0;
      if (cf.cM[j] != 0.0) val += cf.cM[j] * m;
      if (cf.cK[j] != 0.0) val += cf.cK[j] * kk;
      if (cf.cA[j] != 0.0) val += cf.cA[j] * a;
      S64[(size_t)s * kv + j] = val;
      Si[(size_t)s * kv + j] = (TI)val;
    }
  }
  // Eigen DiagonalPreconditioner: 1/a_jj, or 1 when absent or zero
  const int d = diag[r];
  for (int j = 0; j < kv; j++) {
    const double dv = (d >= 0) ? S64[(size_t)d * kv + j] : 0.0;
    const double inv = (dv != 0.0) ? 1.0 / dv : 1.0;
    dinv64[(size_t)r * kv + j] = inv;
    dinvi[(size_t)r * kv + j] = (TI)inv;
  }
}

// Column-scaled inner matrix S_ij D^-1_jj (one value column).
template <typename TI>
__global__ void k_colscale(int n, const int* rp, const int* ci, const TI* Si, const TI* dinvi, TI* SDi) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  for (int s = rp[r]; s < rp[r + 1]; s++) SDi[s] = Si[s] * dinvi[ci[s]];
}

// Barrier epochs restart at 1 every launch.  Barrier e uses buffer e & 3 and
// tag (e >> 2) & 1 in the lowest mantissa bit, so before a launch buffer 0
// must carry tag 0 and buffers 1-3 tag 1 (none of them looks "arrived").
// (A fill kernel rather than a memset node, so the launch sequence can be
// captured in a graph.)
__global__ void k_reset(Slot* slots, int nslots, double* scal, int nteams, int* abort_flag, Inst* insts,
                        int ninst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const double tag1 = __longlong_as_double(1ll);  // smallest denormal: value 0, tag 1
  const int per = kSlotBufs * kMaxRed;
  if (i < nslots + nteams) {
    double* v = (i < nslots) ? &slots[i].vals[0][0] : scal + (size_t)(i - nslots) * per;
    for (int q = 0; q < per; q++) v[q] = (q < kMaxRed) ? 0.0 : tag1;
  }
  if (i == 0) *abort_flag = 0;
  if (i < ninst) {
    insts[i].counters[0] = 0;
    insts[i].counters[1] = 0;
    insts[i].counters[2] = 0;
    *insts[i].status = 0;
  }
}

// Column copy between an [n][nrhs] host-layout array and an [n][k] device array.
__global__ void k_put_cols(double* dst, int k, int col0, const double* src, int ncols, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * ncols) return;
  long long r = i / ncols;
  int j = (int)(i % ncols);
  dst[r * k + col0 + j] = src[i];
}
__global__ void k_get_cols(const double* src, int k, int col0, double* dst, int ncols, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * ncols) return;
  long long r = i / ncols;
  int j = (int)(i % ncols);
  dst[i] = src[r * k + col0 + j];
}

void free_inst(InstHost& h) {
  void* ptrs[] = {h.d_S64, h.d_Si,    h.d_dinv64, h.d_dinvi, h.d_SDi, h.d_u, h.d_v,   h.d_a,
                  h.d_h3,  h.d_h4,    h.d_xnext,  h.d_hpre,  h.d_hvec, h.d_x, h.d_b, h.d_vec,
                  h.d_surf, h.d_h5, h.d_h6};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (h.surf_done) (void)hipEventSynchronize(h.surf_done);
  if (h.h_surf) (void)hipHostFree(h.h_surf);
  if (h.surf_done) (void)hipEventDestroy(h.surf_done);
  fem_ctx* c = h.ctx;
  const int tr = h.transpose;
  h = InstHost();
  h.ctx = c;
  h.transpose = tr;
}

size_t inner_bytes(const fem_batch* b) { return b->opts.precision == FEM_PREC_FP32_IR ? 4 : 8; }

int alloc_inst(fem_batch* b, InstHost& h, int k, int kv) {
  free_inst(h);
  int rc = 0;
  const long long n = h.ctx->n, nnz = h.ctx->nnz;
  const size_t tb = inner_bytes(b);
  h.k = k;
  h.kv = kv;
  h.d_S64 = dalloc<double>((size_t)nnz * kv, &rc);
  h.d_Si = (void*)dalloc<char>((size_t)nnz * kv * tb, &rc);
  h.d_dinv64 = dalloc<double>((size_t)n * kv, &rc);
  h.d_dinvi = (void*)dalloc<char>((size_t)n * kv * tb, &rc);
  if (kv == 1 && stream_vec_layout(k)) h.d_SDi = (void*)dalloc<char>((size_t)nnz * tb, &rc);
  h.d_u = dalloc<double>((size_t)n * k, &rc);
  h.d_v = dalloc<double>((size_t)n * k, &rc);
  h.d_a = dalloc<double>((size_t)n * k, &rc);
  h.d_x = dalloc<double>((size_t)n * k, &rc);
  h.d_b = dalloc<double>((size_t)n * k, &rc);
  h.d_vec = (void*)dalloc<char>((size_t)n * k * tb * 11, &rc);
  h.d_surf = dalloc<float>((size_t)h.ctx->nv * k, &rc);
  if (rc) return rc;
  char* info = b->d_info + (size_t)(&h - b->insts.data()) * kInfoBytes;
  h.d_counters = (long long*)info;
  h.d_status = (int*)(info + kInfoStatus);
  h.d_last_res = (double*)(info + kInfoLastRes);
  h.d_hist = (int*)(info + kInfoHist);
  {
    hipError_t e = hipHostMalloc((void**)&h.h_surf, sizeof(float) * std::max<size_t>(1, (size_t)h.ctx->nv * k),
                                 hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&h.surf_done, hipEventDisableTiming);
    if (e != hipSuccess) return hip_fail(e, "alloc_inst: surface staging");
  }
  const size_t vb = (size_t)n * k * sizeof(double);
  FEM_HIP(hipMemsetAsync(h.d_u, 0, vb, b->stream));
  FEM_HIP(hipMemsetAsync(h.d_v, 0, vb, b->stream));
  FEM_HIP(hipMemsetAsync(h.d_a, 0, vb, b->stream));
  FEM_HIP(hipMemsetAsync(h.d_x, 0, vb, b->stream));
  FEM_HIP(hipMemsetAsync(h.d_b, 0, vb, b->stream));
  FEM_HIP(hipMemsetAsync(h.d_vec, 0, (size_t)n * k * tb * 11, b->stream));
  FEM_HIP(hipMemsetAsync(h.d_status, 0, 4 * sizeof(int), b->stream));
  FEM_HIP(hipMemsetAsync(h.d_counters, 0, 4 * sizeof(long long), b->stream));
  FEM_HIP(hipMemsetAsync(h.d_last_res, 0, kMaxK * sizeof(double), b->stream));
  FEM_HIP(hipMemsetAsync(h.d_hist, 0, sizeof(int), b->stream));
  h.dirty = true;
  return 0;
}

int form_inst(fem_batch* b, InstHost& h) {
  fem_ctx* c = h.ctx;
  Coef cf;
  std::memset(&cf, 0, sizeof(cf));
  const double dt = h.dt;
  switch (h.eq) {
    case FEM_HEAT:  // M/dt + c K                                  CPUSolver.cpp:77
      cf.cM[0] = 1.0 / dt;
      cf.cK[0] = h.p1;
      break;
    case FEM_ADVECTION_DIFFUSION:  // M/dt + c K - A                CPUSolver.cpp:95
      cf.cM[0] = 1.0 / dt;
      cf.cK[0] = h.p1;
      cf.cA[0] = -1.0;
      break;
    case FEM_WAVE:  // M/dt + c^2 K dt                              CPUSolver.cpp:113
      cf.cM[0] = 1.0 / dt;
      cf.cK[0] = h.p1 * h.p1 * dt;
      break;
    case FEM_REACTION_DIFFUSION:  // M/dt + Du K ; M/dt + Dv K      CPUSolver.cpp:133,136
      cf.cM[0] = 1.0 / dt;
      cf.cK[0] = h.p1;
      cf.cM[1] = 1.0 / dt;
      cf.cK[1] = h.p2;
      break;
    case FEM_WAVE_NEWMARK:  // M + beta dt^2 c^2 K
      cf.cM[0] = 1.0;
      cf.cK[0] = h.p2 * dt * dt * h.p1 * h.p1;
      break;
    case EQ_SOLVE:  // one linear solve with caller-given coefficients
      cf.cM[0] = h.p1;
      cf.cK[0] = h.p2;
      cf.cA[0] = h.p3;
      break;
    default:
      set_error("form: equation not set");
      return FEM_ERR_STATE;
  }
  if (c->n > 0) {
    const int* tp = h.transpose ? c->d_tperm : nullptr;
    if (inner_bytes(b) == 4)
      hipLaunchKernelGGL(k_form<float>, dim3(div_up(c->n, kBlock)), dim3(kBlock), 0, b->stream, c->n, c->d_rp,
                         c->d_diag, tp, c->d_M, c->d_K, c->d_A, h.kv, cf, h.d_S64, (float*)h.d_Si, h.d_dinv64,
                         (float*)h.d_dinvi);
    else
      hipLaunchKernelGGL(k_form<double>, dim3(div_up(c->n, kBlock)), dim3(kBlock), 0, b->stream, c->n, c->d_rp,
                         c->d_diag, tp, c->d_M, c->d_K, c->d_A, h.kv, cf, h.d_S64, (double*)h.d_Si, h.d_dinv64,
                         (double*)h.d_dinvi);
    FEM_HIP(hipGetLastError());
    if (h.d_SDi) {
      if (inner_bytes(b) == 4)
        hipLaunchKernelGGL(k_colscale<float>, dim3(div_up(c->n, kBlock)), dim3(kBlock), 0, b->stream, c->n, c->d_rp,
                           c->d_ci, (const float*)h.d_Si, (const float*)h.d_dinvi, (float*)h.d_SDi);
      else
        hipLaunchKernelGGL(k_colscale<double>, dim3(div_up(c->n, kBlock)), dim3(kBlock), 0, b->stream, c->n,
                           c->d_rp, c->d_ci, (const double*)h.d_Si, (const double*)h.d_dinvi, (double*)h.d_SDi);
      FEM_HIP(hipGetLastError());
    }
  }
  h.formed_version = c->version;
  h.dirty = false;
  return 0;
}

void fill_dev(fem_batch* b, InstHost& h) {
  fem_ctx* c = h.ctx;
  Inst& I = h.dev;
  std::memset(&I, 0, sizeof(I));
  const fem_solver_opts& o = b->opts;
  I.n = c->n;
  I.nnz = (int)c->nnz;
  I.k = h.k;
  I.kv = h.kv;
  I.eq = (h.eq_override >= 0) ? h.eq_override : h.eq;
  int kr = o.krylov;
  if (kr == FEM_KRYLOV_AUTO)
    kr = (h.eq == FEM_ADVECTION_DIFFUSION || (h.eq == EQ_SOLVE && h.p3 != 0.0)) ? FEM_KRYLOV_BICGSTAB
                                                                              : FEM_KRYLOV_CG;
  I.krylov = (kr == FEM_KRYLOV_BICGSTAB) ? 1 : (kr == FEM_KRYLOV_CG_CLASSIC) ? 2 : 0;
  I.warm = o.warm_start;
  I.max_inner = o.max_inner > 0 ? o.max_inner : 200;
  I.max_sweeps = o.max_sweeps > 0 ? o.max_sweeps : 30;
  I.max_iters = o.max_iters > 0 ? o.max_iters : std::max(1, 2 * c->n);  // Eigen default 2N
  I.weighting = h.weighting;
  I.max_row = c->max_row;
  const double tol = o.tol > 0 ? o.tol : 1e-10;
  const double itol = o.inner_tol > 0 ? o.inner_tol : 1e-4;
  I.tol2 = tol * tol;
  I.inner_tol2 = itol * itol;
  I.dt = h.dt;
  I.p1 = h.p1;
  I.p2 = h.p2;
  I.p3 = h.p3;
  I.p4 = h.p4;
  I.rp = c->d_rp;
  I.ci = c->d_ci;
  I.M64 = c->d_M;
  I.K64 = c->d_K;
  I.S64 = h.d_S64;
  I.Si = h.d_Si;
  I.dinv64 = h.d_dinv64;
  I.dinvi = h.d_dinvi;
  I.SDi = h.d_SDi;
  I.lumped = c->d_lumped;
  I.u = h.d_u;
  I.v = h.d_v;
  I.a = h.d_a;
  I.h3 = h.d_h3;
  I.h4 = h.d_h4;
  I.h5 = h.d_h5;
  I.h6 = h.d_h6;
  I.xnext = h.d_xnext;
  I.hpre = h.d_hpre;
  I.hvec = h.d_hvec;
  I.x = h.d_x;
  I.b = h.d_b;
  char* base = (char*)h.d_vec;
  const size_t vb = (size_t)c->n * h.k * inner_bytes(b);
  void** slots[] = {&I.xi, &I.r, &I.z, &I.p0, &I.p1b, &I.q, &I.rh, &I.v0, &I.v1, &I.s, &I.t};
  for (int q = 0; q < 11; q++) *slots[q] = (void*)(base + q * vb);
  I.status = h.d_status;
  I.counters = h.d_counters;
  I.last_res = h.d_last_res;
  I.traj = h.traj;
  I.traj_every = h.traj ? std::max(1, h.traj_every) : 1;
  // predictor order: the quadratic one keeps x_{n-2} in `v`, which only the
  // equations without a velocity field leave free (the wave's `v` is its state)
  const bool first_order = h.eq == FEM_HEAT || h.eq == FEM_ADVECTION_DIFFUSION || h.eq == FEM_REACTION_DIFFUSION;
  I.extrap = 0;
  if (o.warm_start && h.eq_override < 0 && (first_order || h.eq == FEM_WAVE)) {
    if (o.predictor == FEM_PREDICTOR_LINEAR) I.extrap = 1;
    else if (o.predictor == FEM_PREDICTOR_QUADRATIC) I.extrap = first_order ? 2 : 1;
    else if (o.predictor == FEM_PREDICTOR_CUBIC) I.extrap = !first_order ? 1 : h.d_h3 ? 3 : 2;
    else if (o.predictor == FEM_PREDICTOR_QUARTIC) I.extrap = !first_order ? 1 : h.d_h4 ? 4 : h.d_h3 ? 3 : 2;
    else if (o.predictor == FEM_PREDICTOR_QUINTIC || o.predictor == FEM_PREDICTOR_SEXTIC) {
      const int want = o.predictor == FEM_PREDICTOR_SEXTIC ? 6 : 5;  // the persistent plan uses 4 at most
      I.extrap = !first_order ? 1 : (want >= 6 && h.d_h6) ? 6 : h.d_h5 ? 5 : h.d_h4 ? 4 : h.d_h3 ? 3 : 2;
    }
  } else if (o.warm_start && h.eq_override < 0 && h.eq == FEM_WAVE_NEWMARK && o.predictor != FEM_PREDICTOR_PREVIOUS) {
    // Newmark extrapolates the acceleration (history in h3, h4): linear or quadratic
    I.extrap = (o.predictor != FEM_PREDICTOR_LINEAR && h.d_h4) ? 2 : h.d_h3 ? 1 : 0;
  }
  I.hist = 0;  // the kernels take the held levels from hist_dev
  I.hist_dev = h.d_hist;
}

void free_plan(fem_batch* b) {
  void* ptrs[] = {b->d_insts, b->d_teams, b->d_wg_team, b->d_team_inst, b->d_slots, b->d_scal};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  b->d_insts = nullptr;
  b->d_teams = nullptr;
  b->d_wg_team = nullptr;
  b->d_team_inst = nullptr;
  b->d_slots = nullptr;
  b->d_scal = nullptr;
  b->uploaded.clear();
  b->plan_ready = false;
}

}  // namespace

// Workgroups of the team of a lone instance, within maxb co-resident
// workgroups: team_size when given and it fits, else about one row per lane
// (at most 1024).
int single_team_size(long long n, int k, int team_size, int maxb) {
  if (team_size > 0 && team_size <= maxb) return team_size;
  const long long lanes = n * k;
  const int g = (int)std::max(1LL, (lanes + kBlock - 1) / kBlock);
  return std::min(g, std::min(maxb, 1024));
}

// A workgroup of a multi-workgroup team whose rows exceed two per lane (the
// kernel's LDS row cache) streams its phases.
bool team_streams(long long n, int k, int size) {
  const long long rows = (n + size - 1) / size;
  return rows * k > 2LL * kBlock;
}

// Exchange mode of a plan: plain for one-workgroup teams only; else forced,
// then the autotuner's choice, then fenced for streaming teams, coherent
// otherwise.
int exchange_mode(bool multi, bool streaming, int forced, int tuned) {
  return !multi ? 0 : forced ? forced : tuned ? tuned : streaming ? kExchangeFenced : kExchangeCoherent;
}

// S = coef[0] M + coef[1] K + coef[2] A (one value column) and its Jacobi
// inverse diagonal, FP64 and inner precision, into caller-given arrays.
int form_system(fem_ctx* c, const double coef[3], int fp32, double* S64, void* Si, double* dinv64, void* dinvi,
                hipStream_t st) {
  if (c->n == 0) return 0;
  Coef cf;
  std::memset(&cf, 0, sizeof(cf));
  cf.cM[0] = coef[0];
  cf.cK[0] = coef[1];
  cf.cA[0] = coef[2];
  if (fp32)
    hipLaunchKernelGGL(k_form<float>, dim3(div_up(c->n, kBlock)), dim3(kBlock), 0, st, c->n, c->d_rp, c->d_diag,
                       nullptr, c->d_M, c->d_K, c->d_A, 1, cf, S64, (float*)Si, dinv64, (float*)dinvi);
  else
    hipLaunchKernelGGL(k_form<double>, dim3(div_up(c->n, kBlock)), dim3(kBlock), 0, st, c->n, c->d_rp, c->d_diag,
                       nullptr, c->d_M, c->d_K, c->d_A, 1, cf, S64, (double*)Si, dinv64, (double*)dinvi);
  FEM_HIP(hipGetLastError());
  return 0;
}

// Barrier slots, team sums, abort flag and the instances' outcome records
// reset before a launch of the stepping kernel (a fill kernel).
int launch_slot_reset(Slot* slots, int nslots, double* scal, int nteams, int* abort_flag, Inst* insts, int ninst,
                      hipStream_t st) {
  hipLaunchKernelGGL(k_reset, dim3(div_up(std::max(nteams + nslots, ninst), kBlock)), dim3(kBlock), 0, st, slots,
                     nslots, scal, nteams, abort_flag, insts, ninst);
  FEM_HIP(hipGetLastError());
  return 0;
}

namespace {

// Team plan.  Work of an instance ~ (nnz + n) * k per Krylov iteration.  With
// one instance the team takes ~4 rows per lane, capped by the co-resident
// grid.  With many, instances heavier than the per-workgroup share get teams
// proportional to their work; the rest are packed longest-first onto
// single-workgroup teams (each walks its instances in turn).

// The persistent kernel's copy of a descriptor: its predictor is quartic at
// most (the history levels beyond h4 belong to the streaming engine).
static Inst persistent_desc(const Inst& d) {
  Inst I = d;
  I.extrap = std::min(I.extrap, kPersistentExtrap);
  return I;
}

int build_plan(fem_batch* b) {
  free_plan(b);
  const int ninst = (int)b->insts.size();
  const int k = b->insts[0].k;
  for (auto& h : b->insts)
    if (h.k != k) {
      set_error("batch instances must share the number of right-hand side columns");
      return FEM_ERR_INVALID;
    }
  const int kr = b->insts[0].dev.krylov;
  for (auto& h : b->insts)
    if (h.dev.krylov != kr) {
      set_error("batch instances must share the Krylov method (set opts.krylov explicitly)");
      return FEM_ERR_INVALID;
    }
  const int inner_fp32 = inner_bytes(b) == 4;
  // co-resident workgroups of each variant: plain (one-workgroup teams),
  // coherent and fenced exchange; they differ in registers
  int maxv[3] = {0, 0, 0};
  int rc = 0;
  for (int m = 0; m < 3 && !rc; m++) rc = step_max_coresident(k, inner_fp32, kr, m, &maxv[m]);
  if (rc) return rc;
  if (std::min(maxv[0], std::min(maxv[1], maxv[2])) <= 0) {
    set_error("persistent kernel cannot be resident on this device");
    return FEM_ERR_HIP;
  }
  std::vector<double> work(ninst);
  double W = 0;
  for (int i = 0; i < ninst; i++) {
    work[i] = (double)(b->insts[i].ctx->nnz + b->insts[i].ctx->n) * k + 1.0;
    W += work[i];
  }
  std::vector<std::vector<int>> teams;  // instance lists
  std::vector<int> sizes;
  // Plan within `maxb` co-resident workgroups.
  auto make_plan = [&](int maxb) {
    teams.clear();
    sizes.clear();
    if (ninst == 1) {
      teams.push_back({0});
      sizes.push_back(single_team_size(b->insts[0].ctx->n, k, b->opts.team_size, maxb));
    } else if (b->opts.team_size > 0 && (long long)b->opts.team_size * ninst <= maxb) {
      for (int i = 0; i < ninst; i++) {
        teams.push_back({i});
        sizes.push_back(b->opts.team_size);
      }
    } else {
      const int budget = (b->grid_budget > 0) ? std::min(b->grid_budget, maxb) : maxb;
      const double share = W / budget;
      std::vector<int> small;
      int used = 0;
      for (int i = 0; i < ninst; i++) {
        int g = (int)std::floor(work[i] / share);
        if (g >= 2) {
          teams.push_back({i});
          sizes.push_back(g);
          used += g;
        } else {
          small.push_back(i);
        }
      }
      if (used > budget) {  // rounding guard
        double f = (double)budget / (used + (double)small.size());
        used = 0;
        for (auto& g : sizes) {
          g = std::max(1, (int)(g * f));
          used += g;
        }
      }
      int bins = std::max(1, std::min((int)small.size(), std::max(1, budget - used)));
      std::sort(small.begin(), small.end(), [&](int x, int y) { return work[x] > work[y]; });
      std::vector<double> load(bins, 0.0);
      std::vector<std::vector<int>> lists(bins);
      for (int i : small) {
        int m = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        lists[m].push_back(i);
        load[m] += work[i];
      }
      for (int m = 0; m < bins; m++)
        if (!lists[m].empty()) {
          teams.push_back(lists[m]);
          sizes.push_back(1);
        }
    }
  };
  // Exchange mode of a plan: a workgroup of a multi-workgroup team whose rows
  // exceed two per lane (the LDS row cache) streams its phases: fenced
  // barriers.  Plans of one-workgroup teams run the plain variant.
  const int forced = (b->opts.exchange == FEM_EXCHANGE_COHERENT || b->opts.exchange == FEM_EXCHANGE_FENCED)
                         ? b->opts.exchange : 0;
  auto plan_mode = [&]() {
    bool multi = false, streaming = false;
    for (size_t t = 0; t < teams.size(); t++) {
      if (sizes[t] <= 1) continue;
      multi = true;
      for (int i : teams[t])
        if (team_streams(b->insts[i].ctx->n, k, sizes[t])) streaming = true;
    }
    return exchange_mode(multi, streaming, forced, b->mode_override);
  };
  auto plan_grid = [&]() {
    int g = 0;
    for (int s_ : sizes) g += s_;
    return g;
  };
  // Size for the coherent variant first; if the variant the plan then picks
  // holds fewer workgroups, re-plan within the smallest limit of all three,
  // which any mode fits.
  make_plan(maxv[kExchangeCoherent]);
  int mode = plan_mode();
  if (plan_grid() > maxv[mode]) {
    make_plan(std::min(maxv[0], std::min(maxv[1], maxv[2])));
    mode = plan_mode();
  }
  const int maxb = maxv[mode];
  const int nteams = (int)teams.size();
  int grid = 0;
  for (int g : sizes) grid += g;
  if (grid > maxb) {
    set_error("team plan exceeds the co-resident grid");
    return FEM_ERR_INVALID;
  }
  std::vector<Team> ht(nteams);
  std::vector<int> wg_team(grid), team_inst;
  int wg = 0;
  for (int t = 0; t < nteams; t++) {
    ht[t].first_wg = wg;
    ht[t].size = sizes[t];
    ht[t].inst_begin = (int)team_inst.size();
    for (int i : teams[t]) team_inst.push_back(i);
    ht[t].inst_end = (int)team_inst.size();
    for (int q = 0; q < sizes[t]; q++) wg_team[wg++] = t;
  }
  b->d_insts = dalloc<Inst>(ninst, &rc);
  b->d_teams = dalloc<Team>(nteams, &rc);
  b->d_wg_team = dalloc<int>(grid, &rc);
  b->d_team_inst = dalloc<int>(team_inst.size(), &rc);
  b->d_slots = dalloc<Slot>(grid, &rc);
  b->d_scal = dalloc<double>((size_t)nteams * kSlotBufs * kMaxRed, &rc);
  if (rc) return rc;
  int poff = 0;
  for (int t = 0; t < nteams; t++) {
    ht[t].slots = b->d_slots + poff;
    ht[t].scal = b->d_scal + (size_t)t * kSlotBufs * kMaxRed;
    ht[t].abort_flag = b->d_abort;
    ht[t].wait_ticks = b->wait_ticks;
    // every workgroup reading every partial costs size^2 * nvals loads per
    // barrier; beyond ~16 KB per workgroup rank 0 reduces and publishes
    ht[t].allpoll = b->poll_override == 1   ? 1
                    : b->poll_override == 2 ? 0
                                            : ((long long)sizes[t] * 3 * k <= 2048) ? 1 : 0;
    poff += sizes[t];
  }
  std::vector<Inst> hi(ninst);
  for (int i = 0; i < ninst; i++) hi[i] = persistent_desc(b->insts[i].dev);
  FEM_HIP(hipMemcpyAsync(b->d_insts, hi.data(), sizeof(Inst) * ninst, hipMemcpyHostToDevice, b->stream));
  b->uploaded = hi;
  FEM_HIP(hipMemcpyAsync(b->d_teams, ht.data(), sizeof(Team) * nteams, hipMemcpyHostToDevice, b->stream));
  FEM_HIP(hipMemcpyAsync(b->d_wg_team, wg_team.data(), sizeof(int) * grid, hipMemcpyHostToDevice, b->stream));
  FEM_HIP(hipMemcpyAsync(b->d_team_inst, team_inst.data(), sizeof(int) * team_inst.size(), hipMemcpyHostToDevice,
                         b->stream));
  FEM_HIP(hipStreamSynchronize(b->stream));
  b->launch.teams = b->d_teams;
  b->launch.wg_team = b->d_wg_team;
  b->launch.insts = b->d_insts;
  b->launch.team_inst = b->d_team_inst;
  b->launch.grid = grid;
  b->launch.k = k;
  b->launch.inner_fp32 = inner_fp32;
  b->launch.kr = kr;
  b->launch.coh = mode;
  b->nteams = nteams;
  b->grid = grid;
  b->team_sizes = sizes;
  b->plan_ready = true;
  b->plan_gen++;
  return 0;
}

}  // namespace

void batch_free(fem_batch* b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (auto& h : b->insts) free_inst(h);
  free_plan(b);
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  if (b->d_plain) (void)hipFree(b->d_plain);
  if (b->splan) stream_plan_free(b->splan);
  if (b->d_sinsts) (void)hipFree(b->d_sinsts);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->d_info) (void)hipFree(b->d_info);
  if (b->h_info) (void)hipHostFree(b->h_info);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int batch_create(fem_ctx* const* ctxs, int count, const fem_solver_opts* opts, fem_batch** out) {
  if (count <= 0 || !ctxs) {
    set_error("batch_create: no contexts");
    return FEM_ERR_INVALID;
  }
  fem_batch* b = new fem_batch();
  fem_opts_default(&b->opts);
  if (opts) b->opts = *opts;
  if (b->opts.nrhs <= 0) b->opts.nrhs = 1;
  const int nr = b->opts.nrhs;
  if (nr != 1 && nr != 2 && nr != 4 && nr != 16) {
    delete b;
    set_error("nrhs must be 1, 2, 4 or 16");
    return FEM_ERR_INVALID;
  }
  if (b->opts.precision != FEM_PREC_FP64 && b->opts.precision != FEM_PREC_FP32_IR) {
    delete b;
    set_error("unknown precision");
    return FEM_ERR_INVALID;
  }
  if (b->opts.engine != FEM_ENGINE_PERSISTENT && b->opts.engine != FEM_ENGINE_PLAIN &&
      b->opts.engine != FEM_ENGINE_STREAM && b->opts.engine != FEM_ENGINE_GL_COMPAT) {
    delete b;
    set_error("unknown engine");
    return FEM_ERR_INVALID;
  }
  b->nrhs = nr;
  hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&b->ev0);
  if (e == hipSuccess) e = hipEventCreate(&b->ev1);
  if (e != hipSuccess) {
    batch_free(b);
    return hip_fail(e, "batch_create");
  }
  b->insts.resize(count);
  // outcome records of every instance, then the abort and surface-change
  // flags: one copy brings all of them back after a launch
  b->info_bytes = (size_t)count * kInfoBytes + sizeof(int);
  {
    int rc = 0;
    b->d_info = dalloc<char>(b->info_bytes, &rc);
    if (!rc) {
      e = hipHostMalloc((void**)&b->h_info, b->info_bytes, hipHostMallocDefault);
      if (e != hipSuccess) rc = hip_fail(e, "batch_create: hipHostMalloc");
    }
    if (!rc) {
      e = hipMemsetAsync(b->d_info, 0, b->info_bytes, b->stream);
      if (e != hipSuccess) rc = hip_fail(e, "batch_create: hipMemsetAsync");
    }
    if (rc) {
      batch_free(b);
      return rc;
    }
    b->d_abort = (int*)(b->d_info + (size_t)count * kInfoBytes);
  }
  for (int i = 0; i < count; i++) {
    if (!ctxs[i]) {
      batch_free(b);
      set_error("batch_create: null context");
      return FEM_ERR_INVALID;
    }
    b->insts[i].ctx = ctxs[i];
  }
  *out = b;
  return 0;
}

int batch_set_equation(fem_batch* b, int inst, int eq, const void* params, size_t size) {
  if (inst < 0 || inst >= (int)b->insts.size() || !params) {
    set_error("set_equation: bad instance or params");
    return FEM_ERR_INVALID;
  }
  InstHost& h = b->insts[inst];
  double dt = 0, p1 = 0, p2 = 0, p3 = 0, p4 = 0;
  int w = 0;
  switch (eq) {
    case FEM_HEAT: {
      if (size != sizeof(fem_heat_params)) goto bad_size;
      auto* p = (const fem_heat_params*)params;
      dt = p->time_step;
      p1 = p->conductivity;
    } break;
    case FEM_ADVECTION_DIFFUSION: {
      if (size != sizeof(fem_advdiff_params)) goto bad_size;
      auto* p = (const fem_advdiff_params*)params;
      dt = p->time_step;
      p1 = p->c;
    } break;
    case FEM_WAVE: {
      if (size != sizeof(fem_wave_params)) goto bad_size;
      auto* p = (const fem_wave_params*)params;
      dt = p->time_step;
      p1 = p->c;
    } break;
    case FEM_REACTION_DIFFUSION: {
      if (size != sizeof(fem_rd_params)) goto bad_size;
      auto* p = (const fem_rd_params*)params;
      dt = p->time_step;
      p1 = p->Du;
      p2 = p->Dv;
      p3 = p->feed_rate;
      p4 = p->kill_rate;
      w = p->reaction_weighting;
      if (b->nrhs != 1) {
        set_error("Gray-Scott needs nrhs == 1 (u and v are the two columns)");
        return FEM_ERR_INVALID;
      }
    } break;
    case FEM_WAVE_NEWMARK: {
      if (size != sizeof(fem_newmark_params)) goto bad_size;
      auto* p = (const fem_newmark_params*)params;
      dt = p->time_step;
      p1 = p->c;
      p2 = p->beta;
      p3 = p->gamma;
    } break;
    default:
      set_error("unknown equation");
      return FEM_ERR_INVALID;
  }
  if (!(dt > 0.0) || !std::isfinite(dt)) {
    set_error("time_step must be positive");
    return FEM_ERR_INVALID;
  }
  {
    const int k = (eq == FEM_REACTION_DIFFUSION) ? 2 : b->nrhs;
    const int kv = (eq == FEM_REACTION_DIFFUSION) ? 2 : 1;
    if (h.k != k || h.kv != kv || !h.d_u) {
      int rc = alloc_inst(b, h, k, kv);
      if (rc) return rc;
      b->plan_ready = false;
    }
    // history vectors beyond `a` and `v` the predictor needs (h3, h4), for the
    // equations that use them
    const int pr = b->opts.predictor;
    int nh = 0;
    if (eq == FEM_HEAT || eq == FEM_ADVECTION_DIFFUSION || eq == FEM_REACTION_DIFFUSION)
      nh = pr == FEM_PREDICTOR_CUBIC     ? 1
           : pr == FEM_PREDICTOR_QUARTIC ? 2
           : pr == FEM_PREDICTOR_QUINTIC ? 3
           : pr == FEM_PREDICTOR_SEXTIC  ? 4
                                         : 0;
    else if (eq == FEM_WAVE_NEWMARK)
      nh = pr == FEM_PREDICTOR_PREVIOUS ? 0 : pr == FEM_PREDICTOR_LINEAR ? 1 : 2;
    for (int q = 0; q < nh; q++) {
      double** hp = q == 0 ? &h.d_h3 : q == 1 ? &h.d_h4 : q == 2 ? &h.d_h5 : &h.d_h6;
      if (*hp) continue;
      int rc = 0;
      *hp = dalloc<double>((size_t)h.ctx->n * k, &rc);
      if (rc) return rc;
      FEM_HIP(hipMemsetAsync(*hp, 0, (size_t)h.ctx->n * k * sizeof(double), b->stream));
    }
    if (h.eq != eq || h.dt != dt || h.p1 != p1 || h.p2 != p2 || h.p3 != p3 || h.p4 != p4 || h.weighting != w) {
      h.dirty = true;
      FEM_HIP(hipMemsetAsync(h.d_hist, 0, sizeof(int), b->stream));  // the history solved another system
    }
    if (h.eq != eq) b->plan_ready = false;
    h.eq = eq;
    h.dt = dt;
    h.p1 = p1;
    h.p2 = p2;
    h.p3 = p3;
    h.p4 = p4;
    h.weighting = w;
    return 0;
  }
bad_size:
  set_error("parameter struct size does not match the equation");
  return FEM_ERR_INVALID;
}

// Instance `inst` solves (coef[0] M + coef[1] K + coef[2] A) x = b once per
// advance; b and the guess x are set by the caller.
int batch_set_linear(fem_batch* b, int inst, const double coef[3]) {
  if (inst < 0 || inst >= (int)b->insts.size()) {
    set_error("bad instance");
    return FEM_ERR_INVALID;
  }
  InstHost& h = b->insts[inst];
  if (h.k != b->nrhs || h.kv != 1 || !h.d_u) {
    const int rc = alloc_inst(b, h, b->nrhs, 1);
    if (rc) return rc;
    b->plan_ready = false;
  }
  if (h.eq != EQ_SOLVE) b->plan_ready = false;
  h.eq = EQ_SOLVE;
  h.dt = 1.0;
  h.p1 = coef[0];
  h.p2 = coef[1];
  h.p3 = coef[2];
  h.p4 = 0.0;
  h.dirty = true;
  return 0;
}

// One solve of (coef[0] M + coef[1] K + coef[2] A) x = b for nrhs columns
// ([N][nrhs] host arrays); x is the initial guess when opts->warm_start.
int linear_solve(fem_ctx* c, const double coef[3], const fem_solver_opts* opts, const double* bh, double* xh,
                 fem_step_stats* stats) {
  fem_batch* b = nullptr;
  int rc = batch_create(&c, 1, opts, &b);
  if (rc) return rc;
  rc = batch_set_linear(b, 0, coef);
  InstHost& h = b->insts[0];
  const size_t bytes = sizeof(double) * (size_t)c->n * b->nrhs;
  if (!rc && c->n > 0) {
    hipError_t e = hipMemcpyAsync(h.d_b, bh, bytes, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess)
      e = b->opts.warm_start ? hipMemcpyAsync(h.d_x, xh, bytes, hipMemcpyHostToDevice, b->stream)
                             : hipMemsetAsync(h.d_x, 0, bytes, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) rc = hip_fail(e, "linear_solve: upload");
  }
  int out = rc;
  if (!rc) {
    out = batch_advance(b, 1, stats);
    if (out >= 0 && c->n > 0) {
      hipError_t e = hipMemcpyAsync(xh, h.d_x, bytes, hipMemcpyDeviceToHost, b->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
      if (e != hipSuccess) out = hip_fail(e, "linear_solve: download");
    }
  }
  batch_free(b);
  return out;
}

// The surface field of the current equation (u; Gray-Scott: v) as float
// values per vertex, boundary vertices 0 (CPUSolver.cpp:28-37).
__global__ void k_state_to_surface(const int* idx, int nv, const double* state, int k, int col, float* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const int r = idx[i];
  out[i] = r >= 0 ? (float)state[(size_t)r * k + col] : 0.0f;
}

int batch_surface_values(fem_batch* b, int inst, float* d_out) {
  InstHost& h = b->insts[inst];
  if (!h.d_u || h.eq < 0) {
    set_error("surface values: set_equation first");
    return FEM_ERR_STATE;
  }
  double* d;
  int col, nc;
  const int field = h.eq == FEM_REACTION_DIFFUSION ? FEM_FIELD_V : FEM_FIELD_U;
  int rc = batch_field_ptr(b, inst, field, &d, &col, &nc);
  if (rc) return rc;
  fem_ctx* c = h.ctx;
  if (c->nv == 0) return 0;
  hipLaunchKernelGGL(k_state_to_surface, dim3(div_up(c->nv, kBlock)), dim3(kBlock), 0, b->stream, c->d_idx, c->nv, d,
                     h.k, col, d_out);
  FEM_HIP(hipGetLastError());
  return 0;
}

// Map a state field to (device array, first column, column count).
int batch_field_ptr(fem_batch* b, int inst, int field, double** ptr, int* col, int* ncols) {
  if (inst < 0 || inst >= (int)b->insts.size()) {
    set_error("bad instance");
    return FEM_ERR_INVALID;
  }
  InstHost& h = b->insts[inst];
  if (!h.d_u) {
    set_error("set_equation must be called before the state is accessed");
    return FEM_ERR_STATE;
  }
  if (h.eq == FEM_REACTION_DIFFUSION) {
    if (field != FEM_FIELD_U && field != FEM_FIELD_V) goto bad;
    *ptr = h.d_u;
    *col = field == FEM_FIELD_U ? 0 : 1;
    *ncols = 1;
    return 0;
  }
  *col = 0;
  *ncols = h.k;
  switch (field) {
    case FEM_FIELD_U: *ptr = h.d_u; return 0;
    case FEM_FIELD_V: *ptr = h.d_v; return 0;
    case FEM_FIELD_A: *ptr = h.d_a; return 0;
    default: break;
  }
bad:
  set_error("bad field");
  return FEM_ERR_INVALID;
}

int batch_set_state_unknowns(fem_batch* b, int inst, int field, const double* x) {
  double* d;
  int col, nc;
  int rc = batch_field_ptr(b, inst, field, &d, &col, &nc);
  if (rc) return rc;
  InstHost& h = b->insts[inst];
  const long long n = h.ctx->n;
  if (n == 0) return 0;
  if (nc == h.k) {  // the field spans every column: copy straight into it
    FEM_HIP(hipMemcpyAsync(d, x, sizeof(double) * n * nc, hipMemcpyHostToDevice, b->stream));
    FEM_HIP(hipMemsetAsync(h.d_hist, 0, sizeof(int), b->stream));  // the history no longer matches the state
    FEM_HIP(hipStreamSynchronize(b->stream));
    return 0;
  }
  double* tmp = dalloc<double>((size_t)n * nc, &rc);
  if (!tmp) return rc;
  hipError_t e = hipMemcpyAsync(tmp, x, sizeof(double) * n * nc, hipMemcpyHostToDevice, b->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_put_cols, dim3(div_up(n * nc, kBlock)), dim3(kBlock), 0, b->stream, d, h.k, col, tmp, nc, n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemsetAsync(h.d_hist, 0, sizeof(int), b->stream);  // the history no longer matches
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return hip_fail(e, "set_state_unknowns");
  return 0;
}

int batch_get_state_unknowns(fem_batch* b, int inst, int field, double* x) {
  double* d;
  int col, nc;
  int rc = batch_field_ptr(b, inst, field, &d, &col, &nc);
  if (rc) return rc;
  InstHost& h = b->insts[inst];
  const long long n = h.ctx->n;
  if (n == 0) return 0;
  if (nc == h.k) {
    FEM_HIP(hipMemcpyAsync(x, d, sizeof(double) * n * nc, hipMemcpyDeviceToHost, b->stream));
    FEM_HIP(hipStreamSynchronize(b->stream));
    return 0;
  }
  double* tmp = dalloc<double>((size_t)n * nc, &rc);
  if (!tmp) return rc;
  hipLaunchKernelGGL(k_get_cols, dim3(div_up(n * nc, kBlock)), dim3(kBlock), 0, b->stream, d, h.k, col, tmp, nc, n);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(x, tmp, sizeof(double) * n * nc, hipMemcpyDeviceToHost, b->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return hip_fail(e, "get_state_unknowns");
  return 0;
}

// Surface values -> state through idx_map (CPUSolver.cpp:13-21).  An entry
// whose state, rounded to float, already equals the value is left alone: a
// caller that writes back what get_state returned (the frame loop of
// Application.cpp:530-540 gathers and scatters every frame) keeps the FP64
// state and the predictor history.  A written entry restarts the predictor:
// the held history levels (*hist) become 0.
__global__ void k_surface_gather(const int* idx, int nv, const float* surf, int nc, double* state, int k, int col,
                                 int* hist) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const int r = idx[i];
  if (r < 0) return;
  bool any = false;
  for (int j = 0; j < nc; j++) {
    const float v = surf[(size_t)i * nc + j];
    double* d = state + (size_t)r * k + col + j;
    if ((float)*d != v) {
      *d = (double)v;
      any = true;
    }
  }
  if (any) *hist = 0;
}

// State -> surface values, boundary vertices 0 (CPUSolver.cpp:28-37).
__global__ void k_surface_scatter(const int* idx, int nv, const double* state, int k, int col, int nc, float* surf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const int r = idx[i];
  for (int j = 0; j < nc; j++) surf[(size_t)i * nc + j] = r >= 0 ? (float)state[(size_t)r * k + col + j] : 0.0f;
}

int batch_set_surface(fem_batch* b, int inst, int field, const float* values_nv) {
  double* d;
  int col, nc;
  int rc = batch_field_ptr(b, inst, field, &d, &col, &nc);
  if (rc) return rc;
  InstHost& h = b->insts[inst];
  fem_ctx* c = h.ctx;
  if (c->nv == 0) return 0;
  // the values go through pinned staging, so the call returns without waiting
  // for the device: the previous upload from the staging must have completed
  FEM_HIP(hipEventSynchronize(h.surf_done));
  std::memcpy(h.h_surf, values_nv, sizeof(float) * c->nv * nc);
  FEM_HIP(hipMemcpyAsync(h.d_surf, h.h_surf, sizeof(float) * c->nv * nc, hipMemcpyHostToDevice, b->stream));
  FEM_HIP(hipEventRecord(h.surf_done, b->stream));
  hipLaunchKernelGGL(k_surface_gather, dim3(div_up(c->nv, kBlock)), dim3(kBlock), 0, b->stream, c->d_idx, c->nv,
                     h.d_surf, nc, d, h.k, col, h.d_hist);
  FEM_HIP(hipGetLastError());
  return 0;
}

int batch_get_surface(fem_batch* b, int inst, int field, float* values_nv) {
  double* d;
  int col, nc;
  int rc = batch_field_ptr(b, inst, field, &d, &col, &nc);
  if (rc) return rc;
  InstHost& h = b->insts[inst];
  fem_ctx* c = h.ctx;
  if (c->nv == 0) return 0;
  hipLaunchKernelGGL(k_surface_scatter, dim3(div_up(c->nv, kBlock)), dim3(kBlock), 0, b->stream, c->d_idx, c->nv,
                     d, h.k, col, nc, h.d_surf);
  FEM_HIP(hipGetLastError());
  // straight into the caller's array (the pinned staging serves uploads only:
  // reading it back on the host is no faster)
  FEM_HIP(hipMemcpyAsync(values_nv, h.d_surf, sizeof(float) * c->nv * nc, hipMemcpyDeviceToHost, b->stream));
  FEM_HIP(hipStreamSynchronize(b->stream));
  return 0;
}

// Restart the predictor of one instance (or all, inst < 0): its history
// levels become 0 (enqueued on the batch stream).
int batch_reset_history(fem_batch* b, int inst) {
  for (int i = 0; i < (int)b->insts.size(); i++) {
    if (inst >= 0 && i != inst) continue;
    if (b->insts[i].d_hist) FEM_HIP(hipMemsetAsync(b->insts[i].d_hist, 0, sizeof(int), b->stream));
  }
  return 0;
}

int batch_clear(fem_batch* b, int inst) {
  for (int i = 0; i < (int)b->insts.size(); i++) {
    if (inst >= 0 && i != inst) continue;
    InstHost& h = b->insts[i];
    if (!h.d_u) continue;
    const size_t vb = (size_t)h.ctx->n * h.k * sizeof(double);
    FEM_HIP(hipMemsetAsync(h.d_u, 0, vb, b->stream));
    FEM_HIP(hipMemsetAsync(h.d_v, 0, vb, b->stream));
    FEM_HIP(hipMemsetAsync(h.d_a, 0, vb, b->stream));
    FEM_HIP(hipMemsetAsync(h.d_x, 0, vb, b->stream));
    FEM_HIP(hipMemsetAsync(h.d_status, 0, sizeof(int), b->stream));
    FEM_HIP(hipMemsetAsync(h.d_hist, 0, sizeof(int), b->stream));
  }
  FEM_HIP(hipStreamSynchronize(b->stream));
  return 0;
}

// Prepare matrices and the plan; shared by both engines.
int batch_prepare(fem_batch* b, bool need_plan) {
  for (auto& h : b->insts) {
    if (h.eq < 0 || !h.d_u) {
      set_error("advance: set_equation was not called for every instance");
      return FEM_ERR_STATE;
    }
    // first-order equations form the next initial guess with each update
    // (Inst::xnext); persistent plans also fuse the update with the next
    // right-hand side, which needs the history term (Inst::hpre, formed from
    // Inst::hvec)
    if (h.eq == FEM_HEAT || h.eq == FEM_ADVECTION_DIFFUSION || h.eq == FEM_REACTION_DIFFUSION) {
      int rc = 0;
      if (!h.d_xnext) h.d_xnext = dalloc<double>((size_t)h.ctx->n * h.k, &rc);
      if (!rc && need_plan && !h.d_hpre) h.d_hpre = dalloc<double>((size_t)h.ctx->n * h.k, &rc);
      if (!rc && need_plan && !h.d_hvec) h.d_hvec = dalloc<double>((size_t)h.ctx->n * h.k, &rc);
      if (rc) return rc;
    }
    // Newmark on the streaming engine: each update forms the next step's
    // displacement predictor (Inst::hvec)
    if (h.eq == FEM_WAVE_NEWMARK && !need_plan && !h.d_hvec) {
      int rc = 0;
      h.d_hvec = dalloc<double>((size_t)h.ctx->n * h.k, &rc);
      if (rc) return rc;
    }
    if (h.dirty || h.formed_version != h.ctx->version) {
      int rc = form_inst(b, h);
      if (rc) return rc;
      b->plan_ready = false;
    }
    fill_dev(b, h);
  }
  if (!need_plan) return 0;
  if (!b->plan_ready) return build_plan(b);
  // descriptors may have changed since the plan was built (trajectory
  // records, equation override, predictor history): upload them when they did
  std::vector<Inst> hi(b->insts.size());
  for (size_t i = 0; i < hi.size(); i++) hi[i] = persistent_desc(b->insts[i].dev);
  bool same = b->uploaded.size() == hi.size();
  for (size_t i = 0; same && i < hi.size(); i++) same = std::memcmp(&b->uploaded[i], &hi[i], sizeof(Inst)) == 0;
  if (same) return 0;
  FEM_HIP(hipMemcpyAsync(b->d_insts, hi.data(), sizeof(Inst) * hi.size(), hipMemcpyHostToDevice, b->stream));
  FEM_HIP(hipStreamSynchronize(b->stream));
  b->uploaded = hi;
  return 0;
}

// Bring the info block (every instance's record and the flags) back in one
// copy and fold it into the call's statistics and status.
int read_info(fem_batch* b) {
  FEM_HIP(hipMemcpyAsync(b->h_info, b->d_info, b->info_bytes, hipMemcpyDeviceToHost, b->stream));
  FEM_HIP(hipStreamSynchronize(b->stream));
  return 0;
}

// Fold the host copy of the info block into the call's statistics and status.
int parse_stats(fem_batch* b, fem_step_stats* stats, float ms) {
  long long its = 0, sw = 0, steps = 0;
  int status = 0;
  double res = 0.0;
  for (size_t i = 0; i < b->insts.size(); i++) {
    const InstHost& h = b->insts[i];
    if (!h.d_counters) continue;
    const char* rec = b->h_info + i * kInfoBytes;
    const long long* cnt = (const long long*)rec;
    const int* st = (const int*)(rec + kInfoStatus);
    const double* lr = (const double*)(rec + kInfoLastRes);
    its += cnt[0];
    sw += cnt[1];
    steps = std::max(steps, cnt[2]);
    status |= st[0];
    for (int j = 0; j < h.k; j++) res = std::max(res, lr[j]);
  }
  if (stats) {
    stats->iterations = its;
    stats->sweeps = sw;
    stats->steps = steps;
    stats->residual = res;
    stats->ms = ms;
    stats->status = status;
  }
  if (status & ST_BARRIER_TIMEOUT) {
    set_error("device barrier wait timed out");
    return FEM_ERR_TIMEOUT;
  }
  if (status & (ST_UNSTABLE | ST_NONFINITE)) return FEM_DIVERGED;
  if (status & ST_NOT_CONVERGED) return FEM_NOT_CONVERGED;
  return FEM_OK;
}

int collect_stats(fem_batch* b, fem_step_stats* stats, float ms) {
  const int rc = read_info(b);
  if (rc) return rc;
  return parse_stats(b, stats, ms);
}

// Abort flag of the last launch, from the host copy read_info made.
int info_abort(const fem_batch* b) { return *(const int*)(b->h_info + b->insts.size() * kInfoBytes); }

int batch_instance_stats(fem_batch* b, int inst, fem_step_stats* stats) {
  if (inst < 0 || inst >= (int)b->insts.size()) {
    set_error("instance index out of range");
    return FEM_ERR_INVALID;
  }
  const InstHost& h = b->insts[inst];
  if (!h.d_counters) {
    set_error("instance has no state yet (set_equation first)");
    return FEM_ERR_STATE;
  }
  int rc = read_info(b);
  if (rc) return rc;
  const char* rec = b->h_info + (size_t)inst * kInfoBytes;
  const long long* cnt = (const long long*)rec;
  const int* st = (const int*)(rec + kInfoStatus);
  const double* lr = (const double*)(rec + kInfoLastRes);
  double res = 0.0;
  for (int j = 0; j < h.k; j++) res = std::max(res, lr[j]);
  stats->iterations = cnt[0];
  stats->sweeps = cnt[1];
  stats->steps = cnt[2];
  stats->residual = res;
  stats->ms = 0.0;  // device time is per call, not per instance
  stats->status = st[0];
  return 0;
}

int batch_set_wait_limit(fem_batch* b, double seconds) {
  if (!(seconds >= 0.0) || seconds > 3600.0) {
    set_error("wait limit must be within [0, 3600] seconds");
    return FEM_ERR_INVALID;
  }
  b->wait_ticks = (long long)(seconds * 1e8);
  b->plan_ready = false;  // the teams carry the limit
  return 0;
}

int batch_plan_info(fem_batch* b, fem_plan_info* info) {
  int rc = batch_prepare(b, true);
  if (rc) return rc;
  int maxb = 0;
  rc = step_max_coresident(b->launch.k, b->launch.inner_fp32, b->launch.kr, b->launch.coh, &maxb);
  if (rc) return rc;
  info->grid = b->grid;
  info->teams = b->nteams;
  info->exchange = b->launch.coh;
  info->coresident = maxb;
  return 0;
}

int batch_advance(fem_batch* b, int nsteps, fem_step_stats* stats) {
  if (nsteps < 0) {
    set_error("nsteps must be >= 0");
    return FEM_ERR_INVALID;
  }
  if (b->opts.engine == FEM_ENGINE_PLAIN) return plain_advance(b, nsteps, stats);
  if (b->opts.engine == FEM_ENGINE_STREAM) return stream_advance(b, nsteps, stats);
  if (b->opts.engine == FEM_ENGINE_GL_COMPAT) return gl_advance(b, nsteps, stats);
  int rc = batch_prepare(b, true);
  if (rc) return rc;
  const int ninst = (int)b->insts.size();
  hipLaunchKernelGGL(k_reset, dim3(div_up(std::max(b->nteams + b->grid, ninst), kBlock)), dim3(kBlock), 0,
                     b->stream, b->d_slots, b->grid, b->d_scal, b->nteams, b->d_abort, b->d_insts, ninst);
  FEM_HIP(hipGetLastError());
  FEM_HIP(hipEventRecord(b->ev0, b->stream));
  if (nsteps > 0) {
    rc = launch_step(b->launch, nsteps, b->stream);
    if (rc) return rc;
  }
  FEM_HIP(hipEventRecord(b->ev1, b->stream));
  // one wait for the launch and its outcome records (the kernel updates the
  // predictor history levels itself)
  rc = read_info(b);
  if (rc) return rc;
  float ms = 0.f;
  FEM_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  rc = parse_stats(b, stats, ms);
  if (info_abort(b)) {
    batch_reset_history(b, -1);
    set_error("device barrier wait timed out; results of this call are invalid");
    return FEM_ERR_TIMEOUT;
  }
  return rc;
}

// Record one advance of nsteps (persistent engine) as a HIP graph: the
// barrier-slot reset kernel (a fill kernel, no memset node) and the stepping
// kernel.  Replays reuse the plan's device arrays; a plan rebuilt since
// (new equation, new matrices, autotuning) makes the graph stale.
static int stream_graph_capture(fem_batch* b, int nsteps);
static int stream_graph_launch(fem_batch* b, fem_step_stats* stats);

int batch_graph_capture(fem_batch* b, int nsteps) {
  if (nsteps <= 0) {
    set_error("graph capture: nsteps must be positive");
    return FEM_ERR_INVALID;
  }
  if (b->opts.engine != FEM_ENGINE_PERSISTENT && b->opts.engine != FEM_ENGINE_STREAM) {
    set_error("graph capture records the persistent or the streaming engine; the plain and GL-compatible engines "
              "are host-driven");
    return FEM_ERR_INVALID;
  }
  if (b->graph) {
    FEM_HIP(hipGraphExecDestroy(b->graph));
    b->graph = nullptr;
  }
  b->graph_splan = nullptr;
  if (b->opts.engine == FEM_ENGINE_STREAM) return stream_graph_capture(b, nsteps);
  int rc = batch_prepare(b, true);
  if (rc) return rc;
  hipGraph_t g = nullptr;
  FEM_HIP(hipGraphCreate(&g, 0));
  int ninst = (int)b->insts.size();
  Slot* slots = b->d_slots;
  int nslots = b->grid, nteams = b->nteams;
  double* scal = b->d_scal;
  int* abort_flag = b->d_abort;
  Inst* insts = b->d_insts;
  void* rargs[] = {(void*)&slots, (void*)&nslots, (void*)&scal, (void*)&nteams, (void*)&abort_flag, (void*)&insts,
                   (void*)&ninst};
  hipKernelNodeParams prm{};
  prm.func = (void*)k_reset;
  prm.gridDim = dim3(div_up(std::max(b->nteams + b->grid, ninst), kBlock));
  prm.blockDim = dim3(kBlock);
  prm.sharedMemBytes = 0;
  prm.kernelParams = rargs;
  prm.extra = nullptr;
  hipGraphNode_t reset = nullptr, step = nullptr;
  hipError_t e = hipGraphAddKernelNode(&reset, g, nullptr, 0, &prm);
  if (e != hipSuccess) rc = hip_fail(e, "hipGraphAddKernelNode(k_reset)");
  if (!rc) rc = step_graph_node(b->launch, nsteps, g, reset, &step);
  if (!rc) {
    e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      b->graph = nullptr;
      rc = hip_fail(e, "hipGraphInstantiate");
    }
  }
  (void)hipGraphDestroy(g);
  if (rc) return rc;
  b->graph_nsteps = nsteps;
  b->graph_gen = b->plan_gen;
  return 0;
}

// Replay the recorded advance: the same as batch_advance(b, graph_nsteps).
int batch_graph_launch(fem_batch* b, fem_step_stats* stats) {
  if (!b->graph) {
    set_error("graph launch: no graph was captured");
    return FEM_ERR_STATE;
  }
  if (b->graph_splan) {
    if (b->opts.engine != FEM_ENGINE_STREAM) {
      set_error("graph launch: the batch no longer runs the streaming engine; capture again");
      return FEM_ERR_STATE;
    }
    return stream_graph_launch(b, stats);
  }
  if (b->opts.engine != FEM_ENGINE_PERSISTENT) {
    set_error("graph launch: the batch no longer runs the persistent engine; capture again");
    return FEM_ERR_STATE;
  }
  int rc = batch_prepare(b, true);  // uploads the current descriptors into the plan's arrays
  if (rc) return rc;
  if (b->graph_gen != b->plan_gen) {
    set_error("graph launch: the launch plan changed since the capture; capture again");
    return FEM_ERR_STATE;
  }
  FEM_HIP(hipEventRecord(b->ev0, b->stream));
  FEM_HIP(hipGraphLaunch(b->graph, b->stream));
  FEM_HIP(hipEventRecord(b->ev1, b->stream));
  rc = read_info(b);
  if (rc) return rc;
  float ms = 0.f;
  FEM_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  rc = parse_stats(b, stats, ms);
  if (info_abort(b)) {
    batch_reset_history(b, -1);
    set_error("device barrier wait timed out; results of this call are invalid");
    return FEM_ERR_TIMEOUT;
  }
  return rc;
}

// Streaming engine: every phase a launch over the row blocks of all
// instances (stream.hip).  Instances must share the column count, the inner
// precision and the Krylov method, as in the persistent plan.  Set-up of a
// run: checks, the descriptors uploaded, the row-block plan (re)built.
static int stream_setup(fem_batch* b, int* max_sweeps, int* max_inner) {
  int rc = batch_prepare(b, false);
  if (rc) return rc;
  const int ninst = (int)b->insts.size();
  const int k = b->insts[0].k, kr = b->insts[0].dev.krylov;
  for (auto& h : b->insts) {
    if (h.k != k || h.dev.krylov != kr) {
      set_error("batch instances must share the number of columns and the Krylov method");
      return FEM_ERR_INVALID;
    }
    if (stream_vec_layout(k) && (h.kv != 1 || !h.d_SDi)) {  // the column layout shares one matrix across columns
      set_error("stream engine: instances of one or of four or more columns need one matrix for all columns");
      return FEM_ERR_INVALID;
    }
  }
  const int fp32 = inner_bytes(b) == 4;
  if (b->sinsts_n != ninst) {
    if (b->d_sinsts) (void)hipFree(b->d_sinsts);
    b->d_sinsts = dalloc<Inst>(ninst, &rc);
    if (rc) return rc;
    b->sinsts_n = ninst;
  }
  std::vector<Inst> hi(ninst);
  for (int i = 0; i < ninst; i++) hi[i] = b->insts[i].dev;
  FEM_HIP(hipMemcpyAsync(b->d_sinsts, hi.data(), sizeof(Inst) * ninst, hipMemcpyHostToDevice, b->stream));
  // rows per row block: team_size when given (the stream engine has no teams)
  const int rows = b->opts.team_size > 0 ? b->opts.team_size : b->stream_rows > 0 ? b->stream_rows : (k >= 16 ? 128 : 256);
  if (!b->splan || b->splan_key != (long long)rows * 1000003LL + k * 131 + kr * 7 + fp32) {
    if (b->splan) stream_plan_free(b->splan);
    b->splan = nullptr;
    rc = stream_plan_build(b->insts, b->d_sinsts, k, fp32, kr, rows, b->stream, &b->splan);
    if (rc) return rc;
    b->splan_key = (long long)rows * 1000003LL + k * 131 + kr * 7 + fp32;
    b->splan_gen++;
    b->sobs_sweeps = 0;  // a new plan: a graph's budget is sized from its own runs
    for (int& v : b->sobs_iters) v = 0;
  }
  *max_sweeps = b->opts.max_sweeps > 0 ? b->opts.max_sweeps : 30;
  *max_inner = b->opts.max_inner > 0 ? b->opts.max_inner : 200;
  return 0;
}

// The device turned the predictor history of first-order instances once per
// step (k_step_end): turn the host's pointers as often.
static void stream_turn_host(fem_batch* b, int turned) {
  for (auto& h : b->insts) {
    Inst& I = h.dev;
    if (!first_order_eq(I.eq) || I.extrap < 2) continue;
    for (int t = 0; t < turned % I.extrap; t++) {
      turn_history(&h.d_a, &h.d_v, &h.d_h3, &h.d_h4, &h.d_h5, &h.d_h6, I.extrap);
      turn_history(&I.a, &I.v, &I.h3, &I.h4, &I.h5, &I.h6, I.extrap);
    }
  }
}

int stream_advance(fem_batch* b, int nsteps, fem_step_stats* stats) {
  int max_sweeps = 0, max_inner = 0;
  int rc = stream_setup(b, &max_sweeps, &max_inner);
  if (rc) return rc;
  FEM_HIP(hipEventRecord(b->ev0, b->stream));
  int turned = 0;
  StreamCtl ctl{};
  rc = stream_run(b->splan, b->d_sinsts, max_sweeps, max_inner, nsteps, b->stream, &turned, &ctl);
  stream_turn_host(b, turned);
  if (rc) return rc;
  // the sweeps and iterations this advance needed: a graph's budget (the
  // most recent steps, so an advance just before a capture sizes it)
  if (nsteps > 0) {
    b->sobs_sweeps = ctl.obs_sweeps;
    for (int q = 0; q < kBudgetSweeps; q++) b->sobs_iters[q] = ctl.obs_iters[q];
  }
  FEM_HIP(hipEventRecord(b->ev1, b->stream));
  rc = read_info(b);
  if (rc) return rc;
  float ms = 0.f;
  FEM_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  return parse_stats(b, stats, ms);
}

// Record one streaming advance of nsteps as a graph: every step issues the
// sweeps and iterations the eager runs of this plan needed, plus a margin
// (a quarter more iterations per sweep, at least 2, and one more sweep),
// over a grid of 8 workgroups per CU; nothing is read on the host.  An
// instance that would need more stops there and reports FEM_NOT_CONVERGED.
static int stream_graph_capture(fem_batch* b, int nsteps) {
  int max_sweeps = 0, max_inner = 0;
  int rc = stream_setup(b, &max_sweeps, &max_inner);
  if (rc) return rc;
  if (b->sobs_sweeps == 0) {
    set_error("graph capture (streaming engine): advance the batch once first, so that the recorded sweeps and "
              "iterations can be sized from what its steps need");
    return FEM_ERR_STATE;
  }
  int dev = 0, cus = 0;
  FEM_HIP(hipGetDevice(&dev));
  FEM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  StreamCtl ctl{};
  ctl.budget = true;
  ctl.grid = 8 * std::max(cus, 1);
  ctl.sweeps = std::min({kBudgetSweeps, max_sweeps, b->sobs_sweeps + 1});
  for (int sw = 0; sw < ctl.sweeps; sw++) {
    const int o = std::max(1, b->sobs_iters[std::min(sw, b->sobs_sweeps - 1)]);
    ctl.iters[sw] = std::min(max_inner, o + std::max(2, (o + 3) / 4));
  }
  // launches the recording would hold (stream_run's schedule: per iteration
  // the Krylov phases with their reductions and one compaction; per sweep its
  // start, a compaction and the two residual passes; per step the right-hand
  // side, the update and the step's end)
  const int kr = b->splan ? b->insts[0].dev.krylov : 0;
  const long long per_it = 2LL * (kr == 1 ? 3 : kr == 2 ? 2 : 1) + 2;
  long long per_step = 2 + 1 + 1;
  for (int sw = 0; sw < ctl.sweeps; sw++) per_step += 3 + (kr == 0 ? 2 : 0) + ctl.iters[sw] * per_it + 8;
  const long long launches = 3 + (long long)nsteps * per_step;
  if (launches > kMaxGraphLaunches) {
    set_error("graph capture (streaming engine): " + std::to_string(nsteps) + " steps would record " +
              std::to_string(launches) + " launches, more than " + std::to_string(kMaxGraphLaunches) +
              "; record fewer steps (at most " + std::to_string(std::max(0LL, (kMaxGraphLaunches - 3) / per_step)) +
              ") and replay the graph several times");
    return FEM_ERR_INVALID;
  }
  FEM_HIP(hipStreamSynchronize(b->stream));  // the upload of stream_setup is done before the capture
  FEM_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
  int turned = 0;
  rc = stream_run(b->splan, b->d_sinsts, max_sweeps, max_inner, nsteps, b->stream, &turned, &ctl);
  hipGraph_t g = nullptr;
  const hipError_t ee = hipStreamEndCapture(b->stream, &g);
  if (!rc && ee != hipSuccess) rc = hip_fail(ee, "hipStreamEndCapture");
  if (!rc) {
    const hipError_t e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      b->graph = nullptr;
      rc = hip_fail(e, "hipGraphInstantiate");
    }
  }
  if (g) (void)hipGraphDestroy(g);
  if (rc) return rc;
  b->graph_nsteps = nsteps;
  b->graph_splan = b->splan;
  b->graph_sgen = b->splan_gen;
  b->graph_sinsts = b->d_sinsts;
  b->graph_budget_sweeps = ctl.sweeps;
  return 0;
}

// Replay of a streaming graph: the descriptors (with the host's turn of the
// history) are uploaded first; the graph turns the device's copy nsteps
// times, and the host's pointers follow.
static int stream_graph_launch(fem_batch* b, fem_step_stats* stats) {
  int max_sweeps = 0, max_inner = 0;
  int rc = stream_setup(b, &max_sweeps, &max_inner);
  if (rc) return rc;
  if (b->splan != b->graph_splan || b->splan_gen != b->graph_sgen || b->d_sinsts != b->graph_sinsts) {
    set_error("graph launch: the streaming plan changed since the capture; capture again");
    return FEM_ERR_STATE;
  }
  FEM_HIP(hipEventRecord(b->ev0, b->stream));
  FEM_HIP(hipGraphLaunch(b->graph, b->stream));
  FEM_HIP(hipEventRecord(b->ev1, b->stream));
  stream_turn_host(b, b->graph_nsteps);
  rc = read_info(b);
  if (rc) return rc;
  float ms = 0.f;
  FEM_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  return parse_stats(b, stats, ms);
}

// Autotuner: run `probe` steps under a few launch plans from the same state
// (snapshotted and restored around every probe) and keep the fastest by
// kernel time.  One instance: the team size around the one-row-per-lane
// default; several: the number of workgroups the plan spreads the batch over.
int batch_autotune(fem_batch* b, int probe, int* chosen, double* best_ms) {
  if (b->opts.engine != FEM_ENGINE_PERSISTENT || probe <= 0) return 0;
  int rc = batch_prepare(b, true);
  if (rc) return rc;
  int maxb = 0;
  rc = step_max_coresident(b->launch.k, b->launch.inner_fp32, b->launch.kr, 1, &maxb);
  if (rc) return rc;
  const bool single = b->insts.size() == 1;
  // Rows x columns per lane of the whole co-resident grid: beyond ~64 every
  // phase of a persistent plan walks thousands of rows per lane at the
  // occupancy of one long-lived grid, and a probe of it runs for seconds;
  // such batches only try the streaming engine.
  double lanes_work = 0.0;
  for (auto& h : b->insts) lanes_work += (double)h.ctx->n * h.k;
  const bool persistent_hopeless = lanes_work > 64.0 * maxb * kBlock;
  std::vector<int> cands;
  if (persistent_hopeless) {
    // no persistent candidates
  } else if (single) {
    const int g0 = b->team_sizes[0];
    for (double f : {0.25, 0.5, 0.75, 1.0, 1.5, 2.0}) {
      const int g = std::max(1, std::min(maxb, (int)std::lround(g0 * f)));
      if (std::find(cands.begin(), cands.end(), g) == cands.end()) cands.push_back(g);
    }
    // one and two workgroups per CU: no CU hosts one workgroup more than
    // another, so no straggler holds up the barriers
    int dev = 0, cus = 0;
    FEM_HIP(hipGetDevice(&dev));
    FEM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    for (int g : {cus / 2, cus, 2 * cus})
      if (g >= 1 && g <= maxb && std::find(cands.begin(), cands.end(), g) == cands.end()) cands.push_back(g);
  } else {
    for (double f : {0.25, 0.5, 0.75, 1.0}) cands.push_back(std::max(1, (int)(maxb * f)));
  }
  // snapshot the state
  std::vector<double*> snap;
  size_t want = 0;
  for (auto& h : b->insts) {
    const size_t vb = (size_t)h.ctx->n * h.k * sizeof(double);
    for (double* src : {h.d_u, h.d_v, h.d_a, h.d_x, h.d_h3, h.d_h4, h.d_h5, h.d_h6}) {
      if (!src) continue;
      want++;
      if (rc) continue;
      double* d = dalloc<double>((size_t)h.ctx->n * h.k, &rc);
      if (!d) continue;
      snap.push_back(d);
      const hipError_t e = hipMemcpyAsync(d, src, vb, hipMemcpyDeviceToDevice, b->stream);
      if (e != hipSuccess) rc = hip_fail(e, "autotune: snapshot");
    }
  }
  // the info block holds each instance's predictor history levels
  char* info0 = rc ? nullptr : dalloc<char>(b->info_bytes, &rc);
  if (info0) FEM_HIP(hipMemcpyAsync(info0, b->d_info, b->info_bytes, hipMemcpyDeviceToDevice, b->stream));
  // restore is only ever run on a complete snapshot
  const bool snap_ok = !rc && info0 && snap.size() == want;
  auto restore = [&]() -> int {
    if (!snap_ok) return FEM_ERR_OOM;
    size_t q = 0;
    for (auto& h : b->insts) {
      const size_t vb = (size_t)h.ctx->n * h.k * sizeof(double);
      for (double* dst : {h.d_u, h.d_v, h.d_a, h.d_x, h.d_h3, h.d_h4, h.d_h5, h.d_h6})
        if (dst) FEM_HIP(hipMemcpyAsync(dst, snap[q++], vb, hipMemcpyDeviceToDevice, b->stream));
    }
    FEM_HIP(hipMemcpyAsync(b->d_info, info0, b->info_bytes, hipMemcpyDeviceToDevice, b->stream));
    return 0;
  };
  const int saved_team = b->opts.team_size, saved_budget = b->grid_budget, saved_mode = b->mode_override;
  const int saved_poll = b->poll_override;
  double best = 1e300;
  int best_c = single ? saved_team : saved_budget, best_m = saved_mode, best_p = saved_poll;
  if (!rc) {
    for (int c : cands) {
      for (int m : {kExchangeCoherent, kExchangeFenced}) {
        if (b->opts.exchange != FEM_EXCHANGE_AUTO && m != kExchangeCoherent) continue;  // mode fixed by opts
        // one instance: both barrier reductions (all workgroups sum, or rank 0 publishes)
        for (int pl : {1, 2}) {
          if (!single && pl == 2) continue;
          if (single) b->opts.team_size = c;
          else b->grid_budget = c;
          b->mode_override = m;
          b->poll_override = single ? pl : saved_poll;
          b->plan_ready = false;
          if ((rc = restore())) break;
          fem_step_stats st{};
          const int r = batch_advance(b, probe, &st);
          if (r < 0) {
            rc = r;
            break;
          }
          if (st.ms < best) {
            best = st.ms;
            best_c = c;
            best_m = m;
            best_p = b->poll_override;
          }
        }
        if (rc) break;
      }
      if (rc) break;
    }
  }
  // the streaming engine as more candidates: rows per row block
  bool stream_best = false;
  const int saved_rows = b->stream_rows;
  int best_rows = saved_rows;
  if (!rc) {
    b->opts.engine = FEM_ENGINE_STREAM;
    const int team = b->opts.team_size;
    b->opts.team_size = 0;  // rows come from stream_rows
    const int k = b->launch.k;
    for (int rows : {k >= 16 ? 64 : 128, k >= 16 ? 128 : 256, k >= 16 ? 256 : 512}) {
      b->stream_rows = rows;
      if ((rc = restore())) break;
      fem_step_stats st{};
      const int r = batch_advance(b, probe, &st);
      if (r < 0) {
        rc = r;
        break;
      }
      if (st.ms < best) {
        best = st.ms;
        stream_best = true;
        best_rows = rows;
      }
    }
    b->opts.team_size = team;
    b->opts.engine = FEM_ENGINE_PERSISTENT;
  }
  const int rr = snap_ok ? restore() : 0;
  if (!rc) rc = rr;
  const hipError_t se = hipStreamSynchronize(b->stream);
  for (double* d : snap) (void)hipFree(d);
  if (info0) (void)hipFree(info0);
  if (se != hipSuccess && !rc) rc = hip_fail(se, "autotune");
  if (single) b->opts.team_size = rc ? saved_team : best_c;
  else b->grid_budget = rc ? saved_budget : best_c;
  b->mode_override = rc ? saved_mode : best_m;
  b->poll_override = rc ? saved_poll : best_p;
  b->stream_rows = (!rc && stream_best) ? best_rows : saved_rows;
  if (!rc && stream_best) {
    b->opts.engine = FEM_ENGINE_STREAM;
    b->opts.team_size = 0;
  }
  b->plan_ready = false;
  if (chosen) *chosen = (!rc && stream_best) ? 0 : best_c;
  if (best_ms) *best_ms = best;
  return rc;
}

}  // namespace femhip
