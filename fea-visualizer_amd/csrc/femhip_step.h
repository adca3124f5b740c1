// femhip_step.h -- host-side batch / solver objects and the launch plan of the
// persistent stepping kernel.
#pragma once

#include "femhip_ctx.h"

namespace femhip {

struct StepLaunch {
  const Team* teams = nullptr;
  const int* wg_team = nullptr;
  const Inst* insts = nullptr;
  const int* team_inst = nullptr;
  int grid = 0;
  int k = 1;
  int inner_fp32 = 1;
  int kr = 0;   // Krylov method: 0 pipelined CG, 1 BiCGSTAB, 2 classic CG
  int coh = 1;  // exchange mode: 0 one-workgroup teams only, 1 coherent, 2 fenced
};

// Exchange mode for teams spanning several workgroups: coherent accesses
// while a workgroup's rows fit the kernel's LDS row cache (latency-bound
// phases), fenced barriers beyond (bandwidth-bound phases).
constexpr int kExchangeCoherent = 1, kExchangeFenced = 2;

int launch_step(const StepLaunch& L, int nsteps, hipStream_t st);

// Plan rules shared by the batch plans and the allocation-free solve (batch.hip).
int single_team_size(long long n, int k, int team_size, int maxb);
bool team_streams(long long n, int k, int size);
int exchange_mode(bool multi, bool streaming, int forced, int tuned);
int launch_slot_reset(Slot* slots, int nslots, double* scal, int nteams, int* abort_flag, Inst* insts, int ninst,
                      hipStream_t st);

// Streaming engine (stream.hip): row-block table and per-instance scalars.
struct StreamPlan;
struct InstHost;
void stream_plan_free(StreamPlan* p);
int stream_plan_build(const std::vector<InstHost>& insts, const Inst* d_insts, int k, int fp32, int kr, int rows,
                      hipStream_t st, StreamPlan** out);
// Control of a streaming run.  Eager runs (budget == false) stop each sweep
// from host reads of the device's counts and record the most sweeps a step
// took and the most iterations each sweep took (obs_*).  A budgeted run, the
// one a graph records, reads nothing on the host: every step issues
// `sweeps` sweeps of iters[sw] iterations over a grid of `grid` workgroups,
// and the launches past an instance's end exit at once on the device.
constexpr int kBudgetSweeps = 8;
// Most kernel launches one streaming graph records.  A 200-step recording of
// the 256-instance batch (~250,000 launches) ended the process with SIGSEGV
// in graph capture or replay, where 20 steps (~25,000) ran; longer advances
// replay a shorter recording several times.
constexpr long long kMaxGraphLaunches = 8192;
struct StreamCtl {
  bool budget = false;
  int sweeps = 0, grid = 0;
  int iters[kBudgetSweeps] = {};
  int obs_sweeps = 0;
  int obs_iters[kBudgetSweeps] = {};
};
int stream_run(const StreamPlan* p, const Inst* d_insts, int max_sweeps, int max_inner, int nsteps, hipStream_t st,
               int* turned, StreamCtl* ctl);
// Column counts whose stream phases use the multi-column layout with the
// staged pattern (one value column, Inst::SDi present).
bool stream_vec_layout(int k);
int step_max_coresident(int k, int inner_fp32, int kr, int coh, int* blocks);
int step_graph_node(const StepLaunch& L, int nsteps, hipGraph_t g, hipGraphNode_t dep, hipGraphNode_t* out);

// Per-instance host bookkeeping.
struct InstHost {
  fem_ctx* ctx = nullptr;
  int eq = -1;
  double dt = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0, p4 = 0.0;
  int weighting = 0;
  int k = 0;                 // columns in the kernel (Gray-Scott: 2)
  int kv = 1;                // value columns
  long long formed_version = -1;
  bool dirty = true;         // matrices need (re)forming
  int transpose = 0;         // form the transposed system matrix (adjoint solves)
  int eq_override = -1;      // >= 0: equation id the next launches run (EQ_SOLVE)
  double* traj = nullptr;    // trajectory record (see Inst::traj)
  int traj_every = 0;
  double* d_h3 = nullptr;    // cubic / quartic predictor history
  double* d_h4 = nullptr;
  double* d_h5 = nullptr;    // quintic / sextic history (streaming engine)
  double* d_h6 = nullptr;
  double* d_xnext = nullptr; // Inst::xnext, Inst::hpre
  double* d_hpre = nullptr;
  double* d_hvec = nullptr;  // Inst::hvec
  double* d_S64 = nullptr;
  void* d_Si = nullptr;
  double* d_dinv64 = nullptr;
  void* d_dinvi = nullptr;
  void* d_SDi = nullptr;     // column-scaled inner matrix (Inst::SDi)
  double* d_u = nullptr;
  double* d_v = nullptr;
  double* d_a = nullptr;
  double* d_x = nullptr;
  double* d_b = nullptr;
  void* d_vec = nullptr;     // slab of inner-precision vectors
  int* d_status = nullptr;         // these three point into the batch's info block
  long long* d_counters = nullptr;
  double* d_last_res = nullptr;
  int* d_hist = nullptr;     // predictor history levels held (Inst::hist_dev), in the info block
  float* d_surf = nullptr;   // surface-space staging of set_state / get_state [nv][k]
  float* h_surf = nullptr;   // pinned host staging of set_state's values [nv][k]
  hipEvent_t surf_done = nullptr;  // the last upload from h_surf has completed
  Inst dev{};                // mirror of the device descriptor
};

// Per-instance record in the batch's info block: counters [0, 32), status
// bits [32, 48), final residuals [48, 48 + 8 kMaxK), predictor history levels
// held (Inst::hist_dev) at 176.  After the instance records: the launch abort
// flag.
constexpr size_t kInfoBytes = 256;
constexpr size_t kInfoStatus = 32, kInfoLastRes = 48, kInfoHist = 176;
static_assert(kInfoLastRes + 8 * kMaxK <= kInfoHist, "info record");

}  // namespace femhip

struct fem_batch {
  std::vector<femhip::InstHost> insts;
  fem_solver_opts opts{};
  int nrhs = 1;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // launch plan (device arrays)
  bool plan_ready = false;
  femhip::StepLaunch launch;
  femhip::Inst* d_insts = nullptr;
  femhip::Team* d_teams = nullptr;
  int* d_wg_team = nullptr;
  int* d_team_inst = nullptr;
  femhip::Slot* d_slots = nullptr;
  double* d_scal = nullptr;
  int grid = 0;
  int* d_abort = nullptr;       // in the info block
  char* d_info = nullptr;       // [ninst][kInfoBytes] + abort flag
  char* h_info = nullptr;       // pinned host copy
  size_t info_bytes = 0;
  std::vector<femhip::Inst> uploaded;  // descriptors last copied to the plan's d_insts
  int nteams = 0;
  long long wait_ticks = femhip::kDefaultWaitTicks;  // longest barrier wait of the persistent kernel
  int grid_budget = 0;          // multi-instance plans: workgroups to spread over (0 = all co-resident)
  int mode_override = 0;        // exchange mode forced by the autotuner (0 = by the row-cache rule)
  int poll_override = 0;        // barrier reduction forced by the autotuner: 1 every workgroup sums all
                                // partials, 2 rank 0 sums and publishes (0 = by the traffic rule)
  std::vector<int> team_sizes;  // for reporting
  // plain-engine scratch
  double* d_plain = nullptr;
  long long plain_n = 0;
  // streaming engine
  femhip::StreamPlan* splan = nullptr;
  femhip::Inst* d_sinsts = nullptr;
  int sinsts_n = 0;
  int stream_rows = 0;  // rows per block (0 = by column count)
  long long splan_key = -1;
  long long splan_gen = 0;  // counts plan rebuilds
  // recorded advance (persistent engine): valid while the plan it was built
  // on is current
  long long plan_gen = 0;
  hipGraphExec_t graph = nullptr;
  int graph_nsteps = 0;
  long long graph_gen = -1;
  // a streaming graph: the plan (and its generation) and descriptor array it
  // was recorded on (graph_splan == nullptr: a persistent graph), and its
  // sweeps per step
  femhip::StreamPlan* graph_splan = nullptr;
  long long graph_sgen = -1;
  femhip::Inst* graph_sinsts = nullptr;
  int graph_budget_sweeps = 0;
  // most sweeps a step and most iterations each sweep took in the last eager
  // streaming advance of the current plan (a streaming graph's budget)
  int sobs_sweeps = 0;
  int sobs_iters[femhip::kBudgetSweeps] = {};
};

struct fem_solver {
  fem_batch* batch = nullptr;
};

namespace femhip {
int batch_create(fem_ctx* const* ctxs, int count, const fem_solver_opts* opts, fem_batch** out);
void batch_free(fem_batch* b);
int batch_set_equation(fem_batch* b, int inst, int equation, const void* params, size_t size);
int batch_set_state_unknowns(fem_batch* b, int inst, int field, const double* x);
int batch_get_state_unknowns(fem_batch* b, int inst, int field, double* x);
int batch_advance(fem_batch* b, int nsteps, fem_step_stats* stats);
int batch_clear(fem_batch* b, int inst);
int batch_set_surface(fem_batch* b, int inst, int field, const float* values_nv);
int batch_reset_history(fem_batch* b, int inst);
int batch_get_surface(fem_batch* b, int inst, int field, float* values_nv);
int batch_field_ptr(fem_batch* b, int inst, int field, double** ptr, int* col, int* k);
int plain_advance(fem_batch* b, int nsteps, fem_step_stats* stats);
int batch_autotune(fem_batch* b, int probe, int* chosen, double* best_ms);
int batch_graph_capture(fem_batch* b, int nsteps);
int batch_instance_stats(fem_batch* b, int inst, fem_step_stats* stats);
int batch_plan_info(fem_batch* b, fem_plan_info* info);
int batch_set_wait_limit(fem_batch* b, double seconds);
int batch_graph_launch(fem_batch* b, fem_step_stats* stats);
int stream_advance(fem_batch* b, int nsteps, fem_step_stats* stats);
int batch_surface_values(fem_batch* b, int inst, float* d_out);
int gl_advance(fem_batch* b, int nsteps, fem_step_stats* stats);
int batch_set_linear(fem_batch* b, int inst, const double coef[3]);
int linear_solve(fem_ctx* c, const double coef[3], const fem_solver_opts* opts, const double* b, double* x,
                 fem_step_stats* stats);
int form_system(fem_ctx* c, const double coef[3], int fp32, double* S64, void* Si, double* dinv64, void* dinvi,
                hipStream_t st);
int ply_export(const float* xyz, int32_t nv, const uint32_t* tri, int32_t ntri, const float* normals,
               const float* values, float vertex_extrusion, float threshold, int32_t mesh_type, const char* path);
int surface_from_obj(const char* path, fem_surface** out);
void surface_free(fem_surface* s);
int surface_get(const fem_surface* s, fem_mesh* mesh, const float** normals, int32_t* num_boundary);
int linear_solve_noalloc(fem_ctx* c, const double coef[3], const fem_solver_opts* opts, const double* b, double* x,
                         fem_step_stats* stats);
}  // namespace femhip
