// femhip_internal.h -- shared host/device definitions of the MI355X (gfx950)
// FEM backend.  Not part of the public C ABI (include/femhip.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "femhip.h"

namespace femhip {

constexpr int kBlock = 256;          // threads per workgroup (4 wave64)
constexpr int kWave = 64;
constexpr int kMaxK = 16;            // right-hand sides per instance
constexpr int kMaxSlots = 4;         // reduction slots per phase
constexpr int kMaxRed = kMaxSlots * kMaxK;

// Equation ids (first four as the reference enum, FEMContext.hpp:9-14).
enum Eq : int {
  EQ_HEAT = 0,
  EQ_WAVE = 1,
  EQ_ADVDIFF = 2,
  EQ_GS = 3,
  EQ_NEWMARK = 4,
  EQ_SOLVE = 5,  // plain solve S x = b (b supplied), used by the adjoint
};

// Equations whose new state is the solve's solution; their predictor
// history (Inst::extrap levels) lives in a, v, h3, h4, newest first.
__host__ __device__ inline bool first_order_eq(int eq) { return eq == EQ_HEAT || eq == EQ_ADVDIFF || eq == EQ_GS; }

// Predictor orders: up to quartic on the persistent kernel (history in a, v,
// h3, h4), up to sextic on the streaming engine (also h5, h6).
constexpr int kMaxExtrap = 6, kPersistentExtrap = 4;

// One turn of that history by pointer: the oldest of the e slots becomes the
// newest (the streaming engine's update has written u_n into it).
__host__ __device__ inline void turn_history(double** a, double** v, double** h3, double** h4, double** h5,
                                             double** h6, int e) {
  double** s[kMaxExtrap] = {a, v, h3, h4, h5, h6};
  if (e < 2 || e > kMaxExtrap) return;
  double* t = *s[e - 1];
  for (int i = e - 1; i > 0; i--) *s[i] = *s[i - 1];
  *s[0] = t;
}

// Status bits accumulated per instance.
enum : int {
  ST_NOT_CONVERGED = 1,
  ST_UNSTABLE = 2,      // reference rule: u > 1e4 || v > 1e4
  ST_NONFINITE = 4,     // NaN / Inf in the state (not caught by the reference)
  ST_BARRIER_TIMEOUT = 8,
};

// Per-workgroup barrier slot: the partial sums of the last four barriers.
// Barrier e writes buffer e & 3, and the lowest mantissa bit of every value
// carries the tag (e >> 2) & 1, so a reader recognises this barrier's values
// without a separate flag.  Arrivals go to the workgroup's own slot and never
// contend on one address.
constexpr int kSlotBufs = 4;
struct alignas(128) Slot {
  double vals[kSlotBufs][kMaxRed];
};


// Device-side description of one instance (a mesh, its parameters and its
// state) advanced by one team of workgroups inside the persistent kernel.
struct Inst {
  int n, nnz, k, kv;         // rows, nonzeros, rhs columns, value columns (1 or k)
  int eq, krylov;            // krylov: 0 CG, 1 BiCGSTAB
  int warm;                  // initial guess = previous primary variable
  int max_inner, max_sweeps, max_iters;
  int weighting;             // Gray-Scott reaction: 0 nodal (reference), 1 lumped mass
  int max_row;               // longest row (nonzeros) of the pattern
  double tol2, inner_tol2;
  double dt, p1, p2, p3, p4; // equation parameters
  const int* rp;
  const int* ci;
  const double* M64;
  const double* K64;
  const double* S64;     // [nnz][kv]
  const void* Si;        // [nnz][kv] inner precision (float or double)
  const double* dinv64;  // [n][kv]
  const void* dinvi;     // [n][kv] inner precision
  const void* SDi;       // [nnz] inner precision S_ij D^-1_jj (one value column, four or more
                         //   right-hand sides; null otherwise)
  const double* lumped;  // [n]
  double* u;             // state [n][k]  (Gray-Scott: u,v interleaved, k = 2)
  double* v;
  double* a;
  double* x;             // FP64 solution [n][k]
  double* b;             // FP64 right-hand side [n][k]
  void* xi;              // inner-precision correction [n][k]
  void* r;
  void* z;
  void* p0;              // direction, double-buffered
  void* p1b;
  void* q;
  void* rh;              // BiCGSTAB shadow residual
  void* v0;              // BiCGSTAB v, double-buffered
  void* v1;
  void* s;
  void* t;
  int* status;           // outcome bits
  long long* counters;   // [0] iterations, [1] sweeps, [2] steps
  double* last_res;      // [k] final relative residual of the last solve
  double* traj;          // optional: primary variable after every traj_every-th
  int traj_every;        //   step of a launch, at traj[(step + 1) / traj_every - 1][n][k]
  int extrap;            // predictor order: 0 x_n, 1 2 x_n - x_{n-1} (x_{n-1} kept in `a`),
                         //   2 3 x_n - 3 x_{n-1} + x_{n-2} (x_{n-2} kept in `v`),
                         //   3 4 x_n - 6 x_{n-1} + 4 x_{n-2} - x_{n-3} (x_{n-3} in `h3`)
  int hist;              // history levels used at launch start (0..extrap)
  int hist_raw;          // history levels held at launch start, before the cap by extrap
  int* hist_dev;         // optional: the held levels in device memory; the stepping kernels
                         //   read it at launch start and store min(order cap, levels + nsteps) at the end,
                         //   and a state write that changes the state stores 0
  double* h3;            // [n][k] x_{n-3} (cubic and quartic predictors)
  double* h4;            // [n][k] x_{n-4} (quartic predictor: 5 x_n - 10 x_{n-1} + 10 x_{n-2} - 5 x_{n-3} + x_{n-4})
  double* xnext;         // [n][k] next initial guess formed with the update (persistent kernel)
  double* hpre;          // [n][k] history part of the next initial residual
  double* hvec;          // [n][k] history part of the next initial guess, formed with the update:
                         //   phase_res gathers it instead of the history levels; Newmark on the
                         //   streaming engine: the next u + dt v + (1/2 - beta) dt^2 a
  double* h5;            // [n][k] x_{n-5}, x_{n-6} (quintic and sextic predictors, streaming engine)
  double* h6;
};

// A team: a contiguous range of workgroups that advances a list of instances
// one after the other, synchronising through its own barrier.
struct Team {
  int first_wg, size;
  int inst_begin, inst_end;  // range in the team-instance list
  int allpoll;               // 1: every workgroup polls all slots and reduces;
                             // 0: rank 0 reduces and releases through gen
  int pad_;
  Slot* slots;               // [size]
  double* scal;             // [kSlotBufs][kMaxRed] tagged sums published by rank 0
  int* abort_flag;
  long long wait_ticks;      // longest barrier wait (wall_clock64 ticks, 100 MHz) before the launch
                             // aborts; 0: abort at the first wait not already satisfied
};

// Default longest barrier wait: 4 s (a correct run never waits that long).
constexpr long long kDefaultWaitTicks = 400000000LL;

// ---------------------------------------------------------------------------
// Host helpers
// ---------------------------------------------------------------------------
void set_error(const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define FEM_HIP(call)                                   \
  do {                                                  \
    hipError_t e_ = (call);                             \
    if (e_ != hipSuccess) return hip_fail(e_, #call);   \
  } while (0)

template <typename T>
inline T* dalloc(size_t count, int* rc) {
  void* p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc(&p, count * sizeof(T));
  if (e != hipSuccess) {
    *rc = hip_fail(e, "hipMalloc");
    return nullptr;
  }
  return static_cast<T*>(p);
}

inline unsigned int div_up(long long a, long long b) { return (unsigned int)((a + b - 1) / b); }

// Device exclusive scan (int32 -> int32 offsets, total returned), scan.hip.
int exclusive_scan_i32(const int* in, int* out, int n, int* total, hipStream_t st);

}  // namespace femhip
