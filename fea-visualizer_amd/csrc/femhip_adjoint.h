// femhip_adjoint.h -- host object of the discrete adjoint (adjoint.hip).
#pragma once

#include "femhip_step.h"

struct fem_adjoint {
  fem_ctx* ctx = nullptr;
  fem_batch* fwd = nullptr;  // forward instance: system matrix S
  fem_batch* adj = nullptr;  // adjoint instance: the same equation with S^T
  int eq = -1;
  double dt = 0.0, coef = 0.0;  // time step, coefficient of K in S
  int nsteps = 0, every = 0;
  long long fwd_version = -1;   // context assembly the forward ran on
  bool have_forward = false;
  double* d_ckpt = nullptr;     // u_0, u_every, u_2every, ...  [nck][n]
  long long ckpt_cap = 0;
  double* d_U = nullptr;        // recomputed states of a segment [every][n]
  double* d_MU = nullptr;       // adjoint states of a segment [every][n]
  int seg_cap = 0;
  double* d_carry = nullptr;    // adjoint state at the segment end
  double* d_gk = nullptr;       // dL/dkappa_e [ntri]
  double* d_tmp = nullptr;
};

namespace femhip {
int adjoint_create(fem_ctx* ctx, const fem_solver_opts* opts, fem_adjoint** out);
void adjoint_free(fem_adjoint* a);
int adjoint_set_equation(fem_adjoint* a, int eq, const void* params, size_t size);
int adjoint_forward(fem_adjoint* a, const double* u0, int nsteps, int every, double* uT, fem_step_stats* stats,
                    bool dev);
int adjoint_backward(fem_adjoint* a, const double* gT, double* gu0, double* gk, fem_step_stats* stats, bool dev);
}  // namespace femhip
