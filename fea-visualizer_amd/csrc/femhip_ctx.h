// femhip_ctx.h -- host-side context object (mesh + CSR pattern + matrices).
#pragma once

#include "femhip_internal.h"

struct fem_ctx {
  int nv = 0, ntri = 0, n = 0, bc = 0, max_row = 0;
  long long nnz = 0;
  int normalize = 1;
  double velocity[3] = {1.0, 0.0, 0.0};
  long long version = 0;  // bumped by every assembly
  hipStream_t stream = nullptr;
  float* d_xyz = nullptr;
  uint32_t* d_tri = nullptr;
  uint8_t* d_onb = nullptr;
  int* d_idx = nullptr;       // idx_map [nv]
  int* d_row2v = nullptr;     // inverse of idx_map [n]
  int* d_inc_ptr = nullptr;   // vertex -> incident triangles [nv + 1]
  int* d_inc = nullptr;
  int* d_rp = nullptr;        // CSR row pointers [n + 1]
  int* d_ci = nullptr;        // CSR column indices [nnz]
  int* d_diag = nullptr;      // diagonal position per row, -1 if absent
  int* d_tperm = nullptr;     // position of (c, r) for entry (r, c)
  double* d_K = nullptr;
  double* d_M = nullptr;
  double* d_A = nullptr;
  double* d_lumped = nullptr; // row sums of M
  double* d_kappa_e = nullptr;
  double* d_vel_e = nullptr;
  std::vector<int> h_rp;
};

struct fem_operator {
  fem_ctx* ctx = nullptr;
  int n = 0;
  long long nnz = 0;
  int precision = FEM_PREC_FP32_IR;  // FP32 values and vectors, or FP64
  void* d_val = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int lanes[17] = {0};         // chosen row-group width per column count (0 = default)
  double* d_host_x = nullptr;  // staging for the host convenience path
  double* d_host_y = nullptr;
  long long host_cap = 0;
};

namespace femhip {
int ctx_create(const fem_mesh* mesh, int bc, fem_ctx** out);
int ctx_assemble(fem_ctx* c, const fem_assembly_params* prm);
void ctx_free(fem_ctx* c);

// smoothed vertex normals of the extruded surface (ctx.hip)
int ctx_smooth_normals(fem_ctx* c, const float* nrm, const float* val, float extrusion, float* out);
int ctx_smooth_normals_device(fem_ctx* c, const float* d_nrm, const float* d_val, float extrusion, float* d_out,
                              hipStream_t st);

// stand-alone CSR operator (operator.hip)
int operator_create(fem_ctx* c, const double coef[3], int precision, int transpose, fem_operator** out);
void operator_free(fem_operator* op);
int operator_apply_device(fem_operator* op, int k, const void* x, void* y);
int operator_apply_host(fem_operator* op, int k, const double* x, double* y);
int operator_time(fem_operator* op, int k, const void* x, void* y, int reps, double* ms);
int operator_autotune(fem_operator* op, int k, const void* x, void* y, int reps, int* lanes, double* ms);
int operator_set_lanes(fem_operator* op, int k, int lanes);
}  // namespace femhip
