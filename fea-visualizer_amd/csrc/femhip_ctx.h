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

namespace femhip {
int ctx_create(const fem_mesh* mesh, int bc, fem_ctx** out);
int ctx_assemble(fem_ctx* c, const fem_assembly_params* prm);
void ctx_free(fem_ctx* c);
}  // namespace femhip
