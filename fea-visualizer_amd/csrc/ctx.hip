// ctx.hip -- mesh context: idx_map, CSR pattern and deterministic P1
// assembly of the stiffness K, mass M and advection A matrices on the GPU.
//
// Reference behaviour restated (tvumcc/fea-visualizer):
//   idx_map                 src/FEM/FEMContext.cpp:224-240
//   K (surface gradients)   src/FEM/FEMContext.cpp:70-113
//   M (|J|/24 [2 1 1;..])   src/FEM/FEMContext.cpp:119-150
//   A (area/3 v.grad phi_j) src/FEM/FEMContext.cpp:156-202
//   max row nonzeros        src/FEM/FEMContext.cpp:207-219
// The pattern is the set of (idx(i), idx(j)) over triangles sharing i and j
// that pass the boundary filter, sorted per row, explicit zeros kept -- the
// pattern Eigen's setFromTriplets builds, which every system matrix of the
// reference shares.  Values are summed per entry in (element, i, j) order, the
// order setFromTriplets sums duplicate triplets in, without atomics: one
// thread owns one row and walks the row vertex's incident triangles in
// ascending element order.
#include "femhip_ctx.h"
#include "femhip_geom.h"

#include <algorithm>
#include <cstring>

namespace femhip {

namespace {

__global__ void k_keep(const uint8_t* onb, int nv, int bc, int* keep) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nv) keep[i] = (bc == FEM_NEUMANN || !onb[i]) ? 1 : 0;
}

__global__ void k_idx_map(const uint8_t* onb, int nv, int bc, const int* offs, int* idx, int* row2v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  bool keep = (bc == FEM_NEUMANN || !onb[i]);
  int r = keep ? offs[i] : -1;
  idx[i] = r;
  if (keep) row2v[r] = i;
}

// Incidence counts: each triangle counted once per distinct vertex.
__global__ void k_inc_count(const uint32_t* tri, int ntri, int* cnt) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntri) return;
  uint32_t a = tri[3 * (size_t)t], b = tri[3 * (size_t)t + 1], c = tri[3 * (size_t)t + 2];
  atomicAdd(&cnt[a], 1);
  if (b != a) atomicAdd(&cnt[b], 1);
  if (c != a && c != b) atomicAdd(&cnt[c], 1);
}

__global__ void k_inc_fill(const uint32_t* tri, int ntri, const int* inc_ptr, int* cursor, int* inc) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntri) return;
  uint32_t v3[3] = {tri[3 * (size_t)t], tri[3 * (size_t)t + 1], tri[3 * (size_t)t + 2]};
  for (int a = 0; a < 3; a++) {
    bool dup = false;
    for (int q = 0; q < a; q++) dup |= (v3[q] == v3[a]);
    if (dup) continue;
    int pos = inc_ptr[v3[a]] + atomicAdd(&cursor[v3[a]], 1);
    inc[pos] = t;
  }
}

// Atomic fill order is arbitrary; sorting each vertex's list restores the
// ascending element order the summation relies on.
__global__ void k_inc_sort(const int* inc_ptr, int nv, int* inc) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  int lo = inc_ptr[v], hi = inc_ptr[v + 1];
  for (int i = lo + 1; i < hi; i++) {
    int key = inc[i];
    int j = i - 1;
    while (j >= lo && inc[j] > key) {
      inc[j + 1] = inc[j];
      j--;
    }
    inc[j + 1] = key;
  }
}

// Distinct columns of row r (vertex i) in ascending order, by repeated
// minimum search over the incident triangles (rows hold ~7 entries, so this
// needs no per-thread list).  mode 0 counts, mode 1 writes col_idx and the
// diagonal position.
__global__ void k_row_cols(const uint32_t* tri, const uint8_t* onb, const int* idx, const int* row2v,
                           const int* inc_ptr, const int* inc, int n, int bc, int mode, int* cnt,
                           const int* rp, int* ci, int* diag) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int i = row2v[r];
  bool bi = onb[i] != 0;
  int lo = inc_ptr[i], hi = inc_ptr[i + 1];
  int last = -1, c = 0, out = (mode == 1) ? rp[r] : 0, dpos = -1;
  while (true) {
    int best = INT32_MAX;
    for (int e = lo; e < hi; e++) {
      int t = inc[e];
      for (int b = 0; b < 3; b++) {
        uint32_t w = tri[3 * (size_t)t + b];
        if (!(bc == FEM_NEUMANN || (!bi && !onb[w]))) continue;
        int col = idx[w];
        if (col > last && col < best) best = col;
      }
    }
    if (best == INT32_MAX) break;
    if (mode == 1) {
      ci[out + c] = best;
      if (best == r) dpos = out + c;
    }
    c++;
    last = best;
  }
  if (mode == 0) cnt[r] = c;
  else diag[r] = dpos;
}

__device__ inline int find_slot(const int* ci, int lo, int hi, int col) {
  for (int s = lo; s < hi; s++)
    if (ci[s] == col) return s;
  return -1;
}

__global__ void k_assemble(const float* xyz, const uint32_t* tri, const uint8_t* onb, const int* idx,
                           const int* row2v, const int* inc_ptr, const int* inc, const int* rp,
                           const int* ci, int n, int bc, double vx, double vy, double vz,
                           const double* vel_e, int normalize, const double* kappa_e, double* K,
                           double* M, double* A, double* lumped) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int i = row2v[r];
  int lo = rp[r], hi = rp[r + 1];
  for (int s = lo; s < hi; s++) {
    K[s] = 0.0;
    M[s] = 0.0;
    A[s] = 0.0;
  }
  bool bi = onb[i] != 0;
  for (int e = inc_ptr[i]; e < inc_ptr[i + 1]; e++) {
    int t = inc[e];
    const uint32_t* tv3 = tri + 3 * (size_t)t;
    Geo G;
    elem_geom(xyz, tv3, G);
    double kap = kappa_e ? kappa_e[t] : 1.0;
    double mo = G.jdet / 24.0;
    double vel[3] = {vx, vy, vz};
    if (vel_e) {
      vel[0] = vel_e[3 * (size_t)t];
      vel[1] = vel_e[3 * (size_t)t + 1];
      vel[2] = vel_e[3 * (size_t)t + 2];
    }
    // FEMContext.cpp:179-180: tangential projection, then normalised
    double nn = G.nrm[0] * G.nrm[0] + G.nrm[1] * G.nrm[1] + G.nrm[2] * G.nrm[2];
    double vd = (vel[0] * G.nrm[0] + vel[1] * G.nrm[1] + vel[2] * G.nrm[2]) / nn;
    double tvel[3];
    for (int q = 0; q < 3; q++) tvel[q] = vel[q] - vd * G.nrm[q];
    if (normalize) {
      double z = tvel[0] * tvel[0] + tvel[1] * tvel[1] + tvel[2] * tvel[2];
      if (z > 0.0) {
        double sz = sqrt(z);
        for (int q = 0; q < 3; q++) tvel[q] = tvel[q] / sz;
      }
    }
    double jd6 = G.jdet / 6.0;
    for (int a = 0; a < 3; a++) {
      if (tv3[a] != (uint32_t)i) continue;
      for (int b = 0; b < 3; b++) {
        uint32_t w = tv3[b];
        bool bw = onb[w] != 0;
        bool keep_km = (bc == FEM_NEUMANN) || (!bi && !bw);  // FEMContext.cpp:99,136
        bool keep_a = !bi && !bw;                              // FEMContext.cpp:188
        if (!keep_km && !keep_a) continue;
        int s = find_slot(ci, lo, hi, idx[w]);
        if (s < 0) continue;
        if (keep_km) {
          double kab = elem_kab(G, a, b);
          K[s] += kap * kab;
          M[s] += (a == b) ? 2.0 * mo : mo;
        }
        if (keep_a) A[s] += jd6 * (tvel[0] * G.g[b][0] + tvel[1] * G.g[b][1] + tvel[2] * G.g[b][2]);
      }
    }
  }
  double acc = 0.0;
  for (int s = lo; s < hi; s++) acc += M[s];
  lumped[r] = acc;
}

// Position of (c, r) for every CSR entry (r, c); the pattern is structurally
// symmetric, so A^T reuses row_ptr / col_idx with permuted values.
__global__ void k_tperm(const int* rp, const int* ci, int n, int* tperm) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  for (int s = rp[r]; s < rp[r + 1]; s++) {
    int c = ci[s];
    int lo = rp[c], hi = rp[c + 1];
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (ci[mid] < r) lo = mid + 1;
      else hi = mid;
    }
    tperm[s] = lo;
  }
}

}  // namespace

void ctx_free(fem_ctx* c) {
  if (!c) return;
  void* ptrs[] = {c->d_xyz, c->d_tri, c->d_onb, c->d_idx, c->d_row2v, c->d_inc_ptr, c->d_inc,
                  c->d_rp, c->d_ci, c->d_diag, c->d_tperm, c->d_K, c->d_M, c->d_A, c->d_lumped,
                  c->d_kappa_e, c->d_vel_e};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int ctx_assemble(fem_ctx* c, const fem_assembly_params* prm) {
  int rc = 0;
  fem_assembly_params p;
  if (prm) {
    p = *prm;
  } else {
    std::memset(&p, 0, sizeof(p));
    p.velocity[0] = 1.0;  // AdvectionDiffusionParameters default (1,0,0)
    p.normalize_velocity = 1;
  }
  if (c->d_kappa_e) { (void)hipFree(c->d_kappa_e); c->d_kappa_e = nullptr; }
  if (c->d_vel_e) { (void)hipFree(c->d_vel_e); c->d_vel_e = nullptr; }
  if (p.kappa_e) {
    c->d_kappa_e = dalloc<double>((size_t)c->ntri, &rc);
    if (!c->d_kappa_e) return rc;
    FEM_HIP(hipMemcpyAsync(c->d_kappa_e, p.kappa_e, sizeof(double) * c->ntri, hipMemcpyHostToDevice, c->stream));
  }
  if (p.velocity_e) {
    c->d_vel_e = dalloc<double>((size_t)c->ntri * 3, &rc);
    if (!c->d_vel_e) return rc;
    FEM_HIP(hipMemcpyAsync(c->d_vel_e, p.velocity_e, sizeof(double) * 3 * c->ntri, hipMemcpyHostToDevice, c->stream));
  }
  c->normalize = p.normalize_velocity;
  for (int q = 0; q < 3; q++) c->velocity[q] = p.velocity[q];
  if (c->n > 0) {
    hipLaunchKernelGGL(k_assemble, dim3(div_up(c->n, kBlock)), dim3(kBlock), 0, c->stream, c->d_xyz, c->d_tri,
                       c->d_onb, c->d_idx, c->d_row2v, c->d_inc_ptr, c->d_inc, c->d_rp, c->d_ci, c->n, c->bc,
                       p.velocity[0], p.velocity[1], p.velocity[2], c->d_vel_e, p.normalize_velocity,
                       c->d_kappa_e, c->d_K, c->d_M, c->d_A, c->d_lumped);
    FEM_HIP(hipGetLastError());
  }
  FEM_HIP(hipStreamSynchronize(c->stream));
  c->version++;
  return 0;
}

int ctx_create(const fem_mesh* mesh, int bc, fem_ctx** out) {
  int rc = 0;
  fem_ctx* c = new fem_ctx();
  c->nv = mesh->num_vertices;
  c->ntri = mesh->num_triangles;
  c->bc = bc;
  const int nv = c->nv, nt = c->ntri;
  for (long long q = 0; q < 3LL * nt; q++) {
    if (mesh->triangles[q] >= (uint32_t)nv) {
      ctx_free(c);
      set_error("fem_ctx_create: triangle index out of range");
      return FEM_ERR_INVALID;
    }
  }
#define CK(x)                  \
  do {                         \
    int rc_ = (x);             \
    if (rc_) {                 \
      ctx_free(c);             \
      return rc_;              \
    }                          \
  } while (0)
#define CKH(x)                                  \
  do {                                          \
    hipError_t e_ = (x);                        \
    if (e_ != hipSuccess) {                     \
      ctx_free(c);                              \
      return hip_fail(e_, #x);                  \
    }                                           \
  } while (0)
  CKH(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  c->d_xyz = dalloc<float>((size_t)nv * 3, &rc);
  c->d_tri = dalloc<uint32_t>((size_t)nt * 3, &rc);
  c->d_onb = dalloc<uint8_t>((size_t)nv, &rc);
  c->d_idx = dalloc<int>((size_t)nv + 1, &rc);
  c->d_inc_ptr = dalloc<int>((size_t)nv + 1, &rc);
  c->d_inc = dalloc<int>((size_t)nt * 3, &rc);
  if (rc) { ctx_free(c); return rc; }
  hipStream_t st = c->stream;
  CKH(hipMemcpyAsync(c->d_xyz, mesh->xyz, sizeof(float) * 3 * nv, hipMemcpyHostToDevice, st));
  CKH(hipMemcpyAsync(c->d_tri, mesh->triangles, sizeof(uint32_t) * 3 * nt, hipMemcpyHostToDevice, st));
  CKH(hipMemcpyAsync(c->d_onb, mesh->on_boundary, (size_t)nv, hipMemcpyHostToDevice, st));
  // idx_map: exclusive scan of the keep flags (FEMContext.cpp:224-240)
  int* keep = dalloc<int>((size_t)nv + 1, &rc);
  if (!keep) { ctx_free(c); return rc; }
  int n = 0;
  if (nv > 0) {
    hipLaunchKernelGGL(k_keep, dim3(div_up(nv, kBlock)), dim3(kBlock), 0, st, c->d_onb, nv, bc, keep);
    rc = exclusive_scan_i32(keep, keep, nv, &n, st);
  }
  if (rc) { (void)hipFree(keep); ctx_free(c); return rc; }
  c->n = n;
  c->d_row2v = dalloc<int>((size_t)n + 1, &rc);
  if (rc) { (void)hipFree(keep); ctx_free(c); return rc; }
  if (nv > 0)
    hipLaunchKernelGGL(k_idx_map, dim3(div_up(nv, kBlock)), dim3(kBlock), 0, st, c->d_onb, nv, bc, keep,
                       c->d_idx, c->d_row2v);
  // vertex -> incident triangles, ascending
  CKH(hipMemsetAsync(c->d_inc_ptr, 0, sizeof(int) * (nv + 1), st));
  if (nt > 0)
    hipLaunchKernelGGL(k_inc_count, dim3(div_up(nt, kBlock)), dim3(kBlock), 0, st, c->d_tri, nt, c->d_inc_ptr);
  int ninc = 0;
  CK(exclusive_scan_i32(c->d_inc_ptr, c->d_inc_ptr, nv + 1, &ninc, st));
  CKH(hipMemsetAsync(keep, 0, sizeof(int) * (nv + 1), st));
  if (nt > 0) {
    hipLaunchKernelGGL(k_inc_fill, dim3(div_up(nt, kBlock)), dim3(kBlock), 0, st, c->d_tri, nt, c->d_inc_ptr,
                       keep, c->d_inc);
  }
  if (nv > 0)
    hipLaunchKernelGGL(k_inc_sort, dim3(div_up(nv, kBlock)), dim3(kBlock), 0, st, c->d_inc_ptr, nv, c->d_inc);
  (void)hipFree(keep);
  // CSR pattern: count, scan, fill
  c->d_rp = dalloc<int>((size_t)n + 1, &rc);
  c->d_diag = dalloc<int>((size_t)n + 1, &rc);
  c->d_lumped = dalloc<double>((size_t)n + 1, &rc);
  if (rc) { ctx_free(c); return rc; }
  int nnz = 0;
  if (n > 0) {
    hipLaunchKernelGGL(k_row_cols, dim3(div_up(n, kBlock)), dim3(kBlock), 0, st, c->d_tri, c->d_onb, c->d_idx,
                       c->d_row2v, c->d_inc_ptr, c->d_inc, n, bc, 0, c->d_rp, nullptr, nullptr, nullptr);
    CKH(hipMemsetAsync(c->d_rp + n, 0, sizeof(int), st));
    CKH(hipGetLastError());
    CK(exclusive_scan_i32(c->d_rp, c->d_rp, n + 1, &nnz, st));
  } else {
    CKH(hipMemsetAsync(c->d_rp, 0, sizeof(int), st));
  }
  c->nnz = nnz;
  c->d_ci = dalloc<int>((size_t)nnz + 1, &rc);
  c->d_tperm = dalloc<int>((size_t)nnz + 1, &rc);
  c->d_K = dalloc<double>((size_t)nnz + 1, &rc);
  c->d_M = dalloc<double>((size_t)nnz + 1, &rc);
  c->d_A = dalloc<double>((size_t)nnz + 1, &rc);
  if (rc) { ctx_free(c); return rc; }
  if (n > 0) {
    hipLaunchKernelGGL(k_row_cols, dim3(div_up(n, kBlock)), dim3(kBlock), 0, st, c->d_tri, c->d_onb, c->d_idx,
                       c->d_row2v, c->d_inc_ptr, c->d_inc, n, bc, 1, nullptr, c->d_rp, c->d_ci, c->d_diag);
    hipLaunchKernelGGL(k_tperm, dim3(div_up(n, kBlock)), dim3(kBlock), 0, st, c->d_rp, c->d_ci, n, c->d_tperm);
    CKH(hipGetLastError());
  }
  // max row nonzeros (FEMContext.cpp:207-219) from the host copy of row_ptr
  c->h_rp.resize((size_t)n + 1);
  CKH(hipMemcpyAsync(c->h_rp.data(), c->d_rp, sizeof(int) * (n + 1), hipMemcpyDeviceToHost, st));
  CKH(hipStreamSynchronize(st));
  c->max_row = 0;
  for (int r = 0; r < n; r++) c->max_row = std::max(c->max_row, c->h_rp[r + 1] - c->h_rp[r]);
  CK(ctx_assemble(c, nullptr));
#undef CK
#undef CKH
  *out = c;
  return 0;
}

}  // namespace femhip

namespace femhip {
namespace {

// smooth_normals.glsl:22-57 without atomics: each vertex sums, over its
// incident triangles in ascending order, the unnormalised normal of the
// triangle whose vertices are moved along their normals by
// min(1, extrusion * max(0, value)).  FP32 as the shader.
__global__ void k_smooth_normals(int nv, const float* xyz, const uint32_t* tri, const int* inc_ptr, const int* inc,
                                 const float* nrm, const float* val, float extrusion, float* out) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  float s0 = 0.f, s1 = 0.f, s2 = 0.f;
  for (int e = inc_ptr[v]; e < inc_ptr[v + 1]; e++) {
    const uint32_t* t = tri + 3 * (size_t)inc[e];
    float p[3][3];
    for (int a = 0; a < 3; a++) {
      const size_t w = t[a];
      const float h = fminf(1.0f, extrusion * fmaxf(0.0f, val[w]));
      for (int q = 0; q < 3; q++) p[a][q] = xyz[3 * w + q] + h * nrm[3 * w + q];
    }
    const float e1[3] = {p[1][0] - p[0][0], p[1][1] - p[0][1], p[1][2] - p[0][2]};
    const float e2[3] = {p[2][0] - p[0][0], p[2][1] - p[0][1], p[2][2] - p[0][2]};
    s0 += e1[1] * e2[2] - e1[2] * e2[1];
    s1 += e1[2] * e2[0] - e1[0] * e2[2];
    s2 += e1[0] * e2[1] - e1[1] * e2[0];
  }
  out[3 * (size_t)v] = s0;
  out[3 * (size_t)v + 1] = s1;
  out[3 * (size_t)v + 2] = s2;
}

}  // namespace

// normals/out: device [nv][3]; values: device [nv].
int ctx_smooth_normals_device(fem_ctx* c, const float* d_nrm, const float* d_val, float extrusion, float* d_out,
                              hipStream_t st) {
  if (c->nv == 0) return 0;
  hipLaunchKernelGGL(k_smooth_normals, dim3(div_up(c->nv, kBlock)), dim3(kBlock), 0, st, c->nv, c->d_xyz, c->d_tri,
                     c->d_inc_ptr, c->d_inc, d_nrm, d_val, extrusion, d_out);
  FEM_HIP(hipGetLastError());
  return 0;
}

int ctx_smooth_normals(fem_ctx* c, const float* nrm, const float* val, float extrusion, float* out) {
  const size_t nv = (size_t)c->nv;
  if (nv == 0) return 0;
  int rc = 0;
  float* d = dalloc<float>(7 * nv, &rc);  // normals, values, output
  if (!d) return rc;
  float *dn = d, *dv = d + 3 * nv, *dout = d + 4 * nv;
  hipError_t e = hipMemcpyAsync(dn, nrm, sizeof(float) * 3 * nv, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dv, val, sizeof(float) * nv, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    rc = ctx_smooth_normals_device(c, dn, dv, extrusion, dout, c->stream);
    if (!rc) e = hipMemcpyAsync(out, dout, sizeof(float) * 3 * nv, hipMemcpyDeviceToHost, c->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return hip_fail(e, "smooth_normals");
  return rc;
}

}  // namespace femhip
