// operator.hip -- the system matrix S = cM M + cK K + cA A of a context as a
// stand-alone CSR operator (values in FP32 or FP64, int32 columns), applied
// to 1, 2, 4 or 16 right-hand sides.  This is the SpMV of the reference's CG
// (Eigen's sparse product at CPUSolver.cpp:78,96,114,134,137; the ELL row
// loop of cgm.glsl:75-102) as a kernel of its own, for callers that need the
// product alone and for measuring it against the HBM roofline.
//
// Schedules.  Row lanes: G lanes per row (G in {1,2,4,8,16}, G >= nrhs); the
// G lanes of a row split into G/K entry slots x K columns, lane (slot, c)
// walks entries slot, slot + G/K, ... of the row, gathers x[col][c], and the
// slots are summed with lane shuffles inside the group; all entry loads of a
// lane are issued before the first is consumed.  Row-block stream: see
// k_spmv_stream.  The schedule is picked per operator and column count by
// fem_operator_autotune.
#include "femhip_ctx.h"

#include <algorithm>

namespace femhip {

namespace {

__global__ void k_op_form(int n, const int* rp, const int* tperm, const double* M, const double* Kk,
                          const double* A, double cM, double cK, double cA, int fp32, void* val) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  for (int s = rp[r]; s < rp[r + 1]; s++) {
    const int t = tperm ? tperm[s] : s;
    double v = 0.0;
    if (cM != 0.0) v += cM * M[t];
    if (cK != 0.0) v += cK * Kk[t];
    if (cA != 0.0) v += cA * A[t];
    if (fp32) ((float*)val)[s] = (float)v;
    else ((double*)val)[s] = v;
  }
}

template <typename T, int K, int G>
__global__ void __launch_bounds__(kBlock) k_spmv(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                 const T* __restrict__ val, const T* __restrict__ x,
                                                 T* __restrict__ y) {
  constexpr int S = G / K;                // entry slots per row
  constexpr int U = (S >= 8) ? 2 : (S >= 4) ? 3 : (S >= 2) ? 5 : 9;  // slots x U >= 9 entries
  const long long gt = (long long)blockIdx.x * kBlock + threadIdx.x;
  const long long row = gt / G;
  if (row >= n) return;  // a whole group leaves together
  const int lane = (int)(threadIdx.x % G);
  const int c = lane % K, slot = lane / K;
  const int lo = rp[row], hi = rp[row + 1];
  T acc = 0;
  for (int s0 = lo + slot; s0 < hi; s0 += S * U) {
    T t[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int s = min(s0 + u * S, hi - 1);
      t[u] = val[s] * x[(size_t)ci[s] * K + c];
    }
#pragma unroll
    for (int u = 0; u < U; u++)
      if (s0 + u * S < hi) acc += t[u];
  }
#pragma unroll
  for (int off = K; off < G; off <<= 1) acc += __shfl_xor(acc, off, G);
  if (slot == 0) y[(size_t)row * K + c] = acc;
}

// Row-block stream schedule.  Each lane owns V consecutive columns of a row
// (V = 16 bytes of the value type, at most K), so K / V lanes share a row and
// a workgroup owns R = 256 * V / K consecutive rows.  The workgroup stages its
// rows' column and value stream [rp[row0], rp[row0 + R]) in LDS with 16-byte
// loads from a 16-byte aligned start (every load fully used, all of them in
// flight at once); then each lane walks its row out of LDS and gathers V
// columns of x per entry with one 16-byte load.  A block whose stream exceeds
// the LDS capacity (rows averaging more than kCapRow entries) reads its rows
// from global memory instead.  NT marks the stream loads non-temporal, so the
// single-use matrix bytes do not displace x from the caches.
constexpr int kCapRow = 10;

typedef int v4i __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));

template <bool NT, typename V>
__device__ inline V load16(const V* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}

template <typename T, int K>
struct StreamShape {
  static constexpr int V = (16 / (int)sizeof(T)) < K ? (16 / (int)sizeof(T)) : K;  // columns per lane
  static constexpr int LPR = K / V;                                                 // lanes per row
  static constexpr int R = kBlock / LPR;                                            // rows per block
};

template <typename T, int K, bool NT>
__global__ void __launch_bounds__(kBlock) k_spmv_stream(int n, long long nnz, const int* __restrict__ rp,
                                                        const int* __restrict__ ci, const T* __restrict__ val,
                                                        const T* __restrict__ x, T* __restrict__ y) {
  using SS = StreamShape<T, K>;
  constexpr int V = SS::V, LPR = SS::LPR, R = SS::R;
  constexpr int CAP = R * kCapRow;  // multiple of 4
  typedef T vec __attribute__((ext_vector_type(V)));
  __shared__ int s_ci[CAP];
  __shared__ T s_val[CAP];
  const int row0 = blockIdx.x * R;
  const int row1 = min(n, row0 + R);
  const int row = row0 + (int)threadIdx.x / LPR;
  const int c = ((int)threadIdx.x % LPR) * V;
  const int lo = rp[row0], hi = rp[row1];
  int rlo = 0, rhi = 0;
  if (row < n) {
    rlo = rp[row];
    rhi = rp[row + 1];
  }
  const int base = lo & ~3;
  const bool staged = hi - base <= CAP;
  if (staged) {
    const int nch = (hi - base + 3) >> 2;  // 4-entry chunks
    for (int q = threadIdx.x; q < nch; q += kBlock) {
      const long long e0 = (long long)base + 4 * q;
      if (e0 + 3 < nnz) {
        const v4i cv = load16<NT>((const v4i*)(ci + e0));
        s_ci[4 * q] = cv.x;
        s_ci[4 * q + 1] = cv.y;
        s_ci[4 * q + 2] = cv.z;
        s_ci[4 * q + 3] = cv.w;
        if constexpr (sizeof(T) == 4) {
          const v4f vv = load16<NT>((const v4f*)(val + e0));
          s_val[4 * q] = vv.x;
          s_val[4 * q + 1] = vv.y;
          s_val[4 * q + 2] = vv.z;
          s_val[4 * q + 3] = vv.w;
        } else {
          const v2d va = load16<NT>((const v2d*)(val + e0));
          const v2d vb = load16<NT>((const v2d*)(val + e0 + 2));
          s_val[4 * q] = va.x;
          s_val[4 * q + 1] = va.y;
          s_val[4 * q + 2] = vb.x;
          s_val[4 * q + 3] = vb.y;
        }
      } else {  // the tail of the arrays: element loads
        for (int u = 0; u < 4; u++)
          if (e0 + u < nnz) {
            s_ci[4 * q + u] = ci[e0 + u];
            s_val[4 * q + u] = val[e0 + u];
          }
      }
    }
  }
  __syncthreads();
  if (row >= n) return;
  // entries come from LDS (offset by the staged start) or, for an
  // oversized block, straight from global memory; separate loops keep LDS
  // and global accesses on their own instructions
  auto walk = [&](const int* cs, const T* vs, int off) {
    vec acc = (vec)(T)0;
    for (int s0 = rlo; s0 < rhi; s0 += 9) {
      vec t[9];
#pragma unroll
      for (int u = 0; u < 9; u++) {
        const int q = min(s0 + u, rhi - 1) - off;
        t[u] = vs[q] * *(const vec*)(x + (size_t)cs[q] * K + c);
      }
#pragma unroll
      for (int u = 0; u < 9; u++)
        if (s0 + u < rhi) acc += t[u];
    }
    return acc;
  };
  const vec acc = staged ? walk(s_ci, s_val, base) : walk(ci, val, 0);
  if constexpr (NT) __builtin_nontemporal_store(acc, (vec*)(y + (size_t)row * K + c));  // y is not re-read here
  else *(vec*)(y + (size_t)row * K + c) = acc;
}

// Pipelined row-block stream: the stream schedule's per-block work, with a
// persistent grid walking the row blocks (block b = blockIdx.x + j gridDim.x,
// gridDim.x a multiple of 8, so every block stays on the XCD the plain
// schedule deals it to).  A block's set-up is two dependent round trips, its
// row pointers and then its column/value stream, which the plain schedule
// waits for before every block's gathers.  Here, while a workgroup gathers
// block j out of LDS, the stream of block j + 1 is in flight into registers
// and the row pointers of block j + 2 behind it; block j + 1's stream goes to
// the other LDS buffer after the gathers, and one barrier per block follows.
template <typename T, int K, bool NT>
__global__ void __launch_bounds__(kBlock) k_spmv_pipe(int n, long long nnz, int nblk, const int* __restrict__ rp,
                                                      const int* __restrict__ ci, const T* __restrict__ val,
                                                      const T* __restrict__ x, T* __restrict__ y) {
  using SS = StreamShape<T, K>;
  constexpr int V = SS::V, LPR = SS::LPR, R = SS::R;
  constexpr int CAP = R * kCapRow;                       // multiple of 4
  constexpr int NCH = (CAP / 4 + kBlock - 1) / kBlock;   // 4-entry chunks per thread
  typedef T vec __attribute__((ext_vector_type(V)));
  typedef T v4t __attribute__((ext_vector_type(4)));
  __shared__ int s_ci[2][CAP];
  __shared__ T s_val[2][CAP];
  const int tid = (int)threadIdx.x;
  const int c = (tid % LPR) * V;
  struct Ptrs {
    int lo, hi, rlo, rhi;
  };
  struct Chunks {
    v4i ci[NCH];
    v4t va[NCH];
  };
  // round trip 1: the block's nonzero range and the lane's row
  auto load_ptrs = [&](int b) {
    Ptrs P{0, 0, 0, 0};
    if (b < nblk) {
      const int row0 = b * R, row1 = min(n, row0 + R);
      const int row = row0 + tid / LPR;
      P.lo = rp[row0];
      P.hi = rp[row1];
      if (row < n) {
        P.rlo = rp[row];
        P.rhi = rp[row + 1];
      }
    }
    return P;
  };
  auto staged = [&](const Ptrs& P) { return P.hi - (P.lo & ~3) <= CAP; };
  // round trip 2: the block's column/value stream, 16-byte loads from a
  // 16-byte aligned start (element loads at the arrays' tail)
  auto load_chunks = [&](const Ptrs& P, bool valid, Chunks& C) {
    if (!valid || !staged(P)) return;
    const int base = P.lo & ~3;
    const int nch = (P.hi - base + 3) >> 2;
#pragma unroll
    for (int u = 0; u < NCH; u++) {
      const int q = tid + u * kBlock;
      if (q >= nch) continue;
      const long long e0 = (long long)base + 4 * q;
      if (e0 + 3 < nnz) {
        C.ci[u] = load16<NT>((const v4i*)(ci + e0));
        C.va[u] = load16<NT>((const v4t*)(val + e0));
      } else {
#pragma unroll
        for (int w = 0; w < 4; w++) {
          C.ci[u][w] = e0 + w < nnz ? ci[e0 + w] : 0;
          C.va[u][w] = e0 + w < nnz ? val[e0 + w] : (T)0;
        }
      }
    }
  };
  auto store_chunks = [&](const Ptrs& P, bool valid, const Chunks& C, int buf) {
    if (!valid || !staged(P)) return;
    const int nch = (P.hi - (P.lo & ~3) + 3) >> 2;
#pragma unroll
    for (int u = 0; u < NCH; u++) {
      const int q = tid + u * kBlock;
      if (q >= nch) continue;
#pragma unroll
      for (int w = 0; w < 4; w++) {
        s_ci[buf][4 * q + w] = C.ci[u][w];
        s_val[buf][4 * q + w] = C.va[u][w];
      }
    }
  };
  auto walk = [&](const int* cs, const T* vs, int off, int rlo, int rhi) {
    vec acc = (vec)(T)0;
    for (int s0 = rlo; s0 < rhi; s0 += 9) {
      vec t[9];
#pragma unroll
      for (int u = 0; u < 9; u++) {
        const int q = min(s0 + u, rhi - 1) - off;
        t[u] = vs[q] * *(const vec*)(x + (size_t)cs[q] * K + c);
      }
#pragma unroll
      for (int u = 0; u < 9; u++)
        if (s0 + u < rhi) acc += t[u];
    }
    return acc;
  };
  const int G = (int)gridDim.x;
  int b = (int)blockIdx.x;
  if (b >= nblk) return;  // the whole workgroup
  Ptrs P0 = load_ptrs(b);
  Chunks C0;
  load_chunks(P0, true, C0);
  Ptrs P1 = load_ptrs(b + G);
  store_chunks(P0, true, C0, 0);
  __syncthreads();
  int buf = 0;
  for (; b < nblk; b += G) {
    const bool has1 = b + G < nblk;
    const Ptrs P2 = load_ptrs(b + 2 * G);  // round trip 1 of block j + 2
    Chunks C1;
    load_chunks(P1, has1, C1);             // round trip 2 of block j + 1
    const int row = b * R + tid / LPR;
    if (row < n) {
      const vec acc = staged(P0) ? walk(s_ci[buf], s_val[buf], P0.lo & ~3, P0.rlo, P0.rhi)
                                 : walk(ci, val, 0, P0.rlo, P0.rhi);
      if constexpr (NT) __builtin_nontemporal_store(acc, (vec*)(y + (size_t)row * K + c));
      else *(vec*)(y + (size_t)row * K + c) = acc;
    }
    store_chunks(P1, has1, C1, buf ^ 1);
    __syncthreads();
    P0 = P1;
    P1 = P2;
    buf ^= 1;
  }
}

// Workgroups of k_spmv_pipe's grid: as many as fit on the device at once,
// a multiple of 8 (one share per XCD), at most one per row block.
template <typename T, int K, bool NT>
int pipe_grid(int nblk, int* grid) {
  static int occ = 0;  // the same for every device of the build's target
  if (occ == 0) {
    int o = 0;
    FEM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, (const void*)k_spmv_pipe<T, K, NT>, kBlock, 0));
    occ = std::max(1, o);
  }
  int dev = 0, cus = 0;
  FEM_HIP(hipGetDevice(&dev));
  FEM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int full = std::max(8, (cus * occ) & ~7);
  *grid = std::max(1, std::min(nblk, full));
  return 0;
}

template <typename T, int K>
const void* pick_g(int g) {
  switch (g) {
    case 1: if constexpr (K <= 1) return (const void*)k_spmv<T, K, 1>; else return nullptr;
    case 2: if constexpr (K <= 2) return (const void*)k_spmv<T, K, 2>; else return nullptr;
    case 4: if constexpr (K <= 4) return (const void*)k_spmv<T, K, 4>; else return nullptr;
    case 8: if constexpr (K <= 8) return (const void*)k_spmv<T, K, 8>; else return nullptr;
    case 16: return (const void*)k_spmv<T, K, 16>;
    case FEM_SPMV_STREAM: return (const void*)k_spmv_stream<T, K, false>;
    case FEM_SPMV_STREAM_NT: return (const void*)k_spmv_stream<T, K, true>;
    case FEM_SPMV_PIPE: return (const void*)k_spmv_pipe<T, K, false>;
    case FEM_SPMV_PIPE_NT: return (const void*)k_spmv_pipe<T, K, true>;
    default: return nullptr;
  }
}

template <typename T>
const void* pick_k(int k, int g) {
  switch (k) {
    case 1: return pick_g<T, 1>(g);
    case 2: return pick_g<T, 2>(g);
    case 4: return pick_g<T, 4>(g);
    case 16: return pick_g<T, 16>(g);
    default: return nullptr;
  }
}

template <typename T>
int stream_rows(int k) {
  switch (k) {
    case 1: return StreamShape<T, 1>::R;
    case 2: return StreamShape<T, 2>::R;
    case 4: return StreamShape<T, 4>::R;
    default: return StreamShape<T, 16>::R;
  }
}

int default_lanes(int k) { return k >= 16 ? 16 : FEM_SPMV_STREAM_NT; }

template <typename T>
int pipe_grid_k(int k, bool nt, int nblk, int* grid) {
  switch (k) {
    case 1: return nt ? pipe_grid<T, 1, true>(nblk, grid) : pipe_grid<T, 1, false>(nblk, grid);
    case 2: return nt ? pipe_grid<T, 2, true>(nblk, grid) : pipe_grid<T, 2, false>(nblk, grid);
    case 4: return nt ? pipe_grid<T, 4, true>(nblk, grid) : pipe_grid<T, 4, false>(nblk, grid);
    default: return nt ? pipe_grid<T, 16, true>(nblk, grid) : pipe_grid<T, 16, false>(nblk, grid);
  }
}

int launch(fem_operator* op, int k, int g, const void* x, void* y) {
  const void* fn = op->precision == FEM_PREC_FP32_IR ? pick_k<float>(k, g) : pick_k<double>(k, g);
  if (!fn) {
    set_error("operator: unsupported (nrhs, lanes) combination");
    return FEM_ERR_INVALID;
  }
  if (op->n == 0) return 0;
  int n = op->n;
  long long nnz = op->nnz;
  const int* rp = op->ctx->d_rp;
  const int* ci = op->ctx->d_ci;
  const void* val = op->d_val;
  if (g >= FEM_SPMV_STREAM) {
    const bool f32 = op->precision == FEM_PREC_FP32_IR;
    const int rows = f32 ? stream_rows<float>(k) : stream_rows<double>(k);
    int nblk = (int)div_up(op->n, rows);
    if (g == FEM_SPMV_PIPE || g == FEM_SPMV_PIPE_NT) {
      int grid = 0;
      const bool nt = g == FEM_SPMV_PIPE_NT;
      const int rc = f32 ? pipe_grid_k<float>(k, nt, nblk, &grid) : pipe_grid_k<double>(k, nt, nblk, &grid);
      if (rc) return rc;
      void* args[] = {(void*)&n, (void*)&nnz, (void*)&nblk, (void*)&rp, (void*)&ci, (void*)&val, (void*)&x, (void*)&y};
      FEM_HIP(hipLaunchKernel(fn, dim3(grid), dim3(kBlock), args, 0, op->stream));
      return 0;
    }
    void* args[] = {(void*)&n, (void*)&nnz, (void*)&rp, (void*)&ci, (void*)&val, (void*)&x, (void*)&y};
    FEM_HIP(hipLaunchKernel(fn, dim3(nblk), dim3(kBlock), args, 0, op->stream));
    return 0;
  }
  const unsigned grid = div_up((long long)op->n * g, kBlock);
  void* args[] = {(void*)&n, (void*)&rp, (void*)&ci, (void*)&val, (void*)&x, (void*)&y};
  FEM_HIP(hipLaunchKernel(fn, dim3(grid), dim3(kBlock), args, 0, op->stream));
  return 0;
}

bool valid_sched(int g) {
  return g == 1 || g == 2 || g == 4 || g == 8 || g == 16 || g == FEM_SPMV_STREAM || g == FEM_SPMV_STREAM_NT ||
         g == FEM_SPMV_PIPE || g == FEM_SPMV_PIPE_NT;
}

bool valid_k(int k) { return k == 1 || k == 2 || k == 4 || k == 16; }

}  // namespace

void operator_free(fem_operator* op) {
  if (!op) return;
  if (op->stream) (void)hipStreamSynchronize(op->stream);
  void* ptrs[] = {op->d_val, op->d_host_x, op->d_host_y};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (op->ev0) (void)hipEventDestroy(op->ev0);
  if (op->ev1) (void)hipEventDestroy(op->ev1);
  if (op->stream) (void)hipStreamDestroy(op->stream);
  delete op;
}

int operator_create(fem_ctx* c, const double coef[3], int precision, int transpose, fem_operator** out) {
  if (precision != FEM_PREC_FP32_IR && precision != FEM_PREC_FP64) {
    set_error("operator: unknown precision");
    return FEM_ERR_INVALID;
  }
  fem_operator* op = new fem_operator();
  op->ctx = c;
  op->n = c->n;
  op->nnz = c->nnz;
  op->precision = precision;
  int rc = 0;
  hipError_t e = hipStreamCreateWithFlags(&op->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&op->ev0);
  if (e == hipSuccess) e = hipEventCreate(&op->ev1);
  if (e != hipSuccess) {
    operator_free(op);
    return hip_fail(e, "operator_create");
  }
  const size_t vb = precision == FEM_PREC_FP32_IR ? 4 : 8;
  op->d_val = (void*)dalloc<char>((size_t)std::max(1LL, c->nnz) * vb, &rc);
  if (rc) {
    operator_free(op);
    return rc;
  }
  if (c->n > 0) {
    hipLaunchKernelGGL(k_op_form, dim3(div_up(c->n, kBlock)), dim3(kBlock), 0, op->stream, c->n, c->d_rp,
                       transpose ? c->d_tperm : nullptr, c->d_M, c->d_K, c->d_A, coef[0], coef[1], coef[2],
                       precision == FEM_PREC_FP32_IR ? 1 : 0, op->d_val);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(op->stream);
    if (e != hipSuccess) {
      operator_free(op);
      return hip_fail(e, "operator_create: form");
    }
  }
  *out = op;
  return 0;
}

int operator_apply_device(fem_operator* op, int k, const void* x, void* y) {
  if (!valid_k(k)) {
    set_error("operator: nrhs must be 1, 2, 4 or 16");
    return FEM_ERR_INVALID;
  }
  const int g = op->lanes[k] ? op->lanes[k] : default_lanes(k);
  return launch(op, k, g, x, y);
}

int operator_apply_host(fem_operator* op, int k, const double* x, double* y) {
  if (!valid_k(k)) {
    set_error("operator: nrhs must be 1, 2, 4 or 16");
    return FEM_ERR_INVALID;
  }
  const long long len = (long long)op->n * k;
  if (len == 0) return 0;
  int rc = 0;
  if (op->host_cap < len) {
    if (op->d_host_x) (void)hipFree(op->d_host_x);
    if (op->d_host_y) (void)hipFree(op->d_host_y);
    op->d_host_x = dalloc<double>((size_t)len, &rc);
    op->d_host_y = dalloc<double>((size_t)len, &rc);
    if (rc) {
      op->host_cap = 0;
      return rc;
    }
    op->host_cap = len;
  }
  const bool f32 = op->precision == FEM_PREC_FP32_IR;
  std::vector<float> xf, yf;
  if (f32) {
    xf.resize(len);
    for (long long i = 0; i < len; i++) xf[i] = (float)x[i];
    FEM_HIP(hipMemcpyAsync(op->d_host_x, xf.data(), len * 4, hipMemcpyHostToDevice, op->stream));
  } else {
    FEM_HIP(hipMemcpyAsync(op->d_host_x, x, len * 8, hipMemcpyHostToDevice, op->stream));
  }
  rc = operator_apply_device(op, k, op->d_host_x, op->d_host_y);
  if (rc) return rc;
  if (f32) {
    yf.resize(len);
    FEM_HIP(hipMemcpyAsync(yf.data(), op->d_host_y, len * 4, hipMemcpyDeviceToHost, op->stream));
    FEM_HIP(hipStreamSynchronize(op->stream));
    for (long long i = 0; i < len; i++) y[i] = (double)yf[i];
  } else {
    FEM_HIP(hipMemcpyAsync(y, op->d_host_y, len * 8, hipMemcpyDeviceToHost, op->stream));
    FEM_HIP(hipStreamSynchronize(op->stream));
  }
  return 0;
}

namespace {
int time_lanes(fem_operator* op, int k, int g, const void* x, void* y, int reps, double* ms) {
  int rc = launch(op, k, g, x, y);  // warm
  if (rc) return rc;
  FEM_HIP(hipEventRecord(op->ev0, op->stream));
  for (int i = 0; i < reps; i++)
    if ((rc = launch(op, k, g, x, y))) return rc;
  FEM_HIP(hipEventRecord(op->ev1, op->stream));
  FEM_HIP(hipEventSynchronize(op->ev1));
  float t = 0.f;
  FEM_HIP(hipEventElapsedTime(&t, op->ev0, op->ev1));
  *ms = (double)t / std::max(1, reps);
  return 0;
}
}  // namespace

int operator_time(fem_operator* op, int k, const void* x, void* y, int reps, double* ms) {
  if (!valid_k(k) || reps < 1) {
    set_error("operator_time: nrhs must be 1, 2, 4 or 16 and reps >= 1");
    return FEM_ERR_INVALID;
  }
  const int g = op->lanes[k] ? op->lanes[k] : default_lanes(k);
  return time_lanes(op, k, g, x, y, reps, ms);
}

int operator_set_lanes(fem_operator* op, int k, int g) {
  if (!valid_k(k) || !valid_sched(g) || g < k) {
    set_error("operator_set_schedule: lanes 1, 2, 4, 8 or 16 (>= nrhs) or a stream schedule");
    return FEM_ERR_INVALID;
  }
  op->lanes[k] = g;
  return 0;
}

int operator_autotune(fem_operator* op, int k, const void* x, void* y, int reps, int* lanes, double* ms) {
  if (!valid_k(k) || reps < 1) {
    set_error("operator_autotune: nrhs must be 1, 2, 4 or 16 and reps >= 1");
    return FEM_ERR_INVALID;
  }
  double best = 1e300;
  int bg = default_lanes(k);
  for (int g : {1, 2, 4, 8, 16, (int)FEM_SPMV_STREAM, (int)FEM_SPMV_STREAM_NT, (int)FEM_SPMV_PIPE,
                (int)FEM_SPMV_PIPE_NT}) {
    if (g < k) continue;
    double t = 0.0;
    const int rc = time_lanes(op, k, g, x, y, reps, &t);
    if (rc) return rc;
    if (t < best) {
      best = t;
      bg = g;
    }
  }
  op->lanes[k] = bg;
  if (lanes) *lanes = bg;
  if (ms) *ms = best;
  return 0;
}

}  // namespace femhip
