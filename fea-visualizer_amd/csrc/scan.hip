// scan.hip -- device exclusive prefix sum over int32 counts (row pointers,
// incidence pointers, idx_map renumbering).  Three passes: per-tile sums,
// one-workgroup scan of the tile sums, per-tile scan with the tile offset.
#include "femhip_internal.h"

namespace femhip {

namespace {

constexpr int kScanItems = 4;
constexpr int kTile = kBlock * kScanItems;  // 1024 values per workgroup

__device__ inline int block_inclusive_scan(int v, int* lds) {
  // Hillis-Steele over 256 lanes in LDS; adequate for the setup-time scans.
  const int t = threadIdx.x;
  lds[t] = v;
  __syncthreads();
  for (int off = 1; off < kBlock; off <<= 1) {
    int add = (t >= off) ? lds[t - off] : 0;
    __syncthreads();
    lds[t] += add;
    __syncthreads();
  }
  int r = lds[t];
  __syncthreads();
  return r;
}

__global__ void __launch_bounds__(kBlock) k_tile_sums(const int* in, int n, int* sums) {
  __shared__ int lds[kBlock];
  long long base = (long long)blockIdx.x * kTile;
  int acc = 0;
  for (int q = 0; q < kScanItems; q++) {
    long long i = base + (long long)threadIdx.x * kScanItems + q;
    if (i < n) acc += in[i];
  }
  int inc = block_inclusive_scan(acc, lds);
  if (threadIdx.x == kBlock - 1) sums[blockIdx.x] = inc;
}

__global__ void __launch_bounds__(kBlock) k_scan_sums(int* sums, int m, int* total) {
  __shared__ int lds[kBlock];
  int carry = 0;
  for (int base = 0; base < m; base += kBlock) {
    int i = base + threadIdx.x;
    int v = (i < m) ? sums[i] : 0;
    int inc = block_inclusive_scan(v, lds);
    if (i < m) sums[i] = carry + inc - v;  // exclusive
    __shared__ int last;
    if (threadIdx.x == kBlock - 1) last = inc;
    __syncthreads();
    carry += last;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ void __launch_bounds__(kBlock) k_tile_scan(const int* in, int n, const int* sums, int* out) {
  __shared__ int lds[kBlock];
  long long base = (long long)blockIdx.x * kTile;
  int vals[kScanItems];
  int acc = 0;
  for (int q = 0; q < kScanItems; q++) {
    long long i = base + (long long)threadIdx.x * kScanItems + q;
    vals[q] = (i < n) ? in[i] : 0;
    acc += vals[q];
  }
  int inc = block_inclusive_scan(acc, lds);
  int run = sums[blockIdx.x] + inc - acc;
  for (int q = 0; q < kScanItems; q++) {
    long long i = base + (long long)threadIdx.x * kScanItems + q;
    if (i < n) out[i] = run;
    run += vals[q];
  }
}

}  // namespace

// out[i] = sum_{j<i} in[i]; out may alias in.  *total receives the grand
// total (host value, synchronous on `st`).
int exclusive_scan_i32(const int* in, int* out, int n, int* total, hipStream_t st) {
  int rc = 0;
  int tiles = (int)div_up(n > 0 ? n : 1, kTile);
  int* sums = dalloc<int>((size_t)tiles + 1, &rc);
  if (!sums) return rc;
  int* dtot = sums + tiles;
  hipLaunchKernelGGL(k_tile_sums, dim3(tiles), dim3(kBlock), 0, st, in, n, sums);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kBlock), 0, st, sums, tiles, dtot);
  hipLaunchKernelGGL(k_tile_scan, dim3(tiles), dim3(kBlock), 0, st, in, n, sums, out);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(total, dtot, sizeof(int), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(sums);
  if (e != hipSuccess) return hip_fail(e, "exclusive_scan_i32");
  return 0;
}

}  // namespace femhip
