// step.hip -- launch of the persistent stepping kernel (step_kernel.h),
// whose instantiations live in one translation unit per exchange mode.
#include "femhip_internal.h"
#include "femhip_step.h"

#include <mutex>
#include <string>

namespace femhip {

const void* step_kernel_plain(int k, int fp32, int kr);
const void* step_kernel_coherent(int k, int fp32, int kr);
const void* step_kernel_fenced(int k, int fp32, int kr);

namespace {

// Kernel instantiation for (columns, inner precision, Krylov method, exchange mode).
const void* pick_kernel(int k, int fp32, int kr, int coh) {
  switch (coh) {
    case 0: return step_kernel_plain(k, fp32, kr);
    case kExchangeCoherent: return step_kernel_coherent(k, fp32, kr);
    case kExchangeFenced: return step_kernel_fenced(k, fp32, kr);
    default: return nullptr;
  }
}

// Workgroups of kernel `fn` that fit the current device at once.  Cached per
// (device, kernel): the answer depends only on the code object.
int kernel_max_coresident(const void* fn, int* blocks) {
  constexpr int kMaxDev = 16, kMaxFn = 128;
  struct Entry { const void* fn; int blocks; };
  static std::mutex mu;
  static Entry cache[kMaxDev][kMaxFn];
  static int used[kMaxDev];
  int dev = 0;
  FEM_HIP(hipGetDevice(&dev));
  if (dev >= 0 && dev < kMaxDev) {
    std::lock_guard<std::mutex> lk(mu);
    for (int i = 0; i < used[dev]; i++)
      if (cache[dev][i].fn == fn) {
        *blocks = cache[dev][i].blocks;
        return 0;
      }
  }
  int per_cu = 0, cus = 0;
  FEM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  FEM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kBlock, 0));
  *blocks = per_cu * cus;
  if (dev >= 0 && dev < kMaxDev) {
    std::lock_guard<std::mutex> lk(mu);
    if (used[dev] < kMaxFn) cache[dev][used[dev]++] = Entry{fn, *blocks};
  }
  return 0;
}

// The spin barriers of the stepping kernel need every workgroup of the grid
// resident at once: refuse a grid the launched variant cannot hold.
int check_resident(const StepLaunch& L, const void* fn) {
  int maxb = 0;
  const int rc = kernel_max_coresident(fn, &maxb);
  if (rc) return rc;
  if (L.grid > maxb) {
    set_error("stepping kernel: grid of " + std::to_string(L.grid) + " workgroups exceeds the " +
              std::to_string(maxb) + " the launched variant can hold co-resident");
    return FEM_ERR_INVALID;
  }
  return 0;
}

}  // namespace

// Launch one persistent cooperative grid for a prepared batch.
int launch_step(const StepLaunch& L, int nsteps, hipStream_t st) {
  void* args[] = {(void*)&L.teams, (void*)&L.wg_team, (void*)&L.insts, (void*)&L.team_inst, (void*)&nsteps};
  const void* fn = pick_kernel(L.k, L.inner_fp32, L.kr, L.coh);
  if (!fn) {
    set_error("unsupported kernel configuration (number of right-hand sides must be 1, 2, 4 or 16)");
    return FEM_ERR_INVALID;
  }
  const int rc = check_resident(L, fn);
  if (rc) return rc;
  hipError_t e = hipLaunchCooperativeKernel(fn, dim3(L.grid), dim3(kBlock), args, 0, st);
  if (e != hipSuccess) return hip_fail(e, "hipLaunchCooperativeKernel(k_step)");
  return 0;
}

// The same launch as a kernel node of a HIP graph, after `dep` (may be null).
// A graph node is an ordinary launch, so the co-residency that the
// cooperative launch checks is checked here against the same limit
// (check_resident); every barrier wait of the kernel is also bounded.
int step_graph_node(const StepLaunch& L, int nsteps, hipGraph_t g, hipGraphNode_t dep, hipGraphNode_t* out) {
  const void* fn = pick_kernel(L.k, L.inner_fp32, L.kr, L.coh);
  if (!fn) {
    set_error("unsupported kernel configuration (number of right-hand sides must be 1, 2, 4 or 16)");
    return FEM_ERR_INVALID;
  }
  const int rc = check_resident(L, fn);
  if (rc) return rc;
  void* args[] = {(void*)&L.teams, (void*)&L.wg_team, (void*)&L.insts, (void*)&L.team_inst, (void*)&nsteps};
  hipKernelNodeParams prm{};
  prm.func = const_cast<void*>(fn);
  prm.gridDim = dim3(L.grid);
  prm.blockDim = dim3(kBlock);
  prm.sharedMemBytes = 0;
  prm.kernelParams = args;
  prm.extra = nullptr;
  FEM_HIP(hipGraphAddKernelNode(out, g, dep ? &dep : nullptr, dep ? 1 : 0, &prm));
  return 0;
}

// coh: 0 plain (one-workgroup teams), 1 coherent, 2 fenced.
int step_max_coresident(int k, int inner_fp32, int kr, int coh, int* blocks) {
  const void* fn = pick_kernel(k, inner_fp32, kr, coh);
  if (!fn) {
    set_error("unsupported kernel configuration (number of right-hand sides must be 1, 2, 4 or 16)");
    return FEM_ERR_INVALID;
  }
  return kernel_max_coresident(fn, blocks);
}

}  // namespace femhip
