// step.hip -- launch of the persistent stepping kernel (step_kernel.h),
// whose instantiations live in one translation unit per exchange mode.
#include "femhip_internal.h"
#include "femhip_step.h"

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

}  // namespace

// Launch one persistent cooperative grid for a prepared batch.
int launch_step(const StepLaunch& L, int nsteps, hipStream_t st) {
  void* args[] = {(void*)&L.teams, (void*)&L.wg_team, (void*)&L.insts, (void*)&L.team_inst, (void*)&nsteps};
  const void* fn = pick_kernel(L.k, L.inner_fp32, L.kr, L.coh);
  if (!fn) {
    set_error("unsupported kernel configuration (number of right-hand sides must be 1, 2, 4 or 16)");
    return FEM_ERR_INVALID;
  }
  hipError_t e = hipLaunchCooperativeKernel(fn, dim3(L.grid), dim3(kBlock), args, 0, st);
  if (e != hipSuccess) return hip_fail(e, "hipLaunchCooperativeKernel(k_step)");
  return 0;
}

// The same launch as a kernel node of a HIP graph, after `dep` (may be null).
// A graph node is an ordinary launch: the plan's grid never exceeds the
// co-resident limit the cooperative launch checks (step_max_coresident), and
// every barrier wait of the kernel is bounded.
int step_graph_node(const StepLaunch& L, int nsteps, hipGraph_t g, hipGraphNode_t dep, hipGraphNode_t* out) {
  const void* fn = pick_kernel(L.k, L.inner_fp32, L.kr, L.coh);
  if (!fn) {
    set_error("unsupported kernel configuration (number of right-hand sides must be 1, 2, 4 or 16)");
    return FEM_ERR_INVALID;
  }
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

int step_max_coresident(int k, int inner_fp32, int kr, int coh, int* blocks) {
  const void* fn = pick_kernel(k, inner_fp32, kr, coh);
  if (!fn) return FEM_ERR_INVALID;
  int per_cu = 0, dev = 0, cus = 0;
  FEM_HIP(hipGetDevice(&dev));
  FEM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  FEM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kBlock, 0));
  *blocks = per_cu * cus;
  return 0;
}

}  // namespace femhip
