// Launch helpers shared by the solver kernel's instantiation units.
#pragma once
#include "solver_device.h"

namespace pgc {

// grid workgroups run njobs jobs: one job per workgroup (queue == nullptr,
// grid == njobs) or a persistent launch taking jobs from *queue (zeroed by
// the caller).
template <typename CT, bool HT, bool MC, int MAXC, int W>
static hipError_t launch_one(const SolverJob* jobs, int njobs, int grid, int* queue, size_t lds, hipStream_t st) {
  static bool attr_set = false;
  if (!attr_set) {
    hipError_t e = hipFuncSetAttribute((const void*)solve_kernel<CT, HT, MC, MAXC, W>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 1024);
    if (e != hipSuccess) return e;
    attr_set = true;
  }
  hipLaunchKernelGGL((solve_kernel<CT, HT, MC, MAXC, W>), dim3(grid), dim3(PGC_BLOCK), lds, st, jobs, njobs, queue);
  return hipGetLastError();
}

template <typename CT, bool HT, bool MC, int W>
static hipError_t launch_maxc(int maxc, const SolverJob* jobs, int njobs, int grid, int* queue, size_t lds,
                              hipStream_t st) {
  switch (maxc) {
    case 0: return launch_one<CT, HT, MC, 0, W>(jobs, njobs, grid, queue, lds, st);
    case 1: return launch_one<CT, HT, MC, 1, W>(jobs, njobs, grid, queue, lds, st);
    case 2: return launch_one<CT, HT, MC, 2, W>(jobs, njobs, grid, queue, lds, st);
    case 3: return launch_one<CT, HT, MC, 3, W>(jobs, njobs, grid, queue, lds, st);
    case 4: return launch_one<CT, HT, MC, 4, W>(jobs, njobs, grid, queue, lds, st);
    default: return launch_one<CT, HT, MC, 8, W>(jobs, njobs, grid, queue, lds, st);
  }
}

template <typename CT, bool HT, bool MC>
static hipError_t launch_width(int maxw, int maxc, const SolverJob* jobs, int njobs, int grid, int* queue,
                               size_t lds, hipStream_t st) {
  if (maxw <= 8) return launch_maxc<CT, HT, MC, 8>(maxc, jobs, njobs, grid, queue, lds, st);
  if (maxw <= 16) return launch_maxc<CT, HT, MC, 16>(maxc, jobs, njobs, grid, queue, lds, st);
  return launch_maxc<CT, HT, MC, 32>(maxc, jobs, njobs, grid, queue, lds, st);
}

}  // namespace pgc
