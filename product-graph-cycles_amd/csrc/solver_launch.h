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
  if (maxw <= 32) return launch_maxc<CT, HT, MC, 32>(maxc, jobs, njobs, grid, queue, lds, st);
  // runtime slot width (in-degrees above 32): the table in global memory, or
  // in LDS with up to eight chunks per wave
  if (maxc == 0) return launch_one<CT, HT, MC, 0, 0>(jobs, njobs, grid, queue, lds, st);
  return launch_one<CT, HT, MC, 8, 0>(jobs, njobs, grid, queue, lds, st);
}

// The device-wide solve of one job: a cooperative launch of nblk workgroups
// (all resident at once, which the runtime checks).
template <typename CT, bool HT, bool MC, int W>
static hipError_t launch_grid_one(const SolverJob* jobs, int nblk, hipStream_t st) {
  void* args[] = {(void*)&jobs};
  return hipLaunchCooperativeKernel((const void*)solve_kernel_grid<CT, HT, MC, W>, dim3(nblk), dim3(PGC_BLOCK), args,
                                    0, st);
}

// Workgroups of the device-wide solve that fit the device at once.
template <typename CT, bool HT, bool MC, int W>
static int grid_capacity_one(int numCU) {
  static int per_cu = -1;
  if (per_cu < 0) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, (const void*)solve_kernel_grid<CT, HT, MC, W>, PGC_BLOCK,
                                                     0) != hipSuccess)
      n = 0;
    per_cu = n;
  }
  return per_cu * numCU;
}

template <typename CT, bool HT, bool MC>
static hipError_t launch_grid(int maxw, const SolverJob* jobs, int nblk, hipStream_t st) {
  if (maxw <= 8) return launch_grid_one<CT, HT, MC, 8>(jobs, nblk, st);
  if (maxw <= 16) return launch_grid_one<CT, HT, MC, 16>(jobs, nblk, st);
  return launch_grid_one<CT, HT, MC, 32>(jobs, nblk, st);
}

template <typename CT, bool HT, bool MC>
static int grid_capacity(int maxw, int numCU) {
  if (maxw <= 8) return grid_capacity_one<CT, HT, MC, 8>(numCU);
  if (maxw <= 16) return grid_capacity_one<CT, HT, MC, 16>(numCU);
  return grid_capacity_one<CT, HT, MC, 32>(numCU);
}

}  // namespace pgc
