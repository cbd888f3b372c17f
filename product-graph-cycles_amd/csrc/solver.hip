// Solver launch dispatch and the edge-lookup kernel (the solver kernel's
// instantiations live in solver_ratio.hip, solver_ratio_time.hip and
// solver_mincost.hip).
#include "solver_device.h"

namespace pgc {

hipError_t launch_solver_ratio(int, int, const SolverJob*, int, int, int*, size_t, hipStream_t);
hipError_t launch_solver_ratio_time(int, int, const SolverJob*, int, int, int*, size_t, hipStream_t);
hipError_t launch_solver_mincost(int, int, const SolverJob*, int, int, int*, size_t, hipStream_t);

hipError_t launch_solver_grid_ratio(int, const SolverJob*, int, hipStream_t);
hipError_t launch_solver_grid_ratio_time(int, const SolverJob*, int, hipStream_t);
hipError_t launch_solver_grid_mincost(int, const SolverJob*, int, hipStream_t);
int solver_grid_capacity_ratio(int, int);
int solver_grid_capacity_ratio_time(int, int);
int solver_grid_capacity_mincost(int, int);

size_t solver_lds_bytes(int npl, int K) { return (size_t)(2 * K + 1) * (npl + 1) * sizeof(double); }

size_t grid_state_bytes() { return sizeof(GridState); }

// Device-wide solve of one job with its table in global memory (J.grid zeroed
// by the caller): nblk workgroups of one cooperative launch.
hipError_t launch_solver_grid(const SolverJob* d_job, int nblk, int maxw, bool mincost, bool hasTime,
                              hipStream_t st) {
  if (nblk < 1 || nblk > kGridMaxBlocks || maxw > 32) return hipErrorInvalidValue;
  if (mincost) return launch_solver_grid_mincost(maxw, d_job, nblk, st);
  if (hasTime) return launch_solver_grid_ratio_time(maxw, d_job, nblk, st);
  return launch_solver_grid_ratio(maxw, d_job, nblk, st);
}

// Workgroups of that launch that the device holds at once.
int solver_grid_capacity(int maxw, bool mincost, bool hasTime, int numCU) {
  if (mincost) return solver_grid_capacity_mincost(maxw, numCU);
  if (hasTime) return solver_grid_capacity_ratio_time(maxw, numCU);
  return solver_grid_capacity_ratio(maxw, numCU);
}

// Solver launch: grid workgroups over njobs jobs (one job per workgroup when
// queue is null, else a persistent launch on the queue counter). maxw = the
// layout's slot width (8, 16, 32, or a multiple of 8 up to kMaxSlotWidth for
// the runtime-width variant), maxChunksPerWave = ceil(ceil(npl/64) / 16) (<= 8)
// for the largest job (8 for any count with the runtime width), or 0 for the
// variant with the distance table in global memory (any number of chunks per
// wave; lds is then 0).
hipError_t launch_solver(const SolverJob* d_jobs, int njobs, int grid, int* queue, int maxChunksPerWave, int maxw,
                         bool mincost, bool hasTime, bool doubleCost, size_t lds, hipStream_t st) {
  if (lds > 160 * 1024 - 1024) return hipErrorInvalidValue;
  if (maxChunksPerWave < 0 || maxChunksPerWave > 8 || maxw > kMaxSlotWidth) return hipErrorInvalidValue;
  if (maxw > 32 && maxChunksPerWave != 0 && maxChunksPerWave != 8) return hipErrorInvalidValue;
  if (mincost) {
    if (!doubleCost) return hipErrorInvalidValue;
    return launch_solver_mincost(maxw, maxChunksPerWave, d_jobs, njobs, grid, queue, lds, st);
  }
  if (hasTime) return launch_solver_ratio_time(maxw, maxChunksPerWave, d_jobs, njobs, grid, queue, lds, st);
  return launch_solver_ratio(maxw, maxChunksPerWave, d_jobs, njobs, grid, queue, lds, st);
}

hipError_t launch_lookup(const SolverJob* d_job, const int32_t* d_nodes, int n, int32_t* d_eids, bool hasTime,
                         bool doubleCost, hipStream_t st) {
  if (n < 2) return hipSuccess;
  dim3 b(256), g((n + 255) / 256);
  if (doubleCost)
    hipLaunchKernelGGL((lookup_kernel<double, false>), g, b, 0, st, d_job, d_nodes, n, d_eids);
  else if (hasTime)
    hipLaunchKernelGGL((lookup_kernel<float, true>), g, b, 0, st, d_job, d_nodes, n, d_eids);
  else
    hipLaunchKernelGGL((lookup_kernel<float, false>), g, b, 0, st, d_job, d_nodes, n, d_eids);
  return hipGetLastError();
}

}  // namespace pgc
