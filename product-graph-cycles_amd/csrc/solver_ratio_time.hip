// Instantiations of the solver kernel for one cost representation (fp32 q-unit cost and time),
// compiled as its own translation unit so the variants build in parallel.
#include "solver_launch.h"

namespace pgc {

hipError_t launch_solver_ratio_time(int maxw, int maxc, const SolverJob* jobs, int njobs, int grid, int* queue,
                               size_t lds, hipStream_t st) {
  return launch_width<float,true,false>(maxw, maxc, jobs, njobs, grid, queue, lds, st);
}

hipError_t launch_solver_grid_ratio_time(int maxw, const SolverJob* jobs, int nblk, hipStream_t st) {
  return launch_grid<float,true,false>(maxw, jobs, nblk, st);
}

int solver_grid_capacity_ratio_time(int maxw, int numCU) { return grid_capacity<float,true,false>(maxw, numCU); }

}  // namespace pgc
