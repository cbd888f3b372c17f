// Instantiations of the solver kernel for one cost representation (fp64 raw cost, minimum-cost mode),
// compiled as its own translation unit so the variants build in parallel.
#include "solver_launch.h"

namespace pgc {

hipError_t launch_solver_mincost(int maxw, int maxc, const SolverJob* jobs, int njobs, int grid, int* queue,
                               size_t lds, hipStream_t st) {
  return launch_width<double,false,true>(maxw, maxc, jobs, njobs, grid, queue, lds, st);
}

hipError_t launch_solver_grid_mincost(int maxw, const SolverJob* jobs, int nblk, hipStream_t st) {
  return launch_grid<double,false,true>(maxw, jobs, nblk, st);
}

int solver_grid_capacity_mincost(int maxw, int numCU) { return grid_capacity<double,false,true>(maxw, numCU); }

}  // namespace pgc
