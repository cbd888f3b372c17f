// Instantiations of the persistent solver kernel for one cost representation
// (compiled as its own translation unit so the variants build in parallel).
#include "solver_device.h"

namespace pgc {

template <typename CT, bool HT, bool MC, int MAXC, int W>
static hipError_t launch_one(const SolverJob* jobs, int njobs, size_t lds, hipStream_t st) {
  static bool attr_set = false;
  if (!attr_set) {
    hipError_t e = hipFuncSetAttribute((const void*)solve_kernel<CT, HT, MC, MAXC, W>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 1024);
    if (e != hipSuccess) return e;
    attr_set = true;
  }
  hipLaunchKernelGGL((solve_kernel<CT, HT, MC, MAXC, W>), dim3(njobs), dim3(PGC_BLOCK), lds, st, jobs);
  return hipGetLastError();
}

template <typename CT, bool HT, bool MC, int W>
static hipError_t launch_maxc(int maxc, const SolverJob* jobs, int njobs, size_t lds, hipStream_t st) {
  switch (maxc) {
    case 1: return launch_one<CT, HT, MC, 1, W>(jobs, njobs, lds, st);
    case 2: return launch_one<CT, HT, MC, 2, W>(jobs, njobs, lds, st);
    case 3: return launch_one<CT, HT, MC, 3, W>(jobs, njobs, lds, st);
    case 4: return launch_one<CT, HT, MC, 4, W>(jobs, njobs, lds, st);
    default: return launch_one<CT, HT, MC, 8, W>(jobs, njobs, lds, st);
  }
}

hipError_t launch_solver_ratio_time(int maxw, int maxc, const SolverJob* jobs, int njobs, size_t lds, hipStream_t st) {
  if (maxw <= 8) return launch_maxc<float, true, false, 8>(maxc, jobs, njobs, lds, st);
  if (maxw <= 16) return launch_maxc<float, true, false, 16>(maxc, jobs, njobs, lds, st);
  return launch_maxc<float, true, false, 32>(maxc, jobs, njobs, lds, st);
}

}  // namespace pgc
