// The library's static device scratch: the memory of the allocation-free
// solve entry point (pgc_solve_ratio_noalloc), for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).
//
// `scratch` is a file-scope device array; every GPU the library runs on holds
// its own copy until the process ends. In host code its name is only a
// host-side stand-in: its device address on each device comes from
// hipGetSymbolAddress, called once per device (at context creation, with that
// device current) and kept per device; host code uses that address, never the
// array's name.
#include <hip/hip_runtime.h>

#include <mutex>

#include "pgc_internal.h"

namespace pgc {

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mutex;
void* g_ptr[kMaxDevices];
hipError_t g_err[kMaxDevices];
bool g_done[kMaxDevices];
std::mutex g_call[kMaxDevices];
}  // namespace

// Resolves the scratch address of `device` (the current device) once.
hipError_t scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mutex);
  if (g_done[device]) return g_err[device];
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  g_done[device] = true;
  g_err[device] = e;
  g_ptr[device] = e == hipSuccess ? p : nullptr;
  if (e != hipSuccess) (void)hipGetLastError();
  return e;
}

// The scratch address of `device`, or nullptr when its setup failed or has not run.
void* scratch_ptr(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_mutex);
  return g_done[device] ? g_ptr[device] : nullptr;
}

// Calls that use the scratch of one device hold this lock for their whole
// duration (each call ends with a stream synchronization).
std::mutex& scratch_call_lock(int device) { return g_call[device < 0 || device >= kMaxDevices ? 0 : device]; }

}  // namespace pgc
