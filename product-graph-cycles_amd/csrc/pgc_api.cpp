// C-ABI of the MI355X product-graph cycle solvers (include/pgc.h).
//
// Host side: graph and scratch management, and the branch-and-bound drivers
// around the device solver kernels:
//   solveWithBnbGPULawler   cost_time_ratio_solver_cuda.cpp:175-331 (+ cost_time_ratio_solver.cpp:741-864)
//   solveWithBnBCuda        dijkstra_solver_cuda.cpp:136-322 (+ dijkstra_solver.cpp:407-559)
// Every Lawler subroutine, sweep, fold, cycle walk and lambda update runs on
// the device (solver.hip); per branch the host reads back one small result
// record. Error paths return a status instead of exiting the process.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <initializer_list>
#include <unordered_set>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/pgc.h"
#include "bnb_heap.h"
#include "pgc_internal.h"

namespace pgc {
// graph_build.hip
hipError_t launch_product_space(const int32_t*, int64_t, const int32_t*, int64_t, int64_t, int, int32_t*, int32_t*,
                                int32_t*, int64_t, hipStream_t);
hipError_t launch_energy(const double*, const double*, const int32_t*, int64_t, const double*, int64_t, int, int,
                         double*, hipStream_t);
hipError_t launch_quantize(double*, int64_t, void*, double*, double*, int64_t, hipStream_t);
hipError_t launch_transpose_fd(const double*, int64_t, int64_t, double*, hipStream_t);
hipError_t launch_quantize_batch(const QItem*, int, void*, void*, double*, hipStream_t);
hipError_t launch_fill_costs_batch(const FItem*, int, hipStream_t);
hipError_t launch_fill_i32_batch(const FillItem*, int, int32_t, hipStream_t);
hipError_t launch_energy_batch(const EItem*, int, hipStream_t);
size_t select_state_bytes();
size_t bselect_state_bytes();
size_t bracket_state_bytes();
hipError_t launch_layout_count(const int32_t*, const int32_t*, int64_t, int64_t, int64_t, int32_t*, int32_t*,
                               hipStream_t);
hipError_t launch_layout_width(const int32_t*, int64_t, int32_t*, hipStream_t);
hipError_t launch_layout_fill(const int32_t*, const int32_t*, int64_t, int64_t, int64_t, int, int32_t*,
                              const int32_t*, const int32_t*, int32_t*, int32_t*, int32_t*, int32_t*, hipStream_t);
hipError_t launch_fill_costs(const int32_t*, const int32_t*, int64_t, const double*, int64_t, int, bool, void*,
                             float*, int, int, hipStream_t);
hipError_t launch_gather_rows(const int32_t*, int, const int32_t*, int64_t, int32_t*, hipStream_t);
hipError_t launch_gather_strided(const int32_t*, int64_t, int64_t, int32_t*, hipStream_t);
hipError_t launch_int_max(const int32_t*, int64_t, int32_t*, hipStream_t);
hipError_t launch_pad_src(int32_t*, int64_t, int32_t, hipStream_t);
hipError_t launch_fill_i32(int32_t*, int64_t, int32_t, hipStream_t);
hipError_t launch_chunk_groups(const int32_t*, int64_t, int64_t, int, int32_t*, hipStream_t);
hipError_t launch_xlayout_count(const int32_t*, int64_t, int64_t, int, int32_t*, hipStream_t);
hipError_t launch_xlayout_fill(const int32_t*, int64_t, int64_t, int, int, const int32_t*, const int32_t*, int32_t*,
                               int32_t*, int32_t*, int32_t*, int32_t*, hipStream_t);
hipError_t launch_xlayout_replicate(const int32_t*, const int32_t*, int64_t, int64_t, int64_t, int, int64_t, int32_t,
                                    int32_t*, int32_t*, hipStream_t);
// scratch.hip
hipError_t scratch_setup(int device);
void* scratch_ptr(int device);
std::mutex& scratch_call_lock(int device);
hipError_t launch_cycle_value(const CycleItem*, int, double*, hipStream_t);
// conj_build.hip
hipError_t launch_conj_product_space(const int32_t*, int64_t, const int32_t*, int64_t, int, const int32_t*,
                                     const int32_t*, const int64_t*, int64_t, int64_t, int2*, int32_t*, int32_t*,
                                     int32_t*, hipStream_t);
hipError_t launch_conj_quat(const double*, const double*, const double*, const double*, const int32_t*, int64_t,
                            const int32_t*, int64_t, int, double*, hipStream_t);
hipError_t launch_conj_energy(const double*, const double*, int64_t, const int32_t*, const int32_t*, const int32_t*,
                              int64_t, const double*, const double*, int64_t, int64_t, int, int, int, double, double,
                              double, double, double, double*, hipStream_t);
hipError_t launch_max_f64(const double*, int64_t, double*, hipStream_t);
hipError_t launch_cycle_grad(const CycleItem*, int, hipStream_t);
hipError_t launch_src_invariant(const int32_t*, int64_t, int64_t, int64_t, int32_t*, hipStream_t);
// solver.hip
size_t solver_lds_bytes(int npl, int K);
size_t grid_state_bytes();
hipError_t launch_solver_grid(const SolverJob*, int, int, bool, bool, hipStream_t);
int solver_grid_capacity(int, bool, bool, int);
hipError_t launch_solver(const SolverJob*, int, int, int*, int, int, bool, bool, bool, size_t, hipStream_t);
hipError_t launch_lookup(const SolverJob*, const int32_t*, int, int32_t*, bool, bool, hipStream_t);
}  // namespace pgc

using namespace pgc;

static thread_local std::string g_last_error;

static int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define HIPCHK(expr)                                                                                       \
  do {                                                                                                     \
    hipError_t e_ = (expr);                                                                                \
    if (e_ != hipSuccess) return fail(PGC_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// ------------------------------------------------ device memory cache
// Graphs, their solver scratch and temporaries take device memory from a
// per-device cache of freed blocks, keyed by rounded size: hipMalloc and
// hipFree cost tens to hundreds of microseconds each (hipFree also waits for
// the device), and a solve on a new graph makes about forty allocations. A
// block returns to the cache only when no queued work can still use it: every
// caller of cache_free has synchronized the stream (or device) that used it.
// Cached blocks are released when an allocation runs out of memory. The cache
// lives until the process ends (no HIP calls from static destructors).
namespace {
struct MemCache {
  std::mutex m;
  std::map<std::pair<int, size_t>, std::vector<void*>> free;
  std::unordered_map<void*, std::pair<int, size_t>> owned;
};
MemCache& memcache() {
  static MemCache* c = new MemCache();
  return *c;
}
size_t round_bytes(size_t b) {
  if (b <= 65536) return (b + 511) & ~(size_t)511;
  size_t p = 65536;
  while (p < b) p <<= 1;
  const size_t q = p >> 4;  // at most 1/16 rounding waste
  return (b + q - 1) / q * q;
}
void cache_trim(int dev) {
  MemCache& c = memcache();
  std::lock_guard<std::mutex> lk(c.m);
  for (auto it = c.free.begin(); it != c.free.end();) {
    if (it->first.first != dev) {
      ++it;
      continue;
    }
    for (void* q : it->second) {
      c.owned.erase(q);
      (void)hipFree(q);
    }
    it = c.free.erase(it);
  }
}
}  // namespace

static hipError_t cache_malloc(void** p, size_t bytes) {
  *p = nullptr;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  const size_t r = round_bytes(bytes ? bytes : 1);
  MemCache& c = memcache();
  {
    std::lock_guard<std::mutex> lk(c.m);
    auto it = c.free.find({dev, r});
    if (it != c.free.end() && !it->second.empty()) {
      *p = it->second.back();
      it->second.pop_back();
      return hipSuccess;
    }
  }
  e = hipMalloc(p, r);
  if (e == hipErrorOutOfMemory) {
    (void)hipGetLastError();
    cache_trim(dev);
    e = hipMalloc(p, r);
  }
  if (e != hipSuccess) {
    *p = nullptr;
    return e;
  }
  std::lock_guard<std::mutex> lk(c.m);
  c.owned[*p] = {dev, r};
  return hipSuccess;
}

// Returns a block to the cache (blocks not from the cache are freed).
static void cache_free(void* p) {
  if (!p) return;
  MemCache& c = memcache();
  {
    std::lock_guard<std::mutex> lk(c.m);
    auto it = c.owned.find(p);
    if (it != c.owned.end()) {
      c.free[it->second].push_back(p);
      return;
    }
  }
  (void)hipFree(p);
}

template <typename T>
static hipError_t dalloc(T** p, size_t n) {
  if (n == 0) n = 1;
  return cache_malloc((void**)p, n * sizeof(T));
}

// Where the device buffers of a solve come from: the memory cache above (the
// default), or, for the allocation-free entry point, consecutive 256-byte
// aligned pieces of the library's static scratch (an arena: nothing is
// returned, the whole arena is reused by the next such call).
struct DevAlloc {
  uint8_t* base = nullptr;  // arena start (device address), nullptr: the memory cache
  size_t cap = 0, used = 0;
  bool overflow = false;  // an arena request did not fit
  template <typename T>
  hipError_t get(T** p, size_t n) {
    if (n == 0) n = 1;
    if (!base) return dalloc(p, n);
    const size_t b = (n * sizeof(T) + 255) & ~(size_t)255;
    if (used + b > cap) {
      *p = nullptr;
      overflow = true;
      return hipErrorOutOfMemory;
    }
    *p = (T*)(base + used);
    used += b;
    return hipSuccess;
  }
  void put(void* p) {
    if (!base && p) cache_free(p);
  }
};

struct pgc_context {
  int device = 0;
  int num_cu = 0;
  pgc_tuning tune = {-1, -1, -1, -1, -1, -1, -1, -1, -1};  // launch choices (-1: the built-in rule)
  hipStream_t stream = nullptr;
  void* select_state = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // batched solves: job table, results, output cycles (grow-only)
  pgc::SolverJob* d_jobs = nullptr;
  int64_t jobs_cap = 0;
  pgc::JobResult* d_res = nullptr;
  int64_t res_cap = 0;
  int32_t* d_out = nullptr;
  int64_t out_cap = 0;
  int32_t* d_rows = nullptr;
  int64_t rows_cap = 0;
  int32_t* d_queue = nullptr;  // one counter per launch group
  // cycle value / gradient calls
  pgc::CycleItem* d_items = nullptr;
  int64_t items_cap = 0;
  int32_t* d_ceids = nullptr;
  int64_t ceids_cap = 0;
  double* d_values = nullptr;
  int64_t values_cap = 0;
  // batched quantization and cost fill
  pgc::QItem* d_qitems = nullptr;
  int64_t qitems_cap = 0;
  pgc::FItem* d_fitems = nullptr;
  int64_t fitems_cap = 0;
  pgc::FillItem* d_fills = nullptr;  // batched -1 fills of global predecessor tables
  int64_t fills_cap = 0;
  pgc::EItem* d_eitems = nullptr;  // batched energies
  int64_t eitems_cap = 0;
  // launch groups of a batch run concurrently on these streams (default priority)
  hipStream_t aux[4] = {nullptr, nullptr, nullptr, nullptr};
  hipEvent_t fork = nullptr, fork2 = nullptr;
  hipEvent_t join[4] = {nullptr, nullptr, nullptr, nullptr};
  uint8_t* d_select = nullptr;  // select states of the batched median
  int64_t select_cap = 0;
  uint8_t* d_bracket = nullptr;  // its per-column brackets
  int64_t bracket_cap = 0;
  double* d_cand = nullptr;  // and the keys gathered inside the brackets
  int64_t cand_cap = 0;
  // the single-column median's gathered keys (grow-only)
  double* d_qcand = nullptr;
  int64_t qcand_cap = 0;
  // pinned host staging of small copies (grow-only, see stage())
  uint8_t* pinned = nullptr;
  size_t pinned_cap = 0;
};

namespace {

// Device scratch of one solver job (sized for one max_depth).
struct JobScratch {
  int K = 0;
  int64_t predN = 0, chainN = 0;
  int cyc_cap = 0;
  int32_t *pred = nullptr, *wnext = nullptr, *wlen = nullptr, *chain = nullptr;
  int32_t *tmp_nodes = nullptr, *tmp_eid = nullptr, *neg_nodes = nullptr, *neg_eid = nullptr, *neg_len = nullptr;
  double *tmp_c = nullptr, *tmp_t = nullptr, *neg_c = nullptr, *neg_t = nullptr;
  int32_t *hits = nullptr, *back = nullptr, *branch = nullptr, *nodes_io = nullptr, *eids_io = nullptr;
  double* lb_key = nullptr;
  double* backup = nullptr;
  double* dtab = nullptr;  // distance table in global memory (graphs beyond the LDS)
  int32_t* grid = nullptr;  // GridState of the device-wide solve (grid_state_bytes())
  JobResult* res = nullptr;
  SolverJob* djob = nullptr;
  // device-resident branch and bound (batched solves)
  double* hkey = nullptr;
  int32_t *hval = nullptr, *hwhere = nullptr, *vslot = nullptr, *vc_eid = nullptr, *vc_len = nullptr;
  double *vc_c = nullptr, *vc_t = nullptr;
  int vc_cap = 0, vc_stride = 0;
  void release() {
    if (pred) (void)hipDeviceSynchronize();  // queued solver work may still use the scratch
    void* ps[] = {pred, wnext, wlen, chain, tmp_nodes, tmp_eid, neg_nodes, neg_eid, neg_len, tmp_c, tmp_t,
                  neg_c, neg_t, hits, back, branch, nodes_io, eids_io, lb_key, backup, dtab, grid, res, djob,
                  hkey, hval, hwhere, vslot, vc_eid, vc_len, vc_c, vc_t};
    for (void* p : ps)
      if (p) cache_free(p);
    *this = JobScratch();
  }
};

}  // namespace

struct pgc_graph {
  pgc_context* ctx = nullptr;
  int64_t E = 0, npl = 0, numNodes = 0, L = 0;
  int ps_cols = 4, ncols = 1;
  int32_t *d_rows = nullptr, *d_src = nullptr, *d_trgt = nullptr;
  double* d_energy = nullptr;
  // layout
  bool layout = false;
  int nChunks = 0;
  int64_t nLayers = 0;
  int64_t nslots[2] = {0, 0};
  int32_t* d_width = nullptr;  // per class: slots per lane, then edge count (4 values)
  int32_t* d_grp = nullptr;    // slot widths above 32: per class and chunk, groups of 8 slots ([2][nChunks])
  int32_t *d_s[2] = {nullptr, nullptr}, *d_e[2] = {nullptr, nullptr};
  int32_t width[2] = {0, 0};
  int32_t lstride[2] = {0, 0};
  int32_t src_inv[2] = {0, 0};  // per class: source slots are the same on every layer
  int32_t cls_deg[2] = {1 << 30, 1 << 30};  // per class: largest in-degree (unknown: above every width)
  int32_t maxw = 0;
  int32_t max_in_degree = 0;
  int32_t* d_cnt = nullptr;
  void* d_cost[2] = {nullptr, nullptr};
  float* d_time[2] = {nullptr, nullptr};
  int cost_kind = 0;  // 0 none, 1 float q-units (+time), 2 double raw
  int64_t dropped = 0;
  int64_t intra_slots_valid = 0, inter_slots_valid = 0;
  // branch graph (host CSR over layer-0 nodes)
  bool has_bg = false;
  std::vector<int64_t> bg_off, bg_val;
  // conjugate graphs: the X data their branch graph is built from at its
  // first use (a min-cost split): EX(j, 0), the X out-edge lists
  std::vector<int32_t> cbg_ex0, cbg_off, cbg_list;
  int cbg_prune = 0;
  JobScratch js;
  // energy inputs, kept to recompute the energy from another feat_diff
  double *d_vx = nullptr, *d_vy = nullptr;
  int64_t nVX = 0, nVY = 0;
  // the X and Y edge lists (non-conjugate graphs): the batched energy decodes
  // an edge's row from its id with them instead of reading the rows
  int32_t *d_ex = nullptr, *d_ey = nullptr;
  int64_t nEX = 0, nEY = 0;
  int prune = 0;
  int64_t last_sweeps = 0;  // Bellman-Ford sweeps of the graph's last ratio solve (batch ordering)
  int32_t xdeg[4] = {-1, -1, -1, -1};  // X in-degrees: largest intralayer, interlayer; their sums (one layer)
  bool conjugate = false;
  // conjugate graphs: edges per layer; every layer has the same in-edge
  // pattern and layer i's edges are [i*rep_B, (i+1)*rep_B) in the same order
  int64_t rep_B = 0;
  int line_integral = 0, length_mode = 0;
  // a derived graph shares base's product space, ids, layout and energy
  // inputs (read-only) and owns its energy, costs and solver scratch
  pgc_graph* base = nullptr;
  int refs = 0;           // derived graphs alive
  bool released = false;  // destroyed by the caller while derived graphs live
  size_t cost_bytes[2] = {0, 0};
  size_t time_bytes[2] = {0, 0};
  // implicit costs (tuning implicit_costs = 1): the graph's feat_diff by
  // columns ([nVY][nVX] float64) and the scale of its last quantization
  double* d_fdt = nullptr;
  double* d_qscale = nullptr;
  bool energy_fresh = false;  // the energy was computed from d_fdt's feat_diff and not quantized since
  bool implicit_now = false;  // inside pgc_solve_ratio: this solve may recompute its costs
  ~pgc_graph() {
    (void)hipDeviceSynchronize();  // queued work may still use the graph's buffers
    js.release();
    if (base) {
      d_rows = d_src = d_trgt = nullptr;
      d_width = d_cnt = d_grp = nullptr;
      d_s[0] = d_s[1] = d_e[0] = d_e[1] = nullptr;
      d_vx = d_vy = nullptr;
      d_ex = d_ey = nullptr;
    }
    if (d_fdt) cache_free(d_fdt);
    if (d_qscale) cache_free(d_qscale);
    if (d_vx) cache_free(d_vx);
    if (d_vy) cache_free(d_vy);
    if (d_ex) cache_free(d_ex);
    if (d_ey) cache_free(d_ey);
    void* ps[] = {d_rows, d_src, d_trgt, d_energy, d_width, d_grp, d_s[0], d_s[1], d_e[0], d_e[1],
                  d_cnt, d_cost[0], d_cost[1], d_time[0], d_time[1]};
    for (void* p : ps)
      if (p) cache_free(p);
  }
};

extern "C" {

const char* pgc_last_error(void) { return g_last_error.c_str(); }
int pgc_abi_version(void) { return 4; }

int pgc_context_create(int device, pgc_context** out) {
  if (!out) return fail(PGC_ERR_ARG, "out is null");
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(PGC_ERR_ARG, "invalid device ordinal");
  HIPCHK(hipSetDevice(device));
  auto* c = new pgc_context();
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&c->select_state, select_state_bytes());
  if (e == hipSuccess) e = hipEventCreate(&c->ev0);
  if (e == hipSuccess) e = hipEventCreate(&c->ev1);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&c->num_cu, hipDeviceAttributeMultiprocessorCount, device);
  if (e == hipSuccess) e = hipMalloc((void**)&c->d_queue, 64 * sizeof(int32_t));
  if (e != hipSuccess) {
    delete c;
    return fail(PGC_ERR_HIP, std::string("context setup: ") + hipGetErrorString(e));
  }
  // the device address of the static scratch, resolved once per device; a
  // failure only disables pgc_solve_ratio_noalloc (which reports it)
  (void)scratch_setup(device);
  *out = c;
  return PGC_OK;
}

int pgc_context_set_tuning(pgc_context* c, const pgc_tuning* t) {
  if (!c) return fail(PGC_ERR_ARG, "context is null");
  if (!t) {
    c->tune = pgc_tuning{-1, -1, -1, -1, -1, -1, -1, -1, -1};
    return PGC_OK;
  }
  if (t->grid_chunks_per_wave < -1 || t->grid_chunks_per_wave > 64)
    return fail(PGC_ERR_ARG, "grid_chunks_per_wave must be -1, 0 or 1..64");
  for (int32_t v : {t->pred_in_lds, t->wrap_in_lds, t->origins, t->cached_sources, t->short_slots})
    if (v != -1 && v != 0) return fail(PGC_ERR_ARG, "placement choices must be -1 (rule) or 0 (off)");
  if (t->global_table < -1 || t->global_table > 1) return fail(PGC_ERR_ARG, "global_table must be -1, 0 or 1");
  if (t->grid_workgroups < -1 || t->grid_workgroups > 256)
    return fail(PGC_ERR_ARG, "grid_workgroups must be -1, 0 or 1..256");
  if (t->implicit_costs < -1 || t->implicit_costs > 1) return fail(PGC_ERR_ARG, "implicit_costs must be -1, 0 or 1");
  c->tune = *t;
  return PGC_OK;
}

int pgc_context_get_tuning(const pgc_context* c, pgc_tuning* out) {
  if (!c || !out) return fail(PGC_ERR_ARG, "invalid arguments");
  *out = c->tune;
  return PGC_OK;
}

void pgc_context_destroy(pgc_context* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->select_state) cache_free(c->select_state);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  void* ps[] = {c->d_jobs, c->d_res, c->d_out, c->d_rows, c->d_queue, c->d_items, c->d_ceids, c->d_values,
                c->d_qitems, c->d_fitems, c->d_eitems, c->d_select, c->d_bracket, c->d_cand, c->d_fills,
                c->d_qcand};
  for (void* p : ps)
    if (p) cache_free(p);
  for (int k = 0; k < 4; ++k) {
    if (c->aux[k]) {
      (void)hipStreamSynchronize(c->aux[k]);
      (void)hipStreamDestroy(c->aux[k]);
    }
    if (c->join[k]) (void)hipEventDestroy(c->join[k]);
  }
  if (c->fork) (void)hipEventDestroy(c->fork);
  if (c->fork2) (void)hipEventDestroy(c->fork2);
  if (c->pinned) (void)hipHostFree(c->pinned);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

}  // extern "C"

namespace {

// Pinned host staging of the small copies of a solve (job upload, result,
// output cycle, the X edge list): a copy from or to pageable memory goes
// through the runtime's own staging and costs tens of microseconds even for
// a few bytes. The context's buffer has an upload half [0, kStageUp) and a
// download half after it; every caller synchronizes the stream before the
// buffer is written again. nullptr: larger than the buffer may grow (the
// caller copies directly).
constexpr size_t kStageUp = (size_t)64 << 10;
constexpr size_t kStageMax = (size_t)8 << 20;
uint8_t* stage(pgc_context* c, size_t down) {
  if (down > kStageMax) return nullptr;
  const size_t need = kStageUp + down;
  if (c->pinned_cap < need) {
    // an upload from the old buffer may still be queued
    if (c->pinned && hipStreamSynchronize(c->stream) != hipSuccess) return nullptr;
    if (c->pinned) (void)hipHostFree(c->pinned);
    c->pinned = nullptr;
    c->pinned_cap = 0;
    const size_t cap = std::max(need, kStageUp + ((size_t)1 << 20));
    void* p = nullptr;
    if (hipHostMalloc(&p, cap, hipHostMallocDefault) != hipSuccess) return nullptr;
    c->pinned = (uint8_t*)p;
    c->pinned_cap = cap;
  }
  return c->pinned;
}

struct Xfer {
  void* dst;
  const void* src;
  size_t bytes;
};

// Uploads of small host buffers (through the upload half when they fit).
int h2d_many(pgc_context* c, std::initializer_list<Xfer> xs) {
  size_t tot = 0;
  for (const Xfer& x : xs) tot += (x.bytes + 255) & ~(size_t)255;
  uint8_t* p = tot <= kStageUp ? stage(c, 0) : nullptr;
  size_t o = 0;
  for (const Xfer& x : xs) {
    if (!x.bytes) continue;
    if (p) {
      std::memcpy(p + o, x.src, x.bytes);
      HIPCHK(hipMemcpyAsync(x.dst, p + o, x.bytes, hipMemcpyHostToDevice, c->stream));
      o += (x.bytes + 255) & ~(size_t)255;
    } else {
      HIPCHK(hipMemcpyAsync(x.dst, x.src, x.bytes, hipMemcpyHostToDevice, c->stream));
    }
  }
  return PGC_OK;
}

int h2d_small(pgc_context* c, void* dst, const void* src, size_t bytes) { return h2d_many(c, {{dst, src, bytes}}); }

// Downloads of device buffers and one stream synchronization.
int d2h_many(pgc_context* c, std::initializer_list<Xfer> xs) {
  size_t tot = 0;
  for (const Xfer& x : xs) tot += (x.bytes + 255) & ~(size_t)255;
  uint8_t* p = stage(c, tot);
  uint8_t* d = p ? p + kStageUp : nullptr;
  size_t o = 0;
  for (const Xfer& x : xs) {
    if (!x.bytes) continue;
    HIPCHK(hipMemcpyAsync(d ? d + o : x.dst, x.src, x.bytes, hipMemcpyDeviceToHost, c->stream));
    o += (x.bytes + 255) & ~(size_t)255;
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  if (d) {
    o = 0;
    for (const Xfer& x : xs) {
      if (!x.bytes) continue;
      std::memcpy(x.dst, d + o, x.bytes);
      o += (x.bytes + 255) & ~(size_t)255;
    }
  }
  return PGC_OK;
}

int d2h_sync(pgc_context* c, void* dst0, const void* src0, size_t b0, void* dst1 = nullptr,
             const void* src1 = nullptr, size_t b1 = 0) {
  return d2h_many(c, {{dst0, src0, b0}, {dst1, src1, b1}});
}

// maxSrc: max(SRC) when the caller knows it in closed form, else -2 (a
// reduction over SRC)
int finish_graph_meta(pgc_graph* g, int64_t maxSrc = -2) {
  // numNodes = max(SRC) + 1, numLayers = numNodes / npl (run(), :870-871)
  int64_t mx = maxSrc;
  if (mx == -2) {
    int32_t* d_max = nullptr;
    HIPCHK(dalloc(&d_max, 1));
    const int32_t init = -1;
    HIPCHK(hipMemcpyAsync(d_max, &init, 4, hipMemcpyHostToDevice, g->ctx->stream));
    HIPCHK(launch_int_max(g->d_src, g->E, d_max, g->ctx->stream));
    int32_t m = -1;
    HIPCHK(hipMemcpyAsync(&m, d_max, 4, hipMemcpyDeviceToHost, g->ctx->stream));
    HIPCHK(hipStreamSynchronize(g->ctx->stream));
    cache_free(d_max);
    mx = m;
  }
  g->numNodes = mx + 1;
  g->L = g->numNodes / g->npl;
  if (g->L < 1) return fail(PGC_ERR_ARG, "product graph has fewer nodes than one layer");
  if (g->numNodes % g->npl != 0)
    return fail(PGC_ERR_UNSUPPORTED, "max(SRC)+1 is not a multiple of the nodes per layer");
  return PGC_OK;
}

constexpr int64_t kSlotPad = 32 * 64;

// both classes get the solver's slot width: the smallest of 8, 16, 32 that
// holds the largest in-degree, above 32 the largest in-degree rounded up to
// a multiple of 8 (the runtime-width variant; beyond kMaxSlotWidth the graph
// is rejected by check_width)
int32_t slot_width(int32_t mx) { return mx <= 8 ? 8 : (mx <= 16 ? 16 : (mx <= 32 ? 32 : (mx + 7) / 8 * 8)); }

// The layout of a generated non-conjugate product graph, in closed form from
// its X graph (k_xcount/k_xfill/k_xreplicate): the same slots, in the same
// order, as the generic build from SRC/TRGT below, at the cost of one pass
// writing them. With three or more Y edges every edge into layer 0 comes from
// the last layer (its source id exceeds npl) and moves to the closing layer;
// with fewer the reference's strict wrap test (Q13) keeps some of them on
// layer 0 or drops them, and the generic build handles that.
int build_layout_closed(pgc_graph* g, DevAlloc& A) {
  hipStream_t st = g->ctx->stream;
  const int64_t npl = g->npl;
  g->nLayers = g->L + 1;
  g->nChunks = (int)((npl + 63) / 64);
  const int nC = g->nChunks;
  int32_t* d_fill = nullptr;
  HIPCHK(A.get(&g->d_cnt, 2 * npl));
  HIPCHK(A.get(&d_fill, 2 * npl));
  HIPCHK(A.get(&g->d_width, 4));
  HIPCHK(launch_fill_i32(g->d_cnt, 2 * npl, 0, st));
  HIPCHK(launch_fill_i32(d_fill, 2 * npl, 0, st));
  HIPCHK(launch_xlayout_count(g->d_ex, g->nEX, npl, g->prune, g->d_cnt, st));
  int32_t wn[4];
  if (g->xdeg[0] >= 0) {
    std::copy(g->xdeg, g->xdeg + 4, wn);  // known from the host copy of EX
  } else {
    HIPCHK(launch_layout_width(g->d_cnt, npl, g->d_width, st));
    HIPCHK(hipMemcpyAsync(wn, g->d_width, 4 * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  const int32_t mx = std::max(wn[0], wn[1]);
  const int32_t W = slot_width(mx);
  g->width[0] = W;
  g->width[1] = W;
  g->max_in_degree = mx;
  g->maxw = mx;
  g->cls_deg[0] = wn[0];
  g->cls_deg[1] = wn[1];
  // each class has in-edges on L layers (intralayer 0..L-1, interlayer 1..L)
  g->intra_slots_valid = (int64_t)wn[2] * g->L;
  g->inter_slots_valid = (int64_t)wn[3] * g->L;
  g->dropped = 0;
  // both classes' width, by a fill kernel: a copy from pageable host memory
  // would wait for the queued work
  HIPCHK(launch_fill_i32(g->d_width, 2, W, st));
  const int64_t tn = (int64_t)nC * W * 64;  // template slots per class
  int32_t* d_t = nullptr;                   // [src0 | src1 | r0 | r1]
  HIPCHK(A.get(&d_t, 4 * tn));
  HIPCHK(launch_fill_i32(d_t, 4 * tn, -1, st));
  HIPCHK(launch_xlayout_fill(g->d_ex, g->nEX, npl, g->prune, nC, g->d_width, g->d_cnt, d_fill, d_t, d_t + tn,
                             d_t + 2 * tn, d_t + 3 * tn, st));
  const int64_t B = (g->prune ? 1 : 2) * g->nEX + g->nVX;  // edges per layer
  for (int c = 0; c < 2; ++c) {
    const int64_t per_layer = tn;
    const int64_t total = per_layer * g->nLayers;
    if (total >= ((int64_t)1 << 31) - 1 - kSlotPad)
      return fail(PGC_ERR_UNSUPPORTED, "in-edge layout exceeds 2^31 slots");
    g->lstride[c] = (int32_t)per_layer;
    g->nslots[c] = total;
    HIPCHK(A.get(&g->d_s[c], total + kSlotPad));
    HIPCHK(A.get(&g->d_e[c], total + kSlotPad));
    HIPCHK(launch_xlayout_replicate(d_t + c * tn, d_t + (2 + c) * tn, per_layer, g->nLayers, total + kSlotPad, c, B,
                                    (int32_t)npl, g->d_s[c], g->d_e[c], st));
  }
  if (W > 32) {
    HIPCHK(A.get(&g->d_grp, 2 * (int64_t)nC));
    HIPCHK(launch_fill_i32(g->d_grp, 2 * (int64_t)nC, 0, st));
    HIPCHK(launch_chunk_groups(g->d_cnt, npl, npl, nC, g->d_grp, st));
  }
  g->src_inv[0] = 1;
  g->src_inv[1] = 1;
  A.put(d_t);
  A.put(d_fill);
  g->layout = true;
  return PGC_OK;
}

// The layout of a conjugate product graph from its first two layers: every
// layer has the same in-edge pattern, with edge ids offset by rep_B per layer
// (k_conj_product_space). The generic build runs on layers 0 and 1 only (a
// three-layer graph whose third layer holds the targets of layer 1's
// interlayer edges); layer 0's intralayer slots and layer 1's interlayer
// slots, whose edge ids are those of layer 0, are then copied to every layer
// with the layer's edge-id base, as for build_layout_closed. With three or
// more layers every edge into layer 0 comes from the last layer and moves to
// the closing layer, so the pattern holds there too. Returns PGC_ERR_UNSUPPORTED
// (nothing built) if the first two layers hold an edge of neither class.
int build_layout_repeat(pgc_graph* g, DevAlloc& A) {
  hipStream_t st = g->ctx->stream;
  const int64_t npl = g->npl, B = g->rep_B;
  const int64_t nTotM = 3 * npl;  // nodes of the three-layer graph
  g->nLayers = g->L + 1;
  g->nChunks = (int)((npl + 63) / 64);
  const int nC = g->nChunks;
  int32_t *d_dropped = nullptr, *d_fill = nullptr;
  HIPCHK(A.get(&g->d_cnt, 2 * nTotM));
  HIPCHK(A.get(&d_fill, 2 * nTotM));
  HIPCHK(A.get(&d_dropped, 1));
  HIPCHK(A.get(&g->d_width, 4));
  HIPCHK(launch_fill_i32(g->d_cnt, 2 * nTotM, 0, st));
  HIPCHK(launch_fill_i32(d_fill, 2 * nTotM, 0, st));
  HIPCHK(launch_fill_i32(d_dropped, 1, 0, st));
  HIPCHK(launch_layout_count(g->d_src, g->d_trgt, 2 * B, npl, 2 * npl, g->d_cnt, d_dropped, st));
  HIPCHK(launch_layout_width(g->d_cnt, nTotM, g->d_width, st));
  int32_t dropped = 0;
  HIPCHK(hipMemcpyAsync(&dropped, d_dropped, 4, hipMemcpyDeviceToHost, st));
  int32_t wn[4];
  HIPCHK(hipMemcpyAsync(wn, g->d_width, 4 * 4, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (dropped) {
    A.put(d_dropped);
    A.put(d_fill);
    A.put(g->d_width);
    A.put(g->d_cnt);
    g->d_width = g->d_cnt = nullptr;
    return PGC_ERR_UNSUPPORTED;
  }
  const int32_t mx = std::max(wn[0], wn[1]);
  const int32_t W = slot_width(mx);
  g->width[0] = W;
  g->width[1] = W;
  g->max_in_degree = mx;
  g->maxw = mx;
  g->cls_deg[0] = wn[0];
  g->cls_deg[1] = wn[1];
  // each class's edges of the two layers are two copies of one layer's
  g->intra_slots_valid = (int64_t)wn[2] / 2 * g->L;
  g->inter_slots_valid = (int64_t)wn[3] / 2 * g->L;
  g->dropped = 0;
  HIPCHK(hipMemcpyAsync(g->d_width, g->width, 2 * 4, hipMemcpyHostToDevice, st));
  const int64_t tn = (int64_t)nC * W * 64;  // slots per layer and class
  int32_t* d_m = nullptr;                   // [src0 | src1 | eid0 | eid1] of the three layers
  HIPCHK(A.get(&d_m, 4 * 3 * tn));
  HIPCHK(launch_fill_i32(d_m, 4 * 3 * tn, -1, st));
  HIPCHK(launch_layout_fill(g->d_src, g->d_trgt, 2 * B, npl, 2 * npl, nC, d_fill, g->d_cnt, g->d_width, d_m,
                            d_m + 3 * tn, d_m + 6 * tn, d_m + 9 * tn, st));
  for (int c = 0; c < 2; ++c) {
    const int64_t total = tn * g->nLayers;
    if (total >= ((int64_t)1 << 31) - 1 - kSlotPad)
      return fail(PGC_ERR_UNSUPPORTED, "in-edge layout exceeds 2^31 slots");
    g->lstride[c] = (int32_t)tn;
    g->nslots[c] = total;
    HIPCHK(A.get(&g->d_s[c], total + kSlotPad));
    HIPCHK(A.get(&g->d_e[c], total + kSlotPad));
    // the template: layer 0 of the intralayer class, layer 1 of the interlayer class
    const int64_t at = c == 0 ? 0 : tn;
    HIPCHK(launch_xlayout_replicate(d_m + c * 3 * tn + at, d_m + (2 + c) * 3 * tn + at, tn, g->nLayers,
                                    total + kSlotPad, c, B, (int32_t)npl, g->d_s[c], g->d_e[c], st));
  }
  if (W > 32) {
    HIPCHK(A.get(&g->d_grp, 2 * (int64_t)nC));
    HIPCHK(launch_fill_i32(g->d_grp, 2 * (int64_t)nC, 0, st));
    HIPCHK(launch_chunk_groups(g->d_cnt, npl, nTotM, nC, g->d_grp, st));
  }
  g->src_inv[0] = 1;
  g->src_inv[1] = 1;
  A.put(d_m);
  A.put(d_dropped);
  A.put(d_fill);
  g->layout = true;
  return PGC_OK;
}

int build_layout(pgc_graph* g, DevAlloc& A) {
  if (g->layout) return PGC_OK;
  const char* generic = std::getenv("PGC_GENERIC_LAYOUT");
  const bool gen = generic && generic[0] == '1';
  if (g->d_ex && !g->conjugate && !g->base && g->nEY >= 3 && g->L == g->nEY && g->npl == g->nVX && !gen)
    return build_layout_closed(g, A);
  if (g->conjugate && g->rep_B > 0 && g->L >= 3 && g->numNodes == g->L * g->npl && g->E == g->L * g->rep_B &&
      !gen) {
    const int rc = build_layout_repeat(g, A);
    if (rc != PGC_ERR_UNSUPPORTED) return rc;
  }
  hipStream_t st = g->ctx->stream;
  const int64_t npl = g->npl;
  g->nLayers = g->L + 1;
  g->nChunks = (int)((npl + 63) / 64);
  const int nC = g->nChunks;
  const int64_t nTot = g->nLayers * npl;
  int32_t *d_dropped = nullptr, *d_fill = nullptr;
  HIPCHK(A.get(&g->d_cnt, 2 * nTot));
  HIPCHK(A.get(&d_fill, 2 * nTot));
  HIPCHK(A.get(&d_dropped, 2));
  HIPCHK(A.get(&g->d_width, 4));
  HIPCHK(launch_fill_i32(g->d_cnt, 2 * nTot, 0, st));
  HIPCHK(launch_fill_i32(d_fill, 2 * nTot, 0, st));
  HIPCHK(launch_fill_i32(d_dropped, 1, 0, st));
  HIPCHK(launch_layout_count(g->d_src, g->d_trgt, g->E, npl, g->numNodes, g->d_cnt, d_dropped, st));
  HIPCHK(launch_layout_width(g->d_cnt, nTot, g->d_width, st));
  int32_t dropped = 0;
  HIPCHK(hipMemcpyAsync(&dropped, d_dropped, 4, hipMemcpyDeviceToHost, st));
  int32_t wn[4];
  HIPCHK(hipMemcpyAsync(wn, g->d_width, 4 * 4, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  const int32_t mx = std::max(wn[0], wn[1]);
  const int32_t W = slot_width(mx);
  g->width[0] = W;
  g->width[1] = W;
  g->max_in_degree = mx;
  g->cls_deg[0] = wn[0];
  g->cls_deg[1] = wn[1];
  g->intra_slots_valid = wn[2];
  g->inter_slots_valid = wn[3];
  HIPCHK(hipMemcpyAsync(g->d_width, g->width, 2 * 4, hipMemcpyHostToDevice, st));
  g->maxw = g->max_in_degree;
  g->dropped = dropped;
  for (int c = 0; c < 2; ++c) {
    const int64_t per_layer = (int64_t)nC * g->width[c] * 64;
    const int64_t total = per_layer * g->nLayers;
    if (total >= ((int64_t)1 << 31) - 1 - kSlotPad)
      return fail(PGC_ERR_UNSUPPORTED, "in-edge layout exceeds 2^31 slots");
    g->lstride[c] = (int32_t)per_layer;
    g->nslots[c] = total;
    // the solver may read up to 32 slots per lane past a chunk start; the
    // padding keeps those reads inside the allocation
    const int64_t alloc = total + kSlotPad;
    HIPCHK(A.get(&g->d_s[c], alloc));
    HIPCHK(A.get(&g->d_e[c], alloc));
    HIPCHK(launch_fill_i32(g->d_s[c], alloc, -1, st));
    HIPCHK(launch_fill_i32(g->d_e[c], alloc, -1, st));
  }
  HIPCHK(launch_layout_fill(g->d_src, g->d_trgt, g->E, npl, g->numNodes, nC, d_fill, g->d_cnt, g->d_width,
                            g->d_s[0], g->d_s[1], g->d_e[0], g->d_e[1], st));
  for (int c = 0; c < 2; ++c) HIPCHK(launch_pad_src(g->d_s[c], g->nslots[c] + kSlotPad, (int32_t)npl, st));
  if (W > 32) {
    HIPCHK(A.get(&g->d_grp, 2 * (int64_t)nC));
    HIPCHK(launch_fill_i32(g->d_grp, 2 * (int64_t)nC, 0, st));
    HIPCHK(launch_chunk_groups(g->d_cnt, npl, nTot, nC, g->d_grp, st));
  }
  // layer-invariant sources: every layer repeats the X adjacency (the
  // interlayer class from layer 1 on, as layer 0 has no interlayer in-edges;
  // the intralayer class up to layer L-1, as the closing layer L has none)
  HIPCHK(launch_fill_i32(d_dropped, 2, -1, st));
  HIPCHK(launch_src_invariant(g->d_s[0], g->lstride[0], 0, g->nLayers - 1, d_dropped, st));
  HIPCHK(launch_src_invariant(g->d_s[1], g->lstride[1], 1, g->nLayers, d_dropped + 1, st));
  int32_t inv[2] = {0, 0};
  HIPCHK(hipMemcpyAsync(inv, d_dropped, 8, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  g->src_inv[0] = inv[0] != 0;
  g->src_inv[1] = inv[1] != 0;
  A.put(d_dropped);
  A.put(d_fill);
  g->layout = true;
  return PGC_OK;
}

int build_layout(pgc_graph* g) {
  DevAlloc A;
  return build_layout(g, A);
}

// kind 1: float q-unit cost (+ float time when a time column exists); kind 2: double raw cost
// Slot cost buffers of the given kind (sized, reused while the size fits);
// fill_costs then writes every slot.
int cost_buffers(pgc_graph* g, int kind, DevAlloc& A) {
  hipStream_t st = g->ctx->stream;
  for (int c = 0; c < 2; ++c) {
    const int64_t n = g->nslots[c];
    const size_t cb = (size_t)(n + kSlotPad) * (kind == 2 ? 8 : 4);
    const bool withTime = kind == 1 && g->ncols > 1;
    const size_t tb = withTime ? (size_t)(n + kSlotPad) * 4 : 0;
    // buffers are reused while their size fits; every slot is rewritten below
    if (g->cost_bytes[c] != cb) {
      if (g->d_cost[c]) {
        HIPCHK(hipStreamSynchronize(st));
        A.put(g->d_cost[c]);
      }
      g->d_cost[c] = nullptr;
      g->cost_bytes[c] = 0;
      int32_t* p = nullptr;
      HIPCHK(A.get(&p, cb / 4));
      HIPCHK(launch_fill_i32(p, (int64_t)(cb / 4), 0, st));
      g->d_cost[c] = p;
      g->cost_bytes[c] = cb;
    }
    if (g->time_bytes[c] != tb) {
      if (g->d_time[c]) {
        HIPCHK(hipStreamSynchronize(st));
        A.put(g->d_time[c]);
      }
      g->d_time[c] = nullptr;
      g->time_bytes[c] = 0;
      if (tb) {
        HIPCHK(A.get(&g->d_time[c], n + kSlotPad));
        HIPCHK(launch_fill_i32((int32_t*)g->d_time[c], n + kSlotPad, 0, st));
        g->time_bytes[c] = tb;
      }
    }
  }
  g->cost_kind = kind;
  return PGC_OK;
}

int cost_buffers(pgc_graph* g, int kind) {
  DevAlloc A;
  return cost_buffers(g, kind, A);
}

int fill_costs(pgc_graph* g, int kind) {
  int rc = cost_buffers(g, kind);
  if (rc) return rc;
  hipStream_t st = g->ctx->stream;
  for (int c = 0; c < 2; ++c) {
    const bool withTime = kind == 1 && g->ncols > 1;
    // the slots past the class's largest in-degree keep the zeros of the
    // buffers' allocation (a graph's layout does not change)
    HIPCHK(launch_fill_costs(g->d_e[c], g->d_s[c], g->nslots[c], g->d_energy, g->E, g->ncols, kind == 2,
                             g->d_cost[c], withTime ? g->d_time[c] : nullptr, g->width[c] <= 32 ? g->width[c] : 0,
                             g->cls_deg[c], st));
  }
  return PGC_OK;
}

// Chunks (64-vertex slices) per wave of the solver kernel for this graph:
// 1, 2, 3, 4 or 8 with the distance table in LDS, 0 for the variant with the
// table in global memory (tables beyond the LDS or more than 8 chunks per wave).
int kernel_maxc(const pgc_graph* g, int K) {
  // PGC_GLOBAL_TABLE=1 selects the global-memory variant for every graph (a
  // validation switch: both variants compute the same results)
  static const bool force_global = [] {
    const char* e = std::getenv("PGC_GLOBAL_TABLE");
    return e && e[0] == '1';
  }();
  if (force_global || g->ctx->tune.global_table == 1) return 0;
  const int per = (int)((g->nChunks + (PGC_BLOCK / PGC_WAVE) - 1) / (PGC_BLOCK / PGC_WAVE));
  if (per > 8 || solver_lds_bytes((int)g->npl, K) > 160 * 1024 - 1024) return 0;
  if (g->maxw > 32) return 8;  // the runtime-width variant has the one LDS instance
  return per <= 4 ? per : 8;
}

// The predecessor table of small graphs lives in LDS after the distance table.
bool pred_fits_lds(const pgc_graph* g, int K) {
  if (kernel_maxc(g, K) == 0 || g->ctx->tune.pred_in_lds == 0) return false;
  const size_t pred = (size_t)(g->L + 1) * K * g->npl * 4;
  return solver_lds_bytes((int)g->npl, K) + pred <= 160 * 1024 - 1024;
}

// The wrap map (two ints per vertex) follows in LDS when it also fits.
bool wrap_fits_lds(const pgc_graph* g, int K) {
  if (kernel_maxc(g, K) == 0 || g->ctx->tune.wrap_in_lds == 0) return false;
  size_t b = solver_lds_bytes((int)g->npl, K);
  if (pred_fits_lds(g, K)) b += (size_t)(g->L + 1) * K * g->npl * 4;
  return b + (size_t)g->npl * 8 <= 160 * 1024 - 1024;
}

// K == 1 graphs of the unrolled sweep (one chunk per wave, 8 slots) with pred
// and the wrap map in LDS: two rows of chain origins follow when they fit.
bool org_fits_lds(const pgc_graph* g, int K) {
  if (K != 1 || kernel_maxc(g, K) != 1 || g->maxw > 8 || !pred_fits_lds(g, K) || !wrap_fits_lds(g, K) ||
      g->ctx->tune.origins == 0)
    return false;
  const size_t b = solver_lds_bytes((int)g->npl, K) + (size_t)(g->L + 1) * K * g->npl * 4 + (size_t)g->npl * 8;
  return b + (size_t)g->npl * 8 <= 160 * 1024 - 1024;
}

size_t job_lds_base(const pgc_graph* g, int K) {
  if (kernel_maxc(g, K) == 0) return 0;
  size_t b = solver_lds_bytes((int)g->npl, K);
  if (pred_fits_lds(g, K)) b += (size_t)(g->L + 1) * K * g->npl * 4;
  if (wrap_fits_lds(g, K)) b += (size_t)g->npl * 8;
  if (org_fits_lds(g, K)) b += (size_t)g->npl * 8;
  return b;
}

// LDS of the implicit-cost sweep's two staged feat_diff columns (after the
// origin rows, 8-byte aligned)
size_t implicit_lds_bytes(const pgc_graph* g) { return 4 + (size_t)(g->npl + 1) * 16; }

// A ratio solve of g at depth K recomputes its slot costs from feat_diff: the
// context asked for it, the energy in the graph is the fresh vanilla energy
// of d_fdt (quantized once, by this solve), and the job runs the max_depth-1
// sweep that holds its rows, predecessors and chain origins in LDS, with
// layer-invariant sources in registers (generated graphs: closed-form layout).
bool implicit_applies(const pgc_graph* g, int K) {
  if (!g->implicit_now || !g->d_fdt || K != 1 || g->conjugate || g->ncols != 1 || g->line_integral) return false;
  const pgc_graph* b = g->base ? g->base : g;  // a derived graph shares its base's Y edges
  if (!b->d_ey || b->nEY != g->L || !g->src_inv[1] || g->ctx->tune.cached_sources == 0) return false;
  if (kernel_maxc(g, K) != 1 || g->maxw > 8 || !org_fits_lds(g, K) || g->npl > 4 * PGC_BLOCK) return false;
  return job_lds_base(g, K) + implicit_lds_bytes(g) <= 160 * 1024 - 1024;
}

// dynamic LDS of a job's launch
size_t job_lds_bytes(const pgc_graph* g, int K) {
  return job_lds_base(g, K) + (implicit_applies(g, K) ? implicit_lds_bytes(g) : 0);
}

// Solver scratch for depth K. min_cost: with the buffers only the min-cost
// solve reads (chains, lower bounds, start vertices, node and edge lists).
int ensure_scratch(pgc_graph* g, int K, DevAlloc& A, bool min_cost = true) {
  JobScratch& s = g->js;
  if (s.K == K && s.pred) return PGC_OK;
  if (!A.base) s.release();
  const int64_t npl = g->npl, L = g->L;
  s.K = K;
  s.predN = (L + 1) * K * npl;
  const int64_t stride = L * K + 2;
  s.chainN = npl * stride;
  const int64_t cap = 100 * L * K + 2;  // the reference's cycle buffer (cost_time_ratio_solver_cuda.cpp:194)
  if (cap >= ((int64_t)1 << 31) || s.predN >= ((int64_t)1 << 31))
    return fail(PGC_ERR_UNSUPPORTED, "expanded graph exceeds int32 node ids");
  s.cyc_cap = (int)cap;
  HIPCHK(A.get(&s.pred, s.predN));
  HIPCHK(A.get(&s.wnext, npl));
  HIPCHK(A.get(&s.wlen, npl));
  HIPCHK(A.get(&s.tmp_nodes, cap));
  HIPCHK(A.get(&s.tmp_eid, cap));
  HIPCHK(A.get(&s.neg_nodes, cap));
  HIPCHK(A.get(&s.neg_eid, cap));
  HIPCHK(A.get(&s.neg_len, 1));
  HIPCHK(A.get(&s.tmp_c, cap));
  HIPCHK(A.get(&s.tmp_t, cap));
  HIPCHK(A.get(&s.neg_c, cap));
  HIPCHK(A.get(&s.neg_t, cap));
  HIPCHK(A.get(&s.hits, npl));
  HIPCHK(A.get(&s.branch, npl));
  HIPCHK(A.get(&s.backup, npl));
  HIPCHK(A.get(&s.res, 1));
  HIPCHK(A.get(&s.djob, 1));
  if (min_cost) {
    HIPCHK(A.get(&s.chain, s.chainN));
    HIPCHK(A.get(&s.back, npl));
    HIPCHK(A.get(&s.nodes_io, cap));
    HIPCHK(A.get(&s.eids_io, cap));
    HIPCHK(A.get(&s.lb_key, npl));
  }
  if (kernel_maxc(g, K) == 0) {
    // rows, the minimum row, and (device-wide K == 2 sweep) the two layers'
    // interleaved depth pairs after them, 16-byte aligned
    HIPCHK(A.get(&s.dtab, (int64_t)(2 * K + 1) * (npl + 1) + 4 * (npl + 1) + 2));
    HIPCHK(A.get(&s.grid, (grid_state_bytes() + 3) / 4));
  }
  return PGC_OK;
}

int ensure_scratch(pgc_graph* g, int K) {
  DevAlloc A;
  return ensure_scratch(g, K, A);
}

// Queue and valid-branch cycle store of the device-resident branch and bound.
constexpr int kValidCycles = 64;

int ensure_bnb(pgc_graph* g, DevAlloc& A, int vc_cap) {
  JobScratch& s = g->js;
  if (s.hkey) return PGC_OK;
  const int64_t npl = g->npl;
  s.vc_cap = vc_cap;
  s.vc_stride = (int)(g->L * s.K + 2);  // a once-wrapping cycle has at most L*K+1 edges
  HIPCHK(A.get(&s.hkey, npl));
  HIPCHK(A.get(&s.hval, npl));
  HIPCHK(A.get(&s.hwhere, npl));
  HIPCHK(A.get(&s.vslot, npl));
  HIPCHK(A.get(&s.vc_eid, (int64_t)s.vc_cap * s.vc_stride));
  HIPCHK(A.get(&s.vc_c, (int64_t)s.vc_cap * s.vc_stride));
  HIPCHK(A.get(&s.vc_t, (int64_t)s.vc_cap * s.vc_stride));
  HIPCHK(A.get(&s.vc_len, s.vc_cap));
  return PGC_OK;
}

int ensure_bnb(pgc_graph* g) {
  DevAlloc A;
  return ensure_bnb(g, A, kValidCycles);
}

SolverJob make_job(pgc_graph* g, int K, int mode, uint32_t flags) {
  JobScratch& s = g->js;
  SolverJob j;
  std::memset(&j, 0, sizeof(j));
  j.npl = (int32_t)g->npl;
  j.L = (int32_t)g->L;
  j.K = K;
  j.nChunks = g->nChunks;
  j.numNodes = g->numNodes;
  j.mode = mode;
  j.flags = flags;
  if (g->ctx->tune.cached_sources == 0) j.flags |= FLAG_NO_CACHED_SOURCES;
  if (g->ctx->tune.short_slots == 0) j.flags |= FLAG_FULL_SLOTS;
  j.branchId = 0;
  j.chain_stride = (int32_t)(g->L * K + 2);
  j.cyc_cap = s.cyc_cap;
  j.ratio_in = 0;
  j.upper_bound = std::numeric_limits<double>::infinity();
  j.time_const = 100000.0 / 2;  // float2intScaling / 2 (cost_time_ratio_solver.cpp:770-777)
  j.tol = (double)(float)1e-8;   // the solver's float tolerance member (:79)
  j.max_sweeps = std::numeric_limits<int64_t>::max();
  if (const char* cap = std::getenv("PGC_SWEEP_CAP")) j.max_sweeps = std::max<int64_t>(1, std::atoll(cap));
  for (int c = 0; c < 2; ++c) {
    SlotClass& S = c == 0 ? j.intra : j.inter;
    S.w = g->width[c];
    S.lstride = g->lstride[c];
    S.src_inv = g->src_inv[c];
    S.deg = g->cls_deg[c];
    S.grp = g->d_grp ? g->d_grp + (int64_t)c * g->nChunks : nullptr;
    S.src = g->d_s[c];
    S.cost = g->d_cost[c];
    S.time = g->d_time[c];
    S.eid = g->d_e[c];
  }
  j.branch = s.branch;
  j.pred = s.pred;
  j.backup = s.backup;
  j.wnext = s.wnext;
  j.wlen = s.wlen;
  j.chain = s.chain;
  j.tmp_nodes = s.tmp_nodes;
  j.tmp_eid = s.tmp_eid;
  j.tmp_c = s.tmp_c;
  j.tmp_t = s.tmp_t;
  j.neg_nodes = s.neg_nodes;
  j.neg_eid = s.neg_eid;
  j.neg_c = s.neg_c;
  j.neg_t = s.neg_t;
  j.neg_len = s.neg_len;
  j.hits = s.hits;
  j.lb_key = s.lb_key;
  j.back = s.back;
  j.res = s.res;
  j.hkey = s.hkey;
  j.hval = s.hval;
  j.hwhere = s.hwhere;
  j.vslot = s.vslot;
  j.dtab = s.dtab;
  j.grid = (GridState*)s.grid;
  j.pred_lds = pred_fits_lds(g, K) ? 1 : 0;
  j.wrap_lds = wrap_fits_lds(g, K) ? 1 : 0;
  j.org_lds = org_fits_lds(g, K) ? 1 : 0;
  j.vc_eid = s.vc_eid;
  j.vc_c = s.vc_c;
  j.vc_t = s.vc_t;
  j.vc_len = s.vc_len;
  j.vc_cap = s.vc_cap;
  // PGC_VALID_STORE=n caps the device store of valid-branch cycles at n (a
  // validation switch for the host fallback of batched solves)
  if (const char* vs = std::getenv("PGC_VALID_STORE")) j.vc_cap = std::min(j.vc_cap, std::max(0, std::atoi(vs)));
  j.vc_stride = s.vc_stride;
  if (mode != MODE_MIN_COST && implicit_applies(g, K)) {
    j.icol = g->d_fdt;
    j.iey = (g->base ? g->base : g)->d_ey;
    j.iscale = g->d_qscale;
  }
  return j;
}

struct Timer {
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  double ms() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
};

// one solver launch: upload job, run, read result (kernel time accumulated)
static thread_local double g_phase_ms[5];
static thread_local int32_t g_launches;  // solver launches of the current call

// Workgroups of the device-wide solve of a job (0: one workgroup). Graphs
// whose table lives in global memory and that have at least 128 chunks of 64
// nodes per layer are solved by ceil(chunks / 16) workgroups, one chunk per
// wave (C2's 720 chunks: 45 workgroups); smaller ones keep one workgroup, as
// a device-wide barrier per step would cost more than their steps take.
// PGC_GRID=0 turns it off; PGC_GRID=n, or the tuning choice grid_workgroups = n,
// forces n workgroups for every such job (a validation switch for small graphs).
int grid_blocks(const pgc_graph* g, int K, bool mincost) {
  if (kernel_maxc(g, K) != 0 || g->maxw > 32) return 0;
  static const int env = [] {
    const char* e = std::getenv("PGC_GRID");
    return e ? std::atoi(e) : -1;
  }();
  if (env == 0) return 0;
  const int cpw = g->ctx->tune.grid_chunks_per_wave;
  int n;
  if (env > 0)
    n = env;
  else if (g->ctx->tune.grid_workgroups > 0)
    n = g->ctx->tune.grid_workgroups;
  else if (cpw == 0)
    n = 0;
  else if (cpw > 0)
    n = (g->nChunks + 16 * cpw - 1) / (16 * cpw);
  else
    n = g->nChunks >= 128 ? (g->nChunks + 15) / 16 : 0;
  if (n <= 0) return 0;
  const int cap = solver_grid_capacity(g->maxw, mincost, g->d_time[0] != nullptr, std::max(1, g->ctx->num_cu));
  n = std::min(n, std::min(cap, 256));
  return n >= 1 ? n : 0;
}

int run_job(pgc_graph* g, const SolverJob& j, bool mincost, JobResult& out, double& kms) {
  pgc_context* c = g->ctx;
  hipStream_t st = c->stream;
  if (int rc = h2d_small(c, g->js.djob, &j, sizeof(j))) return rc;
  const int maxc = kernel_maxc(g, j.K);
  const size_t lds = job_lds_bytes(g, j.K);
  const int nblk = j.grid ? grid_blocks(g, j.K, mincost) : 0;
  if (nblk > 0) HIPCHK(launch_fill_i32(g->js.grid, (int64_t)((grid_state_bytes() + 3) / 4), 0, st));
  HIPCHK(hipEventRecord(c->ev0, st));
  if (nblk > 0)
    HIPCHK(launch_solver_grid(g->js.djob, nblk, g->maxw, mincost, g->d_time[0] != nullptr, st));
  else
    HIPCHK(launch_solver(g->js.djob, 1, 1, nullptr, maxc, g->maxw, mincost, g->d_time[0] != nullptr,
                         g->cost_kind == 2, lds, st));
  HIPCHK(hipEventRecord(c->ev1, st));
  if (int rc = d2h_sync(c, &out, g->js.res, sizeof(out))) return rc;
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  kms += ms;
  const int64_t ticks[5] = {out.t_sweep, out.t_fold, out.t_chase, out.t_walk, out.t_extract};
  for (int i = 0; i < 5; ++i) g_phase_ms[i] += (double)ticks[i] * 1e-5;  // 100 MHz wall clock
  g_launches++;
  return PGC_OK;
}

int check_solvable(pgc_graph* g, int K) {
  if (K < 1) return fail(PGC_ERR_ARG, "max_depth must be >= 1");
  if (g->numNodes % g->npl != 0)
    return fail(PGC_ERR_UNSUPPORTED, "max(SRC)+1 is not a multiple of the nodes per layer (pruned conjugate graph: "
                                     "the reference's numLayers drops the last layer)");
  if ((int64_t)(g->L + 1) * K * g->npl >= ((int64_t)1 << 31))
    return fail(PGC_ERR_UNSUPPORTED, "expanded graph exceeds int32 node ids");
  return PGC_OK;
}

int check_width(pgc_graph* g) {
  if (g->maxw > kMaxSlotWidth)
    return fail(PGC_ERR_UNSUPPORTED, "a node has more than " + std::to_string(kMaxSlotWidth) +
                                         " in-edges of one class");
  return PGC_OK;
}

// E_x: relaxations attempted per sweep = intra slots*(K-1) + inter slots*K
double relax_per_sweep(pgc_graph* g, int K) {
  return (double)g->intra_slots_valid * (K - 1) + (double)g->inter_slots_valid * K;
}

}  // namespace

extern "C" {

int pgc_graph_from_arrays(pgc_context* ctx, const int32_t* productspace, int ps_cols, const int32_t* src,
                          const int32_t* trgt, int64_t n_edges, const double* cost, int n_cols,
                          int64_t nodes_per_layer, pgc_graph** out) {
  if (!ctx || !out || !src || !trgt || !cost || n_edges <= 0 || nodes_per_layer <= 0)
    return fail(PGC_ERR_ARG, "invalid arguments");
  if (n_cols < 1 || n_cols > 2) return fail(PGC_ERR_ARG, "cost must have 1 or 2 columns");
  if (ps_cols != 4 && ps_cols != 6) return fail(PGC_ERR_ARG, "product space must have 4 or 6 columns");
  HIPCHK(hipSetDevice(ctx->device));
  std::unique_ptr<pgc_graph> g(new pgc_graph());
  g->ctx = ctx;
  g->E = n_edges;
  g->npl = nodes_per_layer;
  g->ps_cols = ps_cols;
  g->ncols = n_cols;
  hipStream_t st = ctx->stream;
  HIPCHK(dalloc(&g->d_rows, n_edges * ps_cols));
  HIPCHK(dalloc(&g->d_src, n_edges));
  HIPCHK(dalloc(&g->d_trgt, n_edges));
  HIPCHK(dalloc(&g->d_energy, n_edges * n_cols));
  if (productspace)
    HIPCHK(hipMemcpyAsync(g->d_rows, productspace, n_edges * ps_cols * 4, hipMemcpyHostToDevice, st));
  else
    HIPCHK(hipMemsetAsync(g->d_rows, 0xff, n_edges * ps_cols * 4, st));
  HIPCHK(hipMemcpyAsync(g->d_src, src, n_edges * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(g->d_trgt, trgt, n_edges * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(g->d_energy, cost, n_edges * n_cols * 8, hipMemcpyHostToDevice, st));
  int rc = finish_graph_meta(g.get());
  if (rc) return rc;
  *out = g.release();
  return PGC_OK;
}

}  // extern "C"

namespace {

// Shared body of the two generate entry points; all array arguments are
// device pointers, EX_host (optional) is the host copy used for validation.
// stringToMode of the conjugate energy (higherOrderDeformationEnergy.cpp:393-410),
// including its inverted "penalisedegenerate" test: an unknown name (and "")
// gives PEN_DEG, "penalisedegenerate" itself gives VANILLA
enum ConjMode { CM_NOMODE, CM_VANILLA, CM_FEATURE, CM_LENGTH, CM_SO3_PLAIN, CM_SO3_ROBUST, CM_PEN_DEG };
ConjMode conj_mode(const char* s) {
  const std::string m(s ? s : "");
  if (m == "nomode") return CM_NOMODE;
  if (m == "vanilla") return CM_VANILLA;
  if (m == "feature") return CM_FEATURE;
  if (m == "lengthnormalisation") return CM_LENGTH;
  if (m == "plainso3") return CM_SO3_PLAIN;
  if (m == "robustso3") return CM_SO3_ROBUST;
  if (m != "penalisedegenerate") return CM_PEN_DEG;
  return CM_VANILLA;
}

// Energy table of higherOrderDeformationEnergy.cpp:281-336. (vanilla, nomode)
// aborts in the reference (:216-218, :332-335) and is the one-column mean
// problem here; (length, vanilla) and (penalisedegenerate, vanilla) write a
// second column into a one-column matrix there (:299-313) and have two
// columns here. -1: unsupported pair.
int conj_kind(ConjMode c, ConjMode t) {
  if (c == CM_VANILLA && (t == CM_VANILLA || t == CM_NOMODE)) return CONJ_VV;
  if (c == CM_VANILLA && t == CM_LENGTH) return CONJ_VL;
  if (c == CM_LENGTH && t == CM_VANILLA) return CONJ_LV;
  if (c == CM_VANILLA && t == CM_PEN_DEG) return CONJ_VP;
  if (c == CM_PEN_DEG && t == CM_VANILLA) return CONJ_PV;
  if (c == CM_FEATURE && t == CM_SO3_PLAIN) return CONJ_FSP;
  if (c == CM_FEATURE && t == CM_SO3_ROBUST) return CONJ_FSR;
  if (c == CM_SO3_ROBUST && t == CM_FEATURE) return CONJ_SRF;
  if (c == CM_SO3_PLAIN && t == CM_FEATURE) return CONJ_SPF;
  return -1;
}

double robust_loss(double x, double alpha, double c, double pwr) {
  const double xDivCSqr = std::pow((x / c), pwr);
  const double absAlphaMinusTwo = std::abs(alpha - 2);
  return absAlphaMinusTwo / alpha * (std::pow(xDivCSqr / absAlphaMinusTwo + 1, 0.5 * alpha) - 1);
}

// Conjugate product graph (ProductGraphGenerators::generateConjugate,
// pgg/product_graph_generators.cpp:21-54): product space, branch graph and
// energy. Arrays are device pointers; EX_host is the host copy of EX.
int generate_conj(pgc_context* ctx, const double* dVX, int64_t nVX, const int32_t* dEX, int64_t nEX,
                  const double* dVY, int64_t nVY, const int32_t* dEY, int64_t nEY, const double* dFD,
                  int regularising, int prune, const char* cost_mode, const char* time_mode, const double* dNX,
                  const double* dNY, double rl_alpha, double rl_c, double rl_pwr, const int32_t* EX_host,
                  pgc_graph** out) {
  hipStream_t st = ctx->stream;
  if (!dNX || !dNY)
    return fail(PGC_ERR_ARG, "the conjugate product graph needs normals NX [|VX| x 3] and NY [|VY| x 3]");
  const int kind = conj_kind(conj_mode(cost_mode), conj_mode(time_mode));
  if (kind < 0)
    return fail(PGC_ERR_UNSUPPORTED, std::string("conjugate energy has no (cost, time) mode (") +
                                         (cost_mode ? cost_mode : "") + ", " + (time_mode ? time_mode : "") + ")");
  const ConjMode cm = conj_mode(cost_mode), tm = conj_mode(time_mode);
  // out-edges of every X vertex in ascending edge id; pairs (j, it) per j
  std::vector<int32_t> outOff(nVX + 1, 0), outList(nEX);
  for (int64_t e = 0; e < nEX; ++e) outOff[EX_host[2 * e] + 1]++;
  for (int64_t v = 0; v < nVX; ++v) outOff[v + 1] += outOff[v];
  {
    std::vector<int32_t> fill(outOff.begin(), outOff.end() - 1);
    for (int64_t e = 0; e < nEX; ++e) outList[fill[EX_host[2 * e]]++] = (int32_t)e;
  }
  std::vector<int64_t> pre(nEX + 1, 0);
  for (int64_t j = 0; j < nEX; ++j) {
    int64_t p = 0;
    const int32_t s3 = EX_host[2 * j], m3 = EX_host[2 * j + 1];
    for (int32_t k = outOff[m3]; k < outOff[m3 + 1]; ++k)
      if (EX_host[2 * outList[k] + 1] != s3) p++;
    pre[j + 1] = pre[j] + p;
  }
  const int64_t P = pre[nEX];
  const int64_t intraTotal = prune ? 0 : 2 * P;
  const int64_t perLayer = intraTotal + (prune ? 2 : 3) * P + 2 * nEX;
  const int64_t E = nEY * perLayer;
  if (E >= ((int64_t)1 << 31) - 1) return fail(PGC_ERR_UNSUPPORTED, "conjugate product graph exceeds 2^31 edges");
  std::unique_ptr<pgc_graph> g(new pgc_graph());
  g->ctx = ctx;
  g->E = E;
  g->npl = 3 * nEX;
  g->ps_cols = 6;
  g->conjugate = true;
  g->ncols = kind == CONJ_VV ? 1 : 2;
  int32_t *dOff = nullptr, *dList = nullptr;
  int64_t* dPre = nullptr;
  double *dQuat = nullptr, *dMax = nullptr;
  HIPCHK(dalloc(&g->d_rows, E * 6));
  HIPCHK(dalloc(&g->d_src, E));
  HIPCHK(dalloc(&g->d_trgt, E));
  HIPCHK(dalloc(&g->d_energy, E * g->ncols));
  HIPCHK(dalloc(&dOff, nVX + 1));
  HIPCHK(dalloc(&dList, nEX));
  HIPCHK(dalloc(&dPre, nEX + 1));
  HIPCHK(hipMemcpyAsync(dOff, outOff.data(), (nVX + 1) * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dList, outList.data(), nEX * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dPre, pre.data(), (nEX + 1) * 8, hipMemcpyHostToDevice, st));
  int2* dDesc = nullptr;  // one layer's edge descriptors
  HIPCHK(cache_malloc((void**)&dDesc, (size_t)perLayer * sizeof(int2)));
  HIPCHK(launch_conj_product_space(dEX, nEX, dEY, nEY, prune, dOff, dList, dPre, perLayer, intraTotal, dDesc,
                                   g->d_rows, g->d_src, g->d_trgt, st));
  if (regularising) {
    HIPCHK(dalloc(&dQuat, 4 * nEX * nEY));  // one rotation per (Y edge, X edge) pair
    HIPCHK(launch_conj_quat(dVX, dNX, dVY, dNY, dEX, nEX, dEY, nEY, prune, dQuat, st));
  }
  // maxFeatDiff and maxSO3Val are float in the reference (:211-215)
  float maxFeatDiff = 0;
  if (tm == CM_FEATURE) {
    HIPCHK(dalloc(&dMax, 1));
    HIPCHK(launch_max_f64(dFD, nVX * nVY, dMax, st));
    double m = 0;
    HIPCHK(hipMemcpyAsync(&m, dMax, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    maxFeatDiff = (float)m;
  }
  const float maxSO3Val = (float)robust_loss(2 * M_PI, rl_alpha, rl_c, rl_pwr);
  const int lineIntegral = (cm == CM_LENGTH || tm == CM_LENGTH) ? 1 : 0;
  HIPCHK(launch_conj_energy(dVX, dVY, nVY, g->d_rows, g->d_src, g->d_trgt, E, dFD, dQuat, nEX, nEY, prune, kind,
                            lineIntegral, rl_alpha, rl_c, rl_pwr, (double)maxSO3Val, (double)maxFeatDiff, g->d_energy,
                            st));
  HIPCHK(hipStreamSynchronize(st));
  for (void* p : {(void*)dOff, (void*)dList, (void*)dPre, (void*)dQuat, (void*)dMax, (void*)dDesc})
    if (p) cache_free(p);
  // numNodes = max(SRC) + 1, numLayers = numNodes / npl as run() derives them.
  // max(SRC) in closed form: the last closing edge of the last layer starts
  // at the 3D-degenerate node of the last X edge, off3d + |EX| - 1 in that
  // layer (off3d = 2|EX|, or |EX| when pruned)
  const int64_t maxSrc = E > 0 ? (nEY - 1) * 3 * nEX + (prune ? 2 : 3) * nEX - 1 : -1;
  int rc = finish_graph_meta(g.get(), maxSrc);
  if (rc && rc != PGC_ERR_UNSUPPORTED) return rc;
  g->rep_B = perLayer;
  // the branch graph is built at its first use (build_conj_bg)
  g->cbg_ex0.resize(nEX);
  for (int64_t j = 0; j < nEX; ++j) g->cbg_ex0[j] = EX_host[2 * j];
  g->cbg_off = outOff;
  g->cbg_list = outList;
  g->cbg_prune = prune ? 1 : 0;
  *out = g.release();
  return PGC_OK;
}

// branchGraph of a conjugate product graph (conj_product_space.cpp:170-225;
// both of its loops read EX(j, 0)): node j (non-degenerate) links to j+|EX|
// and to the X out-edges it of EX(j, 0); node j+|EX| to j+2|EX| (to j when
// pruned) and to it+|EX|; node j+2|EX| to j and to it+2|EX|. Values ascending,
// without repeats, as the reference's sets hold them.
void build_conj_bg(pgc_graph* g) {
  const int64_t nEX = (int64_t)g->cbg_ex0.size(), npl = g->npl;
  const bool prune = g->cbg_prune != 0;
  const std::vector<int32_t>& off = g->cbg_off;
  const std::vector<int32_t>& lst = g->cbg_list;
  g->bg_off.assign(npl + 1, 0);
  for (int64_t j = 0; j < nEX; ++j) {
    const int64_t d = 1 + off[g->cbg_ex0[j] + 1] - off[g->cbg_ex0[j]];
    g->bg_off[j + 1] = d;
    g->bg_off[j + nEX + 1] = d;
    if (!prune) g->bg_off[j + 2 * nEX + 1] = d;
  }
  for (int64_t v = 0; v < npl; ++v) g->bg_off[v + 1] += g->bg_off[v];
  g->bg_val.assign(g->bg_off[npl], 0);
  for (int64_t j = 0; j < nEX; ++j) {
    const int32_t v = g->cbg_ex0[j];
    for (int cls = 0; cls < (prune ? 2 : 3); ++cls) {
      const int64_t node = j + cls * nEX;
      int64_t* out = g->bg_val.data() + g->bg_off[node];
      int64_t n = 0;
      out[n++] = cls == 0 ? j + nEX : (cls == 1 ? (prune ? j : j + 2 * nEX) : j);
      for (int32_t k = off[v]; k < off[v + 1]; ++k) out[n++] = (int64_t)lst[k] + cls * nEX;
      std::sort(out, out + n);
    }
  }
  // drop repeats inside each node's list, compacting the CSR
  int64_t w = 0;
  for (int64_t node = 0; node < npl; ++node) {
    const int64_t b = g->bg_off[node], e = g->bg_off[node + 1];
    g->bg_off[node] = w;
    for (int64_t i = b; i < e; ++i)
      if (i == b || g->bg_val[i] != g->bg_val[i - 1]) g->bg_val[w++] = g->bg_val[i];
  }
  g->bg_off[npl] = w;
  g->bg_val.resize(w);
  g->has_bg = true;
}

// Implicit costs: a non-conjugate graph with the vanilla energy (one column,
// no length term) generated in a context with implicit_costs = 1 keeps its
// feat_diff by columns (fd: [nVX x nVY] float64 on the device).
int keep_feat_diff(pgc_graph* g, const double* fd) {
  const pgc_graph* b = g->base ? g->base : g;
  if (g->ctx->tune.implicit_costs != 1 || b->conjugate || g->ncols != 1 || b->line_integral) {
    if (g->d_fdt) cache_free(g->d_fdt);
    g->d_fdt = nullptr;
    return PGC_OK;
  }
  if (!g->d_fdt) HIPCHK(dalloc(&g->d_fdt, b->nVX * b->nVY));
  if (!g->d_qscale) HIPCHK(dalloc(&g->d_qscale, 1));
  HIPCHK(launch_transpose_fd(fd, b->nVX, b->nVY, g->d_fdt, g->ctx->stream));
  return PGC_OK;
}

int generate_impl(pgc_context* ctx, const double* dVX, int64_t nVX, const int32_t* dEX, int64_t nEX,
                  const double* dVY, int64_t nVY, const int32_t* dEY, int64_t nEY, const double* dFD, int conjugate,
                  int regularising, int prune, const char* cost_mode, const char* time_mode, const double* dNX,
                  const double* dNY, double rl_alpha, double rl_c, double rl_pwr, const int32_t* EX_host,
                  pgc_graph** out) {
  hipStream_t st = ctx->stream;
  std::vector<int32_t> exh;
  if (!EX_host) {
    exh.resize(2 * nEX);
    if (int rc = d2h_sync(ctx, exh.data(), dEX, (size_t)(2 * nEX * 4))) return rc;
    EX_host = exh.data();
  }
  int32_t mx = -1;
  for (int64_t i = 0; i < 2 * nEX; ++i) {
    if (EX_host[i] < 0) return fail(PGC_ERR_ARG, "EX holds a negative vertex id");
    mx = std::max(mx, EX_host[i]);
  }
  const int64_t nVXe = (int64_t)mx + 1;  // product_space.cpp:20
  if (nVXe != nVX)
    return fail(PGC_ERR_ARG, "max(EX)+1 differs from |VX|: the product space and the solver would disagree");
  if (conjugate)
    return generate_conj(ctx, dVX, nVX, dEX, nEX, dVY, nVY, dEY, nEY, dFD, regularising, prune, cost_mode, time_mode,
                         dNX, dNY, rl_alpha, rl_c, rl_pwr, EX_host, out);
  std::unique_ptr<pgc_graph> g(new pgc_graph());
  g->ctx = ctx;
  const int64_t E = nEY * ((prune ? 1 : 2) * nEX + nVXe);
  g->E = E;
  g->npl = nVX;
  g->ps_cols = 4;
  const bool lengthMode = cost_mode && time_mode && std::strcmp(cost_mode, "feature") == 0 &&
                          std::strcmp(time_mode, "lengthnormalisation") == 0;
  g->ncols = lengthMode ? 2 : 1;
  HIPCHK(dalloc(&g->d_rows, E * 4));
  HIPCHK(dalloc(&g->d_src, E));
  HIPCHK(dalloc(&g->d_trgt, E));
  HIPCHK(dalloc(&g->d_energy, E * g->ncols));
  HIPCHK(launch_product_space(dEX, nEX, dEY, nEY, nVXe, prune ? 1 : 0, g->d_rows, g->d_src, g->d_trgt, E, st));
  g->nVX = nVX;
  g->nVY = nVY;
  g->line_integral = (regularising || lengthMode) ? 1 : 0;
  g->length_mode = lengthMode ? 1 : 0;
  HIPCHK(dalloc(&g->d_vx, 3 * nVX));
  HIPCHK(dalloc(&g->d_vy, 3 * nVY));
  HIPCHK(hipMemcpyAsync(g->d_vx, dVX, 3 * nVX * 8, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(g->d_vy, dVY, 3 * nVY * 8, hipMemcpyDeviceToDevice, st));
  HIPCHK(dalloc(&g->d_ex, std::max<int64_t>(2 * nEX, 1)));
  HIPCHK(dalloc(&g->d_ey, std::max<int64_t>(2 * nEY, 1)));
  if (nEX) HIPCHK(hipMemcpyAsync(g->d_ex, dEX, 2 * nEX * 4, hipMemcpyDeviceToDevice, st));
  if (nEY) HIPCHK(hipMemcpyAsync(g->d_ey, dEY, 2 * nEY * 4, hipMemcpyDeviceToDevice, st));
  g->nEX = nEX;
  g->nEY = nEY;
  g->prune = prune ? 1 : 0;
  HIPCHK(launch_energy(dVX, dVY, g->d_rows, E, dFD, nVY, g->line_integral, g->length_mode, g->d_energy, st));
  // numNodes / numLayers as the solver derives them (run(), :870-871)
  g->numNodes = nEY * nVXe;
  g->L = nEY;
  int rc = keep_feat_diff(g.get(), dFD);
  if (rc) return rc;
  g->energy_fresh = true;
  // the X graph's in-degrees per class (largest, sum): the closed-form layout
  // takes its slot width from them without reading the device back
  {
    std::vector<int32_t> deg(nVX, 0);
    for (int64_t j = 0; j < nEX; ++j) deg[EX_host[2 * j + 1]]++;
    int32_t mi = 0, mx = 0;
    for (int64_t x = 0; x < nVX; ++x) {
      mi = std::max(mi, prune ? 0 : deg[x]);
      mx = std::max(mx, deg[x] + 1);
    }
    g->xdeg[0] = mi;
    g->xdeg[1] = mx;
    g->xdeg[2] = prune ? 0 : (int32_t)nEX;
    g->xdeg[3] = (int32_t)(nEX + nVX);
  }
  if (prune) {
    // branchGraph[x] = out-neighbour vertex ids (product_space.cpp:122-131)
    std::vector<std::vector<int64_t>> bg(nVX);
    for (int64_t j = 0; j < nEX; ++j) bg[EX_host[2 * j]].push_back(EX_host[2 * j + 1]);
    g->bg_off.assign(nVX + 1, 0);
    g->bg_val.clear();
    for (int64_t v = 0; v < nVX; ++v) {
      std::sort(bg[v].begin(), bg[v].end());
      bg[v].erase(std::unique(bg[v].begin(), bg[v].end()), bg[v].end());
      for (int64_t x : bg[v]) g->bg_val.push_back(x);
      g->bg_off[v + 1] = (int64_t)g->bg_val.size();
    }
    g->has_bg = true;
  }
  *out = g.release();
  return PGC_OK;
}

}  // namespace

extern "C" {

int pgc_graph_generate(pgc_context* ctx, const double* VX, int64_t nVX, const int32_t* EX, int64_t nEX,
                       const double* VY, int64_t nVY, const int32_t* EY, int64_t nEY, const double* feat_diff,
                       int conjugate, int regularising, int prune, const char* cost_mode, const char* time_mode,
                       const double* NX, const double* NY, double rl_alpha, double rl_c, double rl_pwr,
                       pgc_graph** out) {
  if (!ctx || !out || !VX || !EX || !VY || !EY || !feat_diff || nVX <= 0 || nEX <= 0 || nVY <= 0 || nEY <= 0)
    return fail(PGC_ERR_ARG, "invalid arguments");
  for (int64_t i = 0; i < 2 * nEY; ++i)
    if (EY[i] < 0 || EY[i] >= nVY) return fail(PGC_ERR_ARG, "EY holds an id outside [0, |VY|)");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  int32_t *dEX = nullptr, *dEY = nullptr;
  double *dVX = nullptr, *dVY = nullptr, *dFD = nullptr;
  HIPCHK(dalloc(&dEX, 2 * nEX));
  HIPCHK(dalloc(&dEY, 2 * nEY));
  HIPCHK(dalloc(&dVX, 3 * nVX));
  HIPCHK(dalloc(&dVY, 3 * nVY));
  HIPCHK(dalloc(&dFD, nVX * nVY));
  HIPCHK(hipMemcpyAsync(dEX, EX, 2 * nEX * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dEY, EY, 2 * nEY * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dVX, VX, 3 * nVX * 8, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dVY, VY, 3 * nVY * 8, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dFD, feat_diff, nVX * nVY * 8, hipMemcpyHostToDevice, st));
  double *dNX = nullptr, *dNY = nullptr;
  if (NX && NY) {
    HIPCHK(dalloc(&dNX, 3 * nVX));
    HIPCHK(dalloc(&dNY, 3 * nVY));
    HIPCHK(hipMemcpyAsync(dNX, NX, 3 * nVX * 8, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dNY, NY, 3 * nVY * 8, hipMemcpyHostToDevice, st));
  }
  const int rc = generate_impl(ctx, dVX, nVX, dEX, nEX, dVY, nVY, dEY, nEY, dFD, conjugate, regularising, prune,
                               cost_mode, time_mode, dNX, dNY, rl_alpha, rl_c, rl_pwr, EX, out);
  HIPCHK(hipStreamSynchronize(st));
  if (dNX) cache_free(dNX);
  if (dNY) cache_free(dNY);
  cache_free(dEX);
  cache_free(dEY);
  cache_free(dVX);
  cache_free(dVY);
  cache_free(dFD);
  return rc;
}

int pgc_graph_generate_device(pgc_context* ctx, const double* VX, int64_t nVX, const int32_t* EX, int64_t nEX,
                              const double* VY, int64_t nVY, const int32_t* EY, int64_t nEY,
                              const double* feat_diff, int conjugate, int regularising, int prune,
                              const char* cost_mode, const char* time_mode, const double* NX, const double* NY,
                              double rl_alpha, double rl_c, double rl_pwr, pgc_graph** out) {
  if (!ctx || !out || !VX || !EX || !VY || !EY || !feat_diff || nVX <= 0 || nEX <= 0 || nVY <= 0 || nEY <= 0)
    return fail(PGC_ERR_ARG, "invalid arguments");
  HIPCHK(hipSetDevice(ctx->device));
  return generate_impl(ctx, VX, nVX, EX, nEX, VY, nVY, EY, nEY, feat_diff, conjugate, regularising, prune,
                       cost_mode, time_mode, NX, NY, rl_alpha, rl_c, rl_pwr, nullptr, out);
}

int pgc_graph_set_branch_graph(pgc_graph* g, const int64_t* offsets, const int64_t* values) {
  if (!g || !offsets) return fail(PGC_ERR_ARG, "invalid arguments");
  g->bg_off.assign(offsets, offsets + g->npl + 1);
  g->bg_val.assign(values, values + offsets[g->npl]);
  g->has_bg = true;
  return PGC_OK;
}

int pgc_graph_info(const pgc_graph* g, int64_t* n_edges, int32_t* ps_cols, int32_t* n_cost_cols,
                   int64_t* nodes_per_layer, int64_t* n_layers) {
  if (!g) return fail(PGC_ERR_ARG, "graph is null");
  if (n_edges) *n_edges = g->E;
  if (ps_cols) *ps_cols = g->ps_cols;
  if (n_cost_cols) *n_cost_cols = g->ncols;
  if (nodes_per_layer) *nodes_per_layer = g->npl;
  if (n_layers) *n_layers = g->L;
  return PGC_OK;
}

int pgc_graph_get_product_space(pgc_graph* g, int32_t* out) {
  if (!g || !out) return fail(PGC_ERR_ARG, "invalid arguments");
  HIPCHK(hipSetDevice(g->ctx->device));
  HIPCHK(hipMemcpyAsync(out, g->d_rows, g->E * g->ps_cols * 4, hipMemcpyDeviceToHost, g->ctx->stream));
  HIPCHK(hipStreamSynchronize(g->ctx->stream));
  return PGC_OK;
}

int pgc_graph_get_cost(pgc_graph* g, double* out) {
  if (!g || !out) return fail(PGC_ERR_ARG, "invalid arguments");
  HIPCHK(hipSetDevice(g->ctx->device));
  HIPCHK(hipMemcpyAsync(out, g->d_energy, g->E * g->ncols * 8, hipMemcpyDeviceToHost, g->ctx->stream));
  HIPCHK(hipStreamSynchronize(g->ctx->stream));
  return PGC_OK;
}

int pgc_graph_get_layout(pgc_graph* g, int cls, int32_t* src, int32_t* eid, int64_t capacity, int64_t* n_slots,
                         int32_t* width) {
  if (!g || (cls != 0 && cls != 1) || !n_slots || !width) return fail(PGC_ERR_ARG, "invalid arguments");
  HIPCHK(hipSetDevice(g->ctx->device));
  int rc;
  if ((rc = build_layout(g))) return rc;
  *n_slots = g->nslots[cls];
  *width = g->width[cls];
  if (!src || !eid) return PGC_OK;
  if (capacity < g->nslots[cls]) return fail(PGC_ERR_CAPACITY, "layout exceeds capacity");
  hipStream_t st = g->ctx->stream;
  HIPCHK(hipMemcpyAsync(src, g->d_s[cls], g->nslots[cls] * 4, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(eid, g->d_e[cls], g->nslots[cls] * 4, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return PGC_OK;
}

int pgc_graph_set_cost(pgc_graph* g, const double* cost) {
  if (!g || !cost) return fail(PGC_ERR_ARG, "invalid arguments");
  HIPCHK(hipSetDevice(g->ctx->device));
  HIPCHK(hipMemcpyAsync(g->d_energy, cost, g->E * g->ncols * 8, hipMemcpyHostToDevice, g->ctx->stream));
  HIPCHK(hipStreamSynchronize(g->ctx->stream));
  g->cost_kind = 0;
  g->energy_fresh = false;
  return PGC_OK;
}

void pgc_graph_destroy(pgc_graph* g) {
  if (!g) return;
  (void)hipSetDevice(g->ctx->device);
  (void)hipStreamSynchronize(g->ctx->stream);
  if (pgc_graph* b = g->base) {
    delete g;
    if (--b->refs == 0 && b->released) delete b;
    return;
  }
  if (g->refs > 0) {
    g->released = true;  // freed with its last derived graph
    return;
  }
  delete g;
}

static int energy_from(pgc_graph* g, const double* feat_diff, int on_device) {
  const pgc_graph* b = g->base ? g->base : g;
  if (b->conjugate) return fail(PGC_ERR_UNSUPPORTED, "feat_diff updates are available for non-conjugate graphs only");
  if (!b->d_vx) return fail(PGC_ERR_UNSUPPORTED, "the graph was not built by pgc_graph_generate*: no energy inputs");
  hipStream_t st = g->ctx->stream;
  const double* fd = feat_diff;
  double* tmp = nullptr;
  if (!on_device) {
    HIPCHK(dalloc(&tmp, b->nVX * b->nVY));
    HIPCHK(hipMemcpyAsync(tmp, feat_diff, b->nVX * b->nVY * 8, hipMemcpyHostToDevice, st));
    fd = tmp;
  }
  HIPCHK(launch_energy(b->d_vx, b->d_vy, b->d_rows, b->E, fd, b->nVY, b->line_integral, b->length_mode,
                       g->d_energy, st));
  g->cost_kind = 0;
  int rc = keep_feat_diff(g, fd);
  if (rc) return rc;
  g->energy_fresh = true;
  if (tmp) {
    HIPCHK(hipStreamSynchronize(st));
    cache_free(tmp);
  }
  return PGC_OK;
}

int pgc_graph_set_feat_diff(pgc_graph* g, const double* feat_diff, int on_device) {
  if (!g || !feat_diff) return fail(PGC_ERR_ARG, "invalid arguments");
  HIPCHK(hipSetDevice(g->ctx->device));
  return energy_from(g, feat_diff, on_device);
}

int pgc_graph_derive(pgc_graph* base, const double* feat_diff, int on_device, pgc_graph** out) {
  if (!base || !feat_diff || !out) return fail(PGC_ERR_ARG, "invalid arguments");
  pgc_graph* b = base->base ? base->base : base;
  if (b->released) return fail(PGC_ERR_ARG, "base graph was destroyed");
  if (b->conjugate) return fail(PGC_ERR_UNSUPPORTED, "derived graphs are available for non-conjugate graphs only");
  if (!b->d_vx) return fail(PGC_ERR_UNSUPPORTED, "the graph was not built by pgc_graph_generate*: no energy inputs");
  HIPCHK(hipSetDevice(b->ctx->device));
  int rc;
  if ((rc = build_layout(b))) return rc;
  std::unique_ptr<pgc_graph> g(new pgc_graph());
  g->base = b;
  g->ctx = b->ctx;
  g->E = b->E;
  g->npl = b->npl;
  g->numNodes = b->numNodes;
  g->L = b->L;
  g->ps_cols = b->ps_cols;
  g->ncols = b->ncols;
  g->d_rows = b->d_rows;
  g->d_src = b->d_src;
  g->d_trgt = b->d_trgt;
  g->layout = true;
  g->nChunks = b->nChunks;
  g->nLayers = b->nLayers;
  for (int c = 0; c < 2; ++c) {
    g->nslots[c] = b->nslots[c];
    g->d_s[c] = b->d_s[c];
    g->d_e[c] = b->d_e[c];
    g->width[c] = b->width[c];
    g->lstride[c] = b->lstride[c];
    g->src_inv[c] = b->src_inv[c];
    g->cls_deg[c] = b->cls_deg[c];
  }
  g->d_width = b->d_width;
  g->d_grp = b->d_grp;
  g->maxw = b->maxw;
  g->d_cnt = b->d_cnt;
  g->dropped = b->dropped;
  g->intra_slots_valid = b->intra_slots_valid;
  g->inter_slots_valid = b->inter_slots_valid;
  g->has_bg = b->has_bg;
  g->bg_off = b->bg_off;
  g->bg_val = b->bg_val;
  g->nVX = b->nVX;
  g->nVY = b->nVY;
  g->line_integral = b->line_integral;
  g->length_mode = b->length_mode;
  HIPCHK(dalloc(&g->d_energy, g->E * g->ncols));
  // the derived graph counts as a user of b from here on, so that a failure
  // below releases it through the normal path
  b->refs++;
  pgc_graph* raw = g.release();
  raw->base = b;
  if ((rc = energy_from(raw, feat_diff, on_device))) {
    pgc_graph_destroy(raw);
    return rc;
  }
  *out = raw;
  return PGC_OK;
}

}  // extern "C"

namespace {

// The host branch and bound of solveWithBnbGPULawler
// (cost_time_ratio_solver_cuda.cpp:175-331, the CPU twin
// cost_time_ratio_solver.cpp:794-862) around the device Lawler launches, on a
// graph whose costs are quantized and filled and whose scratch is sized for K.
int ratio_bnb_host(pgc_context* ctx, pgc_graph* g, int K, uint32_t flags, int32_t* edge_ids, int32_t* rows,
                   int64_t capacity, pgc_result* res, const Timer& wall) {
  int rc;
  for (double& x : g_phase_ms) x = 0;
  g_launches = 0;
  hipStream_t st = ctx->stream;
  const int64_t npl = g->npl;
  JobScratch& s = g->js;
  HIPCHK(hipMemsetAsync(s.pred, 0xff, s.predN * 4, st));
  HIPCHK(hipMemsetAsync(s.neg_len, 0, 4, st));
  double kms = 0;
  int64_t sweeps = 0, bfc = 0, its = 0, branches = 0;

  // initial upper bound (findInitalUpperBound)
  SolverJob job = make_job(g, K, MODE_UPPER_BOUND, flags);
  JobResult r;
  if ((rc = run_job(g, job, false, r, kms))) return rc;
  if (r.error) return fail(PGC_ERR_SOLVER, "upper-bound search failed (error " + std::to_string(r.error) + ")");
  sweeps += r.sweeps;
  bfc += r.bf_calls;
  const double ratioUpperBound = r.ratio_ub;

  // branch and bound over layer-0 vertex partitions (:794-851)
  std::vector<int32_t> branch(npl, 0);
  int maxBranchId = 0;
  BnbHeap heap(npl);
  heap.push(0, 0);
  std::vector<char> valid(npl, 0);
  struct Cyc {
    std::vector<int32_t> eid;
    std::vector<double> c, t;
    std::vector<int32_t> rows;  // product-space rows, when the caller asked for them
  };
  // the m edges of the cycle in the job's buffers, with their rows gathered
  // on the device from the same ids: one synchronization for all of it
  auto read_cycle = [&](int m, Cyc& cy) -> int {
    cy.eid.resize(m);
    cy.c.resize(m);
    cy.t.resize(m);
    cy.rows.resize(rows ? (size_t)m * g->ps_cols : 0);
    if (!m) return PGC_OK;
    int32_t* d_out = nullptr;
    if (rows) {
      HIPCHK(dalloc(&d_out, (int64_t)m * g->ps_cols));
      HIPCHK(launch_gather_rows(g->d_rows, g->ps_cols, s.neg_eid, m, d_out, st));
      HIPCHK(hipMemcpyAsync(cy.rows.data(), d_out, (size_t)m * g->ps_cols * 4, hipMemcpyDeviceToHost, st));
    }
    HIPCHK(hipMemcpyAsync(cy.eid.data(), s.neg_eid, m * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(cy.c.data(), s.neg_c, m * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(cy.t.data(), s.neg_t, m * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (d_out) cache_free(d_out);
    return PGC_OK;
  };
  std::map<int, Cyc> validCycles;
  Cyc finalCyc;
  bool fromValid = false;
  bool branchDirty = true;
  std::vector<int32_t> hits;
  job.mode = MODE_RATIO;
  while (!heap.empty()) {
    const auto top = heap.pop();
    const int bid = (int)top.second;
    if (valid[bid]) {
      finalCyc = validCycles.at(bid);
      fromValid = true;
      break;
    }
    branches++;
    if (branchDirty) {
      HIPCHK(hipMemcpyAsync(s.branch, branch.data(), npl * 4, hipMemcpyHostToDevice, st));
      branchDirty = false;
    }
    job.branchId = bid;
    job.ratio_in = ratioUpperBound;
    if ((rc = run_job(g, job, false, r, kms))) return rc;
    if (r.error) return fail(PGC_ERR_SOLVER, "Lawler subroutine failed (error " + std::to_string(r.error) + ")");
    sweeps += r.sweeps;
    bfc += r.bf_calls;
    its += r.lawler_iters;
    const bool wraps = r.num_wraps > 1;
    const double lbBranch = r.ratio;
    if (ratioUpperBound < heap.peak() && !wraps) break;
    if (!wraps) {
      heap.push(lbBranch, bid);
      valid[bid] = 1;
      Cyc cy;
      if ((rc = read_cycle(std::max(0, r.neg_len - 1), cy))) return rc;
      validCycles.emplace(bid, std::move(cy));
      continue;
    }
    if (r.num_in_branch == 1) continue;
    // branch(): peel zero-layer hits 2.. into their own branches (:142-179)
    hits.resize(std::min<int64_t>(r.n_hits, npl));
    if (!hits.empty()) {
      HIPCHK(hipMemcpyAsync(hits.data(), s.hits, hits.size() * 4, hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
    }
    int offset = 0;
    for (int32_t h : hits) {
      if (offset > 0 && h >= 0 && h < npl) branch[h] = maxBranchId + offset;
      offset++;
    }
    for (int64_t v = 0; v < npl; ++v)
      if (branch[v] == bid) branch[v] = maxBranchId;
    maxBranchId += offset;
    branchDirty = true;
    for (int32_t h : hits)
      if (h >= 0 && h < npl) heap.push(lbBranch, branch[h]);
  }
  if (!fromValid) {
    int32_t nl = 0;
    HIPCHK(hipMemcpyAsync(&nl, s.neg_len, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if ((rc = read_cycle(std::max(0, nl - 1), finalCyc))) return rc;
  }
  // output (:857-862): ratio of the cycle and its product-space rows
  double cs = 0, ts = 0;
  for (size_t i = 0; i < finalCyc.c.size(); ++i) {
    cs += finalCyc.c[i];
    ts += finalCyc.t[i];
  }
  const int64_t n = (int64_t)finalCyc.eid.size();
  if (n > capacity) return fail(PGC_ERR_CAPACITY, "output capacity too small: need " + std::to_string(n));
  if (edge_ids) std::copy(finalCyc.eid.begin(), finalCyc.eid.end(), edge_ids);
  if (rows) std::copy(finalCyc.rows.begin(), finalCyc.rows.end(), rows);
  res->passes_layers_once = 1;
  res->objective = (float)(cs / ts);
  res->n_edges = n;
  res->sweeps = sweeps;
  g->last_sweeps = sweeps;
  res->bf_calls = bfc;
  res->lawler_iters = its;
  res->branches = branches;
  res->relaxations = relax_per_sweep(g, K) * (double)sweeps;
  res->kernel_ms = kms;
  res->total_ms = wall.ms();
  for (int i = 0; i < 5; ++i) res->phase_ms[i] = g_phase_ms[i];
  res->launches = g_launches;
  res->implicit_costs = implicit_applies(g, K) ? 1 : 0;
  res->status = PGC_OK;
  return PGC_OK;
}

// The same solve as ratio_bnb_host with the branch and bound on the device:
// upper bound, branch and bound, Lawler iterations and output in one launch of
// the one-workgroup solver (Solver::full_ratio), two synchronizations in all.
// The store of valid-branch cycles has one slot per branch id below npl (the
// heap drops higher ids, Q9), so every valid branch fits, unless the graph's
// store was sized before by a batch (64 slots) or PGC_VALID_STORE caps it:
// then an overflow, or the branch cap, sets `redo`, and the caller runs the
// host branch and bound on the same quantized costs (same result).
int ratio_bnb_device(pgc_context* ctx, pgc_graph* g, int K, uint32_t flags, int32_t* edge_ids, int32_t* rows,
                     int64_t capacity, pgc_result* res, const Timer& wall, bool& redo) {
  redo = false;
  int rc;
  hipStream_t st = ctx->stream;
  DevAlloc A;
  if ((rc = ensure_bnb(g, A, (int)g->npl))) return rc;
  JobScratch& s = g->js;
  const int64_t outCap = std::max<int64_t>(0, std::min<int64_t>(capacity, s.cyc_cap));
  int32_t *d_out = nullptr, *d_rows = nullptr;
  HIPCHK(dalloc(&d_out, std::max<int64_t>(outCap, 1)));
  SolverJob job = make_job(g, K, MODE_FULL_RATIO, flags | FLAG_FRESH_JOB);
  job.res = s.res;
  job.out_eid = d_out;
  job.out_cap = (int32_t)outCap;
  if (!pred_fits_lds(g, K)) HIPCHK(launch_fill_i32(s.pred, s.predN, -1, st));
  if ((rc = h2d_small(ctx, s.djob, &job, sizeof(job)))) return rc;
  HIPCHK(hipEventRecord(ctx->ev0, st));
  HIPCHK(launch_solver(s.djob, 1, 1, nullptr, kernel_maxc(g, K), g->maxw, false, g->d_time[0] != nullptr, false,
                       job_lds_bytes(g, K), st));
  HIPCHK(hipEventRecord(ctx->ev1, st));
  JobResult r;
  if ((rc = d2h_sync(ctx, &r, s.res, sizeof(r)))) return rc;
  float kms = 0;
  HIPCHK(hipEventElapsedTime(&kms, ctx->ev0, ctx->ev1));
  if (r.error) {
    cache_free(d_out);
    if (r.error == ERR_VALID_STORE || r.error == ERR_BRANCH_CAP) {
      redo = true;
      return PGC_OK;
    }
    if (r.error == ERR_CYCLE_TOO_LONG && r.out_len > outCap)
      return fail(PGC_ERR_CAPACITY, "output capacity too small: need " + std::to_string(r.out_len));
    return fail(PGC_ERR_SOLVER, "solver error " + std::to_string(r.error));
  }
  const int64_t m = r.out_len;
  if (rows && m) {
    HIPCHK(dalloc(&d_rows, m * g->ps_cols));
    HIPCHK(launch_gather_rows(g->d_rows, g->ps_cols, d_out, m, d_rows, st));
  }
  if ((rc = d2h_sync(ctx, edge_ids, d_out, edge_ids && m ? (size_t)m * 4 : 0, rows, d_rows,
                     rows && m ? (size_t)(m * g->ps_cols * 4) : 0)))
    return rc;
  cache_free(d_out);
  if (d_rows) cache_free(d_rows);
  std::memset(res, 0, sizeof(*res));
  res->passes_layers_once = 1;
  res->objective = (float)(r.out_cs / r.out_ts);
  res->n_edges = m;
  res->sweeps = r.sweeps;
  g->last_sweeps = r.sweeps;
  res->bf_calls = r.bf_calls;
  res->lawler_iters = r.lawler_iters;
  res->branches = r.branches;
  res->relaxations = relax_per_sweep(g, K) * (double)r.sweeps;
  res->kernel_ms = kms;
  res->total_ms = wall.ms();
  const int64_t ticks[5] = {r.t_sweep, r.t_fold, r.t_chase, r.t_walk, r.t_extract};
  for (int k = 0; k < 5; ++k) res->phase_ms[k] = (double)ticks[k] * 1e-5;  // 100 MHz wall clock
  res->launches = 1;
  res->implicit_costs = implicit_applies(g, K) ? 1 : 0;
  res->status = PGC_OK;
  return PGC_OK;
}

// Everything a ratio solve of g at depth K needs besides its energy: the
// in-edge layout, the slot cost buffers and the solver scratch. A graph that
// cannot be solved is refused here, before its energy is quantized in place.
int prepare_ratio(pgc_graph* g, int K) {
  int rc;
  if ((rc = check_solvable(g, K))) return rc;
  if ((rc = build_layout(g))) return rc;
  if ((rc = check_width(g))) return rc;
  if ((rc = cost_buffers(g, 1))) return rc;
  return ensure_scratch(g, K);
}

}  // namespace

extern "C" {

// ------------------------------------------------------------ Lawler + BnB
int pgc_solve_ratio(pgc_context* ctx, pgc_graph* g, int K, uint32_t flags, int32_t* edge_ids, int32_t* rows,
                    int64_t capacity, pgc_result* res) {
  Timer wall;
  if (!ctx || !g || !res) return fail(PGC_ERR_ARG, "invalid arguments");
  if (g->ctx != ctx) return fail(PGC_ERR_ARG, "graph belongs to another context");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  int rc = prepare_ratio(g, K);
  if (rc) return rc;
  // a fresh energy quantized once can be recomputed from feat_diff by the sweep
  struct Implicit {
    pgc_graph* g;
    ~Implicit() { g->implicit_now = false; }
  } implicit_scope{g};
  g->implicit_now = g->energy_fresh && g->d_fdt;
  g->energy_fresh = false;
  // robustFloatToIntConversion of every column, in place (:875-882). From
  // kGatherMin keys on, the median's later radix passes read the keys its
  // third pass gathers (those sharing the median's top 16 bits), each
  // workgroup from its own region: sized for every key of the column below
  // 8 M keys (no region can overflow), an eighth of it above (a region that
  // overflows sends the later passes back to the column). Below kGatherMin
  // a pass costs mostly its fixed overhead and every pass reads the column.
  constexpr int64_t kGatherMin = (int64_t)1 << 20;
  double* qc = nullptr;
  int64_t ccap = 0;
  if (g->E >= kGatherMin) {
    // 512 workgroups, each with at most one tile of 2048 keys beyond E / 512
    ccap = g->E >= ((int64_t)8 << 20) ? g->E / 8 + 1024 : g->E + 512 * 2048;
    if (ctx->qcand_cap < ccap) {
      if (ctx->d_qcand) {
        HIPCHK(hipStreamSynchronize(st));
        cache_free(ctx->d_qcand);
      }
      ctx->d_qcand = nullptr;
      ctx->qcand_cap = 0;
      HIPCHK(dalloc(&ctx->d_qcand, ccap));
      ctx->qcand_cap = ccap;
    }
    qc = ctx->d_qcand;
  }
  for (int c = 0; c < g->ncols; ++c)
    HIPCHK(launch_quantize(g->d_energy + (int64_t)c * g->E, g->E, ctx->select_state,
                           g->implicit_now ? g->d_qscale : nullptr, qc, ccap, st));
  if ((rc = fill_costs(g, 1))) return rc;
  // the device branch and bound wherever the one-workgroup solver runs the
  // job (the device-wide solver has no device branch and bound);
  // PGC_HOST_BNB=1 selects the host driver everywhere
  const char* hb = std::getenv("PGC_HOST_BNB");
  if (grid_blocks(g, K, false) == 0 && !(hb && hb[0] == '1')) {
    bool redo = false;
    if ((rc = ratio_bnb_device(ctx, g, K, flags, edge_ids, rows, capacity, res, wall, redo)) || !redo) return rc;
  }
  return ratio_bnb_host(ctx, g, K, flags, edge_ids, rows, capacity, res, wall);
}

// ------------------------------------------- allocation-free ratio solve
}  // extern "C"

namespace {

// Lends a graph layout, slot-cost and solver-scratch state that lives in the
// static scratch for one call, and gives the graph its own state back when it
// ends (nothing of the scratch is ever freed or kept by the graph).
struct ScratchView {
  pgc_graph* g;
  bool layout;
  int nChunks;
  int64_t nLayers, nslots[2], dropped, isv, xsv;
  int32_t *d_width, *d_grp;
  int32_t *d_s[2], *d_e[2], *d_cnt;
  int32_t width[2], lstride[2], src_inv[2], cls_deg[2], maxw, mid;
  void* d_cost[2];
  float* d_time[2];
  int cost_kind;
  size_t cost_bytes[2], time_bytes[2];
  JobScratch js;
  explicit ScratchView(pgc_graph* gr) : g(gr) {
    layout = g->layout;
    nChunks = g->nChunks;
    nLayers = g->nLayers;
    dropped = g->dropped;
    isv = g->intra_slots_valid;
    xsv = g->inter_slots_valid;
    d_width = g->d_width;
    d_grp = g->d_grp;
    d_cnt = g->d_cnt;
    maxw = g->maxw;
    mid = g->max_in_degree;
    cost_kind = g->cost_kind;
    js = g->js;
    for (int c = 0; c < 2; ++c) {
      nslots[c] = g->nslots[c];
      d_s[c] = g->d_s[c];
      d_e[c] = g->d_e[c];
      width[c] = g->width[c];
      lstride[c] = g->lstride[c];
      src_inv[c] = g->src_inv[c];
      cls_deg[c] = g->cls_deg[c];
      d_cost[c] = g->d_cost[c];
      d_time[c] = g->d_time[c];
      cost_bytes[c] = g->cost_bytes[c];
      time_bytes[c] = g->time_bytes[c];
      g->d_cost[c] = nullptr;
      g->d_time[c] = nullptr;
      g->cost_bytes[c] = g->time_bytes[c] = 0;
    }
    g->layout = false;
    g->d_grp = nullptr;
    g->js = JobScratch();
  }
  ~ScratchView() {
    g->layout = layout;
    g->nChunks = nChunks;
    g->nLayers = nLayers;
    g->dropped = dropped;
    g->intra_slots_valid = isv;
    g->inter_slots_valid = xsv;
    g->d_width = d_width;
    g->d_grp = d_grp;
    g->d_cnt = d_cnt;
    g->maxw = maxw;
    g->max_in_degree = mid;
    g->cost_kind = cost_kind;
    g->js = js;
    for (int c = 0; c < 2; ++c) {
      g->nslots[c] = nslots[c];
      g->d_s[c] = d_s[c];
      g->d_e[c] = d_e[c];
      g->width[c] = width[c];
      g->lstride[c] = lstride[c];
      g->src_inv[c] = src_inv[c];
      g->cls_deg[c] = cls_deg[c];
      g->d_cost[c] = d_cost[c];
      g->d_time[c] = d_time[c];
      g->cost_bytes[c] = cost_bytes[c];
      g->time_bytes[c] = time_bytes[c];
    }
  }
};

}  // namespace

extern "C" {

int pgc_solve_ratio_noalloc(pgc_context* ctx, pgc_graph* g, int K, uint32_t flags, int32_t* edge_ids, int32_t* rows,
                            int64_t capacity, pgc_result* res) {
  Timer wall;
  if (!ctx || !g || !res) return fail(PGC_ERR_ARG, "invalid arguments");
  if (g->ctx != ctx) return fail(PGC_ERR_ARG, "graph belongs to another context");
  HIPCHK(hipSetDevice(ctx->device));
  uint8_t* base = (uint8_t*)scratch_ptr(ctx->device);
  if (!base) return fail(PGC_ERR_HIP, "the static scratch of this device has no device address");
  // the scratch is shared by every call on this device: calls are serialized
  std::lock_guard<std::mutex> lock(scratch_call_lock(ctx->device));
  int rc = check_solvable(g, K);
  if (rc) return rc;
  hipStream_t st = ctx->stream;
  ScratchView view(g);
  DevAlloc A;
  A.base = base;
  A.cap = kScratchBytes;
  auto fit = [&](int code) {
    if (A.overflow)
      return fail(PGC_ERR_CAPACITY, "this solve needs more device memory than the " +
                                        std::to_string(kScratchBytes >> 20) + " MiB static scratch holds");
    return code;
  };
  const int64_t npl = g->npl;
  if ((rc = build_layout(g, A))) return fit(rc);
  if ((rc = check_width(g))) return rc;
  if ((rc = cost_buffers(g, 1, A))) return fit(rc);
  if ((rc = ensure_scratch(g, K, A, false))) return fit(rc);
  // one store slot per branch id below npl: every valid branch fits
  if ((rc = ensure_bnb(g, A, (int)npl))) return fit(rc);
  JobScratch& s = g->js;
  const int64_t outCap = std::max<int64_t>(0, std::min<int64_t>(capacity, s.cyc_cap));
  int32_t *d_out = nullptr, *d_rows = nullptr;
  void* sel = nullptr;
  uint8_t* selb = nullptr;
  if (A.get(&d_out, outCap) != hipSuccess) return fit(PGC_ERR_HIP);
  if (rows && A.get(&d_rows, outCap * g->ps_cols) != hipSuccess) return fit(PGC_ERR_HIP);
  if (A.get(&selb, select_state_bytes()) != hipSuccess) return fit(PGC_ERR_HIP);
  sel = selb;
  // robustFloatToIntConversion of every column, in place (:875-882), then
  // the slot costs
  g->energy_fresh = false;
  for (int c = 0; c < g->ncols; ++c)
    HIPCHK(launch_quantize(g->d_energy + (int64_t)c * g->E, g->E, sel, nullptr, nullptr, 0, st));
  for (int c = 0; c < 2; ++c)
    HIPCHK(launch_fill_costs(g->d_e[c], g->d_s[c], g->nslots[c], g->d_energy, g->E, g->ncols, false, g->d_cost[c],
                             g->ncols > 1 ? g->d_time[c] : nullptr, 0, 0, st));
  // the complete solve in one launch: upper bound, branch and bound, output
  SolverJob job = make_job(g, K, MODE_FULL_RATIO, flags | FLAG_FRESH_JOB);
  job.vc_cap = s.vc_cap;
  job.res = s.res;
  job.out_eid = d_out;
  job.out_cap = (int32_t)outCap;
  if (!pred_fits_lds(g, K)) HIPCHK(launch_fill_i32(s.pred, s.predN, -1, st));
  HIPCHK(hipMemcpyAsync(s.djob, &job, sizeof(job), hipMemcpyHostToDevice, st));
  HIPCHK(hipEventRecord(ctx->ev0, st));
  HIPCHK(launch_solver(s.djob, 1, 1, nullptr, kernel_maxc(g, K), g->maxw, false, g->d_time[0] != nullptr, false,
                       job_lds_bytes(g, K), st));
  HIPCHK(hipEventRecord(ctx->ev1, st));
  JobResult r;
  HIPCHK(hipMemcpyAsync(&r, s.res, sizeof(r), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  float kms = 0;
  HIPCHK(hipEventElapsedTime(&kms, ctx->ev0, ctx->ev1));
  if (r.error) {
    if (r.error == ERR_CYCLE_TOO_LONG && r.out_len > outCap)
      return fail(PGC_ERR_CAPACITY, "output capacity too small: need " + std::to_string(r.out_len));
    return fail(PGC_ERR_SOLVER, "solver error " + std::to_string(r.error));
  }
  const int64_t m = r.out_len;
  if (edge_ids && m) HIPCHK(hipMemcpyAsync(edge_ids, d_out, m * 4, hipMemcpyDeviceToHost, st));
  if (rows && m) {
    HIPCHK(launch_gather_rows(g->d_rows, g->ps_cols, d_out, m, d_rows, st));
    HIPCHK(hipMemcpyAsync(rows, d_rows, m * g->ps_cols * 4, hipMemcpyDeviceToHost, st));
  }
  HIPCHK(hipStreamSynchronize(st));
  std::memset(res, 0, sizeof(*res));
  res->passes_layers_once = 1;
  res->objective = (float)(r.out_cs / r.out_ts);
  res->n_edges = m;
  res->sweeps = r.sweeps;
  res->bf_calls = r.bf_calls;
  res->lawler_iters = r.lawler_iters;
  res->branches = r.branches;
  res->relaxations = relax_per_sweep(g, K) * (double)r.sweeps;
  res->kernel_ms = kms;
  res->total_ms = wall.ms();
  const int64_t ticks[5] = {r.t_sweep, r.t_fold, r.t_chase, r.t_walk, r.t_extract};
  for (int k = 0; k < 5; ++k) res->phase_ms[k] = (double)ticks[k] * 1e-5;
  res->launches = 1;
  res->status = PGC_OK;
  return PGC_OK;
}

// ------------------------------------------------- batched complete solves
extern "C++" {
template <typename T>
static int grow(T** p, int64_t& cap, int64_t need) {
  if (need <= cap && *p) return PGC_OK;
  if (*p) {
    (void)hipDeviceSynchronize();  // queued work may still use the arena
    cache_free(*p);
  }
  *p = nullptr;
  cap = 0;
  const int64_t n = std::max<int64_t>(need, 64) * 3 / 2;
  HIPCHK(dalloc(p, n));
  cap = n;
  return PGC_OK;
}
}  // extern "C++"

int pgc_graphs_set_feat_diff(pgc_graph* const* graphs, const void* const* feat_diff, int64_t n, int dtype) {
  if (n < 0 || (n > 0 && (!graphs || !feat_diff)) || (dtype != 0 && dtype != 1))
    return fail(PGC_ERR_ARG, "invalid arguments");
  if (n == 0) return PGC_OK;
  pgc_context* ctx = graphs[0] ? graphs[0]->ctx : nullptr;
  if (!ctx) return fail(PGC_ERR_ARG, "graph 0 is null");
  HIPCHK(hipSetDevice(ctx->device));
  std::vector<EItem> items((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    pgc_graph* g = graphs[i];
    if (!g || g->ctx != ctx || !feat_diff[i])
      return fail(PGC_ERR_ARG, "graph " + std::to_string(i) + ": null, of another context, or no feat_diff");
    const pgc_graph* b = g->base ? g->base : g;
    if (b->conjugate)
      return fail(PGC_ERR_UNSUPPORTED, "feat_diff updates are available for non-conjugate graphs only");
    if (!b->d_vx) return fail(PGC_ERR_UNSUPPORTED, "the graph was not built by pgc_graph_generate*: no energy inputs");
    items[i] = {b->d_vx, b->d_vy, b->d_rows, b->E, feat_diff[i], b->nVY, b->line_integral, b->length_mode, dtype,
                g->d_energy, b->d_ex, b->d_ey, b->nEX, b->nEY, b->nVX, b->prune};
  }
  int rc;
  if ((rc = grow(&ctx->d_eitems, ctx->eitems_cap, n))) return rc;
  hipStream_t st = ctx->stream;
  HIPCHK(hipMemcpyAsync(ctx->d_eitems, items.data(), n * sizeof(EItem), hipMemcpyHostToDevice, st));
  // the host table must outlive the copy (the energies themselves stay asynchronous)
  HIPCHK(hipStreamSynchronize(st));
  const int64_t chunk = 65535;
  for (int64_t a = 0; a < n; a += chunk)
    HIPCHK(launch_energy_batch(ctx->d_eitems + a, (int)std::min<int64_t>(chunk, n - a), st));
  // the energies now come from these feat_diffs: a kept feat_diff is stale
  for (int64_t i = 0; i < n; ++i) {
    graphs[i]->cost_kind = 0;
    graphs[i]->energy_fresh = false;
    if (graphs[i]->d_fdt) cache_free(graphs[i]->d_fdt);
    graphs[i]->d_fdt = nullptr;
  }
  return PGC_OK;
}

int pgc_solve_ratio_batch(pgc_context* ctx, pgc_graph* const* graphs, const int32_t* max_depth, int64_t n,
                          uint32_t flags, int32_t* edge_ids, int32_t* rows, const int64_t* offsets,
                          pgc_result* res) {
  Timer wall;
  if (!ctx || n < 0 || (n > 0 && (!graphs || !max_depth || !offsets || !res)))
    return fail(PGC_ERR_ARG, "invalid arguments");
  if (n == 0) return PGC_OK;
  if (n > (int64_t)1 << 30) return fail(PGC_ERR_ARG, "too many graphs in one batch");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  int rc;
  {
    // every graph once: two items would quantize one energy in place twice
    // and share one solver scratch
    std::unordered_set<const pgc_graph*> seen;
    seen.reserve((size_t)n * 2);
    for (int64_t i = 0; i < n; ++i) {
      pgc_graph* g = graphs[i];
      if (!g) return fail(PGC_ERR_ARG, "graph " + std::to_string(i) + " is null");
      if (g->ctx != ctx) return fail(PGC_ERR_ARG, "graph " + std::to_string(i) + " belongs to another context");
      if (offsets[i + 1] < offsets[i]) return fail(PGC_ERR_ARG, "offsets must be non-decreasing");
      if (rows && g->ps_cols != graphs[0]->ps_cols)
        return fail(PGC_ERR_ARG, "rows output needs one product-space width for the whole batch");
      if (!seen.insert(g).second)
        return fail(PGC_ERR_ARG, "graph " + std::to_string(i) + " appears twice in one batch");
    }
  }
  // per graph: layout, cost buffers, scratch (a graph that cannot be solved
  // fails the call here, before any energy is quantized); then the in-place
  // quantization of every energy column and the fp32 slot costs of every
  // graph in a few batched launches (each column as pgc_solve_ratio does it)
  for (int64_t i = 0; i < n; ++i) {
    if ((rc = prepare_ratio(graphs[i], max_depth[i])) || (rc = ensure_bnb(graphs[i])))
      return fail(rc, "graph " + std::to_string(i) + ": " + g_last_error);
  }
  std::vector<SolverJob> jobs(n);
  for (int64_t i = 0; i < n; ++i) {
    graphs[i]->energy_fresh = false;  // quantized in place below
    jobs[i] = make_job(graphs[i], max_depth[i], MODE_FULL_RATIO, flags);
    jobs[i].flags |= FLAG_FRESH_JOB;
  }
  // Heavy items first: when every graph was solved before and a few of them
  // ran far more sweeps (times their size) than the rest, those are prepared
  // and launched first, and the other items are prepared while they run (the
  // batch's makespan is its longest item; in repeated batches over the same
  // graphs, a solve's sweep count is the best predictor of the next one's).
  std::vector<char> heavy(n, 0);
  int64_t nHeavy = 0;
  {
    std::vector<double> est(n);
    double mx = 0;
    bool known = n >= 16;
    for (int64_t i = 0; i < n && known; ++i) {
      if (graphs[i]->last_sweeps <= 0) known = false;
      est[i] = (double)graphs[i]->last_sweeps * (double)graphs[i]->E * max_depth[i];
      mx = std::max(mx, est[i]);
    }
    if (known)
      for (int64_t i = 0; i < n; ++i)
        if (est[i] * 8 >= mx) {
          heavy[i] = 1;
          nHeavy++;
        }
    if (nHeavy * 16 > n) {
      std::fill(heavy.begin(), heavy.end(), 0);
      nHeavy = 0;
    }
  }
  // batch tables: results and output cycles in one arena each
  if ((rc = grow(&ctx->d_jobs, ctx->jobs_cap, n))) return rc;
  if ((rc = grow(&ctx->d_res, ctx->res_cap, n))) return rc;
  const int64_t outN = std::max<int64_t>(offsets[n] - offsets[0], 1);
  if ((rc = grow(&ctx->d_out, ctx->out_cap, outN))) return rc;
  for (int64_t i = 0; i < n; ++i) {
    jobs[i].res = ctx->d_res + i;
    jobs[i].out_eid = ctx->d_out + (offsets[i] - offsets[0]);
    jobs[i].out_cap = (int32_t)std::min<int64_t>(offsets[i + 1] - offsets[i], INT32_MAX);
  }
  // preparation scratch, sized for the whole batch (the phases use it in turn
  // on the context's stream)
  const int64_t chunk = 65535;  // columns per launch (the launch grid's y limit)
  {
    int64_t nq = 0, candN = 0;
    for (int64_t i = 0; i < n; ++i) {
      nq += graphs[i]->ncols;
      for (int c = 0; c < graphs[i]->ncols; ++c) candN += median_cand_cap(graphs[i]->E);
    }
    if ((rc = grow(&ctx->d_qitems, ctx->qitems_cap, nq))) return rc;
    if ((rc = grow(&ctx->d_fitems, ctx->fitems_cap, 2 * n))) return rc;
    if ((rc = grow(&ctx->d_fills, ctx->fills_cap, n))) return rc;
    const int64_t nsel = std::min<int64_t>(nq, chunk);
    if ((rc = grow(&ctx->d_select, ctx->select_cap, nsel * (int64_t)bselect_state_bytes()))) return rc;
    if ((rc = grow(&ctx->d_bracket, ctx->bracket_cap, nsel * (int64_t)bracket_state_bytes()))) return rc;
    // exact size (no growth margin): it is 1/8 of the batch's energies
    if (candN > ctx->cand_cap || !ctx->d_cand) {
      if (ctx->d_cand) {
        (void)hipDeviceSynchronize();  // queued work may still use the buffer
        cache_free(ctx->d_cand);
        ctx->d_cand = nullptr;
        ctx->cand_cap = 0;
      }
      HIPCHK(dalloc(&ctx->d_cand, std::max<int64_t>(candN, 1)));
      ctx->cand_cap = std::max<int64_t>(candN, 1);
    }
  }
  for (int k = 0; k < 4; ++k) {
    if (!ctx->aux[k]) HIPCHK(hipStreamCreateWithFlags(&ctx->aux[k], hipStreamNonBlocking));
    if (!ctx->join[k]) HIPCHK(hipEventCreateWithFlags(&ctx->join[k], hipEventDisableTiming));
  }
  for (hipEvent_t* e : {&ctx->fork, &ctx->fork2})
    if (!*e) HIPCHK(hipEventCreateWithFlags(e, hipEventDisableTiming));
  HIPCHK(hipMemsetAsync(ctx->d_queue, 0, 64 * sizeof(int32_t), st));
  // launch groups: jobs sharing one kernel instantiation run in one
  // persistent launch, ordered largest first inside the group
  struct Key {
    int maxc, maxw;
    bool time;
    bool operator<(const Key& o) const {
      return std::tie(maxc, maxw, time) < std::tie(o.maxc, o.maxw, o.time);
    }
  };
  int nLaunch = 0;
  bool usedAux[4] = {false, false, false, false};
  bool timing = false;
  int64_t jobBase = 0;
  // one phase: the quantization and slot costs of its items, then its
  // launch groups on the auxiliary streams (the heavy phase on stream 3, the
  // other on streams 0-2, so that no launch queues behind a heavy one)
  auto phase = [&](const std::vector<int64_t>& items, int firstAux, int nAux) -> int {
    if (items.empty()) return PGC_OK;
    std::vector<QItem> qitems;
    std::vector<FItem> fitems;
    std::vector<FillItem> fills;
    int64_t candN = 0;
    for (int64_t i : items) {
      pgc_graph* g = graphs[i];
      for (int c = 0; c < g->ncols; ++c) {
        const int64_t cap = median_cand_cap(g->E);
        qitems.push_back({g->d_energy + (int64_t)c * g->E, g->E, candN, cap});
        candN += cap;
      }
      for (int c = 0; c < 2; ++c)
        fitems.push_back({g->d_e[c], g->d_s[c], g->nslots[c], g->d_energy, g->E, (float*)g->d_cost[c],
                          g->ncols > 1 ? g->d_time[c] : nullptr});
      // a batch job is one launch: the kernel starts it with an empty cycle
      // and, for a predecessor table in LDS, with -1 predecessors (no global
      // table to reset, load or write back); global tables are reset by one
      // fill launch per phase
      if (!pred_fits_lds(g, max_depth[i])) fills.push_back({g->js.pred, g->js.predN});
    }
    HIPCHK(hipMemcpyAsync(ctx->d_qitems, qitems.data(), qitems.size() * sizeof(QItem), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(ctx->d_fitems, fitems.data(), fitems.size() * sizeof(FItem), hipMemcpyHostToDevice, st));
    if (!fills.empty()) {
      HIPCHK(hipMemcpyAsync(ctx->d_fills, fills.data(), fills.size() * sizeof(FillItem), hipMemcpyHostToDevice, st));
      for (int64_t a = 0; a < (int64_t)fills.size(); a += chunk)
        HIPCHK(launch_fill_i32_batch(ctx->d_fills + a, (int)std::min<int64_t>(chunk, (int64_t)fills.size() - a), -1,
                                     st));
    }
    for (int64_t a = 0; a < (int64_t)qitems.size(); a += chunk)
      HIPCHK(launch_quantize_batch(ctx->d_qitems + a, (int)std::min<int64_t>(chunk, (int64_t)qitems.size() - a),
                                   ctx->d_select, ctx->d_bracket, ctx->d_cand, st));
    for (int64_t a = 0; a < (int64_t)fitems.size(); a += chunk)
      HIPCHK(launch_fill_costs_batch(ctx->d_fitems + a, (int)std::min<int64_t>(chunk, (int64_t)fitems.size() - a),
                                     st));
    std::map<Key, std::vector<int64_t>> groups;
    for (int64_t i : items) {
      pgc_graph* g = graphs[i];
      const int mc = kernel_maxc(g, max_depth[i]);
      const int mw = g->maxw <= 8 ? 8 : (g->maxw <= 16 ? 16 : (g->maxw <= 32 ? 32 : kMaxSlotWidth));
      groups[{mc, mw, g->d_time[0] != nullptr}].push_back(i);
    }
    if (nLaunch + (int)groups.size() > 64) return fail(PGC_ERR_UNSUPPORTED, "too many launch groups");
    std::vector<SolverJob> ordered;
    ordered.reserve(items.size());
    struct Launch {
      Key k;
      int64_t first, count;
      size_t lds;
    };
    std::vector<Launch> launches;
    for (auto& kv : groups) {
      auto& idx = kv.second;
      std::stable_sort(idx.begin(), idx.end(), [&](int64_t x, int64_t y) {
        return graphs[x]->E * max_depth[x] > graphs[y]->E * max_depth[y];
      });
      size_t lds = 0;
      for (int64_t i : idx) lds = std::max(lds, job_lds_bytes(graphs[i], max_depth[i]));
      launches.push_back({kv.first, jobBase + (int64_t)ordered.size(), (int64_t)idx.size(), lds});
      for (int64_t i : idx) ordered.push_back(jobs[i]);
    }
    HIPCHK(hipMemcpyAsync(ctx->d_jobs + jobBase, ordered.data(), ordered.size() * sizeof(SolverJob),
                          hipMemcpyHostToDevice, st));
    jobBase += (int64_t)ordered.size();
    hipEvent_t fork = firstAux == 3 ? ctx->fork : ctx->fork2;
    if (!timing) {
      HIPCHK(hipEventRecord(ctx->ev0, st));
      timing = true;
    }
    HIPCHK(hipEventRecord(fork, st));
    for (int k = firstAux; k < firstAux + nAux; ++k) HIPCHK(hipStreamWaitEvent(ctx->aux[k], fork, 0));
    // the groups run concurrently: a group's persistent workgroups hold CUs
    // only while it has jobs, and the CUs a group leaves (while its longest
    // jobs finish) take the other groups' workgroups
    for (size_t li = 0; li < launches.size(); ++li) {
      const Launch& L = launches[li];
      const int k = firstAux + (int)(li % nAux);
      usedAux[k] = true;
      const int grid = (int)std::min<int64_t>(L.count, std::max(1, ctx->num_cu));
      HIPCHK(launch_solver(ctx->d_jobs + L.first, (int)L.count, grid, ctx->d_queue + nLaunch, L.k.maxc, L.k.maxw,
                           false, L.k.time, false, L.lds, ctx->aux[k]));
      nLaunch++;
    }
    return PGC_OK;
  };
  {
    std::vector<int64_t> first, rest;
    for (int64_t i = 0; i < n; ++i) (heavy[i] ? first : rest).push_back(i);
    if ((rc = phase(first, 3, 1))) return rc;
    if ((rc = phase(rest, 0, nHeavy > 0 ? 3 : 4))) return rc;
  }
  for (int k = 0; k < 4; ++k) {
    if (!usedAux[k]) continue;
    HIPCHK(hipEventRecord(ctx->join[k], ctx->aux[k]));
    HIPCHK(hipStreamWaitEvent(st, ctx->join[k], 0));
  }
  HIPCHK(hipEventRecord(ctx->ev1, st));
  std::vector<JobResult> out(n);
  HIPCHK(hipMemcpyAsync(out.data(), ctx->d_res, n * sizeof(JobResult), hipMemcpyDeviceToHost, st));
  if (edge_ids)
    HIPCHK(hipMemcpyAsync(edge_ids + offsets[0], ctx->d_out, (offsets[n] - offsets[0]) * 4, hipMemcpyDeviceToHost,
                          st));
  HIPCHK(hipStreamSynchronize(st));
  float kms = 0;
  HIPCHK(hipEventElapsedTime(&kms, ctx->ev0, ctx->ev1));
  // per-item outcome: solved on the device; solved again by the host branch
  // and bound when the device's valid-branch store or branch cap ran out
  // (same quantized costs, same cycle as pgc_solve_ratio); or failed with its
  // own status
  std::vector<int> status(n, PGC_OK);
  std::vector<char> redo(n, 0);
  std::vector<std::string> why(n);
  for (int64_t i = 0; i < n; ++i) {
    const JobResult& r = out[i];
    if (!r.error) continue;
    if (r.error == ERR_VALID_STORE || r.error == ERR_BRANCH_CAP) {
      redo[i] = 1;
      continue;
    }
    status[i] = r.error == ERR_CYCLE_TOO_LONG && r.out_len > jobs[i].out_cap ? PGC_ERR_CAPACITY : PGC_ERR_SOLVER;
    why[i] = "graph " + std::to_string(i) + ": solver error " + std::to_string(r.error) +
             (r.error == ERR_CYCLE_TOO_LONG ? " (cycle of " + std::to_string(r.out_len) + " edges, capacity " +
                                                  std::to_string(jobs[i].out_cap) + ")"
                                            : "");
  }
  const int64_t pc = graphs[0]->ps_cols;
  if (rows) {
    int64_t tot = 0;
    for (int64_t i = 0; i < n; ++i)
      if (status[i] == PGC_OK && !redo[i]) tot += out[i].out_len * (int64_t)graphs[i]->ps_cols;
    if ((rc = grow(&ctx->d_rows, ctx->rows_cap, std::max<int64_t>(tot, 1)))) return rc;
    int64_t at = 0;
    std::vector<int64_t> rowAt(n, 0);
    for (int64_t i = 0; i < n; ++i) {
      if (status[i] != PGC_OK || redo[i]) continue;
      rowAt[i] = at;
      if (out[i].out_len)
        HIPCHK(launch_gather_rows(graphs[i]->d_rows, graphs[i]->ps_cols, jobs[i].out_eid, out[i].out_len,
                                  ctx->d_rows + at, st));
      at += out[i].out_len * (int64_t)graphs[i]->ps_cols;
    }
    std::vector<int32_t> h(std::max<int64_t>(tot, 1));
    HIPCHK(hipMemcpyAsync(h.data(), ctx->d_rows, tot * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    // rows of graph i at rows + (offsets[i] - offsets[0]) * ps_cols
    for (int64_t i = 0; i < n; ++i)
      if (status[i] == PGC_OK && !redo[i])
        std::memcpy(rows + (offsets[i] - offsets[0]) * pc, h.data() + rowAt[i],
                    out[i].out_len * (int64_t)graphs[i]->ps_cols * 4);
  }
  std::map<int64_t, pgc_result> host_res;
  for (int64_t i = 0; i < n; ++i) {
    if (!redo[i]) continue;
    pgc_result one;
    std::memset(&one, 0, sizeof(one));
    Timer t1;
    rc = ratio_bnb_host(ctx, graphs[i], max_depth[i], flags, edge_ids ? edge_ids + offsets[i] : nullptr,
                        rows ? rows + (offsets[i] - offsets[0]) * pc : nullptr, offsets[i + 1] - offsets[i], &one,
                        t1);
    if (rc) {
      status[i] = rc;
      why[i] = "graph " + std::to_string(i) + ": " + g_last_error;
      continue;
    }
    host_res[i] = one;
  }
  const double total = wall.ms();
  int64_t first_bad = -1;
  for (int64_t i = 0; i < n; ++i) {
    const JobResult& r = out[i];
    pgc_result& o = res[i];
    std::memset(&o, 0, sizeof(o));
    o.status = status[i];
    if (status[i] != PGC_OK) {
      if (first_bad < 0) first_bad = i;
      continue;
    }
    if (redo[i]) {
      o = host_res[i];
      o.kernel_ms += kms;
      o.total_ms = total;
      continue;
    }
    o.passes_layers_once = 1;
    o.objective = (float)(r.out_cs / r.out_ts);
    o.n_edges = r.out_len;
    o.sweeps = r.sweeps;
    graphs[i]->last_sweeps = r.sweeps;
    o.bf_calls = r.bf_calls;
    o.lawler_iters = r.lawler_iters;
    o.branches = r.branches;
    o.relaxations = relax_per_sweep(graphs[i], max_depth[i]) * (double)r.sweeps;
    o.kernel_ms = kms;
    o.total_ms = total;
    const int64_t ticks[5] = {r.t_sweep, r.t_fold, r.t_chase, r.t_walk, r.t_extract};
    for (int k = 0; k < 5; ++k) o.phase_ms[k] = (double)ticks[k] * 1e-5;
  }
  if (first_bad >= 0) return fail(status[first_bad], why[first_bad]);
  return PGC_OK;
}

// ------------------------------------- raw-unit cycle ratio and its gradient
static int cycle_items(pgc_context* ctx, pgc_graph* const* graphs, int64_t n, const double* const* fd,
                       double* const* dfd, const int32_t* edge_ids, const int64_t* offsets, const int64_t* lengths,
                       const double* grad_out, std::vector<CycleItem>& items) {
  if (!ctx || n < 0 || (n > 0 && (!graphs || !edge_ids || !offsets || !lengths)))
    return fail(PGC_ERR_ARG, "invalid arguments");
  if (n > INT32_MAX) return fail(PGC_ERR_ARG, "too many cycles in one call");
  int64_t tot = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (!graphs[i] || graphs[i]->ctx != ctx) return fail(PGC_ERR_ARG, "graph " + std::to_string(i) + " is invalid");
    const pgc_graph* b = graphs[i]->base ? graphs[i]->base : graphs[i];
    if (!b->d_vx) return fail(PGC_ERR_UNSUPPORTED, "graph " + std::to_string(i) + " has no energy inputs");
    if (b->ps_cols != 4) return fail(PGC_ERR_UNSUPPORTED, "conjugate graphs have no raw-unit gradient here");
    if (lengths[i] < 0 || lengths[i] > INT32_MAX) return fail(PGC_ERR_ARG, "invalid cycle length");
    tot += lengths[i];
  }
  int rc;
  if ((rc = grow(&ctx->d_items, ctx->items_cap, std::max<int64_t>(n, 1)))) return rc;
  if ((rc = grow(&ctx->d_ceids, ctx->ceids_cap, std::max<int64_t>(tot, 1)))) return rc;
  std::vector<int32_t> cat(std::max<int64_t>(tot, 1));
  items.resize(n);
  int64_t at = 0;
  for (int64_t i = 0; i < n; ++i) {
    const pgc_graph* g = graphs[i];
    const pgc_graph* b = g->base ? g->base : g;
    for (int64_t k = 0; k < lengths[i]; ++k) {
      const int32_t e = edge_ids[offsets[i] + k];
      if (e < 0 || e >= g->E) return fail(PGC_ERR_ARG, "edge id outside the graph");
      cat[at + k] = e;
    }
    CycleItem& it = items[i];
    std::memset(&it, 0, sizeof(it));
    it.rows = b->d_rows;
    it.vx = b->d_vx;
    it.vy = b->d_vy;
    it.fd = fd ? fd[i] : nullptr;
    it.dfd = dfd ? dfd[i] : nullptr;
    it.eids = ctx->d_ceids + at;
    it.nVY = b->nVY;
    it.m = (int32_t)lengths[i];
    it.line = b->line_integral;
    it.length = b->length_mode;
    it.gout = grad_out ? grad_out[i] : 0.0;
    at += lengths[i];
  }
  hipStream_t st = ctx->stream;
  HIPCHK(hipMemcpyAsync(ctx->d_ceids, cat.data(), std::max<int64_t>(tot, 1) * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(ctx->d_items, items.data(), n * sizeof(CycleItem), hipMemcpyHostToDevice, st));
  return PGC_OK;
}

int pgc_cycle_ratio_values(pgc_context* ctx, pgc_graph* const* graphs, int64_t n, const double* const* feat_diff,
                           const int32_t* edge_ids, const int64_t* offsets, const int64_t* lengths, double* values) {
  if (!feat_diff || (n > 0 && !values)) return fail(PGC_ERR_ARG, "invalid arguments");
  if (n == 0) return PGC_OK;
  HIPCHK(hipSetDevice(ctx ? ctx->device : 0));
  std::vector<CycleItem> items;
  int rc = cycle_items(ctx, graphs, n, feat_diff, nullptr, edge_ids, offsets, lengths, nullptr, items);
  if (rc) return rc;
  if ((rc = grow(&ctx->d_values, ctx->values_cap, n))) return rc;
  HIPCHK(launch_cycle_value(ctx->d_items, (int)n, ctx->d_values, ctx->stream));
  HIPCHK(hipMemcpyAsync(values, ctx->d_values, n * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return PGC_OK;
}

int pgc_cycle_ratio_grad(pgc_context* ctx, pgc_graph* const* graphs, int64_t n, const int32_t* edge_ids,
                         const int64_t* offsets, const int64_t* lengths, const double* grad_out,
                         double* const* dfeat_diff) {
  if (!dfeat_diff || (n > 0 && !grad_out)) return fail(PGC_ERR_ARG, "invalid arguments");
  if (n == 0) return PGC_OK;
  HIPCHK(hipSetDevice(ctx ? ctx->device : 0));
  std::vector<CycleItem> items;
  int rc = cycle_items(ctx, graphs, n, nullptr, dfeat_diff, edge_ids, offsets, lengths, grad_out, items);
  if (rc) return rc;
  HIPCHK(launch_cycle_grad(ctx->d_items, (int)n, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return PGC_OK;
}

// ---------------------------------------------------------- min-cost + BnB
int pgc_solve_min_cost(pgc_context* ctx, pgc_graph* g, int K, uint32_t flags, int32_t* edge_ids, int32_t* rows,
                       int64_t capacity, pgc_result* res) {
  Timer wall;
  for (double& x : g_phase_ms) x = 0;
  g_launches = 0;
  if (!ctx || !g || !res) return fail(PGC_ERR_ARG, "invalid arguments");
  if (g->ctx != ctx) return fail(PGC_ERR_ARG, "graph belongs to another context");
  int rc = check_solvable(g, K);
  if (rc) return rc;
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int64_t npl = g->npl;
  const bool cpuRules = (flags & PGC_MINCOST_CPU_RULES) != 0;
  if ((rc = build_layout(g))) return rc;
  if ((rc = check_width(g))) return rc;
  if ((rc = fill_costs(g, 2))) return rc;
  if ((rc = ensure_scratch(g, K))) return rc;
  JobScratch& s = g->js;
  HIPCHK(hipMemsetAsync(s.pred, 0xff, s.predN * 4, st));
  // branch graph: the given one, else the in-edges of layer-0 nodes (ascending
  // edge id), and the SRC ids of its values (the BFS reads SRCIds(value)).
  // Only branch() reads them, so they are built at the first split.
  std::vector<int32_t> bgSrc;
  bool bgReady = false;
  auto ensure_bg = [&]() -> int {
    if (bgReady) return PGC_OK;
    if (!g->has_bg && g->conjugate && !g->cbg_off.empty()) build_conj_bg(g);
    if (!g->has_bg) {
      const int64_t n0 = g->lstride[0];  // layer-0 block of the intralayer class
      std::vector<int32_t> sl(n0), el(n0);
      if (n0) {
        HIPCHK(hipMemcpy(sl.data(), g->d_s[0], n0 * 4, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(el.data(), g->d_e[0], n0 * 4, hipMemcpyDeviceToHost));
      }
      g->bg_off.assign(npl + 1, 0);
      g->bg_val.clear();
      for (int64_t v = 0; v < npl; ++v) {
        const int c = (int)(v >> 6), lane = (int)(v & 63);
        const int w = g->width[0];
        std::vector<int64_t> ids;
        for (int j = 0; j < w; ++j) {
          const int64_t idx = ((int64_t)c * w + j) * 64 + lane;
          if (sl[idx] < 0 || sl[idx] >= npl) break;
          ids.push_back(el[idx]);
        }
        std::sort(ids.begin(), ids.end());
        g->bg_val.insert(g->bg_val.end(), ids.begin(), ids.end());
        g->bg_off[v + 1] = (int64_t)g->bg_val.size();
      }
      g->has_bg = true;
      g->bg_val.push_back(0);
      g->bg_val.pop_back();
    }
    bgSrc.assign(g->bg_val.size(), -1);
    if (!g->bg_val.empty()) {
      std::vector<int32_t> ids(g->bg_val.size());
      for (size_t i = 0; i < ids.size(); ++i)
        ids[i] = (g->bg_val[i] >= 0 && g->bg_val[i] < g->E) ? (int32_t)g->bg_val[i] : 0;
      int32_t *d_ids = nullptr, *d_out = nullptr;
      HIPCHK(dalloc(&d_ids, ids.size()));
      HIPCHK(dalloc(&d_out, ids.size()));
      HIPCHK(hipMemcpyAsync(d_ids, ids.data(), ids.size() * 4, hipMemcpyHostToDevice, st));
      HIPCHK(launch_gather_rows(g->d_src, 1, d_ids, (int64_t)ids.size(), d_out, st));
      HIPCHK(hipMemcpyAsync(bgSrc.data(), d_out, ids.size() * 4, hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      cache_free(d_ids);
      cache_free(d_out);
      for (size_t i = 0; i < ids.size(); ++i)
        if (!(g->bg_val[i] >= 0 && g->bg_val[i] < g->E)) bgSrc[i] = -1;
    }
    bgReady = true;
    return PGC_OK;
  };

  double kms = 0;
  int64_t sweeps = 0, bfc = 0, branches = 0;
  const double INF = std::numeric_limits<double>::infinity();
  double upperBound = INF;
  std::vector<int32_t> branch(npl, 0);
  BnbHeap heap(npl);
  heap.push(0.0, 0);
  int maxBranchId = 0;
  std::vector<char> valid(npl, 0);
  std::map<int, std::vector<int32_t>> validCycles;
  std::vector<int32_t> cycle;  // best path, layer L-1 ... source (original ids)
  std::vector<double> keys(npl);
  std::vector<int32_t> back(npl), clen(npl);
  std::vector<std::pair<double, int64_t>> lbs;
  SolverJob job = make_job(g, K, MODE_MIN_COST, flags);
  bool branchDirty = true;
  while (!heap.empty()) {
    const auto top = heap.pop();
    const int bid = (int)top.second;
    if (valid[bid]) {
      upperBound = top.first;
      cycle = validCycles.at(bid);
      break;
    }
    branches++;
    if (branchDirty) {
      HIPCHK(hipMemcpyAsync(s.branch, branch.data(), npl * 4, hipMemcpyHostToDevice, st));
      branchDirty = false;
    }
    job.branchId = bid;
    job.upper_bound = upperBound;
    JobResult r;
    if ((rc = run_job(g, job, true, r, kms))) return rc;
    if (r.error) return fail(PGC_ERR_SOLVER, "min-cost sweep failed (error " + std::to_string(r.error) + ")");
    sweeps += r.sweeps;
    bfc += r.bf_calls;
    if (r.status != ST_CAN_CONTAIN) continue;
    if ((rc = d2h_many(ctx, {{keys.data(), s.lb_key, (size_t)npl * 8},
                             {back.data(), s.back, (size_t)npl * 4},
                             {clen.data(), s.wlen, (size_t)npl * 4}})))
      return rc;
    // lower bounds in ascending (key, vertex) order (dijkstra_bnb_subroutine.cu:391-393)
    lbs.clear();
    bool anyFinite = false;
    for (int64_t v = 0; v < npl; ++v) {
      if (branch[v] != bid) continue;
      lbs.push_back({keys[v], v});
      if (!std::isinf(keys[v])) anyFinite = true;
    }
    if (!anyFinite) continue;
    std::sort(lbs.begin(), lbs.end());
    bool first = true, foundFirst = false;
    double lbCur = 0;
    int bSrc = -1, bTrgt = -1;
    bool pruned = false;
    for (auto& lb : lbs) {
      const float lowerBound = (float)lb.first;
      if (std::isinf(lowerBound)) {
        first = false;
        continue;
      }
      const int64_t v = lb.second;
      if (first) lbCur = lowerBound;
      if (v == back[v]) {
        if (first) foundFirst = true;
        const float newUB = lowerBound;
        if (newUB > upperBound && first) {
          pruned = true;
          break;
        }
        if (newUB < upperBound) {
          upperBound = newUB;
          cycle.resize(clen[v]);
          if (clen[v]) {
            // column v of the chain table (stride npl), gathered on the device:
            // a strided copy of 4-byte rows costs the copy engine ~2 us per row
            HIPCHK(launch_gather_strided(s.chain + v, npl, clen[v], s.nodes_io, st));
            if ((rc = d2h_sync(ctx, cycle.data(), s.nodes_io, (size_t)clen[v] * 4))) return rc;
          }
        }
        break;
      } else if (first) {
        if (cpuRules) {
          int32_t front = -1;
          if (clen[v] > 0) HIPCHK(hipMemcpy(&front, s.chain + v, 4, hipMemcpyDeviceToHost));
          bSrc = clen[v] > 0 ? (int)(front - g->numNodes) : -1;
        } else {
          bSrc = (int)v;
        }
        bTrgt = back[v];
      }
      first = false;
    }
    if (pruned) continue;
    if (upperBound < heap.peak() && foundFirst) break;
    if (foundFirst) {
      heap.push(lbCur, bid);
      valid[bid] = 1;
      validCycles.emplace(bid, cycle);
      continue;
    }
    // branch(): split by a multi-source unit-weight BFS (dijkstra_solver_cuda.cpp:30-117)
    if ((rc = ensure_bg())) return rc;
    maxBranchId++;
    {
      BnbHeap bh(npl);
      std::vector<double> tbl(npl, -1);
      for (int64_t i = 0; i < npl; ++i) {
        tbl[i] = -1;
        if (i == bSrc || i == bTrgt) {
          tbl[i] = 1;
          continue;
        }
        if (branch[i] == bid) {
          bh.push(INF, i);
          tbl[i] = -1;
        }
      }
      bh.push(0.0, bSrc);
      bh.push(0.0, bTrgt);
      if (bTrgt >= 0 && bTrgt < npl) branch[bTrgt] = maxBranchId;
      int counter = 0;
      int numInBranch = r.num_in_branch;
      for (auto& lb : lbs) {
        const int64_t v = lb.second;
        if (v == bSrc || v == bTrgt) continue;
        if (branch[v] != bid) continue;
        if (counter > std::max(1, numInBranch / 4)) {
          bh.push(0.1, v);
          branch[v] = maxBranchId;
        }
        counter++;
      }
      while (!bh.empty()) {
        const auto cur = bh.pop();
        const int64_t node = cur.second;
        const float curCost = (float)cur.first;
        tbl[node] = 1.0;
        for (int64_t p = g->bg_off[node]; p < g->bg_off[node + 1]; ++p) {
          const int64_t e = g->bg_val[p];
          if (e == -1) continue;
          const int32_t sId = bgSrc[p];
          if (sId < 0 || sId >= npl) continue;
          if (tbl[sId] < 0) {
            const float oldV = (float)bh.key_value(sId);
            const float newV = curCost + 1.0;
            if (newV < oldV) {
              bh.decrease(sId, newV);
              branch[sId] = branch[node];
            }
          }
        }
      }
    }
    branchDirty = true;
    heap.push(lbCur, bid);
    heap.push(lbCur, maxBranchId);
  }
  // path output (dijkstra_solver_cuda.cpp:261-305)
  const int64_t cycleLength = (int64_t)cycle.size();
  std::vector<int32_t> cyc(cycle.rbegin(), cycle.rend());
  if (!cyc.empty()) cyc.push_back(cyc.front() + (int32_t)g->numNodes);
  std::vector<int32_t> eids(cyc.size() > 1 ? cyc.size() - 1 : 0, -2);
  int loops = 0;
  if (cyc.size() > 1) {
    if ((int64_t)cyc.size() > s.cyc_cap) return fail(PGC_ERR_CAPACITY, "path longer than the cycle buffer");
    SolverJob lj = make_job(g, K, MODE_MIN_COST, flags);
    if ((rc = h2d_many(ctx, {{s.djob, &lj, sizeof(lj)}, {s.nodes_io, cyc.data(), cyc.size() * 4}}))) return rc;
    HIPCHK(launch_lookup(s.djob, s.nodes_io, (int)cyc.size(), s.eids_io, false, true, st));
    if ((rc = d2h_sync(ctx, eids.data(), s.eids_io, eids.size() * 4))) return rc;
  }
  std::vector<int32_t> outE;
  const int64_t loopmax = cycleLength + 1;
  for (int64_t i = 0; i < loopmax && !cyc.empty(); ++i) {
    const int64_t a = cyc[i % loopmax], b = cyc[(i + 1) % loopmax];
    if (b >= g->numNodes) loops++;
    if (b == a - g->numNodes) continue;  // closing step through the dummy sentinel
    if (i + 1 >= (int64_t)cyc.size()) continue;
    const int32_t e = eids[i];
    if (e < 0) continue;
    outE.push_back(e);
  }
  const int64_t n = (int64_t)outE.size();
  if (n > capacity) return fail(PGC_ERR_CAPACITY, "output capacity too small: need " + std::to_string(n));
  if (edge_ids) std::copy(outE.begin(), outE.end(), edge_ids);
  if (rows && n) {
    if ((rc = h2d_small(ctx, s.eids_io, outE.data(), (size_t)n * 4))) return rc;
    int32_t* d_out = nullptr;
    HIPCHK(dalloc(&d_out, n * g->ps_cols));
    HIPCHK(launch_gather_rows(g->d_rows, g->ps_cols, s.eids_io, n, d_out, st));
    if ((rc = d2h_sync(ctx, rows, d_out, (size_t)(n * g->ps_cols * 4)))) return rc;
    cache_free(d_out);
  }
  res->passes_layers_once = loops < 2 ? 1 : 0;
  res->objective = (float)upperBound;
  res->n_edges = n;
  res->sweeps = sweeps;
  res->bf_calls = bfc;
  res->lawler_iters = 0;
  res->branches = branches;
  res->relaxations = relax_per_sweep(g, K) * (double)sweeps;
  res->kernel_ms = kms;
  res->total_ms = wall.ms();
  for (int i = 0; i < 5; ++i) res->phase_ms[i] = g_phase_ms[i];
  res->launches = g_launches;
  res->status = PGC_OK;
  return PGC_OK;
}

}  // extern "C"
