// Layer-synchronous Bellman-Ford / Lawler solver kernels for gfx950.
//
// One workgroup (1024 threads, 16 wave64s) owns one job (one product graph,
// one branch) for the whole Lawler subroutine: every sweep, fold, negative
// cycle walk, cycle extraction and lambda update happens on the device; the
// host only runs the branch-and-bound bookkeeping between launches.
//
// Semantics follow the reference's CPU path so that results are identical:
//   mooreBellmanFordSubroutineCPU  cost_time_ratio_solver.cpp:251-542
//   lawlerSoubroutine              cost_time_ratio_solver.cpp:558-635
//   findInitalUpperBound           cost_time_ratio_solver.cpp:652-707
//   computeRatioForCycle           cost_time_ratio_solver.cpp:88-133
//   dijkstra BF + lower bounds     dijkstra_solver.cpp:203-387, dijkstra_bnb_subroutine.cu:366-393
// The distance table keeps the reference's row layout ((2K) rolling rows:
// even layers in rows [0,K), odd layers in rows [K,2K)) in LDS, so that the
// reference's fold row (numLayers % 2) and the stale layer-0 depth rows behave
// identically; the zero-layer backup row lives in registers of the vertex
// owners. The per-source minimum over depths of the previous layer is kept
// in an extra LDS row so that an interlayer relaxation reads one value; the
// depth of the winning source is then recovered exactly (first depth whose
// rounded sum equals the minimum, the reference's loop order).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "bnb_device.h"
#include "fastdiv.h"
#include "pgc_internal.h"

namespace pgc {

#define PGC_INF __builtin_huge_val()

// The job's arrays are reached through pointers stored in device memory, so
// the compiler cannot infer their address space and would emit FLAT
// accesses, which also count against the LDS counter: every LDS wait and
// barrier would then wait for all outstanding global loads. Hot paths
// access them through explicit global-address-space pointers instead.
#define PGC_GLOBAL __attribute__((address_space(1)))
// The job table is read-only for a launch: read through the constant address
// space with a wave-uniform index, its fields come in scalar loads and stay
// in scalar registers (a job reference in the generic address space would
// make every field a per-lane vector load).
#define PGC_CONST __attribute__((address_space(4)))
template <typename T>
__device__ __forceinline__ PGC_GLOBAL T* gp(T* p) {
  return (PGC_GLOBAL T*)p;
}
template <typename T>
__device__ __forceinline__ const PGC_GLOBAL T* gp(const T* p) {
  return (const PGC_GLOBAL T*)p;
}


// std::round then double->int as on x86-64 (INT_MIN outside the int range)
__device__ __forceinline__ int round_to_int(double w) {
  const double r = round(w);
  if (!(r >= -2147483648.0 && r < 2147483648.0)) return (int)0x80000000;
  return (int)r;
}

struct Shared {
  int orCount[2];  // barrier_or counters, alternating by parity
  int gridFlag;    // device-wide solve: this workgroup has a true predicate (GridOr)
  int gridOk;      // device-wide solve: no barrier has timed out (set by thread 0)
  int i0, i1, i2, i3;
  double d0, d1, d2;
  int status;
  int numInBranch;
  int indexNode;
  int pickV;
  double pickD;
  int redI[PGC_BLOCK / PGC_WAVE];
  double redD[PGC_BLOCK / PGC_WAVE];
};

// Block-wide OR of a predicate with a single barrier: waves with a true
// predicate bump the counter of the current parity; a thread compares the
// counter after the barrier with its previous reading. Counters alternate
// by parity so a fast wave's next increment cannot reach a counter a slow
// wave has still to read (that wave must first pass the next barrier).
struct BarrierOr {
  int prev0 = 0, prev1 = 0;
  bool odd = false;
  __device__ __forceinline__ bool operator()(Shared& sh, bool pred) {
    int* cnt = odd ? &sh.orCount[1] : &sh.orCount[0];
    const unsigned long long b = __ballot(pred);
    if (b && (threadIdx.x & 63) == 0) atomicAdd(cnt, 1);
    __syncthreads();
    const int c = *cnt;
    bool any;
    if (odd) {
      any = c != prev1;
      prev1 = c;
    } else {
      any = c != prev0;
      prev0 = c;
    }
    odd = !odd;
    return any;
  }
};

__device__ __forceinline__ uint64_t wclock() { return wall_clock64(); }

// ------------------------------------------------------------ device-wide solve
// Graphs whose distance rows exceed the LDS can be solved by all workgroups of
// a cooperative launch together: every (layer, depth) step of a sweep is
// spread over the waves of the whole grid (one 64-vertex chunk per wave), and
// the steps are separated by device-wide barriers. The control flow (Lawler
// iterations, bounds, cycle walks) runs in every workgroup alike from values
// that one thread writes to GridState::sh and all threads read after a
// barrier.
constexpr int kGridMaxBlocks = 256;

struct GridState {
  unsigned long long arrive;  // barrier arrivals (monotonic, zero at launch)
  unsigned int orc[2];        // device-wide OR counters, alternating by parity
  unsigned int abort;         // a barrier waited past its time limit: every workgroup stops
  unsigned int pad;
  Shared sh;                  // broadcast values (written by the grid's first thread)
  int redI[2][kGridMaxBlocks];     // per-workgroup partials of a reduction, two banks
  double redD[2][kGridMaxBlocks];
};

// Device-wide barrier: thread 0 of every workgroup adds one arrival to a
// monotonic counter and waits until all workgroups have arrived at this
// barrier (its ordinal times the grid size). Every wave first waits for its
// own vector memory operations (the workgroup barrier does not), so that all
// stores of the workgroup have reached its L2. FULL: a device-scope release
// before the arrival (it writes back the XCD's L2, whose lines are not
// coherent with the other XCDs' otherwise) and an acquire after the wait (it
// invalidates the L2). Without FULL only the acquire: the barrier then orders
// the data stored with device-scope (write-through) stores before it (the
// distance rows between the steps of a sweep); plain stores (predecessors)
// become visible to other workgroups at the next full barrier. A wait of more than two seconds (the wall
// clock runs at 100 MHz) sets the abort flag instead of hanging; every later
// barrier then returns at once. Returns false once the solve is aborted
// (thread 0 reads the flag, the workgroup gets it through okFlag in LDS).
template <bool FULL = true>
__device__ __forceinline__ bool grid_sync(GridState* gs, uint64_t& bgen, uint32_t nblk, int* okFlag) {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  bgen += 1;
  if (threadIdx.x == 0) {
    const unsigned long long target = (unsigned long long)bgen * nblk;
    if (FULL) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_fetch_add(&gs->arrive, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t t0 = wall_clock64();
    bool ok = __hip_atomic_load(&gs->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
    while (ok && __hip_atomic_load(&gs->arrive, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (wall_clock64() - t0 > 200000000ull) {
        __hip_atomic_store(&gs->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = false;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      ok = __hip_atomic_load(&gs->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    *(volatile int*)okFlag = ok ? 1 : 0;
  }
  __syncthreads();
  return *(volatile int*)okFlag != 0;
}

// Device-wide OR of a predicate: a workgroup with a true predicate in any wave
// adds one to the counter of the current parity before the barrier; after it
// every thread compares the counter with its previous reading. Counters
// alternate by parity, as in BarrierOr.
struct GridOr {
  unsigned int prev0 = 0, prev1 = 0;
  bool odd = false;
  // ldsFlag, okFlag: two ints of the workgroup's LDS; ok: false once the solve is aborted
  __device__ __forceinline__ bool operator()(GridState* gs, uint64_t& bgen, uint32_t nblk, int* ldsFlag, int* okFlag,
                                             bool pred, bool& ok) {
    if (__ballot(pred) && (threadIdx.x & 63) == 0) *(volatile int*)ldsFlag = 1;
    __syncthreads();
    if (threadIdx.x == 0 && *(volatile int*)ldsFlag) {
      __hip_atomic_fetch_add(&gs->orc[odd ? 1 : 0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      *(volatile int*)ldsFlag = 0;
    }
    ok = grid_sync(gs, bgen, nblk, okFlag) && ok;
    const unsigned int c = __hip_atomic_load(&gs->orc[odd ? 1 : 0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    bool any;
    if (odd) {
      any = c != prev1;
      prev1 = c;
    } else {
      any = c != prev0;
      prev0 = c;
    }
    odd = !odd;
    return any;
  }
};

}  // namespace pgc

#include "sweep_device.h"
#include "wrap_device.h"

namespace pgc {

// ---------------------------------------------------------------- reductions
__device__ __forceinline__ int block_sum(Shared& sh, int x) {
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh.redI[w] = x;
  __syncthreads();
  int s = 0;
  for (int i = 0; i < (PGC_BLOCK / PGC_WAVE); ++i) s += sh.redI[i];
  __syncthreads();
  return s;
}

__device__ __forceinline__ int block_max(Shared& sh, int x) {
  for (int o = 32; o > 0; o >>= 1) x = max(x, __shfl_xor(x, o));
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh.redI[w] = x;
  __syncthreads();
  int s = sh.redI[0];
  for (int i = 1; i < (PGC_BLOCK / PGC_WAVE); ++i) s = max(s, sh.redI[i]);
  __syncthreads();
  return s;
}

// Ascending sort of N (key, value) pairs in registers by Batcher's odd-even
// merge network (N a power of two; keys distinct, or equal keys whose order
// does not matter).
template <int N>
__device__ __forceinline__ void sort_pairs(int (&k)[N], double (&x)[N]) {
#pragma unroll
  for (int p = 1; p < N; p <<= 1)
#pragma unroll
    for (int q = p; q >= 1; q >>= 1)
#pragma unroll
      for (int j = q % p; j + q < N; j += 2 * q)
#pragma unroll
        for (int i = 0; i < q; ++i) {
          const int a = i + j, b = i + j + q;
          if (b < N && a / (2 * p) == b / (2 * p)) {
            const bool sw = k[b] < k[a];
            const int ka = k[a], kb = k[b];
            const double xa = x[a], xb = x[b];
            k[a] = sw ? kb : ka;
            k[b] = sw ? ka : kb;
            x[a] = sw ? xb : xa;
            x[b] = sw ? xa : xb;
          }
        }
}

// lexicographic minimum of (key, v); v = INT_MAX means none
__device__ __forceinline__ void block_min_key(Shared& sh, double& key, int& v) {
  for (int o = 32; o > 0; o >>= 1) {
    const double k2 = __shfl_xor(key, o);
    const int v2 = __shfl_xor(v, o);
    if (k2 < key || (k2 == key && v2 < v)) {
      key = k2;
      v = v2;
    }
  }
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
    sh.redD[w] = key;
    sh.redI[w] = v;
  }
  __syncthreads();
  key = sh.redD[0];
  v = sh.redI[0];
  for (int i = 1; i < (PGC_BLOCK / PGC_WAVE); ++i) {
    if (sh.redD[i] < key || (sh.redD[i] == key && sh.redI[i] < v)) {
      key = sh.redD[i];
      v = sh.redI[i];
    }
  }
  __syncthreads();
}

// ------------------------------------------------------------ edge lookup
// Edge joining consecutive cycle nodes a -> b (original node ids): the first
// in-edge of b (remapped to the closing copy when it wraps) whose source is a.
template <typename CT, bool HAS_TIME, typename JT>
__device__ __forceinline__ bool lookup_edge(const JT& J, int a, int b, int& eid, double& c, double& t) {
  const int64_t npl = J.npl;
  int64_t bb = b;
  if (bb < npl && a > npl) bb += J.numNodes;
  if (a < 0 || bb < 0 || bb >= J.numNodes + npl) return false;
  const int lb = (int)(bb / npl);
  const int vi = (int)(bb - (int64_t)lb * npl);
  const bool same = (int64_t)a / npl == lb;
  const auto* S = same ? &J.intra : &J.inter;
  const int sl = same ? (int)(a - (int64_t)lb * npl) : (int)(a - (int64_t)(lb - 1) * npl);
  if (!same && (lb == 0 || sl < 0 || sl >= npl)) return false;
  const int ch = vi >> 6, lane = vi & 63;
  const int width = S->w;
  const int base = lb * S->lstride + ch * width * 64;
  for (int j = 0; j < width; ++j) {
    const int idx = base + j * 64 + lane;
    const int s = gp(S->src)[idx];
    if (s < 0 || s >= npl) break;
    if (s == sl) {
      eid = gp(S->eid)[idx];
      c = (double)gp((const CT*)S->cost)[idx];
      t = HAS_TIME ? (double)gp(S->time)[idx] : J.time_const;
      return true;
    }
  }
  return false;
}

// ---------------------------------------------------------------- kernel
// MAXC = chunks (64-vertex slices) per wave, unrolled; MAXC == 0: any number
// of chunks per wave in runtime loops, with the distance table in global
// memory (graphs whose table exceeds the LDS).
template <typename CT, bool HAS_TIME, bool MINCOST, int MAXC, int W, bool GRID = false>
struct Solver {
  static constexpr bool DYN = MAXC == 0;
  static_assert(!GRID || DYN, "the device-wide solve keeps its table in global memory");
  // The unrolled sweep (slots of the next block prefetched into registers)
  // pays off with one chunk per wave and 8 slots per vertex; with more, its
  // register sets spill, and the sweep that takes one chunk at a time (rows
  // in LDS) is faster.
  static constexpr bool UNROLLED = MAXC == 1 && W == 8;
  const PGC_CONST SolverJob& J;
  double* D;     // [2K][ds] distance rows; entry npl of every row is +inf (padding slots point there)
  double* Dmin;  // [ds]
  Shared& sh;
  PGC_LDS int32_t* Pl;  // pred in LDS (J.pred_lds), else J.pred in global memory
  bool predLds;
  PGC_LDS int32_t* WnL;  // wrap map in LDS after pred (J.wrap_lds), else J.wnext/J.wlen
  PGC_LDS int32_t* WlL;
  bool wrapLds;
  PGC_LDS int32_t* OrgL;  // chain-origin rows after the wrap map (J.org_lds), see sweep_layers
  bool orgOn;
  uint32_t inb = 0;  // unrolled variants: bit i set when owned vertex i is in the current branch
  FastDiv fdNpl, fdPer;  // division by npl and by npl*K
  int npl, ds, K, L, nC, wave, lane, nw, npass;
  BarrierOr bor;
  uint64_t tSweep = 0, tFold = 0, tChase = 0, tWalk = 0, tExtract = 0;
  // device-wide solve (GRID): sh is GridState::sh in global memory, lsh the
  // workgroup's own LDS block
  GridState* gs = nullptr;
  PGC_LDS Shared* lsh = nullptr;
  uint64_t bgen = 0;
  uint32_t nblk = 1;
  GridOr gor;
  int rbank = 0;
  bool alive = true;

  // workgroup barrier, or the device-wide one
  __device__ __forceinline__ void sync() {
    if constexpr (GRID)
      alive = grid_sync(gs, bgen, nblk, (int*)&lsh->gridOk) && alive;
    else
      __syncthreads();
  }
  // distance rows outside the sweep (the solver's barriers write them back
  // and invalidate the caches)
  __device__ __forceinline__ double dld(const double* p) const { return *p; }
  __device__ __forceinline__ void dst(double* p, double v) const { *p = v; }
  // the thread that runs the sequential parts
  __device__ __forceinline__ bool leader() const {
    if constexpr (GRID) return blockIdx.x == 0 && threadIdx.x == 0;
    return threadIdx.x == 0;
  }
  // strided loops over all threads of the solve
  __device__ __forceinline__ int tfirst() const {
    if constexpr (GRID) return (int)(blockIdx.x * PGC_BLOCK + threadIdx.x);
    return (int)threadIdx.x;
  }
  __device__ __forceinline__ int tstride() const {
    if constexpr (GRID) return (int)(gridDim.x * PGC_BLOCK);
    return PGC_BLOCK;
  }
  // OR of a predicate over all threads of the solve
  __device__ __forceinline__ bool any(bool pred) {
    if constexpr (GRID) return gor(gs, bgen, nblk, (int*)&lsh->gridFlag, (int*)&lsh->gridOk, pred, alive);
    return bor(sh, pred);
  }
  // reductions over all threads of the solve: the workgroup's partial, then
  // (GRID) the partials of all workgroups in workgroup order, two banks in turn
  __device__ __forceinline__ int rsum(int x) {
    if constexpr (GRID) {
      const int part = block_sum(*(Shared*)lsh, x);
      if (threadIdx.x == 0) gs->redI[rbank][blockIdx.x] = part;
      sync();
      int t = 0;
      for (uint32_t b = 0; b < nblk; ++b) t += gs->redI[rbank][b];
      rbank ^= 1;
      return t;
    }
    return block_sum(sh, x);
  }
  __device__ __forceinline__ int rmax(int x) {
    if constexpr (GRID) {
      const int part = block_max(*(Shared*)lsh, x);
      if (threadIdx.x == 0) gs->redI[rbank][blockIdx.x] = part;
      sync();
      int t = gs->redI[rbank][0];
      for (uint32_t b = 1; b < nblk; ++b) t = max(t, gs->redI[rbank][b]);
      rbank ^= 1;
      return t;
    }
    return block_max(sh, x);
  }
  __device__ __forceinline__ void rmin_key(double& key, int& v) {
    if constexpr (GRID) {
      block_min_key(*(Shared*)lsh, key, v);
      if (threadIdx.x == 0) {
        gs->redD[rbank][blockIdx.x] = key;
        gs->redI[rbank][blockIdx.x] = v;
      }
      sync();
      key = gs->redD[rbank][0];
      v = gs->redI[rbank][0];
      for (uint32_t b = 1; b < nblk; ++b) {
        const double k2 = gs->redD[rbank][b];
        const int v2 = gs->redI[rbank][b];
        if (k2 < key || (k2 == key && v2 < v)) {
          key = k2;
          v = v2;
        }
      }
      rbank ^= 1;
      return;
    }
    block_min_key(sh, key, v);
  }

  __device__ __forceinline__ Solver(const PGC_CONST SolverJob& j, double* d, double* dm, Shared& s, int32_t* pl)
      : J(j), D(d), Dmin(dm), sh(s) {
    Pl = (PGC_LDS int32_t*)pl;
    predLds = !DYN && J.pred_lds != 0;
    wrapLds = !DYN && J.wrap_lds != 0;
    WnL = Pl + (predLds ? (int64_t)(J.L + 1) * J.K * J.npl : 0);
    WlL = WnL + J.npl;
    orgOn = UNROLLED && !MINCOST && J.K == 1 && predLds && wrapLds && J.org_lds != 0;
    OrgL = WnL + 2 * J.npl;
    npl = J.npl;
    ds = npl + 1;
    K = J.K;
    L = J.L;
    nC = J.nChunks;
    lane = threadIdx.x & 63;
    if constexpr (GRID) {
      wave = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (PGC_BLOCK / PGC_WAVE) + (threadIdx.x >> 6)));
      nw = (int)gridDim.x * (PGC_BLOCK / PGC_WAVE);
    } else {
      wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
      nw = PGC_BLOCK / PGC_WAVE;
    }
    npass = DYN ? (nC + nw - 1) / nw : MAXC;
    fdNpl.init((uint32_t)npl);
    fdPer.init((uint32_t)(npl * K));
  }

  // original graph index (layer * npl + vertex) of an expanded node id
  // (getOriginalGraphIndex, cost_time_ratio_solver.cpp:240-250)
  __device__ __forceinline__ int orig(int64_t p) const {
    const uint32_t u = (uint32_t)p;
    const uint32_t layer = fdPer.div(u);
    const uint32_t rem = u - layer * (uint32_t)(npl * K);
    const uint32_t depth = fdNpl.div(rem);
    return (int)(rem - depth * (uint32_t)npl + layer * (uint32_t)npl);
  }

  // While the leader walks the upper bound's path back from the last layer,
  // waves 1 to kWarmWaves read the predecessor table (global memory) from its
  // end backwards, one load per 128-byte line, so that the leader's dependent
  // loads find their lines in the L2 (after a sweep most of the table has
  // left it). The values are discarded. (The cycle extraction's walks follow
  // chains the chase has just read: warming there only adds contention.)
  static constexpr int kWarmWaves = 4;
  __device__ __forceinline__ void warm_pred_reverse() {
    const PGC_GLOBAL int32_t* P = gp((const int32_t*)J.pred);
    const int64_t n = (int64_t)(L + 1) * K * npl;
    const int64_t lines = (n + 31) / 32;
    const int t = (int)threadIdx.x - 64;
    constexpr int nt = kWarmWaves * 64;
    int acc = 0;
    for (int64_t i = lines - 1 - t; i >= 0; i -= 2 * nt) {
      const int a = P[i * 32];
      const int b = i - nt >= 0 ? P[(i - nt) * 32] : 0;
      acc ^= a ^ b;
    }
    if (acc == 0x7ffffffe && npl < 0) sh.i3 = acc;  // never taken: keeps the loads
  }

  // predecessor table (expanded node ids) in LDS or global memory
  __device__ __forceinline__ int pget(int64_t i) const { return predLds ? Pl[i] : gp(J.pred)[i]; }
  __device__ __forceinline__ void pset(int64_t i, int v) const {
    if (predLds)
      Pl[i] = v;
    else
      gp(J.pred)[i] = v;
  }

  // passes over the owned chunks (a constant when MAXC > 0)
  __device__ __forceinline__ int np() const { return DYN ? npass : MAXC; }

  // geometry of owned chunk i (wave-uniform: scalar loads)
  __device__ __forceinline__ int chunk_of(int i) const { return wave + i * nw; }

  __device__ __forceinline__ int vert(int i) const {
    const int c = wave + i * nw;
    if (c >= nC) return -1;
    const int v = c * 64 + lane;
    return v < npl ? v : -1;
  }

  // One pass over layers 0..L (the sweep runs in its own function, which
  // gets the whole register file; the caller's state is saved once per sweep).
  __device__ __forceinline__ int run_layers(double ratio, double ub) {
    SweepState r;
    if constexpr (GRID) {
      constexpr int SX = W == 8 ? 7 : (W == 16 ? 12 : W), SI = W == 8 || W == 16 ? 6 : W;
      // with at most one chunk per wave the sources of a wave's chunk stay in
      // registers for the whole sweep (layer-invariant sources)
      const bool one = nC <= nw && !(J.flags & FLAG_NO_CACHED_SOURCES);
      const bool shortSlots = SX < W && J.inter.deg <= SX && J.intra.deg <= SI && !(J.flags & FLAG_FULL_SLOTS);
      int p0 = (int)gor.prev0, p1 = (int)gor.prev1, od = gor.odd ? 1 : 0;
      const bool k2 = W != 0 && one && K == 2 && J.inter.src_inv && J.intra.src_inv;
      // conjugate graphs of valence-6 meshes: interlayer in-degree <= 10
      constexpr int SX10 = W == 16 ? 10 : SX;
      if (k2 && shortSlots && J.inter.deg <= SX10)
        r = sweep_layers_global<CT, HAS_TIME, MINCOST, W, false, 1, SX10, SI, true, true>(
            (const SolverJob*)&J, D, Dmin, (Shared*)lsh, nullptr, 0, p0, p1, od, ratio, ub, gs, &bgen);
      else if (k2 && shortSlots)
        r = sweep_layers_global<CT, HAS_TIME, MINCOST, W, false, 1, SX, SI, true, true>(
            (const SolverJob*)&J, D, Dmin, (Shared*)lsh, nullptr, 0, p0, p1, od, ratio, ub, gs, &bgen);
      else if (k2)
        r = sweep_layers_global<CT, HAS_TIME, MINCOST, W, false, 1, W, W, true, true>(
            (const SolverJob*)&J, D, Dmin, (Shared*)lsh, nullptr, 0, p0, p1, od, ratio, ub, gs, &bgen);
      else if (one && shortSlots)
        r = sweep_layers_global<CT, HAS_TIME, MINCOST, W, false, 1, SX, SI, true>(
            (const SolverJob*)&J, D, Dmin, (Shared*)lsh, nullptr, 0, p0, p1, od, ratio, ub, gs, &bgen);
      else if (one)
        r = sweep_layers_global<CT, HAS_TIME, MINCOST, W, false, 1, W, W, true>(
            (const SolverJob*)&J, D, Dmin, (Shared*)lsh, nullptr, 0, p0, p1, od, ratio, ub, gs, &bgen);
      else if (shortSlots)
        r = sweep_layers_global<CT, HAS_TIME, MINCOST, W, false, 0, SX, SI, true>(
            (const SolverJob*)&J, D, Dmin, (Shared*)lsh, nullptr, 0, p0, p1, od, ratio, ub, gs, &bgen);
      else
        r = sweep_layers_global<CT, HAS_TIME, MINCOST, W, false, 0, W, W, true>(
            (const SolverJob*)&J, D, Dmin, (Shared*)lsh, nullptr, 0, p0, p1, od, ratio, ub, gs, &bgen);
      gor.prev0 = (unsigned)r.p0;
      gor.prev1 = (unsigned)r.p1;
      gor.odd = r.odd != 0;
      alive = alive && lsh->gridOk != 0;
      // the sweep's last steps end in light barriers: a full one makes the
      // predecessors they stored visible to every workgroup before the
      // walks read them
      sync();
      return r.complete;
    } else if constexpr (DYN) {
      // short slots for the interlayer / intralayer classes: 7 / 6 of 8 (icosphere
      // product graphs), 12 / 6 of 16 (conjugate product graphs: in-degrees <= 10 / 5)
      constexpr int SX = W == 8 ? 7 : (W == 16 ? 12 : W), SI = W == 8 || W == 16 ? 6 : W;
      if (SX < W && J.inter.deg <= SX && J.intra.deg <= SI && !(J.flags & FLAG_FULL_SLOTS))
        r = sweep_layers_global<CT, HAS_TIME, MINCOST, W, false, 0, SX, SI>((const SolverJob*)&J, D, Dmin, &sh,
                                                                            nullptr, 0, bor.prev0, bor.prev1,
                                                                            bor.odd ? 1 : 0, ratio, ub);
      else
        r = sweep_layers_global<CT, HAS_TIME, MINCOST, W>((const SolverJob*)&J, D, Dmin, &sh, nullptr, 0, bor.prev0,
                                                          bor.prev1, bor.odd ? 1 : 0, ratio, ub);
    }
    else if constexpr (!UNROLLED) {
      constexpr int MC = W == 8 && MAXC <= (HAS_TIME ? 2 : 3) ? MAXC : 0;
      // icosphere-like graphs: interlayer in-degree <= 7 and intralayer <= 6
      // leave the last one and two slots padding on every node
      if (MC > 0 && J.inter.deg <= W - 1 && J.intra.deg <= W - 2 && !(J.flags & FLAG_FULL_SLOTS))
        r = sweep_layers_global<CT, HAS_TIME, MINCOST, W, true, MC, (MC > 0 ? W - 1 : W), (MC > 0 ? W - 2 : W)>(
            (const SolverJob*)&J, D, Dmin, &sh, (int32_t*)Pl, predLds ? 1 : 0, bor.prev0, bor.prev1, bor.odd ? 1 : 0,
            ratio, ub);
      else
        r = sweep_layers_global<CT, HAS_TIME, MINCOST, W, true, MC>((const SolverJob*)&J, D, Dmin, &sh, (int32_t*)Pl,
                                                                    predLds ? 1 : 0, bor.prev0, bor.prev1,
                                                                    bor.odd ? 1 : 0, ratio, ub);
    }
    else if (K != 1)
      r = sweep_layers<CT, HAS_TIME, MINCOST, MAXC, W>((const SolverJob*)&J, D, Dmin, &sh, (int32_t*)Pl,
                                                       predLds ? 1 : 0, bor.prev0, bor.prev1, bor.odd ? 1 : 0, ratio,
                                                       ub);
    else if (!MINCOST && !HAS_TIME && orgOn && J.icol) {
      // implicit costs: the staging columns follow the origin rows (8-byte aligned)
      PGC_LDS int32_t* st32 = OrgL + 2 * npl;
      st32 += ((uint32_t)(uintptr_t)st32 & 4u) ? 1 : 0;
      double* stage = (double*)(PGC_LDS double*)st32;
      if (J.inter.deg <= W - 1 && !(J.flags & FLAG_FULL_SLOTS))
        r = sweep_layers_k1_implicit<CT, MAXC, W, W - 1>((const SolverJob*)&J, D, Dmin, &sh, (int32_t*)Pl,
                                                         bor.prev0, bor.prev1, bor.odd ? 1 : 0, ratio, ub,
                                                         (int32_t*)OrgL, stage);
      else
        r = sweep_layers_k1_implicit<CT, MAXC, W, W>((const SolverJob*)&J, D, Dmin, &sh, (int32_t*)Pl, bor.prev0,
                                                     bor.prev1, bor.odd ? 1 : 0, ratio, ub, (int32_t*)OrgL, stage);
    }
    else if (J.inter.deg <= W - 1 && !(J.flags & FLAG_FULL_SLOTS))  // the last interlayer slot is padding on every node
      r = sweep_layers<CT, HAS_TIME, MINCOST, MAXC, W, true, W - 1>((const SolverJob*)&J, D, Dmin, &sh,
                                                                    (int32_t*)Pl, predLds ? 1 : 0, bor.prev0,
                                                                    bor.prev1, bor.odd ? 1 : 0, ratio, ub,
                                                                    (int32_t*)OrgL, orgOn ? 1 : 0);
    else
      r = sweep_layers<CT, HAS_TIME, MINCOST, MAXC, W, true>((const SolverJob*)&J, D, Dmin, &sh, (int32_t*)Pl,
                                                             predLds ? 1 : 0, bor.prev0, bor.prev1, bor.odd ? 1 : 0,
                                                             ratio, ub, (int32_t*)OrgL, orgOn ? 1 : 0);
    bor.prev0 = r.p0;
    bor.prev1 = r.p1;
    bor.odd = r.odd != 0;
    return r.complete;
  }

  // D <- inf, branch sources <- 0, pred of sources <- -1; returns #sources.
  __device__ __forceinline__ int setup(int bid, int onlyV) {
    for (int i = tfirst(); i < 2 * K * ds; i += tstride()) dst(D + i, PGC_INF);
    for (int i = tfirst(); i < ds; i += tstride()) dst(Dmin + i, PGC_INF);
    sync();
    int cnt = 0, last = -1;
    inb = 0;
#pragma unroll
    for (int i = 0; i < np(); ++i) {
      const int v = vert(i);
      if (v < 0) continue;
      const bool in = onlyV >= 0 ? (v == onlyV) : (gp(J.branch)[v] == bid);
      if (!DYN && in) inb |= 1u << i;
      if (in) {
        dst(D + v, 0.0);
        pset(v, -1);
        cnt++;
        last = max(last, v);
      }
    }
    const int n = rsum(cnt);
    const int lv = rmax(last);
    if (leader()) {
      sh.numInBranch = n;
      sh.indexNode = lv;
    }
    sync();
    return n;
  }

  __device__ __forceinline__ bool in_branch(int v, int bid, int onlyV) const {
    return onlyV >= 0 ? (v == onlyV) : (gp(J.branch)[v] == bid);
  }

  // Fold of the closing layer into layer 0; returns whether any value improved.
  // The reference reads row L % 2: the closing layer's depth-0 row only for
  // even L or K == 1 (for odd L it is row 1 of layer L-1). The corrected
  // semantics read row (L % 2) * K and, for odd L, make the closing layer's
  // depth rows 1..K-1 layer 0's, as the shared rows do for even L.
  __device__ __forceinline__ int closing_row() const {
    return (J.flags & FLAG_CORRECTED_ODD_L) != 0 ? (L % 2) * K : L % 2;
  }

  __device__ __forceinline__ bool fold(int bid, int onlyV) {
    const int lastRow = closing_row();
    if ((J.flags & FLAG_CORRECTED_ODD_L) != 0 && (L % 2) == 1) {
#pragma unroll
      for (int i = 0; i < np(); ++i) {
        const int v = vert(i);
        if (v < 0) continue;
        for (int k = 1; k < K; ++k) dst(D + k * ds + v, dld(D + (K + k) * ds + v));
      }
    }
    bool imp = false;
#pragma unroll
    for (int i = 0; i < np(); ++i) {
      const int v = vert(i);
      if (v < 0 || !(DYN ? in_branch(v, bid, onlyV) : ((inb >> i) & 1u) != 0)) continue;
      const double dl = dld(D + lastRow * ds + v);
      const double bk = gp(J.backup)[v];
      if (dl < bk) {
        imp = true;
        dst(D + v, dl);
        pset(v, pget((int64_t)npl * K * L + v));
      } else {
        dst(D + v, bk);
      }
    }
    return any(imp);
  }

  // Chase every layer-0 vertex's predecessor chain back to the previous
  // layer-0 vertex (one wrap): the wrap map wnext/wlen. The chains' node ids
  // are not stored; the extraction re-walks the few chains on its cycle.
  __device__ __forceinline__ void chase_wraps() {
    const int stride = J.chain_stride;
    if constexpr (DYN) {
      // one chain at a time per thread
      for (int i = 0; i < npass; ++i) {
        const int v = vert(i);
        if (v < 0) continue;
        int pp = pget(v);
        int pos = 0;
        while (true) {
          if (pp == -1) {
            gp(J.wnext)[v] = -1;
            gp(J.wlen)[v] = pos;
            break;
          }
          if (pp < npl) {
            pos++;
            gp(J.wnext)[v] = pp;
            gp(J.wlen)[v] = pos;
            break;
          }
          if (pos >= stride - 1) {
            gp(J.wnext)[v] = -2;
            gp(J.wlen)[v] = pos;
            break;
          }
          pos++;
          pp = pget(pp);
        }
      }
    } else if (orgOn) {
      chase_from_origins();
    } else {
      chase_wraps_fn<DYN ? 1 : MAXC>(predLds ? (int32_t*)Pl : J.pred, predLds ? 1 : 0,
                                     wrapLds ? (int32_t*)WnL : J.wnext, wrapLds ? (int32_t*)WlL : J.wlen,
                                     wrapLds ? 1 : 0, npl, nC, stride);
    }
  }

  // The wrap map from the chain origins the K == 1 sweep kept (the result of
  // the chase, without walking the chains): the predecessor of layer-0 vertex
  // v is a node y of layer L-1 (the fold copies it from the closing copy), and
  // its chain visits one node per layer: it ends at vertex o of layer 0 after
  // L steps, or after L - m steps at a node of layer m without predecessor.
  __device__ __forceinline__ void chase_from_origins() {
    const PGC_LDS int32_t* last = OrgL + ((L - 1) & 1) * npl;
#pragma unroll
    for (int i = 0; i < np(); ++i) {
      const int v = vert(i);
      if (v < 0) continue;
      const int p = pget(v);
      int wn, wl;
      if (p == -1) {
        wn = -1;
        wl = 0;
      } else if (p < npl) {
        wn = p;
        wl = 1;
      } else if (p >= (L - 1) * npl && p < L * npl) {
        const int o = last[p - (L - 1) * npl];
        wn = o >= 0 ? o : -1;
        wl = o >= 0 ? L : L - (-1 - o);
      } else {
        // not a node of layer L-1 (no path sets such a predecessor): walk
        // the chain as the chase does
        int pp = p, pos = 0;
        while (true) {
          if (pp < npl) {
            wn = pp;
            wl = pos + 1;
            break;
          }
          if (pos >= J.chain_stride - 1) {
            wn = -2;
            wl = pos;
            break;
          }
          ++pos;
          pp = pget(pp);
          if (pp == -1) {
            wn = -1;
            wl = pos;
            break;
          }
        }
      }
      WnL[v] = wn;
      WlL[v] = wl;
    }
  }

  // wrap map entries (either placement; for the rare sequential paths)
  __device__ __forceinline__ int wn_get(int v) const { return wrapLds ? WnL[v] : gp(J.wnext)[v]; }
  __device__ __forceinline__ int wl_get(int v) const { return wrapLds ? WlL[v] : gp(J.wlen)[v]; }

  // Walk of the reference (cost_time_ratio_solver.cpp:443-489) expressed on
  // the wrap map W: first repeated layer-0 vertex from v and its step count.
  template <typename WT>
  __device__ __forceinline__ bool walk_in(WT Wn, WT Ln, int v, int64_t capSteps, int& vstar) const {
    const int lim = 4 * npl + 8;
    int t = Wn[v];
    if (t < 0) return false;
    int h = Wn[t];
    int it = 0;
    while (t != h && h >= 0 && it < lim) {
      t = Wn[t];
      h = Wn[h];
      if (h >= 0) h = Wn[h];
      ++it;
    }
    if (h < 0 || it >= lim) return false;
    int a = v;
    it = 0;
    while (a != h && it < lim) {
      a = Wn[a];
      h = Wn[h];
      ++it;
    }
    const int e = a;
    int64_t P = 0;
    int x = v;
    it = 0;
    while (x != e && it < lim) {
      P += Ln[x];
      x = Wn[x];
      ++it;
    }
    x = e;
    do {
      P += Ln[x];
      x = Wn[x];
      ++it;
    } while (x != e && it < 2 * lim);
    if (P > capSteps) return false;
    vstar = e;
    return true;
  }
  __device__ __forceinline__ bool walk_from(int v, int64_t capSteps, int& vstar) const {
    if (wrapLds) return walk_in((const PGC_LDS int32_t*)WnL, (const PGC_LDS int32_t*)WlL, v, capSteps, vstar);
    return walk_in(gp((const int32_t*)J.wnext), gp((const int32_t*)J.wlen), v, capSteps, vstar);
  }

  // Best candidate of the negative-cycle search: branch vertices with a
  // negative (or, searching zero cycles, |D| < tol) layer-0 value whose walk
  // closes within the step cap; key -v (pick last) or D.
  template <typename WT>
  __device__ __forceinline__ void candidates(WT Wn, WT Ln, int bid, int onlyV, bool zeroSearch, double& bestKey,
                                             int& bestV) const {
    const bool pickLast = (J.flags & FLAG_PICK_LAST_VERTEX) != 0;
    const int64_t capSteps = J.numNodes - 1;
#pragma unroll
    for (int i = 0; i < np(); ++i) {
      const int v = vert(i);
      if (v < 0 || !in_branch(v, bid, onlyV)) continue;
      const double lb = dld(D + v);
      if (!zeroSearch && lb >= 0) continue;
      if (zeroSearch && (lb <= -J.tol || lb >= J.tol)) continue;
      int vs;
      if (!walk_in(Wn, Ln, v, capSteps, vs)) continue;
      const double key = pickLast ? -(double)v : lb;
      if (key < bestKey || (key == bestKey && v < bestV)) {
        bestKey = key;
        bestV = v;
      }
    }
  }

  // sequential sums over a cycle's pairs (thread 0 only)
  __device__ __forceinline__ void pair_sums(const double* c_, const double* t_, int m, double ratio, double& cs,
                                            double& ts, double& ws) const {
    const PGC_GLOBAL double* c = gp(c_);
    const PGC_GLOBAL double* t = gp(t_);
    cs = 0;
    ts = 0;
    ws = 0;
    // unrolled so that the loads of eight pairs are in flight together; the
    // additions stay in index order
#pragma unroll 8
    for (int i = 0; i < m; ++i) {
      const double ci = c[i], ti = t[i];
      cs += ci;
      ts += ti;
      ws += __dsub_rn(ci, __dmul_rn(ratio, ti));
    }
  }

  // edge lookups for tmp_nodes[0..n) -> tmp_eid/c/t[0..n-1); returns ok (block-uniform)
  __device__ __forceinline__ bool lookup_pairs(int n) {
    bool bad = false;
    for (int i = tfirst(); i + 1 < n; i += tstride()) {
      int eid = -1;
      double c = 0, t = 0;
      if (!lookup_edge<CT, HAS_TIME>(J, gp(J.tmp_nodes)[i], gp(J.tmp_nodes)[i + 1], eid, c, t)) bad = true;
      gp(J.tmp_eid)[i] = eid;
      gp(J.tmp_c)[i] = c;
      gp(J.tmp_t)[i] = t;
    }
    return !any(bad);
  }

  __device__ __forceinline__ void copy_tmp_to_neg(int n) {
    for (int i = tfirst(); i < n; i += tstride()) {
      gp(J.neg_nodes)[i] = gp(J.tmp_nodes)[i];
      if (i + 1 < n) {
        gp(J.neg_eid)[i] = gp(J.tmp_eid)[i];
        gp(J.neg_c)[i] = gp(J.tmp_c)[i];
        gp(J.neg_t)[i] = gp(J.tmp_t)[i];
      }
    }
    if (leader()) *gp(J.neg_len) = n;
    sync();
  }

  // ZERO_CYCLE if the current cycle has (rounded) zero reweighted cost, else NO_NEGATIVE_CYCLE
  __device__ __forceinline__ int zero_or_none(double ratio) {
    if (leader()) {
      const int n = *gp(J.neg_len);
      double cs, ts, ws;
      pair_sums(J.neg_c, J.neg_t, n > 0 ? n - 1 : 0, ratio, cs, ts, ws);
      sh.status = round_to_int(ws) == 0 ? ST_ZERO_CYCLE : ST_NO_NEGATIVE_CYCLE;
    }
    sync();
    return sh.status;
  }

  // mooreBellmanFordSubroutine (ratio and upper-bound modes).
  // Returns status; on FOUND the new cycle is in neg_* and its ratio in sh.d0.
  __device__ __forceinline__ int bellman_ford(double ratio, bool zeroSearch, bool ubSearch, int bid, int onlyV, int64_t& sweeps,
                              int& err) {
    const int n = setup(bid, onlyV);
    if (n == 0) {
      err = ERR_EMPTY_BRANCH;
      return ST_NO_NEGATIVE_CYCLE;
    }
    const double ub = PGC_INF;
    int64_t iter = 0;
    while (true) {
#pragma unroll
      for (int i = 0; i < np(); ++i) {
        const int v = vert(i);
        if (v >= 0) gp(J.backup)[v] = dld(D + v);
      }
      sweeps++;
      uint64_t t0 = wclock();
      const int complete = run_layers(ratio, ub);
      uint64_t t1 = wclock();
      tSweep += t1 - t0;
      if (GRID && !alive) {
        err = ERR_GRID_SYNC;
        return ST_NO_NEGATIVE_CYCLE;
      }
      if (!complete) return zero_or_none(ratio);

      if (ubSearch) {
        // path from the closing copy of the start vertex back to it (:370-397)
        if (leader()) {
          const int start = sh.indexNode;
          const int64_t maxLength = (int64_t)L * K;
          int cnt = 0;
          int64_t firstV = start + (int64_t)npl * L * K;
          gp(J.tmp_nodes)[cnt++] = orig(firstV);
          int64_t p = pget(firstV);
          gp(J.tmp_nodes)[cnt++] = orig(p);
          for (int64_t kk = maxLength - 1; kk > 0 && cnt < J.cyc_cap; --kk) {
            if (p == -1) break;
            p = pget(p);
            gp(J.tmp_nodes)[cnt++] = orig(p);
            if (p == start) break;
          }
          sh.i0 = cnt;
        } else if (!GRID && !predLds && (threadIdx.x >> 6) >= 1 && (threadIdx.x >> 6) <= kWarmWaves) {
          warm_pred_reverse();
        }
        sync();
        const int cnt = sh.i0;
        // reverse the path in place: each pair (i, cnt-1-i) belongs to one thread
        for (int i = tfirst(); i < cnt / 2; i += tstride()) {
          const int a = gp(J.tmp_nodes)[i], b = gp(J.tmp_nodes)[cnt - 1 - i];
          gp(J.tmp_nodes)[i] = b;
          gp(J.tmp_nodes)[cnt - 1 - i] = a;
        }
        sync();
        if (!lookup_pairs(cnt)) err = ERR_EDGE_LOOKUP;
        copy_tmp_to_neg(cnt);
        if (leader()) {
          double cs, ts, ws;
          pair_sums(J.neg_c, J.neg_t, cnt - 1, ratio, cs, ts, ws);
          const int w = round_to_int(ws);
          sh.status = w == 0 ? ST_ZERO_CYCLE : (w < 0 ? ST_FOUND_NEGATIVE_CYCLE : ST_NO_NEGATIVE_CYCLE);
        }
        sync();
        return sh.status;
      }

      t0 = wclock();
      const bool folded = fold(bid, onlyV);
      t1 = wclock();
      tFold += t1 - t0;
      if (!folded) return zero_or_none(ratio);

      chase_wraps();
      sync();
      t0 = wclock();
      tChase += t0 - t1;

      const int64_t capSteps = J.numNodes - 1;
      double bestKey = PGC_INF;
      int bestV = 0x7fffffff;
      if constexpr (DYN) {
        candidates(gp((const int32_t*)J.wnext), gp((const int32_t*)J.wlen), bid, onlyV, zeroSearch, bestKey, bestV);
      } else {
        const Candidate c = candidates_fn<DYN ? 1 : MAXC>(
            D, inb, zeroSearch ? 1 : 0, J.tol, (J.flags & FLAG_PICK_LAST_VERTEX) != 0 ? 1 : 0, capSteps,
            wrapLds ? (int32_t*)WnL : J.wnext, wrapLds ? (int32_t*)WlL : J.wlen, wrapLds ? 1 : 0, npl, nC);
        bestKey = c.key;
        bestV = c.v;
      }
      // the block minimum (shuffles, three barriers and a scan of the wave
      // minima) only when some thread has a candidate: in long Bellman-Ford
      // calls no walk closes in most sweeps, and then every thread holds
      // (+inf, none) already, which is the minimum
      if (any(bestV != 0x7fffffff)) rmin_key(bestKey, bestV);
      t1 = wclock();
      tWalk += t1 - t0;
      if (bestV != 0x7fffffff) {
        // extract the cycle through v* (first repeated vertex of bestV's walk)
        if (leader()) {
          int vs = -1;
          walk_from(bestV, capSteps, vs);
          int64_t m = 0;
          int x = vs, it = 0;
          do {
            m += wl_get(x);
            x = wn_get(x);
          } while (x != vs && ++it < 4 * npl + 8);
          sh.i0 = vs;
          sh.i1 = (m + 1 <= J.cyc_cap) ? (int)m : -1;
        }
        sync();
        const int vs = sh.i0;
        const int m = sh.i1;
        if (m < 0) {
          err = ERR_CYCLE_TOO_LONG;
          return ST_NO_NEGATIVE_CYCLE;
        }
        if (leader()) {
          // re-walk each wrap's predecessor chain (unchanged since the
          // chase), writing original node ids back to front
          int x = vs, offp = 0, guard = 0;
          do {
            const int len = wl_get(x);
            int pp = pget(x);
            for (int j = 0; j < len && offp + j < m; ++j) {
              gp(J.tmp_nodes)[m - 1 - (offp + j)] = pp < npl ? pp : orig(pp);
              if (j + 1 < len) pp = pget(pp);
            }
            offp += len;
            x = wn_get(x);
          } while (x != vs && x >= 0 && offp < m && ++guard < 4 * npl + 8);
          gp(J.tmp_nodes)[m] = vs + (int)J.numNodes;
        }
        sync();
        const int nn = m + 1;
        if (!lookup_pairs(nn)) {
          err = ERR_EDGE_LOOKUP;
          return ST_NO_NEGATIVE_CYCLE;
        }
        if (leader()) {
          double cs, ts, ws;
          pair_sums(J.tmp_c, J.tmp_t, nn - 1, ratio, cs, ts, ws);
          sh.i2 = round_to_int(ws);
          sh.d0 = cs / ts;
        }
        sync();
        tExtract += wclock() - t1;
        const int w = sh.i2;
        if (w < 0) {
          copy_tmp_to_neg(nn);
          return ST_FOUND_NEGATIVE_CYCLE;
        }
        // a zero-cost cycle ends the sweep loop through its condition and
        // reaches the final NO_NEGATIVE_CYCLE return (:506-518, :282)
        if (w == 0) return ST_NO_NEGATIVE_CYCLE;
        return zero_or_none(ratio);
      }
      iter++;
      if (GRID && !alive) {
        err = ERR_GRID_SYNC;
        return ST_NO_NEGATIVE_CYCLE;
      }
      if (iter >= J.numNodes - 1) return ST_NO_NEGATIVE_CYCLE;
      if (sweeps >= J.max_sweeps) {
        err = ERR_SWEEP_CAP;
        return ST_NO_NEGATIVE_CYCLE;
      }
    }
  }

  // wraps and zero-layer hits of the current cycle (:606-634); thread 0
  __device__ __forceinline__ void wraps_and_hits(int& numWraps, int& nHits) const {
    const int n = *gp(J.neg_len);
    numWraps = 0;
    nHits = 0;
    bool entered = true, left = false;
    const PGC_GLOBAL int32_t* nodes = gp(J.neg_nodes);
    // eight nodes are loaded before they are scanned (hits is a different
    // array, so its stores cannot change them): one load latency per eight
    for (int i0 = 0; i0 < n; i0 += 8) {
      int vv[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) vv[k] = nodes[i0 + k < n ? i0 + k : i0];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int v = vv[k];
        if (i0 + k < n) {
          if (left && v < npl) {
            entered = true;
            left = false;
          }
          if (entered) {
            if (nHits < npl) gp(J.hits)[nHits] = v;
            nHits++;
            entered = false;
          }
          if (v >= npl && !left) {
            numWraps++;
            left = true;
          }
        }
      }
    }
  }

  // ------------------------------------------------------- modes
  // lawlerSoubroutine (:558-635) on branch bid starting at lambda = ratio_in;
  // returns the last Bellman-Ford status, the final lambda in `ratio`.
  __device__ __forceinline__ int lawler_core(int bid, double ratio_in, double& ratio, int64_t& sweeps, int64_t& bfc,
                                             int64_t& its, int& err) {
    ratio = ratio_in;
    double oldRatio = ratio;
    const double tol = J.tol;
    int i = 0;
    int lastStatus = ST_NO_NEGATIVE_CYCLE;
    while (true) {
      its++;
      bfc++;
      const int st = bellman_ford(ratio, false, false, bid, -1, sweeps, err);
      lastStatus = st;
      if (GRID && !alive) err = ERR_GRID_SYNC;
      if (err != ERR_NONE) break;
      if (st == ST_ZERO_CYCLE || st == ST_NO_NEGATIVE_CYCLE) break;
      ratio = sh.d0;
      if (fabs(oldRatio - ratio) < tol) break;
      if (ratio - tol > oldRatio) break;
      oldRatio = ratio;
      if (i > 1000) break;
      i++;
    }
    return lastStatus;
  }

  __device__ __forceinline__ void write_result(int status, int err, double ratio, double ratio_ub, int numWraps,
                                               int nHits, int bestV, int64_t sweeps, int64_t bfc, int64_t its) {
    PGC_GLOBAL JobResult& r = *gp(J.res);
    r.status = status;
    r.error = err;
    r.ratio = ratio;
    r.ratio_ub = ratio_ub;
    r.num_wraps = numWraps;
    r.num_in_branch = sh.numInBranch;
    r.n_hits = nHits;
    r.neg_len = *gp(J.neg_len);
    r.best_v = bestV;
    r.sweeps = sweeps;
    r.bf_calls = bfc;
    r.lawler_iters = its;
    r.relax = 0;
    r.t_sweep = tSweep;
    r.t_fold = tFold;
    r.t_chase = tChase;
    r.t_walk = tWalk;
    r.t_extract = tExtract;
  }

  __device__ __forceinline__ void lawler() {
    int64_t sweeps = 0, bfc = 0, its = 0;
    int err = ERR_NONE;
    double ratio;
    const int st = lawler_core(J.branchId, J.ratio_in, ratio, sweeps, bfc, its, err);
    if (leader()) {
      int nw_ = 0, nh = 0;
      wraps_and_hits(nw_, nh);
      write_result(st, err, ratio, 0, nw_, nh, -1, sweeps, bfc, its);
      PGC_GLOBAL JobResult& r = *gp(J.res);
      r.branches = 0;
    }
  }

  // findInitalUpperBound (:652-707)
  // Returns lambda_UB (float-rounded) in every thread.
  __device__ __forceinline__ double upper_bound_core(int64_t& sweeps, int& err, int& startV) {
    // start vertex: minimal float mean of the in-edge costs of its layer-1
    // copy (accumulated in ascending edge id), among vertices with > 3 in-edges
    double bestKey = PGC_INF;
    int bestV = 0x7fffffff;
    for (int v = tfirst(); v < npl; v += tstride()) {
      const int node_layer = 1;
      float mean = 0.0f;
      int n = 0;
      if constexpr (W == 8) {
        // the 2 x 8 slots of the vertex's layer-1 copy, loaded once; a class's
        // valid slots are a prefix (padding holds the sentinel npl), the
        // others get key INT_MAX and sort last; then the mean in ascending
        // edge id, as the scan below computes it
        int es[16];
        double cv[16];
        const int ch = v >> 6, ln = v & 63;
#pragma unroll
        for (int cls = 0; cls < 2; ++cls) {
          const auto& S = cls == 0 ? J.intra : J.inter;
          const int base = node_layer * S.lstride + ch * 8 * 64 + ln;
          bool ok = true;
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            const int idx = base + j * 64;
            const int sv = gp(S.src)[idx];
            const int e = gp(S.eid)[idx];
            const double c = (double)gp((const CT*)S.cost)[idx];
            ok = ok && sv >= 0 && sv < npl;
            es[cls * 8 + j] = ok ? e : 0x7fffffff;
            cv[cls * 8 + j] = c;
          }
        }
        sort_pairs<16>(es, cv);
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          if (es[j] != 0x7fffffff) {
            mean = (float)((double)mean + cv[j]);
            n++;
          }
        }
      }
      int lastE = -1;
      while (W != 8) {
        int nextE = 0x7fffffff;
        double nextC = 0;
        for (int cls = 0; cls < 2; ++cls) {
          const auto& S = cls == 0 ? J.intra : J.inter;
          const int ch = v >> 6, ln = v & 63;
          const int width = S.w;
          const int base = node_layer * S.lstride + ch * width * 64;
          for (int j = 0; j < width; ++j) {
            const int idx = base + j * 64 + ln;
            const int sv = gp(S.src)[idx];
            if (sv < 0 || sv >= npl) break;
            const int e = gp(S.eid)[idx];
            if (e > lastE && e < nextE) {
              nextE = e;
              nextC = (double)gp((const CT*)S.cost)[idx];
            }
          }
        }
        if (nextE == 0x7fffffff) break;
        mean = (float)((double)mean + nextC);
        n++;
        lastE = nextE;
      }
      mean = mean / (float)n;
      if (n > 3 && mean < (float)bestKey) {
        // strict '<' in ascending v: keep the first minimiser
        bestKey = (double)mean;
        bestV = v;
      }
    }
    rmin_key(bestKey, bestV);
    const int start = (bestV == 0x7fffffff) ? 0 : bestV;
    startV = start;
    double ratio = 10.0;
    bellman_ford(ratio, true, true, 0, start, sweeps, err);
    // The reference then doubles currentRatio until the cycle's reweighted
    // cost is <= 0 (:690-695) and takes lambda_UB from the last
    // computeRatioForCurrentCycle: cs / ts of the same cycle, which does not
    // depend on the doubled ratio. The loop only decides when to stop, so
    // lambda_UB is that quotient, computed once.
    if (leader()) {
      const int n = *gp(J.neg_len);
      double cs, ts, ws;
      pair_sums(J.neg_c, J.neg_t, n > 0 ? n - 1 : 0, ratio, cs, ts, ws);
      sh.d1 = (double)(float)(cs / ts);
    }
    sync();
    return sh.d1;
  }

  __device__ __forceinline__ void upper_bound() {
    int64_t sweeps = 0;
    int err = ERR_NONE, start = 0;
    const double r_ub = upper_bound_core(sweeps, err, start);
    if (leader()) {
      write_result(ST_NO_NEGATIVE_CYCLE, err, r_ub, r_ub, 0, 0, start, sweeps, 1, 0);
      PGC_GLOBAL JobResult& r = *gp(J.res);
      r.num_in_branch = 1;
      r.branches = 0;
    }
  }

  // Complete lawlergpu solve on the device: upper bound, then the branch and
  // bound of solveWithBnbGPULawler (cost_time_ratio_solver_cuda.cpp:175-331,
  // the CPU twin cost_time_ratio_solver.cpp:794-862) with its queue in global
  // memory, then the output cycle (edge ids, fp64 cost and time sums in walk
  // order) in out_eid / res.
  __device__ __forceinline__ void full_ratio() {
    int64_t sweeps = 0, bfc = 0, its = 0, branches = 0;
    int err = ERR_NONE, start = 0;
    const double rUB = upper_bound_core(sweeps, err, start);
    bfc++;
    // branch ids, queue and valid-branch slots
    for (int v = threadIdx.x; v < npl; v += PGC_BLOCK) {
      gp(J.branch)[v] = 0;
      gp(J.vslot)[v] = -1;
      gp(J.hkey)[v] = -1.0;
      gp(J.hval)[v] = -1;
      gp(J.hwhere)[v] = -1;
    }
    __syncthreads();
    DevHeap heap;
    int maxBranchId = 0, nValid = 0;
    if (threadIdx.x == 0) {
      heap.init(J.hkey, J.hval, J.hwhere, npl);
      heap.push(0.0, 0);
    }
    int finalSlot = -1;  // -1: the current cycle; >= 0: a stored valid branch
    int lastStatus = ST_NO_NEGATIVE_CYCLE, numWraps = 0, nHits = 0;
    double lbBranch = 0;
    while (err == ERR_NONE) {
      // 0 = solve branch i1, 1 = finish with the current cycle, 2 = finish with slot i1
      if (threadIdx.x == 0) {
        if (heap.empty()) {
          sh.i0 = 1;
        } else {
          double k;
          const int bid = heap.pop(&k);
          const int slot = gp(J.vslot)[bid];
          sh.i0 = slot >= 0 ? 2 : 0;
          sh.i1 = slot >= 0 ? slot : bid;
        }
      }
      __syncthreads();
      const int cmd = sh.i0, arg = sh.i1;
      __syncthreads();
      if (cmd == 1) break;
      if (cmd == 2) {
        finalSlot = arg;
        break;
      }
      const int bid = arg;
      if (++branches > 4 * (int64_t)npl + 64) {
        err = ERR_BRANCH_CAP;
        break;
      }
      lastStatus = lawler_core(bid, rUB, lbBranch, sweeps, bfc, its, err);
      if (err != ERR_NONE) break;
      // 0 = continue, 1 = stop with the current cycle, 2 = store the current cycle in slot i1,
      // 3 = relabel branch bid to i1 (peel), 4 = error
      if (threadIdx.x == 0) {
        wraps_and_hits(numWraps, nHits);
        const bool wraps = numWraps > 1;
        if (rUB < heap.peak() && !wraps) {
          sh.i0 = 1;
        } else if (!wraps) {
          heap.push(lbBranch, bid);
          if (nValid >= J.vc_cap) {
            sh.i0 = 4;
          } else {
            gp(J.vslot)[bid] = nValid;
            sh.i0 = 2;
            sh.i1 = nValid++;
          }
        } else if (sh.numInBranch == 1) {
          sh.i0 = 0;
        } else {
          // branch() (:142-179): hits after the first get their own branch
          const int nh = nHits < npl ? nHits : npl;
          int offset = 0;
          for (int h = 0; h < nh; ++h) {
            const int x = gp(J.hits)[h];
            if (offset > 0 && x >= 0 && x < npl) gp(J.branch)[x] = maxBranchId + offset;
            offset++;
          }
          sh.i0 = 3;
          sh.i1 = maxBranchId;
          sh.i2 = offset;
        }
      }
      __syncthreads();
      const int act = sh.i0, a1 = sh.i1, a2 = sh.i2;
      __syncthreads();
      if (act == 1) break;
      if (act == 4) {
        err = ERR_VALID_STORE;
        break;
      }
      if (act == 2) {
        const int m = max(0, *gp(J.neg_len) - 1);
        const int64_t base = (int64_t)a1 * J.vc_stride;
        if (m > J.vc_stride) {
          err = ERR_CYCLE_TOO_LONG;
          break;
        }
        for (int i = threadIdx.x; i < m; i += PGC_BLOCK) {
          gp(J.vc_eid)[base + i] = gp(J.neg_eid)[i];
          gp(J.vc_c)[base + i] = gp(J.neg_c)[i];
          gp(J.vc_t)[base + i] = gp(J.neg_t)[i];
        }
        if (threadIdx.x == 0) gp(J.vc_len)[a1] = m;
        __syncthreads();
        continue;
      }
      if (act == 3) {
        for (int v = threadIdx.x; v < npl; v += PGC_BLOCK)
          if (gp(J.branch)[v] == bid) gp(J.branch)[v] = a1;
        __syncthreads();
        if (threadIdx.x == 0) {
          maxBranchId = a1 + a2;
          const int nh = nHits < npl ? nHits : npl;
          for (int h = 0; h < nh; ++h) {
            const int x = gp(J.hits)[h];
            if (x >= 0 && x < npl) heap.push(lbBranch, gp(J.branch)[x]);
          }
        }
        __syncthreads();
      }
    }
    // output cycle
    const int m = finalSlot >= 0 ? gp(J.vc_len)[finalSlot] : max(0, *gp(J.neg_len) - 1);
    const PGC_GLOBAL int32_t* se = finalSlot >= 0 ? gp(J.vc_eid) + (int64_t)finalSlot * J.vc_stride : gp(J.neg_eid);
    const PGC_GLOBAL double* sc = finalSlot >= 0 ? gp(J.vc_c) + (int64_t)finalSlot * J.vc_stride : gp(J.neg_c);
    const PGC_GLOBAL double* stt = finalSlot >= 0 ? gp(J.vc_t) + (int64_t)finalSlot * J.vc_stride : gp(J.neg_t);
    const bool fits = m <= J.out_cap;
    if (fits)
      for (int i = threadIdx.x; i < m; i += PGC_BLOCK) gp(J.out_eid)[i] = se[i];
    if (threadIdx.x == 0) {
      double cs = 0, ts = 0;
#pragma unroll 8
      for (int i = 0; i < m; ++i) {
        cs += sc[i];
        ts += stt[i];
      }
      if (!fits && err == ERR_NONE) err = ERR_CYCLE_TOO_LONG;
      write_result(lastStatus, err, lbBranch, rUB, numWraps, nHits, start, sweeps, bfc, its);
      PGC_GLOBAL JobResult& r = *gp(J.res);
      r.branches = branches;
      r.out_len = m;
      r.out_cs = cs;
      r.out_ts = ts;
    }
    __syncthreads();
  }

  // one sweep of the minimum-cost solver + lower bounds and paths
  __device__ __forceinline__ void min_cost() {
    int64_t sweeps = 0;
    const int bid = J.branchId;
    const int n = setup(bid, -1);
    int status = ST_CAN_CONTAIN;
    int err = ERR_NONE;
    if (n == 0) {
      err = ERR_EMPTY_BRANCH;
      status = ST_CANNOT_CONTAIN;
    } else {
      sweeps++;
      const uint64_t t0 = wclock();
      if (!run_layers(0.0, J.upper_bound)) status = ST_CANNOT_CONTAIN;
      tSweep += wclock() - t0;
    }
    const uint64_t tc0 = wclock();
    if (status == ST_CAN_CONTAIN) {
      const int lastRow = closing_row();
      const bool cpuRules = (J.flags & FLAG_MINCOST_CPU_RULES) != 0;
      const int stride = J.chain_stride;
      const int64_t maxLen = (int64_t)L * K;
#pragma unroll
      for (int i = 0; i < np(); ++i) {
        const int v = vert(i);
        if (v < 0) continue;
        if (gp(J.branch)[v] != bid) {
          gp(J.lb_key)[v] = PGC_INF;
          gp(J.back)[v] = -1;
          gp(J.wlen)[v] = 0;
          continue;
        }
        double key = dld(D + lastRow * ds + v);
        if (!cpuRules) key = (double)(float)key;
        gp(J.lb_key)[v] = key;
        int back = -1, cnt = 0;
        if (!isinf((float)key)) {
          int64_t p = pget((int64_t)v + (int64_t)npl * L * K);
          if (p != -1) {
            gp(J.chain)[(int64_t)(cnt++) * npl + v] = orig(p);
            for (int64_t kk = maxLen; kk > 0; --kk) {
              const int64_t q = pget(p);
              if (q == -1) break;
              p = q;
              if (cnt < stride) gp(J.chain)[(int64_t)cnt * npl + v] = orig(p);
              cnt++;
            }
            back = orig(p);
          }
        }
        gp(J.back)[v] = back;
        gp(J.wlen)[v] = cnt < stride ? cnt : stride;
      }
    }
    sync();
    tChase += wclock() - tc0;
    if (GRID && !alive) err = ERR_GRID_SYNC;
    if (leader()) {
      PGC_GLOBAL JobResult& r = *gp(J.res);
      r.status = status;
      r.error = err;
      r.ratio = 0;
      r.ratio_ub = 0;
      r.num_wraps = 0;
      r.num_in_branch = sh.numInBranch;
      r.n_hits = 0;
      r.neg_len = 0;
      r.best_v = -1;
      r.sweeps = sweeps;
      r.bf_calls = 1;
      r.lawler_iters = 0;
      r.relax = 0;
      r.t_sweep = tSweep;
      r.t_fold = tFold;
      r.t_chase = tChase;
      r.t_walk = tWalk;
      r.t_extract = tExtract;
    }
  }
};

// One workgroup per job. With a queue counter the launch is persistent:
// every workgroup takes job indices from the counter until all njobs are done
// (jobs are independent; the dynamic LDS fits the largest job of the launch).
template <typename CT, bool HAS_TIME, bool MINCOST, int MAXC, int W>
__global__ void __launch_bounds__(PGC_BLOCK) solve_kernel(const SolverJob* __restrict__ jobs, int njobs,
                                                          int* __restrict__ queue) {
  extern __shared__ double lds[];
  __shared__ Shared sh;
  int next = blockIdx.x;
  while (true) {
    __syncthreads();
    if (threadIdx.x == 0) {
      sh.orCount[0] = 0;
      sh.orCount[1] = 0;
      sh.i3 = queue ? atomicAdd(queue, 1) : next;
    }
    __syncthreads();
    const int j = __builtin_amdgcn_readfirstlane(sh.i3);
    if (j >= njobs) break;
    const PGC_CONST SolverJob& J = ((const PGC_CONST SolverJob*)jobs)[j];
    double* D = MAXC == 0 ? J.dtab : lds;
    double* Dmin = D + 2 * J.K * (J.npl + 1);
    // Small graphs keep pred in LDS after the table. It is loaded from and
    // written back to the global table around the job: predecessors persist
    // across launches as in the reference (with an odd layer count the fold
    // and the lower bounds read a row whose node may not have been relaxed in
    // the current call, and then follow its predecessor from an earlier one).
    int32_t* P = (int32_t*)(lds + (int64_t)(2 * J.K + 1) * (J.npl + 1));
    const bool predLds = MAXC != 0 && J.pred_lds;
    const bool fresh = (J.flags & FLAG_FRESH_JOB) != 0;
    const int64_t nPred = (int64_t)(J.L + 1) * J.K * J.npl;
    if (predLds) {
      if (fresh)
        for (int64_t i = threadIdx.x; i < nPred; i += PGC_BLOCK) ((PGC_LDS int32_t*)P)[i] = -1;
      else
        for (int64_t i = threadIdx.x; i < nPred; i += PGC_BLOCK) ((PGC_LDS int32_t*)P)[i] = gp(J.pred)[i];
    }
    if (fresh && threadIdx.x == 0) *gp(J.neg_len) = 0;
    __syncthreads();
    Solver<CT, HAS_TIME, MINCOST, MAXC, W> s(J, D, Dmin, sh, P);
    if (MINCOST) {
      s.min_cost();
    } else if (J.mode == MODE_UPPER_BOUND) {
      s.upper_bound();
    } else if (J.mode == MODE_FULL_RATIO) {
      s.full_ratio();
    } else {
      s.lawler();
    }
    if (predLds && !fresh) {
      __syncthreads();
      for (int64_t i = threadIdx.x; i < nPred; i += PGC_BLOCK) gp(J.pred)[i] = ((PGC_LDS int32_t*)P)[i];
    }
    if (!queue) break;
    next += gridDim.x;
  }
}

// The device-wide solve of one job (upper bound, Lawler subroutine or
// min-cost sweep) by every workgroup of a cooperative launch; J.grid is
// zeroed before the launch.
template <typename CT, bool HAS_TIME, bool MINCOST, int W>
__global__ void __launch_bounds__(PGC_BLOCK) solve_kernel_grid(const SolverJob* __restrict__ jobs) {
  __shared__ Shared lsh;
  if (threadIdx.x == 0) {
    lsh.orCount[0] = 0;
    lsh.orCount[1] = 0;
    lsh.gridFlag = 0;
    lsh.gridOk = 1;
  }
  __syncthreads();
  const PGC_CONST SolverJob& J = ((const PGC_CONST SolverJob*)jobs)[0];
  GridState* gs = J.grid;
  double* D = J.dtab;
  double* Dmin = D + 2 * J.K * (J.npl + 1);
  Solver<CT, HAS_TIME, MINCOST, 0, W, true> s(J, D, Dmin, gs->sh, nullptr);
  s.gs = gs;
  s.lsh = (PGC_LDS Shared*)&lsh;
  s.nblk = gridDim.x;
  if (MINCOST)
    s.min_cost();
  else if (J.mode == MODE_UPPER_BOUND)
    s.upper_bound();
  else
    s.lawler();
}

// Edge lookups for an arbitrary node list (min-cost path output).
template <typename CT, bool HAS_TIME>
__global__ void lookup_kernel(const SolverJob* __restrict__ jobs, const int32_t* nodes, int n, int32_t* eids) {
  const SolverJob& J = jobs[0];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i + 1 >= n) return;
  int eid = -2;
  double c, t;
  if (!lookup_edge<CT, HAS_TIME>(J, nodes[i], nodes[i + 1], eid, c, t)) eid = -2;
  eids[i] = eid;
}

}  // namespace pgc

