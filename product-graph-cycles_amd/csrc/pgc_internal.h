// Internal host/device structures of the MI355X product-graph cycle solvers.
//
// Device layout of a layered product graph (one "job" = one graph):
//   * node (layer l, vertex v) has id l*npl + v; layer L (= numNodes/npl) is
//     the copy of layer 0 that closes the cylinder (the reference's dummy
//     nodes, cost_time_ratio_solver.cpp:42-66);
//   * the in-edges of every node are split into two classes, intralayer
//     (source on the same layer) and interlayer (source on the previous
//     layer), and stored per layer as sliced ELL with 64-vertex slices (one
//     wave per slice): every slice of a class has w slots per lane (the
//     class's maximum in-degree) and slot j of lane i of slice c on layer l
//     sits at l*stride + (c*w + j)*64 + i, so a wave's loads are contiguous
//     and all addressing is arithmetic on wave-uniform values;
//   * inside a vertex the in-edges are ordered by (source id, edge id), the
//     documented deterministic replacement for the reference's hash-set
//     iteration order; padding slots carry src = npl (a sentinel whose
//     distance is +inf in the solver) and eid = -1.
#pragma once
#include <stddef.h>
#include <stdint.h>

#define PGC_WAVE 64
#define PGC_BLOCK 1024

namespace pgc {

// Largest slot width of the in-edge layout: nodes with up to this many
// in-edges of one class (the runtime-width solver variant above 32).
constexpr int kMaxSlotWidth = 1024;

// Bytes of the static device scratch of the allocation-free solve (per device).
constexpr size_t kScratchBytes = (size_t)128 << 20;

struct SlotClass {
  const int32_t* src;    // source vertex index inside its layer; padding = npl (sentinel)
  const void* cost;      // float (q-units, exact) or double (raw)
  const float* time;     // nullptr: constant time
  const int32_t* eid;    // product-graph edge id (-1 for padding)
  int32_t w;             // slots per lane of every chunk (the class's maximum in-degree)
  int32_t lstride;       // slots per layer = nChunks * w * 64
  int32_t src_inv;       // 1: the source slots repeat on every layer: the intralayer class on
                         // layers 0..L-1 (layer L has no intralayer edges), the interlayer
                         // class on layers 1..L (the same X adjacency on every layer)
  int32_t deg;           // the largest in-degree of the class: slots deg..w-1 are padding on
                         // every node
  const int32_t* grp;    // slot widths above 32 (runtime width): per chunk, the groups of eight
                         // slots that hold its largest in-degree; nullptr otherwise
};

// Conjugate energy (cost, time) kinds of the device kernel.
enum ConjKind : int32_t { CONJ_VV, CONJ_VL, CONJ_LV, CONJ_VP, CONJ_PV, CONJ_FSP, CONJ_FSR, CONJ_SRF, CONJ_SPF };

enum JobMode : int32_t { MODE_RATIO = 0, MODE_UPPER_BOUND = 1, MODE_MIN_COST = 2, MODE_FULL_RATIO = 3 };

enum Status : int32_t {
  ST_CAN_CONTAIN = 0,
  ST_CANNOT_CONTAIN = 1,
  ST_NO_NEGATIVE_CYCLE = 2,
  ST_FOUND_NEGATIVE_CYCLE = 3,
  ST_ZERO_CYCLE = 4,
};

enum JobError : int32_t {
  ERR_NONE = 0,
  ERR_EMPTY_BRANCH = 1,
  ERR_CYCLE_TOO_LONG = 2,
  ERR_EDGE_LOOKUP = 3,
  ERR_CHAIN = 4,
  ERR_SWEEP_CAP = 5,
  ERR_VALID_STORE = 6,  // more valid branches than the job's cycle store holds
  ERR_BRANCH_CAP = 7,   // safety cap on branches of one solve (4*npl + 64)
  ERR_GRID_SYNC = 8,    // a device-wide barrier waited past its time limit
};

struct GridState;

// flags
enum : uint32_t {
  FLAG_PICK_LAST_VERTEX = 1u,  // candidate pick of the reference's GPU path (last reporting vertex)
  FLAG_MINCOST_CPU_RULES = 2u, // min-cost: double lower-bound keys
  FLAG_CORRECTED_ODD_L = 4u,   // odd L, K >= 2: the closing layer's depth-0 row (not row L % 2)
  FLAG_FRESH_JOB = 1u << 16,   // the solve's only launch: the kernel clears neg_len and an LDS pred table
  FLAG_NO_CACHED_SOURCES = 1u << 17,  // tuning: load the source slots every step (no register copy)
  FLAG_FULL_SLOTS = 1u << 18,         // tuning: read every slot of the layout's width (no short instances)
};

struct JobResult {
  int32_t status;
  int32_t error;
  double ratio;         // lambda at the end of the Lawler loop (lowerBoundOfCurrentBranch)
  double ratio_ub;      // upper-bound mode: lambda_UB (float-rounded)
  int32_t num_wraps;
  int32_t num_in_branch;
  int32_t n_hits;
  int32_t neg_len;      // nodes in the current cycle
  int32_t best_v;       // upper-bound mode: start vertex
  int32_t pad;
  int64_t sweeps;
  int64_t bf_calls;
  int64_t lawler_iters;
  int64_t relax;        // candidate relaxations (finite source value)
  // device wall-clock ticks (wall_clock64, 100 MHz) spent per phase
  int64_t t_sweep, t_fold, t_chase, t_walk, t_extract;
  // full solve (MODE_FULL_RATIO)
  int64_t branches;
  double out_cs, out_ts;  // fp64 sums of the output cycle's cost and time (walk order)
  int32_t out_len;        // edges of the output cycle (in out_eid)
  int32_t pad2;
};

struct SolverJob {
  int32_t npl, L, K, nChunks;
  int64_t numNodes;
  int32_t mode;
  uint32_t flags;
  int32_t branchId;
  int32_t chain_stride;  // L*K + 2
  int32_t cyc_cap;       // capacity of cycle buffers (nodes)
  int32_t pad0;
  double ratio_in;
  double upper_bound;
  double time_const;
  double tol;
  int64_t max_sweeps;     // safety cap on sweeps per launch (error when reached)
  SlotClass intra, inter;
  int32_t* branch;        // [npl] branch id per layer-0 vertex
  int32_t* pred;          // [(L+1)*K*npl] predecessor (expanded node id)
  // per-sweep scratch
  double* backup;         // [npl] layer-0 values at the start of a sweep
  int32_t* wnext;         // [npl]
  int32_t* wlen;          // [npl]
  int32_t* chain;         // [chain_stride*npl] original node ids along each wrap; entry j of v at j*npl+v
  int32_t* tmp_nodes;     // [cyc_cap]
  int32_t* tmp_eid;       // [cyc_cap]
  double* tmp_c;          // [cyc_cap]
  double* tmp_t;          // [cyc_cap]
  // persistent current cycle (the reference's negativeCycle)
  int32_t* neg_nodes;     // [cyc_cap]
  int32_t* neg_eid;       // [cyc_cap]
  double* neg_c;          // [cyc_cap]
  double* neg_t;          // [cyc_cap]
  int32_t* neg_len;       // [1]
  int32_t* hits;          // [npl] zero-layer hits
  // min-cost outputs
  double* lb_key;         // [npl]
  int32_t* back;          // [npl] start vertex of the path ending at (L, v)
  JobResult* res;
  // device-resident branch and bound (MODE_FULL_RATIO)
  double* hkey;           // [npl] queue slots
  int32_t* hval;          // [npl]
  int32_t* hwhere;        // [npl]
  int32_t* vslot;         // [npl] cycle-store slot of a valid branch (-1: none)
  int32_t* vc_eid;        // [vc_cap * vc_stride] cycles of valid branches
  double* vc_c;
  double* vc_t;
  int32_t* vc_len;        // [vc_cap]
  int32_t* out_eid;       // [out_cap] output cycle
  int32_t vc_cap, vc_stride, out_cap;
  int32_t pred_lds;       // 1: pred lives in LDS after the distance table (small graphs)
  int32_t wrap_lds;       // 1: the wrap map (wnext, wlen) lives in LDS after the table and pred
  int32_t org_lds;        // 1 (K == 1, pred and wrap map in LDS): two rows of chain origins in LDS
                          // after the wrap map, kept by the sweep, replace the wrap chase
  double* dtab;           // [(2K+1)*(npl+1)] distance table in global memory (graphs beyond the LDS)
  GridState* grid;        // barrier and broadcast state of a device-wide solve (cooperative launch)
  // implicit costs (max_depth 1, org_lds): the sweep recomputes every slot's
  // q-cost from feat_diff instead of loading it; nullptr: the slot costs
  const double* icol;     // [nVY][npl] feat_diff (float64), column y of Y vertex y contiguous
  const int32_t* iey;     // [L][2] Y edges: layer l's interlayer edges go from column iey[2(l-1)]
                          // (sources) to column iey[2(l-1)+1] (targets)
  const double* iscale;   // [1] the quantization scale s of the graph's energy
};

// One cycle of the raw-unit ratio value / gradient kernels.
struct CycleItem {
  const int32_t* rows;  // E x 4 product-space rows
  const double* vx;
  const double* vy;
  const double* fd;     // forward: the item's feat_diff
  double* dfd;          // backward: gradient accumulator (same shape)
  const int32_t* eids;  // m edge ids
  int64_t nVY;
  int32_t m, line, length, pad;
  double gout;          // backward: upstream gradient
};

// One int32 array of a batched fill: p[0, n) <- the launch's value.
struct FillItem {
  int32_t* p;
  int64_t n;
};

// One column of a batched quantization. Its median bracket's candidates go
// to cand[cand_off, cand_off + cand_cap) of the batch's candidate buffer.
struct QItem {
  double* col;
  int64_t n;
  int64_t cand_off;
  int64_t cand_cap;
};

// Keys sampled per column, and the bracket's margin in sample positions on
// each side of the median's expected position (about four standard
// deviations of the sample rank of the median).
constexpr int kMedianSample = 8192;
constexpr int kMedianMargin = 192;

// Candidate capacity of a column of n keys: the bracket holds about
// (2 * margin + 1) / sample = 4.7% of the keys; a column whose bracket
// overflows this falls back to the full radix select.
inline int64_t median_cand_cap(int64_t n) {
  if (n <= kMedianSample) return 0;
  const int64_t c = n / 8 > 16384 ? n / 8 : 16384;
  return c < n ? c : n;
}

// One slot class of a batched fp32 cost fill.
// One graph of the batched energy kernel (k_benergy).
struct EItem {
  const double* vx;
  const double* vy;
  const int32_t* rows;
  int64_t E;
  const void* fd;  // [nVX x nVY] feat_diff, float64 or float32 (f32)
  int64_t nVY;
  int32_t line, lengthMode, f32;
  double* energy;
  // the X and Y edge lists of the product space (non-conjugate): an edge's
  // row is decoded from its id instead of read from rows (nullptr: read it)
  const int32_t* ex;
  const int32_t* ey;
  int64_t nEX, nEY, nVX;
  int32_t prune;
};

struct FItem {
  const int32_t* eid;
  const int32_t* s;
  int64_t nslots;
  const double* energy;
  int64_t E;
  float* cost;
  float* time;  // nullptr without a time column
};

}  // namespace pgc
