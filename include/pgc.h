/*
 * pgc.h - C-ABI of the MI355X product-graph cycle solvers (libpgc.so).
 *
 * Drop-in boundary for the reference's GPU solver seam. Plain pointers and
 * sizes only; every entry point returns PGC_OK (0) or a negative status and
 * never exits the process (the reference calls exit(-1) on an unknown solver
 * name, cost_time_ratio_solver.cpp:896-899). pgc_last_error() describes the
 * last failure of the calling thread.
 *
 * Threading: one context per host thread; a context owns one HIP stream on
 * one device. Inputs are caller-owned and copied; device memory is owned by
 * the context/graph objects.
 */
#ifndef PGC_H_
#define PGC_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pgc_context pgc_context;
typedef struct pgc_graph pgc_graph;

#define PGC_OK 0
#define PGC_ERR_ARG (-1)
#define PGC_ERR_HIP (-2)
#define PGC_ERR_SOLVER (-3)
#define PGC_ERR_CAPACITY (-4)
#define PGC_ERR_UNSUPPORTED (-5)

/* Solver rule flags. lawlergpu / dijkstragpu are the defaults (0 for
 * dijkstragpu); the CPU solvers' rules are selectable for parity studies. */
#define PGC_LAWLER_PICK_LAST 1u  /* negative-cycle candidate of the reference's GPU path:
                                    last reporting vertex (ctr_bnb_subroutine.cu:489-500);
                                    without it: lowest (D, vertex) (cost_time_ratio_solver.cpp:430-437) */
#define PGC_MINCOST_CPU_RULES 2u /* dijkstracpu: double lower-bound keys, branching seed
                                    cycle.front()-numNodes (dijkstra_solver.cpp:318-381) */
#define PGC_CORRECTED_ODD_L 4u   /* odd layer count with max_depth >= 2: fold (ratio) and
                                    lower bounds (min-cost) read the closing layer's depth-0
                                    row, and its depth rows 1..K-1 become layer 0's, as they
                                    do for even L. Without it: the reference's row L % 2
                                    (cost_time_ratio_solver.cpp:400, dijkstra_solver.cpp:318) */

typedef struct pgc_result {
  int32_t passes_layers_once; /* the reference's passes_layers_just_once */
  float objective;            /* ratio in q-units (Lawler) or raw cost (min-cost), as fp32 */
  int64_t n_edges;            /* edges of the returned cycle */
  int64_t sweeps;             /* Bellman-Ford sweeps over all layers */
  int64_t bf_calls;           /* Bellman-Ford calls */
  int64_t lawler_iters;       /* Lawler iterations (0 for min-cost) */
  int64_t branches;           /* branch-and-bound branches explored */
  double relaxations;         /* edge relaxations attempted (E_x per sweep, summed) */
  double kernel_ms;           /* device time of the solver kernels (HIP events) */
  double total_ms;            /* wall time of the call */
  double phase_ms[5];         /* device time inside the solver kernels by phase: layer sweeps,
                                 closing-layer fold, wrap chase, cycle walks, cycle extraction */
  int32_t status;             /* PGC_OK, or this solve's own error (batched solves: per graph) */
  int32_t launches;           /* solver kernel launches of the call (batched solves: 0 per graph) */
  int32_t implicit_costs;     /* 1: the sweeps recomputed the slot costs from feat_diff
                                 (pgc_tuning.implicit_costs), 0: they loaded them */
  int32_t reserved;
} pgc_result;

/* Launch choices of a context's solves. Every field defaults to -1, the
 * built-in rule; the autotuner (cyclic_product_graphs/autotune.py) times the
 * alternatives on a workload and keeps the fastest. None changes a result. */
typedef struct pgc_tuning {
  int32_t grid_chunks_per_wave; /* graphs whose distance table exceeds the LDS: 0 solves them in one
                                   workgroup, n >= 1 in one cooperative launch with n 64-node chunks
                                   per wave; -1: one chunk per wave from 128 chunks per layer, else
                                   one workgroup */
  int32_t pred_in_lds;          /* predecessor table in LDS when it fits: -1 yes, 0 no */
  int32_t wrap_in_lds;          /* wrap map in LDS when it fits: -1 yes, 0 no */
  int32_t origins;              /* max_depth 1: rows of chain origins in LDS instead of the wrap
                                   chase, when they fit: -1 yes, 0 no */
  int32_t cached_sources;       /* layer-invariant source slots held in registers: -1 yes, 0 no */
  int32_t short_slots;          /* skip slots that are padding on every node: -1 yes, 0 no */
  int32_t global_table;         /* 1: every graph keeps its distance table in global memory (a
                                   validation choice: the same results); other values: the rule */
  int32_t grid_workgroups;      /* n >= 1: graphs with the table in global memory are solved by n
                                   workgroups of one cooperative launch (at most what the device
                                   holds at once); other values: grid_chunks_per_wave decides */
  int32_t implicit_costs;       /* 1: graphs generated in this context keep their feat_diff, and a
                                   max_depth-1 solve whose sweep holds its rows, predecessors and
                                   chain origins in LDS recomputes every slot's q-cost from it
                                   instead of loading it (same results); other values: off */
} pgc_tuning;

const char* pgc_last_error(void);
int pgc_abi_version(void);

/* Context: device ordinal + one stream. */
int pgc_context_create(int device, pgc_context** out);
void pgc_context_destroy(pgc_context* ctx);

/* The context's launch choices (NULL: the built-in rules) and a copy of them.
 * No counterpart in the reference (its launch geometry is fixed,
 * cost_time_ratio_solver_cuda.cpp:219-231). */
int pgc_context_set_tuning(pgc_context* ctx, const pgc_tuning* tuning);
int pgc_context_get_tuning(const pgc_context* ctx, pgc_tuning* out);

/* Builds the layered product graph and its per-edge cost/time on the device.
 * Replaces ProductGraphGenerators::generate() / generateConjugate()
 * (product_graph_generators.cpp:21-89), i.e. ProductSpace / ConjProductSpace
 * (product_spaces/product_space.cpp:19-132, conj_product_space.cpp:70-228) and
 * DeformationEnergy / HigherOrderDeformationEnergy
 * (energy/deformationEnergy.cpp:11-103, higherOrderDeformationEnergy.cpp:209-339).
 * VX [nVX x 3], VY [nVY x 3], feat_diff [nVX x nVY] row-major float64; EX
 * [nEX x 2], EY [nEY x 2] int32. NX/NY [n x 3] (conjugate only, else NULL). */
int pgc_graph_generate(pgc_context* ctx, const double* VX, int64_t nVX, const int32_t* EX, int64_t nEX,
                       const double* VY, int64_t nVY, const int32_t* EY, int64_t nEY, const double* feat_diff,
                       int conjugate, int regularising, int prune, const char* cost_mode, const char* time_mode,
                       const double* NX, const double* NY, double rl_alpha, double rl_c, double rl_pwr,
                       pgc_graph** out);

/* Same as pgc_graph_generate with every array already in device memory of the
 * context's device (e.g. torch tensors); the caller orders its own writes to
 * these buffers before the call. */
int pgc_graph_generate_device(pgc_context* ctx, const double* VX, int64_t nVX, const int32_t* EX, int64_t nEX,
                              const double* VY, int64_t nVY, const int32_t* EY, int64_t nEY,
                              const double* feat_diff, int conjugate, int regularising, int prune,
                              const char* cost_mode, const char* time_mode, const double* NX, const double* NY,
                              double rl_alpha, double rl_c, double rl_pwr, pgc_graph** out);

/* Takes the solver inputs of the reference's seam as host arrays: the
 * arguments of CostTimeRatioSolver / DijkstraSolver
 * (cost_time_ratio_solver.hpp:107, dijkstra_solver.hpp:83): productspace
 * [n_edges x ps_cols] row-major, SRC/TRGT [n_edges], cost [n_edges x n_cols]
 * column-major (Eigen layout), numNodesPerLayer. */
int pgc_graph_from_arrays(pgc_context* ctx, const int32_t* productspace, int ps_cols, const int32_t* src,
                          const int32_t* trgt, int64_t n_edges, const double* cost, int n_cols,
                          int64_t nodes_per_layer, pgc_graph** out);

/* Optional branch graph of the min-cost branching (the reference's
 * branchGraph, product_space.cpp:122-131, conj_product_space.cpp:177-225):
 * CSR over layer-0 nodes, values as the reference stores them. */
int pgc_graph_set_branch_graph(pgc_graph* g, const int64_t* offsets, const int64_t* values);

int pgc_graph_info(const pgc_graph* g, int64_t* n_edges, int32_t* ps_cols, int32_t* n_cost_cols,
                   int64_t* nodes_per_layer, int64_t* n_layers);
int pgc_graph_get_product_space(pgc_graph* g, int32_t* out);   /* n_edges x ps_cols row-major */
int pgc_graph_get_cost(pgc_graph* g, double* out);             /* n_edges x n_cost_cols column-major */
int pgc_graph_set_cost(pgc_graph* g, const double* cost);      /* same layout; replaces the energy */
void pgc_graph_destroy(pgc_graph* g);

/* The solver's in-edge layout of one class (0 intralayer, 1 interlayer):
 * the device form of the reference's VECSET in-edge sets
 * (ctr_bnb_subroutine.cuh:232-287, built by buildingInGraph,
 * cost_time_ratio_solver.cpp:42-66), for inspection. Builds it if needed.
 * Writes *n_slots and *width (slots per lane); if src and eid are not NULL
 * and capacity >= *n_slots, copies the slots (slot j of lane i of chunk c on
 * layer l at l*chunks*width*64 + (c*width + j)*64 + i): the source's index
 * inside its layer (npl for padding) and the edge id (-1 for padding). */
int pgc_graph_get_layout(pgc_graph* g, int cls, int32_t* src, int32_t* eid, int64_t capacity, int64_t* n_slots,
                         int32_t* width);

/* Lawler minimum cost/time-ratio cycle with branch and bound ("lawlergpu").
 * Replaces CostTimeRatioSolver::run -> solveWithBnbGPULawler
 * (cost_time_ratio_solver.cpp:868-900, cost_time_ratio_solver_cuda.cpp:175-331).
 * Quantizes the graph's cost in place to q-units, as the reference does
 * (cost_time_ratio_solver.cpp:875-882). Writes the cycle's edge ids (walk
 * order) and, if rows != NULL, their product-space rows. */
int pgc_solve_ratio(pgc_context* ctx, pgc_graph* g, int max_depth, uint32_t flags, int32_t* edge_ids,
                    int32_t* rows, int64_t capacity, pgc_result* res);

/* Complete "lawlergpu" solves of n graphs, each exactly as pgc_solve_ratio
 * (same quantization in place, same cycle, objective and counters), with the
 * whole solve of a graph (upper bound, branch and bound, Lawler iterations)
 * inside one workgroup and all graphs in persistent launches (one per kernel
 * variant) that take graphs from a device queue. Batched form of the solve loop
 * over CostTimeRatioSolver::run (cost_time_ratio_solver.cpp:868-900); the
 * reference has no batched entry point. max_depth[i] is graph i's depth; graph
 * i's cycle goes to edge_ids[offsets[i] .. offsets[i+1]) and, if rows != NULL,
 * its product-space rows to rows + (offsets[i] - offsets[0]) * ps_cols (all
 * graphs must then have the same ps_cols). A graph may appear once per batch.
 * Argument errors and graphs that cannot be solved fail the call before any
 * energy is quantized. A graph whose device branch and bound outgrows its
 * store of valid-branch cycles (64) is solved again by the host branch and
 * bound of pgc_solve_ratio on the same quantized costs (same result). Every
 * res[i].status holds graph i's own outcome; the call returns PGC_OK when all
 * are PGC_OK, else the first failing graph's status, and the graphs that
 * solved still have their cycles and results written. */
int pgc_solve_ratio_batch(pgc_context* ctx, pgc_graph* const* graphs, const int32_t* max_depth, int64_t n,
                          uint32_t flags, int32_t* edge_ids, int32_t* rows, const int64_t* offsets,
                          pgc_result* res);

/* pgc_solve_ratio without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). No counterpart in the reference. Same quantization in
 * place, cycle, objective and counters as pgc_solve_ratio; the whole solve
 * runs in one launch (the device-resident branch and bound), and every buffer
 * beyond the graph's own arrays and the caller's outputs lives in the
 * library's static device scratch: 128 MiB per device, held from the first
 * context on that device until the process ends (graphs up to about the size
 * of an icosphere of subdivision 4 against a 128-point curve at max_depth 3).
 * A solve that needs more returns PGC_ERR_CAPACITY. The scratch is shared by
 * all calls on a device: calls on one device are serialized by the library
 * (without that they would race), and the scratch keeps the last call's
 * contents; every call rebuilds what it reads. */
int pgc_solve_ratio_noalloc(pgc_context* ctx, pgc_graph* g, int max_depth, uint32_t flags, int32_t* edge_ids,
                            int32_t* rows, int64_t capacity, pgc_result* res);

/* Recomputes the graph's per-edge energy from another feat_diff ([nVX x nVY]
 * row-major float64, device memory when on_device != 0, else host) with the
 * graph's construction modes; the product space and layout are unchanged.
 * The analogue of building a new generator for perturbed features
 * (product_graph_generators.cpp:21-89) without rebuilding the graph. */
int pgc_graph_set_feat_diff(pgc_graph* g, const double* feat_diff, int on_device);

/* pgc_graph_set_feat_diff for n graphs of one context in one launch:
 * feat_diff[i] is graph i's [nVX x nVY] row-major feat_diff in device memory,
 * float64 (dtype 0) or float32 (dtype 1, widened exactly). No counterpart in
 * the reference (batched perturbation samples). */
int pgc_graphs_set_feat_diff(pgc_graph* const* graphs, const void* const* feat_diff, int64_t n, int dtype);

/* A new graph that shares base's product space, ids and in-edge layout
 * (read-only) and owns its energy (from feat_diff, as above), quantized costs
 * and solver scratch: the per-sample graphs of a perturbation batch. base may
 * be destroyed first; its shared memory is freed with its last derived graph. */
int pgc_graph_derive(pgc_graph* base, const double* feat_diff, int on_device, pgc_graph** out);

/* Cycle ratio in raw energy units, lambda = sum_e c_e / sum_e t_e over edge_ids[
 * offsets[i] .. offsets[i] + lengths[i]) of graph i (walk order, fp64), with c_e
 * the raw edge energy (deformationEnergy.cpp:11-103 as computed on the device)
 * for feat_diff[i] (device pointer, [nVX x nVY] float64) and t_e the time column
 * (1 per edge without one). No counterpart in the reference (the value whose
 * gradient the training interface needs). */
int pgc_cycle_ratio_values(pgc_context* ctx, pgc_graph* const* graphs, int64_t n, const double* const* feat_diff,
                           const int32_t* edge_ids, const int64_t* offsets, const int64_t* lengths, double* values);

/* Adds grad_out[i] * d lambda_i / d feat_diff into dfeat_diff[i] (device
 * pointers, [nVX x nVY] float64) for the same cycles: every cycle edge adds
 * dc_e/dFD / T(C) at its two (x, y) entries. No counterpart in the reference. */
int pgc_cycle_ratio_grad(pgc_context* ctx, pgc_graph* const* graphs, int64_t n, const int32_t* edge_ids,
                         const int64_t* offsets, const int64_t* lengths, const double* grad_out,
                         double* const* dfeat_diff);

/* Minimum-cost once-wrapping cycle with branch and bound ("dijkstragpu").
 * Replaces DijkstraSolver::run -> solveWithBnBCuda
 * (dijkstra_solver.cpp:563-579, dijkstra_solver_cuda.cpp:136-322). */
int pgc_solve_min_cost(pgc_context* ctx, pgc_graph* g, int max_depth, uint32_t flags, int32_t* edge_ids,
                       int32_t* rows, int64_t capacity, pgc_result* res);

#ifdef __cplusplus
}
#endif

#endif /* PGC_H_ */
