// The MATLAB front-end of the solvers, without MATLAB: everything the MEX
// gateway product_graph_generator_mex does between reading its right-hand
// sides and writing its left-hand sides, over the C-ABI of include/pgc.h.
// Reference: mex/product_graph_generator_mex.cpp:29-151.
//
// The gateway (product_graph_generator_mex.cpp beside this file) only wraps
// mxArrays as Arg views and copies the Output into new mxArrays; the argument
// rules (accepted nrhs, 1-based ids, the cost/time and solver id tables, the
// dispatch on the solver id) live here, in plain C++ that builds and is
// tested without MATLAB.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

namespace pgc_mex {

// A right-hand side as MATLAB holds it: a real double matrix, column-major.
struct Arg {
  const double* data;
  int64_t rows, cols;
};

enum Route : int { ROUTE_NONE = 0, ROUTE_RATIO = 1, ROUTE_MIN_COST = 2 };

// What the gateway does with one call, from its scalar arguments.
struct Plan {
  int conjugate = 0;        // prhs[5], (int)
  bool regulariser = false; // prhs[6], (bool)
  bool prune = false;       // prhs[8] when nrhs >= 9
  bool set_depth = false;   // nrhs >= 10
  int max_depth = 4;        // prhs[9]; the generator's default otherwise
  bool set_modes = false;   // nrhs >= 12
  std::string cost_name = "vanilla", time_name = "";
  bool normals = false;     // nrhs >= 14: NX = prhs[12], NY = prhs[13]
  bool robust = false;      // nrhs >= 17 (and a conjugate graph)
  double rl_alpha = 0.7, rl_c = 0.6, rl_pwr = 4.0;
  int solver_id = 0;        // prhs[7]
  std::string solver;
  Route route = ROUTE_NONE;
  std::string log;          // the lines the reference prints on std::cout
};

// nrhs in {8, 9, 10, 12, 14, 17} and nlhs <= 3 (mex/product_graph_generator_mex.cpp:31-33).
bool arity_ok(int nrhs, int nlhs);
const char* usage();

// Cost/time id -> mode name (:8-27): -1 nomode, 0 vanilla, 1 feature,
// 2 lengthnormalisation, 3 plainso3, 4 robustso3, 5 tbd. As in the reference,
// -1 also prints the "id not supported" line (its if/else chain starts over
// at 0), and an unknown id gives "" with that line.
std::string cost_time_name(int id, std::string* log);

// Solver id -> name (:86-98): 5 lawlercpu, 6 lawlergpu, 7 dijkstracpu,
// 8 dijkstragpu, -1 export, anything else hartmannorlin.
std::string solver_name(int id);

// Dispatch (:123-131): id > 6 min-cost, id > 4 ratio, else none (the
// reference prints "solver not supported" and returns false, 0, an empty
// matching; so "export" (-1) is never reached through the gateway).
Route route(int id);

Plan make_plan(int nrhs, const Arg* prhs);

// Column-major double matrix -> row-major double (VX, VY, feat_diff, normals).
std::vector<double> to_row_major(const Arg& a);
// Column-major 1-based double ids -> row-major 0-based int32 (EX, EY): the
// reference's cast<int>() (truncation) then minus one (:40-52).
std::vector<int32_t> ids_to_zero_based(const Arg& a);
// Row-major int32 matching [n x cols] -> column-major 1-based doubles (:140-150).
void matching_to_matlab(const int32_t* rows, int64_t n, int cols, double* out);

struct Output {
  double passes = 0.0;
  double mean = 0.0;  // the solver's fp32 objective, widened
  int64_t rows = 0, cols = 0;
  std::vector<double> matching;  // column-major, 1-based
};

// The whole call: parse, generate, set the modes, solve, convert. Returns
// PGC_OK or a negative PGC_ERR_* with *err set; *log gets what the reference
// prints. One library context on device 0 is kept across calls.
int run(int nrhs, int nlhs, const Arg* prhs, Output* out, std::string* log, std::string* err);

// Releases the kept context (the gateway registers it with mexAtExit).
void shutdown();

}  // namespace pgc_mex

// C entry points of the same logic, for tests through ctypes.
extern "C" {
int pgc_mex_arity_ok(int nrhs, int nlhs);
// scalars[i] stands for a 1x1 prhs[i] (i < nrhs); writes the plan as JSON.
int pgc_mex_plan_json(int nrhs, const double* scalars, char* buf, int64_t cap);
void pgc_mex_ids_to_zero_based(const double* colmajor, int64_t rows, int64_t cols, int32_t* out);
void pgc_mex_matching_to_matlab(const int32_t* rows, int64_t n, int32_t cols, double* out);
// The whole call on host arrays: data[i] / rows[i] / cols[i] describe prhs[i].
// matching gets the column-major result (capacity cap doubles; PGC_ERR_CAPACITY
// if it does not fit); log and err get NUL-terminated text (truncated to cap).
int pgc_mex_call(int nrhs, int nlhs, const double* const* data, const int64_t* rows, const int64_t* cols,
                 double* passes, double* mean, double* matching, int64_t cap, int64_t* m_rows, int64_t* m_cols,
                 char* log, int64_t log_cap, char* err, int64_t err_cap);
}
