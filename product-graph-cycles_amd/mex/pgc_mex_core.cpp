// The MATLAB front-end's logic over include/pgc.h (see pgc_mex_core.h).
// Reference: mex/product_graph_generator_mex.cpp:29-151, and for the state it
// drives, ProductGraphGenerators (product_graph_generators.cpp:93-333).
#include "pgc_mex_core.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "pgc.h"

namespace pgc_mex {

namespace {

pgc_context* g_ctx = nullptr;
std::mutex g_mu;

double scalar(const Arg& a) { return a.rows * a.cols > 0 ? a.data[0] : 0.0; }

}  // namespace

bool arity_ok(int nrhs, int nlhs) {
  const bool n = nrhs == 8 || nrhs == 9 || nrhs == 10 || nrhs == 12 || nrhs == 14 || nrhs == 17;
  return n && nlhs <= 3;
}

const char* usage() {
  return "Usage: [passes_layers_once, minimum_mean_mean, minimum_mean_matching] = "
         "product_graph_generator_mex(VX, EX, VY, EY, feat_diff, use_conjugate_graph, use_regulariser, "
         "mean_cycle_algo, prune_intralayer, maxDepth, ctCostName, ctTimeName, NX, NY, rl_alpha, rl_c, "
         "rl_gamma)";
}

std::string cost_time_name(int id, std::string* log) {
  static const char* names[] = {"vanilla", "feature", "lengthnormalisation", "plainso3", "robustso3", "tbd"};
  std::string name = id == -1 ? "nomode" : "";
  if (id >= 0 && id <= 5)
    name = names[id];
  else if (log)
    *log += "[CPG-MEX] error: id not supported\n";
  return name;
}

std::string solver_name(int id) {
  switch (id) {
    case 5: return "lawlercpu";
    case 6: return "lawlergpu";
    case 7: return "dijkstracpu";
    case 8: return "dijkstragpu";
    case -1: return "export";
    default: return "hartmannorlin";
  }
}

Route route(int id) {
  if (id > 6) return ROUTE_MIN_COST;
  if (id > 4) return ROUTE_RATIO;
  return ROUTE_NONE;
}

Plan make_plan(int nrhs, const Arg* prhs) {
  Plan p;
  if (nrhs >= 17) {
    p.rl_alpha = scalar(prhs[14]);
    p.rl_c = scalar(prhs[15]);
    p.rl_pwr = scalar(prhs[16]);
  }
  if (nrhs >= 10) {
    p.set_depth = true;
    p.max_depth = (int)scalar(prhs[9]);
  }
  if (nrhs >= 9) {
    p.prune = scalar(prhs[8]) != 0.0;
    p.log += std::string("pruneIntralayerEdges ") + (p.prune ? "1" : "0") + "\n";
  }
  if (nrhs >= 12) {
    p.set_modes = true;
    p.cost_name = cost_time_name((int)scalar(prhs[10]), &p.log);
    p.time_name = cost_time_name((int)scalar(prhs[11]), &p.log);
  }
  p.normals = nrhs >= 14;
  p.conjugate = (int)scalar(prhs[5]);
  p.regulariser = scalar(prhs[6]) != 0.0;
  p.solver_id = (int)scalar(prhs[7]);
  p.solver = solver_name(p.solver_id);
  p.robust = p.conjugate != 0 && nrhs >= 17;
  p.route = route(p.solver_id);
  return p;
}

std::vector<double> to_row_major(const Arg& a) {
  std::vector<double> out((size_t)(a.rows * a.cols));
  for (int64_t r = 0; r < a.rows; ++r)
    for (int64_t c = 0; c < a.cols; ++c) out[(size_t)(r * a.cols + c)] = a.data[c * a.rows + r];
  return out;
}

std::vector<int32_t> ids_to_zero_based(const Arg& a) {
  std::vector<int32_t> out((size_t)(a.rows * a.cols));
  for (int64_t r = 0; r < a.rows; ++r)
    for (int64_t c = 0; c < a.cols; ++c) out[(size_t)(r * a.cols + c)] = (int32_t)a.data[c * a.rows + r] - 1;
  return out;
}

void matching_to_matlab(const int32_t* rows, int64_t n, int cols, double* out) {
  for (int64_t r = 0; r < n; ++r)
    for (int c = 0; c < cols; ++c) out[c * n + r] = (double)rows[r * cols + c] + 1.0;
}

int run(int nrhs, int nlhs, const Arg* prhs, Output* out, std::string* log, std::string* err) {
  if (!arity_ok(nrhs, nlhs)) {
    *err = usage();
    return PGC_ERR_ARG;
  }
  const Plan p = make_plan(nrhs, prhs);
  *log += p.log;
  const Arg &VX = prhs[0], &EX = prhs[1], &VY = prhs[2], &EY = prhs[3], &FD = prhs[4];
  if (VX.cols != 3 || VY.cols != 3 || EX.cols != 2 || EY.cols != 2) {
    *err = "VX and VY must have 3 columns, EX and EY 2";
    return PGC_ERR_ARG;
  }
  if (FD.rows != VX.rows || FD.cols != VY.rows) {
    *err = "feat_diff must be |VX| x |VY|";
    return PGC_ERR_ARG;
  }
  const bool withNormals = p.conjugate != 0 && p.normals;
  if (withNormals && (prhs[12].rows != VX.rows || prhs[12].cols != 3 || prhs[13].rows != VY.rows ||
                      prhs[13].cols != 3)) {
    // the reference prints a shape error and keeps no normals
    // (product_graph_generators.cpp:241-251); a conjugate graph needs them
    *err = "NX must be |VX| x 3 and NY |VY| x 3";
    return PGC_ERR_ARG;
  }
  const std::vector<double> vx = to_row_major(VX), vy = to_row_major(VY), fd = to_row_major(FD);
  const std::vector<int32_t> ex = ids_to_zero_based(EX), ey = ids_to_zero_based(EY);
  std::vector<double> nx, ny;
  if (withNormals) {
    nx = to_row_major(prhs[12]);
    ny = to_row_major(prhs[13]);
  }

  std::lock_guard<std::mutex> lock(g_mu);
  int st = PGC_OK;
  if (!g_ctx && (st = pgc_context_create(0, &g_ctx)) != PGC_OK) {
    *err = pgc_last_error();
    g_ctx = nullptr;
    return st;
  }
  pgc_graph* g = nullptr;
  const double rlA = p.robust ? p.rl_alpha : 0.7, rlC = p.robust ? p.rl_c : 0.6, rlP = p.robust ? p.rl_pwr : 4.0;
  st = pgc_graph_generate(g_ctx, vx.data(), VX.rows, ex.data(), EX.rows, vy.data(), VY.rows, ey.data(), EY.rows,
                          fd.data(), p.conjugate != 0, p.regulariser, p.prune, p.cost_name.c_str(),
                          p.time_name.c_str(), withNormals ? nx.data() : nullptr, withNormals ? ny.data() : nullptr,
                          rlA, rlC, rlP, &g);
  if (st != PGC_OK) {
    *err = pgc_last_error();
    return st;
  }
  *out = Output();
  if (p.route == ROUTE_NONE) {
    *log += "solver not supported\n";
    pgc_graph_destroy(g);
    return PGC_OK;
  }
  int64_t nE = 0, npl = 0, nL = 0;
  int32_t psCols = 0, nCost = 0;
  pgc_graph_info(g, &nE, &psCols, &nCost, &npl, &nL);
  const int K = p.max_depth;
  const int64_t cap = std::max<int64_t>(64, 101 * nL * (int64_t)std::max(K, 1) + 8);
  std::vector<int32_t> eids((size_t)cap), rows((size_t)(cap * psCols));
  pgc_result res;
  std::memset(&res, 0, sizeof(res));
  if (p.route == ROUTE_RATIO) {
    const uint32_t flags = p.solver == "lawlergpu" ? PGC_LAWLER_PICK_LAST : 0u;
    st = pgc_solve_ratio(g_ctx, g, K, flags, eids.data(), rows.data(), cap, &res);
  } else {
    uint32_t flags = 0u;
    if (p.solver != "dijkstragpu") {
      if (p.solver != "dijkstracpu")
        *log += "Solver >>" + p.solver + "<< is not implemented\nFalling back to >> dijkstracpu <<\n";
      flags = PGC_MINCOST_CPU_RULES;
    }
    st = pgc_solve_min_cost(g_ctx, g, K, flags, eids.data(), rows.data(), cap, &res);
  }
  if (st != PGC_OK) {
    *err = pgc_last_error();
    pgc_graph_destroy(g);
    return st;
  }
  out->passes = res.passes_layers_once ? 1.0 : 0.0;
  out->mean = (double)res.objective;
  out->rows = res.n_edges;
  out->cols = psCols;
  out->matching.assign((size_t)(res.n_edges * psCols), 0.0);
  matching_to_matlab(rows.data(), res.n_edges, psCols, out->matching.data());
  pgc_graph_destroy(g);
  return PGC_OK;
}

void shutdown() {
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_ctx) pgc_context_destroy(g_ctx);
  g_ctx = nullptr;
}

}  // namespace pgc_mex

namespace {

void put(const std::string& s, char* buf, int64_t cap) {
  if (!buf || cap <= 0) return;
  const size_t n = std::min<size_t>(s.size(), (size_t)cap - 1);
  std::memcpy(buf, s.data(), n);
  buf[n] = 0;
}

std::string json_str(const std::string& s) {
  std::string o = "\"";
  for (char c : s) {
    if (c == '"' || c == '\\') o += '\\';
    if (c == '\n') {
      o += "\\n";
      continue;
    }
    o += c;
  }
  return o + "\"";
}

}  // namespace

extern "C" {

int pgc_mex_arity_ok(int nrhs, int nlhs) { return pgc_mex::arity_ok(nrhs, nlhs) ? 1 : 0; }

int pgc_mex_plan_json(int nrhs, const double* scalars, char* buf, int64_t cap) {
  if (nrhs < 8 || nrhs > 17 || !scalars) return PGC_ERR_ARG;
  std::vector<pgc_mex::Arg> a((size_t)nrhs);
  for (int i = 0; i < nrhs; ++i) a[(size_t)i] = pgc_mex::Arg{scalars + i, 1, 1};
  const pgc_mex::Plan p = pgc_mex::make_plan(nrhs, a.data());
  char num[256];
  std::snprintf(num, sizeof(num),
                "\"conjugate\": %d, \"regulariser\": %d, \"prune\": %d, \"set_depth\": %d, \"max_depth\": %d, "
                "\"set_modes\": %d, \"normals\": %d, \"robust\": %d, \"rl\": [%.17g, %.17g, %.17g], "
                "\"solver_id\": %d, \"route\": %d",
                p.conjugate, (int)p.regulariser, (int)p.prune, (int)p.set_depth, p.max_depth, (int)p.set_modes,
                (int)p.normals, (int)p.robust, p.rl_alpha, p.rl_c, p.rl_pwr, p.solver_id, (int)p.route);
  const std::string s = std::string("{") + num + ", \"cost_name\": " + json_str(p.cost_name) +
                        ", \"time_name\": " + json_str(p.time_name) + ", \"solver\": " + json_str(p.solver) +
                        ", \"log\": " + json_str(p.log) + "}";
  if ((int64_t)s.size() + 1 > cap) return PGC_ERR_CAPACITY;
  put(s, buf, cap);
  return PGC_OK;
}

void pgc_mex_ids_to_zero_based(const double* colmajor, int64_t rows, int64_t cols, int32_t* out) {
  const std::vector<int32_t> v = pgc_mex::ids_to_zero_based(pgc_mex::Arg{colmajor, rows, cols});
  std::copy(v.begin(), v.end(), out);
}

void pgc_mex_matching_to_matlab(const int32_t* rows, int64_t n, int32_t cols, double* out) {
  pgc_mex::matching_to_matlab(rows, n, cols, out);
}

int pgc_mex_call(int nrhs, int nlhs, const double* const* data, const int64_t* rows, const int64_t* cols,
                 double* passes, double* mean, double* matching, int64_t cap, int64_t* m_rows, int64_t* m_cols,
                 char* log, int64_t log_cap, char* err, int64_t err_cap) {
  if (nrhs < 0 || nrhs > 17 || (nrhs > 0 && (!data || !rows || !cols))) return PGC_ERR_ARG;
  std::vector<pgc_mex::Arg> a((size_t)nrhs);
  for (int i = 0; i < nrhs; ++i) a[(size_t)i] = pgc_mex::Arg{data[i], rows[i], cols[i]};
  pgc_mex::Output o;
  std::string lg, er;
  const int st = pgc_mex::run(nrhs, nlhs, a.data(), &o, &lg, &er);
  put(lg, log, log_cap);
  put(er, err, err_cap);
  if (st != PGC_OK) return st;
  if ((int64_t)o.matching.size() > cap) {
    put("matching does not fit the output capacity", err, err_cap);
    return PGC_ERR_CAPACITY;
  }
  *passes = o.passes;
  *mean = o.mean;
  *m_rows = o.rows;
  *m_cols = o.cols;
  std::copy(o.matching.begin(), o.matching.end(), matching);
  return PGC_OK;
}

}  // extern "C"
