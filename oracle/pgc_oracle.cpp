// TEST INFRASTRUCTURE ONLY (oracle).
//
// CPU restatement of the reference's product-graph cycle path, used as the
// parity checker for the HIP implementation and as the CPU baseline that the
// benchmark times beside it. Only tests/, the smoke check and the benchmark's
// CPU-baseline leg load this library; the product never links it.
//
// Reference files restated here (paths relative to the reference root):
//   product_graph_generators/product_spaces/product_space.cpp:19-132   (orc_product_space)
//   product_graph_generators/energy/deformationEnergy.cpp:11-103       (orc_energy)
//   product_graph_generators/cost_time_ratio_solver/cost_time_ratio_solver.cpp:42-905
//       + ctr_bnb_subroutine.cu:231-276,455-504 (GPU candidate pick)    (orc_ctr_solve)
//   product_graph_generators/dijkstra_solver/dijkstra_solver.cpp:41-579
//       + dijkstra_solver_cuda.cpp:30-322, dijkstra_bnb_subroutine.cu:366-451 (orc_dijkstra_solve)
//   helper/minheap.cpp:16-277 (minheap_oracle.hpp), helper/utils.hpp:196-209 (median)
//
// Deterministic choices where the reference depends on hash-set iteration
// order or on OpenMP races (tsl::robin_set order of in-edges, racy candidate
// walks): in-edges of a node are visited in ascending (source id, edge id)
// order; the initial-upper-bound mean is accumulated in ascending edge id.
//
// Pinning: product space and energy against closed forms; quantization
// against numpy's median and round-half-away; both solvers against brute-force
// enumeration of once-wrapping cycles on tiny graphs (tests/test_oracle.py).
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off -fopenmp -shared).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "minheap_oracle.hpp"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace oracle {

static const double kInf = std::numeric_limits<double>::infinity();

// std::round followed by the implicit double->int conversion of the
// reference (`const int x = std::round(w)`); x86-64 cvttsd2si yields INT_MIN
// for values outside the int range.
static inline int round_to_int(double w) {
  const double r = std::round(w);
  if (!(r >= -2147483648.0 && r < 2147483648.0)) return std::numeric_limits<int>::min();
  return (int)r;
}

// getOriginalGraphIndex, cost_time_ratio_solver.cpp:240-250
static inline int orig_index(long npl, long K, long pred) {
  const long per = npl * K;
  const long layer = pred / per;
  const long rem = pred - layer * per;
  const long depth = rem / npl;
  const long v = rem - depth * npl;
  return (int)(v + layer * npl);
}

// utils::median, helper/utils.hpp:196-209
static double median_of(std::vector<double> v) {
  std::sort(v.begin(), v.end());
  const size_t n = v.size();
  if (n % 2 == 0) return (v[(n - 2) / 2] + v[(n - 2) / 2 + 1]) / 2.0;
  return v[n / 2];
}

// robustFloatToIntConversion, cost_time_ratio_solver.cpp:709-721
static void quantize_column(double* col, long n, double scaling) {
  std::vector<double> copy(col, col + n);
  const double med = std::max(median_of(copy), 0.001);
  const double s = scaling / (2.0 * med);
  const double maxVal = 100 * scaling;
  for (long i = 0; i < n; ++i) {
    double q = std::round(s * col[i]);
    if (q > maxVal) q = maxVal;
    col[i] = q;
  }
}

// Per-node in-edge lists (buildingInGraph, cost_time_ratio_solver.cpp:42-66):
// edges whose target is on layer 0 and whose source id is > npl are moved to
// the dummy copy numNodes + target.
struct InGraph {
  std::vector<long> off;
  std::vector<long> edges;  // sorted by (SRC, e) per node
  void build(const int32_t* SRC, const int32_t* TRGT, long E, long npl, long numNodes) {
    const long nTot = numNodes + npl;
    std::vector<long> cnt(nTot + 1, 0);
    auto tgt = [&](long e) -> long {
      long v = TRGT[e];
      if (v < npl && SRC[e] > npl) v = numNodes + v;
      return v;
    };
    for (long e = 0; e < E; ++e) cnt[tgt(e) + 1]++;
    for (long i = 0; i < nTot; ++i) cnt[i + 1] += cnt[i];
    off = cnt;
    edges.assign(E, 0);
    std::vector<long> pos(cnt.begin(), cnt.end() - 1);
    for (long e = 0; e < E; ++e) edges[pos[tgt(e)]++] = e;
    for (long v = 0; v < nTot; ++v) {
      std::sort(edges.begin() + off[v], edges.begin() + off[v + 1], [&](long a, long b) {
        if (SRC[a] != SRC[b]) return SRC[a] < SRC[b];
        return a < b;
      });
    }
  }
};

enum BFStatus { NO_NEGATIVE_CYCLE = 2, FOUND_NEGATIVE_CYCLE = 3, ZERO_CYCLE = 4 };

struct Stats {
  long bf_calls = 0, sweeps = 0, lawler_iters = 0, branches = 0, relaxations = 0;
  // Events of the ratio branch and bound at which the reference's algorithm
  // can end above the once-wrapping optimum (counted, never changed):
  long q_push_id = 0;       // a push with a branch id outside [0, npl): the heap ignores it
                            // (helper/minheap.cpp:137-139; branch ids grow past npl, Q9)
  long q_push_present = 0;  // a push of a branch id already queued: ignored, its key is kept
  long q_single = 0;        // a multi-wrap branch with one vertex is dropped
                            // (cost_time_ratio_solver_cuda.cpp:294-297)
  long q_rise = 0;          // Lawler stopped on a rising ratio (:620-623): its cycle is not the
                            // branch's best one found
  long q_stale = 0;         // a non-root branch's first Bellman-Ford call found no negative
                            // cycle: the branch keeps lambda_UB and the previous branch's cycle
  long q_cap = 0;           // a Bellman-Ford call stopped at the sweep cap numNodes - 1
  long q_walk = 0;          // the first closing candidate walk gave a cycle whose reweighted
                            // cost is not negative, and the call stopped there (:506-518)
  long q_empty = 0;         // the queue ran empty: the output is the last Lawler cycle
  long q_maxiter = 0;       // Lawler's iteration cap (:624-627)
  long q_ub_break = 0;      // the lambda_UB < peak stop (:282-286), lambda_UB never lowered (Q8)
  long valid_pops = 0;      // the stop on a popped branch whose stored cycle is valid
};

// ---------------------------------------------------------------------------
// Shared layered DP sweep (the body of mooreBellmanFordSubroutineCPU,
// cost_time_ratio_solver.cpp:298-368 and dijkstra_solver.cpp:247-311).
// Returns false when a layer 0<l<L had no improvement.
// ---------------------------------------------------------------------------
struct Layered {
  long npl, L, K, numNodes, E;
  const int32_t* SRC;
  const double* c;  // cost column
  const double* t;  // time column (nullptr: weight = cost)
  double ratio = 0;
  double upperBound = kInf;
  InGraph g;
  std::vector<double> D;  // (2K+1) x npl, row-major
  std::vector<int32_t> pred;
  long relax_count = 0;

  double& d(long row, long v) { return D[row * npl + v]; }

  bool sweep_layers() {
    bool improved = true;
    for (long layer = 0; layer < L + 1; ++layer) {
      const long prevOff = ((layer + 1) % 2) * K;
      const long curOff = (layer % 2) * K;
      if (layer != 0) {
        std::fill(D.begin() + curOff * npl, D.begin() + (curOff + K) * npl, kInf);
        improved = false;
      }
      for (long k = 0; k < K; ++k) {
        if (layer == 0 && k == 0) continue;
        int any = 0;
        long rc = 0;
#pragma omp parallel for reduction(| : any) reduction(+ : rc) schedule(static)
        for (long vi = 0; vi < npl; ++vi) {
          const long v = vi + layer * npl * K + k * npl;
          const long actualV = vi + layer * npl;
          for (long p = g.off[actualV]; p < g.off[actualV + 1]; ++p) {
            const long e = g.edges[p];
            const double w = t ? (c[e] - ratio * t[e]) : c[e];
            const long src = SRC[e];
            const bool intra = (src / npl) == layer;
            if (intra && k != 0) {
              if (K <= 1) continue;
              const long sl = src - layer * npl;
              const double dv = D[(curOff + k - 1) * npl + sl];
              if (std::isinf(dv)) continue;
              ++rc;
              const double nv = dv + w;
              double& cur = D[(curOff + k) * npl + vi];
              if (nv < cur && nv < upperBound) {
                any = 1;
                pred[v] = (int32_t)(sl + layer * npl * K + (k - 1) * npl);
                cur = nv;
              }
            }
            if (!intra && k == 0) {
              const long sl = src - (layer - 1) * npl;
              if (sl < 0 || sl >= npl) continue;
              for (long kk = 0; kk < K; ++kk) {
                const double dv = D[(prevOff + kk) * npl + sl];
                if (std::isinf(dv)) continue;
                ++rc;
                const double nv = dv + w;
                double& cur = D[curOff * npl + vi];
                if (nv < cur && nv < upperBound) {
                  any = 1;
                  pred[v] = (int32_t)(sl + (layer - 1) * npl * K + kk * npl);
                  cur = nv;
                }
              }
            }
          }
        }
        relax_count += rc;
        if (any) improved = true;
      }
      if (!improved && layer < L && layer > 0) return false;
    }
    return true;
  }

  // Edge joining consecutive cycle nodes (computeRatioForCycle lookup,
  // cost_time_ratio_solver.cpp:109-122): first in-edge of trgt with SRC==src.
  long find_edge(long src, long trgt) const {
    if (trgt < npl && src > npl) trgt += numNodes;
    if (trgt < 0 || trgt >= numNodes + npl) return -2;
    for (long p = g.off[trgt]; p < g.off[trgt + 1]; ++p)
      if (SRC[g.edges[p]] == src) return g.edges[p];
    return -2;
  }
};

// ---------------------------------------------------------------------------
// Lawler cost/time-ratio solver with branch and bound
// ---------------------------------------------------------------------------
struct CtrSolver {
  Layered S;
  std::vector<double> cost, time;
  std::vector<int32_t> branch;
  int branchId = 0, maxBranchId = 0, numInBranch = 0, indexNodeInLayer = -1;
  double tol;
  std::vector<int> negCycle, hits;
  double lbCurrentBranch = 0;
  bool gpu_pick = false;
  bool corrected = false;  // fold the closing layer's depth-0 row for odd L (Q1 fixed)
  Stats st;

  struct Ratio {
    double ratio, w;
    bool ok;
  };
  // computeRatioForCycle, cost_time_ratio_solver.cpp:88-133
  Ratio ratio_of(const std::vector<int>& cyc, std::vector<long>* edges_out = nullptr) const {
    double cs = 0, ts = 0, ws = 0;
    bool ok = true;
    for (size_t i = 0; i + 1 < cyc.size(); ++i) {
      const long e = S.find_edge(cyc[i], cyc[i + 1]);
      if (e < 0) {
        ok = false;
        continue;
      }
      cs += cost[e];
      ts += time[e];
      ws += cost[e] - S.ratio * time[e];
      if (edges_out) edges_out->push_back(e);
    }
    return {cs / ts, ws, ok};
  }

  BFStatus zero_or_none() {
    const Ratio r = ratio_of(negCycle);
    return round_to_int(r.w) == 0 ? ZERO_CYCLE : NO_NEGATIVE_CYCLE;
  }

  // mooreBellmanFordSubroutineCPU, cost_time_ratio_solver.cpp:251-542
  BFStatus bellman_ford(bool searchForZeroCycle, bool searchForUpperBound) {
    const long npl = S.npl, K = S.K, L = S.L;
    st.bf_calls++;
    std::fill(S.D.begin(), S.D.end(), kInf);
    numInBranch = 0;
    for (long i = 0; i < npl; ++i) {
      if (branch[i] == branchId) {
        indexNodeInLayer = (int)i;
        numInBranch++;
        S.d(0, i) = 0;
        S.pred[i] = -1;
      }
    }
    if (numInBranch == 0) throw std::runtime_error("no nodes in branch");
    const long backupRow = 2 * K;
    long iter = 0;
    std::vector<char> seen(npl);
    while (true) {
      std::copy(S.D.begin(), S.D.begin() + npl, S.D.begin() + backupRow * npl);
      st.sweeps++;
      if (!S.sweep_layers()) return zero_or_none();

      if (searchForUpperBound) {
        const long maxLength = L * K;
        std::vector<int> cyc;
        const long firstV = indexNodeInLayer + npl * L * K;
        cyc.push_back(orig_index(npl, K, firstV));
        long p = S.pred[firstV];
        cyc.push_back(orig_index(npl, K, p));
        for (long kk = maxLength - 1; kk > 0; --kk) {
          if (p == -1) break;
          p = S.pred[p];
          cyc.push_back(orig_index(npl, K, p));
          if (p == indexNodeInLayer) break;
        }
        std::reverse(cyc.begin(), cyc.end());
        negCycle = cyc;
        const int w = round_to_int(ratio_of(cyc).w);
        if (w == 0) return ZERO_CYCLE;
        return w < 0 ? FOUND_NEGATIVE_CYCLE : NO_NEGATIVE_CYCLE;
      }

      // fold of the dummy layer into layer 0 (:399-417). The reference reads
      // row L % 2, which is the closing layer's depth-0 row only for even L
      // or K == 1 (Q1); `corrected` reads that row, (L % 2) * K. With even L
      // the closing layer's rows are layer 0's rows, so its depths 1..K-1
      // become layer 0's; `corrected` copies them for odd L likewise.
      const long lastRow = corrected ? (L % 2) * K : L % 2;
      if (corrected && L % 2 == 1)
        for (long k = 1; k < K; ++k)
          for (long vi = 0; vi < npl; ++vi) S.d(k, vi) = S.d(K + k, vi);
      bool any = false;
      for (long vi = 0; vi < npl; ++vi) {
        if (branch[vi] != branchId) continue;
        const long vLast = npl * K * L + vi;
        if (S.d(lastRow, vi) < S.d(backupRow, vi)) {
          any = true;
          S.d(0, vi) = S.d(lastRow, vi);
          S.pred[vi] = S.pred[vLast];
        } else {
          S.d(0, vi) = S.d(backupRow, vi);
        }
      }
      if (!any) return zero_or_none();

      // candidate order: ascending (D, v) (CPU rule, :430-437) or the last
      // reporting vertex (GPU rule, ctr_bnb_subroutine.cu:489-500)
      std::vector<std::pair<double, long>> cand;
      for (long vi = 0; vi < npl; ++vi)
        if (branch[vi] == branchId) cand.push_back({S.d(0, vi), vi});
      if (gpu_pick)
        std::sort(cand.begin(), cand.end(), [](auto& a, auto& b) { return a.second > b.second; });
      else
        std::sort(cand.begin(), cand.end());
      long vStar = -1, cycLen = -1;
      const long maxCycleLength = S.numNodes;
      for (auto& cd : cand) {
        const long vi = cd.second;
        const double lb = S.d(0, vi);
        if (!searchForZeroCycle && lb >= 0) continue;
        if (searchForZeroCycle && (lb <= -tol || lb >= tol)) continue;
        std::fill(seen.begin(), seen.end(), 0);
        long p = S.pred[vi];
        seen[vi] = 1;
        long cl = 1;
        bool found = false;
        for (long kk = maxCycleLength - 1; kk > 0; --kk) {
          if (p < npl && p >= 0) {
            if (seen[p]) {
              found = true;
              vStar = p;
              cycLen = cl;
              break;
            }
            seen[p] = 1;
          }
          p = S.pred[p];
          if (p == -1) break;
          cl++;
        }
        if (found) break;
      }
      if (vStar >= 0) {
        std::vector<int> cyc;
        long p = S.pred[vStar];
        cyc.push_back(orig_index(npl, K, p));
        for (long kk = cycLen - 1; kk > 0; --kk) {
          p = S.pred[p];
          cyc.push_back(orig_index(npl, K, p));
          if (p == vStar) break;
        }
        std::reverse(cyc.begin(), cyc.end());
        cyc.push_back(cyc.front() + (int)S.numNodes);
        const int w = round_to_int(ratio_of(cyc).w);
        if (w < 0) {
          negCycle = cyc;
          return FOUND_NEGATIVE_CYCLE;
        }
        st.q_walk++;
        // a zero-cost cycle leaves the sweep loop through its condition and
        // reaches the final NO_NEGATIVE_CYCLE return (:506-518, :282)
        if (w == 0) return NO_NEGATIVE_CYCLE;
        return zero_or_none();
      }
      iter++;
      if (iter >= S.numNodes - 1) {
        st.q_cap++;
        return NO_NEGATIVE_CYCLE;
      }
    }
  }

  // lawlerSoubroutine, cost_time_ratio_solver.cpp:558-635
  bool lawler() {
    int i = 0;
    double oldRatio = S.ratio;
    while (true) {
      st.lawler_iters++;
      const BFStatus s = bellman_ford(false, false);
      if (s == ZERO_CYCLE || s == NO_NEGATIVE_CYCLE) {
        if (i == 0 && s == NO_NEGATIVE_CYCLE && branchId != 0) st.q_stale++;
        break;
      }
      S.ratio = ratio_of(negCycle).ratio;
      if (std::fabs(oldRatio - S.ratio) < tol) break;
      if (S.ratio - tol > oldRatio) {
        st.q_rise++;
        break;
      }
      oldRatio = S.ratio;
      if (i > 1000) {
        st.q_maxiter++;
        break;
      }
      i++;
    }
    lbCurrentBranch = S.ratio;
    int numWraps = 0;
    hits.clear();
    bool entered = true, left = false;
    for (int vtx : negCycle) {
      if (left && vtx < S.npl) {
        entered = true;
        left = false;
      }
      if (entered) {
        hits.push_back(vtx);
        entered = false;
      }
      if (vtx >= S.npl && !left) {
        numWraps++;
        left = true;
      }
    }
    return numWraps > 1;
  }

  // findInitalUpperBound, cost_time_ratio_solver.cpp:652-707
  void initial_upper_bound() {
    const long npl = S.npl;
    float best = std::numeric_limits<float>::infinity();
    int bestV = 0;
    for (long v = 0; v < npl; ++v) {
      const long node = v + npl;
      std::vector<long> ids(S.g.edges.begin() + S.g.off[node], S.g.edges.begin() + S.g.off[node + 1]);
      std::sort(ids.begin(), ids.end());
      float mean = 0;
      int n = 0;
      for (long e : ids) {
        mean += cost[e];
        n++;
      }
      mean /= n;
      if (mean < best && n > 3) {
        best = mean;
        bestV = (int)v;
      }
    }
    S.upperBound = (double)std::numeric_limits<float>::infinity();
    S.ratio = 10;
    std::fill(branch.begin(), branch.end(), 1);
    branch[bestV] = 0;
    branchId = 0;
    bellman_ford(true, true);
    Ratio r = ratio_of(negCycle);
    double cycleCost = (double)(float)r.w;
    while (cycleCost > 0 || std::isnan(cycleCost)) {
      S.ratio += S.ratio;
      r = ratio_of(negCycle);
      cycleCost = (double)(float)r.w;
    }
    S.ratio = (double)(float)r.ratio;
    std::fill(branch.begin(), branch.end(), 0);
  }

  // branch(), cost_time_ratio_solver.cpp:142-179 (the code after its return is dead)
  void branch_split() {
    int offset = 0;
    for (int h : hits) {
      if (offset > 0 && h >= 0 && h < S.npl) branch[h] = maxBranchId + offset;
      offset++;
    }
    for (long v = 0; v < S.npl; ++v)
      if (branch[v] == branchId) branch[v] = maxBranchId;
    maxBranchId += offset;
  }

  // solveWithBnbCPULawler, cost_time_ratio_solver.cpp:741-864
  bool solve(std::vector<long>& out_edges, float& out_obj) {
    const long npl = S.npl;
    initial_upper_bound();
    branchId = 0;
    double ratioUpperBound = S.ratio;
    BnbHeap heap(npl);
    heap.push(0, branchId);
    std::vector<char> valid(npl, 0);
    std::map<int, std::vector<int>> validCycles;
    std::fill(branch.begin(), branch.end(), 0);
    // the reference's push, with the cases it ignores counted
    auto push = [&](double key, long id) {
      if (id < 0 || id >= npl)
        st.q_push_id++;
      else if (heap.contains(id))
        st.q_push_present++;
      heap.push(key, id);
    };
    bool stopped = false;
    while (!heap.empty()) {
      const auto top = heap.pop();
      branchId = (int)top.second;
      if (valid[branchId]) {
        ratioUpperBound = top.first;
        negCycle = validCycles.at(branchId);
        st.valid_pops++;
        stopped = true;
        break;
      }
      st.branches++;
      S.ratio = ratioUpperBound;
      const bool wraps = lawler();
      if (ratioUpperBound < heap.peak() && !wraps) {
        st.q_ub_break++;
        stopped = true;
        break;
      }
      if (!wraps) {
        // the reference pushes ratioLowerBoundOfCurrentBranch, a member it
        // never assigns (cost_time_ratio_solver_cuda.cpp:288,
        // cost_time_ratio_solver.cpp:835; declared at
        // cost_time_ratio_solver.hpp:51): the branch's Lawler lambda here
        push(lbCurrentBranch, branchId);
        valid[branchId] = 1;
        validCycles.emplace(branchId, negCycle);
        continue;
      }
      if (numInBranch == 1) {
        st.q_single++;
        continue;
      }
      branch_split();
      for (int h : hits) {
        if (h >= 0 && h < npl) push(lbCurrentBranch, branch[h]);
      }
    }
    if (!stopped) st.q_empty++;
    out_edges.clear();
    Ratio r = ratio_of(negCycle, &out_edges);
    out_obj = (float)r.ratio;
    return r.ok;
  }
};

// ---------------------------------------------------------------------------
// Minimum-cost once-wrapping cycle ("dijkstra") with branch and bound
// ---------------------------------------------------------------------------
struct MinCostSolver {
  Layered S;
  std::vector<int32_t> branch;
  int branchId = 0, numInBranch = 0;
  std::vector<int> cycle;  // best valid path (walk order reversed, as in the reference)
  double lbCurrentBranch = 0;
  bool foundFirst = false;
  int bSrc = -1, bTrgt = -1;
  std::vector<std::pair<double, long>> lbs;
  bool gpu_rules = true;
  bool corrected = false;  // lower bounds from the closing layer's depth-0 row for odd L (Q1 fixed)
  // branch graph: CSR of values (edge ids, or vertex ids used as edge ids)
  std::vector<long> bg_off, bg_val;
  Stats st;

  // mooreBellmanFordSubroutineCPU (dijkstra_solver.cpp:203-387); lower-bound
  // keys are float-rounded and ties go to the lower vertex id
  // (dijkstra_bnb_subroutine.cu:173-195,391-393)
  bool bellman_ford() {
    const long npl = S.npl, K = S.K, L = S.L;
    st.bf_calls++;
    std::fill(S.D.begin(), S.D.end(), kInf);
    numInBranch = 0;
    for (long i = 0; i < npl; ++i) {
      if (branch[i] == branchId) {
        numInBranch++;
        S.d(0, i) = 0;
        S.pred[i] = -1;
      }
    }
    if (numInBranch == 0) throw std::runtime_error("no nodes in branch");
    std::copy(S.D.begin(), S.D.begin() + npl, S.D.begin() + 2 * K * npl);
    st.sweeps++;
    if (!S.sweep_layers()) return false;
    const long lastRow = corrected ? (L % 2) * K : L % 2;  // Q1, as in the ratio fold
    lbs.clear();
    bool anyFinite = false;
    for (long vi = 0; vi < npl; ++vi) {
      if (branch[vi] != branchId) continue;
      double key = S.d(lastRow, vi);
      if (gpu_rules) key = (double)(float)key;
      lbs.push_back({key, vi});
      if (!std::isinf(key)) anyFinite = true;
    }
    if (!anyFinite) return false;
    std::sort(lbs.begin(), lbs.end());
    bool first = true;
    const long maxCycleLength = L * K;
    foundFirst = false;
    for (auto& lb : lbs) {
      const float lowerBound = (float)lb.first;
      if (std::isinf(lowerBound)) {
        first = false;
        continue;
      }
      const long v = lb.second;
      const long vLast = v + npl * L * K;
      std::vector<int> cyc;
      long p = S.pred[vLast];
      int back = -1;
      if (p != -1) {
        cyc.push_back(orig_index(npl, K, p));
        for (long kk = maxCycleLength; kk > 0; --kk) {
          p = S.pred[p];
          if (p == -1) break;
          cyc.push_back(orig_index(npl, K, p));
        }
        back = cyc.back();
      }
      if (first) lbCurrentBranch = lowerBound;
      if (v == back) {
        if (first) foundFirst = true;
        const float newUB = lowerBound;
        if (newUB > S.upperBound && first) return false;
        if (newUB < S.upperBound) {
          S.upperBound = newUB;
          cycle = cyc;
        }
        break;
      } else if (first) {
        bSrc = gpu_rules ? (int)v : (cyc.empty() ? -1 : cyc.front() - (int)S.numNodes);
        bTrgt = back;
      }
      first = false;
    }
    return true;
  }

  // branch(), dijkstra_solver.cpp:87-174 (GPU seed: dijkstra_solver_cuda.cpp:30-117)
  void branch_split(int maxBranchId) {
    const long npl = S.npl;
    BnbHeap heap(npl);
    std::vector<double> tbl(npl, -1);
    for (long i = 0; i < npl; ++i) {
      tbl[i] = -1;
      if (i == bSrc || i == bTrgt) {
        tbl[i] = 1;
        continue;
      }
      if (branch[i] == branchId) {
        heap.push(kInf, i);
        tbl[i] = -1;
      }
    }
    heap.push(0.0, bSrc);
    heap.push(0.0, bTrgt);
    if (bTrgt >= 0 && bTrgt < npl) branch[bTrgt] = maxBranchId;
    int counter = 0;
    for (auto& lb : lbs) {
      const long v = lb.second;
      if (v == bSrc || v == bTrgt) continue;
      if (branch[v] != branchId) continue;
      if (counter > std::max(1, numInBranch / 4)) {
        heap.push(0.1, v);
        branch[v] = maxBranchId;
      }
      counter++;
    }
    while (!heap.empty()) {
      const auto cur = heap.pop();
      const long node = cur.second;
      const float curCost = (float)cur.first;
      tbl[node] = 1.0;
      for (long p = bg_off[node]; p < bg_off[node + 1]; ++p) {
        const long e = bg_val[p];
        if (e == -1) continue;
        if (e < 0 || e >= S.E) continue;
        const long s = S.SRC[e];
        if (s < 0 || s >= npl) continue;
        if (tbl[s] < 0) {
          const float oldV = (float)heap.key_value(s);
          const float newV = curCost + 1.0;
          if (newV < oldV) {
            heap.decrease(s, newV);
            branch[s] = branch[node];
          }
        }
      }
    }
    bSrc = -1;
    bTrgt = -1;
  }

  // solveWithBnB / solveWithBnBCuda, dijkstra_solver.cpp:407-559
  bool solve(std::vector<long>& out_edges, float& out_obj, bool& passes) {
    const long npl = S.npl;
    S.upperBound = kInf;
    std::fill(branch.begin(), branch.end(), 0);
    BnbHeap heap(npl);
    branchId = 0;
    heap.push(0.0, branchId);
    int maxBranchId = 0;
    std::vector<char> valid(npl, 0);
    std::map<int, std::vector<int>> validCycles;
    while (!heap.empty()) {
      const auto top = heap.pop();
      branchId = (int)top.second;
      if (valid[branchId]) {
        S.upperBound = top.first;
        cycle = validCycles.at(branchId);
        break;
      }
      st.branches++;
      if (!bellman_ford()) continue;
      if (S.upperBound < heap.peak() && foundFirst) break;
      if (foundFirst) {
        heap.push(lbCurrentBranch, branchId);
        valid[branchId] = 1;
        validCycles.emplace(branchId, cycle);
        continue;
      }
      maxBranchId++;
      branch_split(maxBranchId);
      heap.push(lbCurrentBranch, branchId);
      heap.push(lbCurrentBranch, maxBranchId);
    }
    // path output, dijkstra_solver.cpp:511-557
    out_edges.clear();
    const long cycleLength = (long)cycle.size();
    std::vector<int> cyc(cycle.rbegin(), cycle.rend());
    if (!cyc.empty()) cyc.push_back(cyc.front() + (int)S.numNodes);
    const long loopmax = cycleLength + 1;
    int loops = 0;
    for (long i = 0; i < loopmax && !cyc.empty(); ++i) {
      const long s = cyc[i % loopmax];
      const long tr = cyc[(i + 1) % loopmax];
      if (tr >= S.numNodes) loops++;
      if (tr == s - S.numNodes) continue;  // dummy closing step (sentinel -1 edge)
      long e = -2;
      if (tr >= 0 && tr < S.numNodes + npl)
        for (long p = S.g.off[tr]; p < S.g.off[tr + 1]; ++p)
          if (S.SRC[S.g.edges[p]] == s) {
            e = S.g.edges[p];
            break;
          }
      if (e == -2) continue;
      out_edges.push_back(e);
    }
    passes = loops < 2;
    out_obj = (float)S.upperBound;
    return true;
  }
};

static void setup_layered(Layered& S, const int32_t* SRC, const int32_t* TRGT, long E, long npl, long K) {
  long mx = 0;
  for (long e = 0; e < E; ++e) mx = std::max<long>(mx, SRC[e]);
  S.numNodes = mx + 1;
  S.L = S.numNodes / npl;
  S.npl = npl;
  S.K = K;
  S.E = E;
  S.SRC = SRC;
  S.g.build(SRC, TRGT, E, npl, S.numNodes);
  S.D.assign((2 * K + 1) * npl, kInf);
  S.pred.assign(npl * (S.L + 1) * K, -1);
}

static thread_local std::string g_err;

}  // namespace oracle

using namespace oracle;

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

// Non-conjugate product space, product_space.cpp:19-132.
// Row layout [y_src, y_trgt, x_src, x_trgt]; rows/src/trgt must hold
// nEY*(2*nEX+nVX) entries (nEY*(nEX+nVX) with pruning).
long orc_product_space(const int32_t* EX, long nEX, const int32_t* EY, long nEY, int prune, int32_t* rows,
                       int32_t* src, int32_t* trgt) {
  long nVX = 0;
  for (long i = 0; i < 2 * nEX; ++i) nVX = std::max<long>(nVX, EX[i] + 1);
  long n = 0;
  for (long i = 0; i < nEY; ++i) {
    const long t = (i + 1) % nEY;
    for (long j = 0; j < nEX; ++j) {
      const int x0 = EX[2 * j], x1 = EX[2 * j + 1];
      if (!prune) {
        int32_t* r = rows + 4 * n;
        r[0] = EY[2 * i];
        r[1] = EY[2 * i];
        r[2] = x0;
        r[3] = x1;
        src[n] = (int32_t)(i * nVX + x0);
        trgt[n] = (int32_t)(i * nVX + x1);
        n++;
      }
      int32_t* r = rows + 4 * n;
      r[0] = EY[2 * i];
      r[1] = EY[2 * i + 1];
      r[2] = x0;
      r[3] = x1;
      src[n] = (int32_t)(i * nVX + x0);
      trgt[n] = (int32_t)(t * nVX + x1);
      n++;
    }
    for (long x = 0; x < nVX; ++x) {
      int32_t* r = rows + 4 * n;
      r[0] = EY[2 * i];
      r[1] = EY[2 * i + 1];
      r[2] = (int32_t)x;
      r[3] = (int32_t)x;
      src[n] = (int32_t)(i * nVX + x);
      trgt[n] = (int32_t)(t * nVX + x);
      n++;
    }
  }
  return n;
}

// Non-conjugate energy, deformationEnergy.cpp:11-103. FD is |VX| x |VY|
// row-major; out is column-major E x ncols. Returns ncols.
int orc_energy(const double* VX, long nVX, const double* VY, long nVY, const int32_t* rows, long E, const double* FD,
               int reg, const char* cost_mode, const char* time_mode, double* out) {
  (void)nVX;
  const bool lengthMode = std::strcmp(cost_mode, "feature") == 0 && std::strcmp(time_mode, "lengthnormalisation") == 0;
  const bool lineIntegral = reg || lengthMode;
  const int ncols = lengthMode ? 2 : 1;
  for (long i = 0; i < E; ++i) {
    const long y0 = rows[4 * i], y1 = rows[4 * i + 1], x0 = rows[4 * i + 2], x1 = rows[4 * i + 3];
    double len = 1.0;
    if (lineIntegral) {
      len = 0;
      for (int j = 0; j < 3; ++j) {
        len += std::pow(VY[3 * y0 + j] - VY[3 * y1 + j], 2);
        len += std::pow(VX[3 * x0 + j] - VX[3 * x1 + j], 2);
      }
      len = std::sqrt(len);
    }
    const double fd = FD[x0 * nVY + y0] + FD[x1 * nVY + y1];
    if (!lengthMode) {
      out[i] = 0.5 * fd * len;
    } else {
      out[i] = fd;
      out[E + i] = len;
    }
  }
  return ncols;
}

double orc_median(const double* x, long n) { return median_of(std::vector<double>(x, x + n)); }

// In-place q-unit quantization of every column (run(), :875-882).
void orc_quantize(double* cost, long E, int ncols) {
  for (int c = 0; c < ncols; ++c) quantize_column(cost + (long)c * E, E, 100000.0);
}

// Lawler ratio solve. cost is column-major E x ncols, quantized in place.
// flags bit0: candidate pick rule of the GPU reference (last reporting
// vertex) instead of the CPU order; bit1: corrected odd-L fold (Q1). Outputs
// the cycle's edge ids in walk order. stats[16]: [bf_calls, sweeps, lawler_iters,
// branches, relaxations], then the branch-and-bound events of Stats (q_push_id,
// q_push_present, q_single, q_rise, q_stale, q_cap, q_walk, q_empty, q_maxiter,
// q_ub_break, valid_pops).
int orc_ctr_solve(const int32_t* SRC, const int32_t* TRGT, long E, double* cost, int ncols, long npl, int K,
                  int flags, int32_t* out_edges, long cap, long* out_len, float* out_obj, int* passes,
                  long* stats) {
  try {
    orc_quantize(cost, E, ncols);
    CtrSolver s;
    setup_layered(s.S, SRC, TRGT, E, npl, K);
    s.cost.assign(cost, cost + E);
    if (ncols > 1)
      s.time.assign(cost + E, cost + 2 * E);
    else
      s.time.assign(E, 100000.0 / 2);
    s.S.c = s.cost.data();
    s.S.t = s.time.data();
    s.tol = (double)(float)1e-8;
    s.gpu_pick = (flags & 1) != 0;
    s.corrected = (flags & 2) != 0;
    s.branch.assign(npl, 0);
    std::vector<long> edges;
    float obj = 0;
    const bool ok = s.solve(edges, obj);
    if (!ok) {
      g_err = "cycle edge lookup failed";
      return -2;
    }
    if ((long)edges.size() > cap) {
      g_err = "output capacity too small";
      return -3;
    }
    for (size_t i = 0; i < edges.size(); ++i) out_edges[i] = (int32_t)edges[i];
    *out_len = (long)edges.size();
    *out_obj = obj;
    *passes = 1;
    s.st.relaxations = s.S.relax_count;
    if (stats) {
      stats[0] = s.st.bf_calls;
      stats[1] = s.st.sweeps;
      stats[2] = s.st.lawler_iters;
      stats[3] = s.st.branches;
      stats[4] = s.st.relaxations;
      const long q[] = {s.st.q_push_id, s.st.q_push_present, s.st.q_single, s.st.q_rise,  s.st.q_stale,
                        s.st.q_cap,     s.st.q_walk,         s.st.q_empty,  s.st.q_maxiter, s.st.q_ub_break,
                        s.st.valid_pops};
      for (int i = 0; i < 11; ++i) stats[5 + i] = q[i];
    }
    return 0;
  } catch (const std::exception& ex) {
    g_err = ex.what();
    return -1;
  }
}

// Minimum-cost solve. cost column 0 is used raw. Branch graph given as CSR
// over layer-0 nodes (bg_off[npl+1], bg_val); pass bg_off == nullptr to use
// the in-edges of layer-0 nodes (the reference's default). flags bit0: CPU
// rules (float-free keys, negative branching seed) instead of the GPU ones;
// bit1: corrected odd-L lower bounds (Q1).
int orc_dijkstra_solve(const int32_t* SRC, const int32_t* TRGT, long E, const double* cost, long npl, int K,
                       const long* bg_off, const long* bg_val, int flags, int32_t* out_edges, long cap,
                       long* out_len, float* out_obj, int* passes, long* stats) {
  try {
    MinCostSolver s;
    setup_layered(s.S, SRC, TRGT, E, npl, K);
    s.S.c = cost;
    s.S.t = nullptr;
    s.gpu_rules = (flags & 1) == 0;
    s.corrected = (flags & 2) != 0;
    s.branch.assign(npl, 0);
    if (bg_off) {
      s.bg_off.assign(bg_off, bg_off + npl + 1);
      s.bg_val.assign(bg_val, bg_val + bg_off[npl]);
    } else {
      s.bg_off.assign(npl + 1, 0);
      for (long v = 0; v < npl; ++v) {
        std::vector<long> ids(s.S.g.edges.begin() + s.S.g.off[v], s.S.g.edges.begin() + s.S.g.off[v + 1]);
        std::sort(ids.begin(), ids.end());
        for (long e : ids) s.bg_val.push_back(e);
        s.bg_off[v + 1] = (long)s.bg_val.size();
      }
    }
    std::vector<long> edges;
    float obj = 0;
    bool ps = false;
    s.solve(edges, obj, ps);
    if ((long)edges.size() > cap) {
      g_err = "output capacity too small";
      return -3;
    }
    for (size_t i = 0; i < edges.size(); ++i) out_edges[i] = (int32_t)edges[i];
    *out_len = (long)edges.size();
    *out_obj = obj;
    *passes = ps ? 1 : 0;
    s.st.relaxations = s.S.relax_count;
    if (stats) {
      stats[0] = s.st.bf_calls;
      stats[1] = s.st.sweeps;
      stats[2] = s.st.lawler_iters;
      stats[3] = s.st.branches;
      stats[4] = s.st.relaxations;
    }
    return 0;
  } catch (const std::exception& ex) {
    g_err = ex.what();
    return -1;
  }
}

}  // extern "C"
