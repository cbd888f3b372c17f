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
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "pgc_internal.h"

namespace pgc {

#define PGC_INF __builtin_huge_val()

__device__ __forceinline__ int orig_index(int64_t npl, int64_t K, int64_t p) {
  const int64_t per = npl * K;
  const int64_t layer = p / per;
  const int64_t rem = p - layer * per;
  const int64_t depth = rem / npl;
  return (int)(rem - depth * npl + layer * npl);
}

// std::round then double->int as on x86-64 (INT_MIN outside the int range)
__device__ __forceinline__ int round_to_int(double w) {
  const double r = round(w);
  if (!(r >= -2147483648.0 && r < 2147483648.0)) return (int)0x80000000;
  return (int)r;
}

template <typename CT, bool HAS_TIME, bool MINCOST>
__device__ __forceinline__ double slot_weight(const SlotClass& S, int idx, double ratio, double tconst) {
  const double c = (double)((const CT*)S.cost)[idx];
  if (MINCOST) return c;
  const double t = HAS_TIME ? (double)S.time[idx] : tconst;
  return __dsub_rn(c, __dmul_rn(ratio, t));
}

struct Shared {
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

// ---------------------------------------------------------------- reductions
__device__ int block_sum(Shared& sh, int x) {
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh.redI[w] = x;
  __syncthreads();
  int s = 0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += sh.redI[i];
  __syncthreads();
  return s;
}

__device__ int block_max(Shared& sh, int x) {
  for (int o = 32; o > 0; o >>= 1) x = max(x, __shfl_xor(x, o));
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh.redI[w] = x;
  __syncthreads();
  int s = sh.redI[0];
  for (int i = 1; i < (int)(blockDim.x >> 6); ++i) s = max(s, sh.redI[i]);
  __syncthreads();
  return s;
}

// lexicographic minimum of (key, v); v = INT_MAX means none
__device__ void block_min_key(Shared& sh, double& key, int& v) {
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
  for (int i = 1; i < (int)(blockDim.x >> 6); ++i) {
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
template <typename CT, bool HAS_TIME>
__device__ bool lookup_edge(const SolverJob& J, int a, int b, int& eid, double& c, double& t) {
  const int64_t npl = J.npl;
  int64_t bb = b;
  if (bb < npl && a > npl) bb += J.numNodes;
  if (a < 0 || bb < 0 || bb >= J.numNodes + npl) return false;
  const int lb = (int)(bb / npl);
  const int vi = (int)(bb - (int64_t)lb * npl);
  const SlotClass* S;
  int sl;
  if ((int64_t)a / npl == lb) {
    S = &J.intra;
    sl = (int)(a - (int64_t)lb * npl);
  } else {
    S = &J.inter;
    sl = (int)(a - (int64_t)(lb - 1) * npl);
    if (lb == 0 || sl < 0 || sl >= npl) return false;
  }
  const int ch = vi >> 6, lane = vi & 63;
  const int base = S->off[lb * J.nChunks + ch];
  const int width = (S->off[lb * J.nChunks + ch + 1] - base) >> 6;
  for (int j = 0; j < width; ++j) {
    const int idx = base + j * 64 + lane;
    const int s = S->src[idx];
    if (s < 0) break;
    if (s == sl) {
      eid = S->eid[idx];
      c = (double)((const CT*)S->cost)[idx];
      t = HAS_TIME ? (double)S->time[idx] : J.time_const;
      return true;
    }
  }
  return false;
}

// ---------------------------------------------------------------- kernel
template <typename CT, bool HAS_TIME, bool MINCOST, int MAXC>
struct Solver {
  const SolverJob& J;
  double* D;     // [2K*npl]
  double* Dmin;  // [npl]
  Shared& sh;
  int npl, K, L, nC, wave, lane, nw;
  double backup[MAXC];

  __device__ Solver(const SolverJob& j, double* d, double* dm, Shared& s)
      : J(j), D(d), Dmin(dm), sh(s) {
    npl = J.npl;
    K = J.K;
    L = J.L;
    nC = J.nChunks;
    wave = threadIdx.x >> 6;
    lane = threadIdx.x & 63;
    nw = blockDim.x >> 6;
  }

  __device__ __forceinline__ int vert(int i) const {
    const int c = wave + i * nw;
    if (c >= nC) return -1;
    const int v = c * 64 + lane;
    return v < npl ? v : -1;
  }

  // One pass over layers 0..L. Returns 0 when a layer 0<l<L had no improvement.
  __device__ int run_layers(double ratio, double ub) {
    const double tconst = J.time_const;
    double dmin_r[MAXC];
    for (int layer = 0; layer <= L; ++layer) {
      const int cur = (layer & 1) * K;
      const int prev = K - cur;
      bool imp = false;
      if (layer == 0) {
#pragma unroll
        for (int i = 0; i < MAXC; ++i) {
          const int v = vert(i);
          if (v >= 0) dmin_r[i] = D[v];
        }
      } else {
        const int* off = J.inter.off + layer * nC;
#pragma unroll
        for (int i = 0; i < MAXC; ++i) {
          const int c = wave + i * nw;
          if (c >= nC) break;
          const int v = c * 64 + lane;
          const int base = off[c];
          const int width = (off[c + 1] - base) >> 6;
          if (v < npl) {
            double best = PGC_INF, bw = 0;
            int bs = -1;
            for (int j = 0; j < width; ++j) {
              const int idx = base + j * 64 + lane;
              const int s = J.inter.src[idx];
              if (s < 0) break;
              const double w = slot_weight<CT, HAS_TIME, MINCOST>(J.inter, idx, ratio, tconst);
              const double dv = (K == 1) ? D[prev * npl + s] : Dmin[s];
              const double nv = dv + w;
              if (nv < best) {
                best = nv;
                bs = s;
                bw = w;
              }
            }
            double res = PGC_INF;
            if (best < PGC_INF && best < ub) {
              int ks = 0;
              if (K > 1) {
                for (int kk = 0; kk < K; ++kk) {
                  if (D[(prev + kk) * npl + bs] + bw == best) {
                    ks = kk;
                    break;
                  }
                }
              }
              J.pred[(int64_t)layer * npl * K + v] = bs + (layer - 1) * npl * K + ks * npl;
              res = best;
              imp = true;
            }
            D[cur * npl + v] = res;
            dmin_r[i] = res;
          }
        }
        if (K == 1) {
          const int any = __syncthreads_or(imp);
          if (!any && layer < L) return 0;
          continue;
        }
        __syncthreads();
      }
      for (int k = 1; k < K; ++k) {
        const int* off = J.intra.off + layer * nC;
#pragma unroll
        for (int i = 0; i < MAXC; ++i) {
          const int c = wave + i * nw;
          if (c >= nC) break;
          const int v = c * 64 + lane;
          const int base = off[c];
          const int width = (off[c + 1] - base) >> 6;
          if (v < npl) {
            double best = PGC_INF;
            int bs = -1;
            for (int j = 0; j < width; ++j) {
              const int idx = base + j * 64 + lane;
              const int s = J.intra.src[idx];
              if (s < 0) break;
              const double w = slot_weight<CT, HAS_TIME, MINCOST>(J.intra, idx, ratio, tconst);
              const double nv = D[(cur + k - 1) * npl + s] + w;
              if (nv < best) {
                best = nv;
                bs = s;
              }
            }
            const double curv = (layer == 0) ? D[(cur + k) * npl + v] : PGC_INF;
            double res = curv;
            if (best < curv && best < ub) {
              J.pred[(int64_t)layer * npl * K + (int64_t)k * npl + v] = bs + layer * npl * K + (k - 1) * npl;
              res = best;
              imp = true;
            }
            D[(cur + k) * npl + v] = res;
            dmin_r[i] = res < dmin_r[i] ? res : dmin_r[i];
            if (k == K - 1) Dmin[v] = dmin_r[i];
          }
        }
        if (k == K - 1) {
          const int any = __syncthreads_or(imp);
          if (layer > 0 && layer < L && !any) return 0;
        } else {
          __syncthreads();
        }
      }
    }
    return 1;
  }

  // D <- inf, branch sources <- 0, pred of sources <- -1; returns #sources.
  __device__ int setup(int bid, int onlyV) {
    for (int i = threadIdx.x; i < 2 * K * npl; i += blockDim.x) D[i] = PGC_INF;
    for (int i = threadIdx.x; i < npl; i += blockDim.x) Dmin[i] = PGC_INF;
    __syncthreads();
    int cnt = 0, last = -1;
#pragma unroll
    for (int i = 0; i < MAXC; ++i) {
      const int v = vert(i);
      if (v < 0) continue;
      const bool in = onlyV >= 0 ? (v == onlyV) : (J.branch[v] == bid);
      if (in) {
        D[v] = 0;
        J.pred[v] = -1;
        cnt++;
        last = max(last, v);
      }
    }
    const int n = block_sum(sh, cnt);
    const int lv = block_max(sh, last);
    if (threadIdx.x == 0) {
      sh.numInBranch = n;
      sh.indexNode = lv;
    }
    __syncthreads();
    return n;
  }

  __device__ bool in_branch(int v, int bid, int onlyV) const {
    return onlyV >= 0 ? (v == onlyV) : (J.branch[v] == bid);
  }

  // Fold of the closing layer into layer 0; returns whether any value improved.
  __device__ bool fold(int bid, int onlyV) {
    const int lastRow = L % 2;
    bool imp = false;
#pragma unroll
    for (int i = 0; i < MAXC; ++i) {
      const int v = vert(i);
      if (v < 0 || !in_branch(v, bid, onlyV)) continue;
      const double dl = D[lastRow * npl + v];
      if (dl < backup[i]) {
        imp = true;
        D[v] = dl;
        J.pred[v] = J.pred[(int64_t)npl * K * L + v];
      } else {
        D[v] = backup[i];
      }
    }
    return __syncthreads_or(imp) != 0;
  }

  // Chase every layer-0 vertex's predecessor chain back to the previous
  // layer-0 vertex (one wrap), recording the original node ids.
  __device__ void chase_wraps() {
    const int stride = J.chain_stride;
    int p[MAXC], pos[MAXC];
    bool act[MAXC];
#pragma unroll
    for (int i = 0; i < MAXC; ++i) {
      const int v = vert(i);
      act[i] = v >= 0;
      pos[i] = 0;
      p[i] = act[i] ? J.pred[v] : -1;
      if (act[i] && p[i] == -1) {
        J.wnext[v] = -1;
        J.wlen[v] = 0;
        act[i] = false;
      }
    }
    bool any = true;
    while (any) {
      any = false;
#pragma unroll
      for (int i = 0; i < MAXC; ++i) {
        if (!act[i]) continue;
        const int v = vert(i);
        const int pp = p[i];
        if (pp < npl) {
          J.chain[(int64_t)v * stride + pos[i]] = pp;
          pos[i]++;
          J.wnext[v] = pp;
          J.wlen[v] = pos[i];
          act[i] = false;
          continue;
        }
        if (pos[i] >= stride - 1) {
          J.wnext[v] = -2;
          J.wlen[v] = pos[i];
          act[i] = false;
          continue;
        }
        J.chain[(int64_t)v * stride + pos[i]] = orig_index(npl, K, pp);
        pos[i]++;
        const int np = J.pred[pp];
        p[i] = np;
        if (np == -1) {
          J.wnext[v] = -1;
          J.wlen[v] = pos[i];
          act[i] = false;
          continue;
        }
        any = true;
      }
    }
  }

  // Walk of the reference (cost_time_ratio_solver.cpp:443-489) expressed on
  // the wrap map W: first repeated layer-0 vertex from v and its step count.
  __device__ bool walk_from(int v, int64_t capSteps, int& vstar) const {
    const int* W = J.wnext;
    const int* Ln = J.wlen;
    const int lim = 4 * npl + 8;
    int t = W[v];
    if (t < 0) return false;
    int h = W[t];
    int it = 0;
    while (t != h && h >= 0 && it < lim) {
      t = W[t];
      h = W[h];
      if (h >= 0) h = W[h];
      ++it;
    }
    if (h < 0 || it >= lim) return false;
    int a = v;
    it = 0;
    while (a != h && it < lim) {
      a = W[a];
      h = W[h];
      ++it;
    }
    const int e = a;
    int64_t P = 0;
    int x = v;
    it = 0;
    while (x != e && it < lim) {
      P += Ln[x];
      x = W[x];
      ++it;
    }
    x = e;
    do {
      P += Ln[x];
      x = W[x];
      ++it;
    } while (x != e && it < 2 * lim);
    if (P > capSteps) return false;
    vstar = e;
    return true;
  }

  // sequential sums over a cycle's pairs (thread 0 only)
  __device__ void pair_sums(const double* c, const double* t, int m, double ratio, double& cs, double& ts,
                            double& ws) const {
    cs = 0;
    ts = 0;
    ws = 0;
    for (int i = 0; i < m; ++i) {
      const double ci = c[i], ti = t[i];
      cs += ci;
      ts += ti;
      ws += __dsub_rn(ci, __dmul_rn(ratio, ti));
    }
  }

  // edge lookups for tmp_nodes[0..n) -> tmp_eid/c/t[0..n-1); returns ok (block-uniform)
  __device__ bool lookup_pairs(int n) {
    bool bad = false;
    for (int i = threadIdx.x; i + 1 < n; i += blockDim.x) {
      int eid = -1;
      double c = 0, t = 0;
      if (!lookup_edge<CT, HAS_TIME>(J, J.tmp_nodes[i], J.tmp_nodes[i + 1], eid, c, t)) bad = true;
      J.tmp_eid[i] = eid;
      J.tmp_c[i] = c;
      J.tmp_t[i] = t;
    }
    return __syncthreads_or(bad) == 0;
  }

  __device__ void copy_tmp_to_neg(int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      J.neg_nodes[i] = J.tmp_nodes[i];
      if (i + 1 < n) {
        J.neg_eid[i] = J.tmp_eid[i];
        J.neg_c[i] = J.tmp_c[i];
        J.neg_t[i] = J.tmp_t[i];
      }
    }
    if (threadIdx.x == 0) *J.neg_len = n;
    __syncthreads();
  }

  // ZERO_CYCLE if the current cycle has (rounded) zero reweighted cost, else NO_NEGATIVE_CYCLE
  __device__ int zero_or_none(double ratio) {
    if (threadIdx.x == 0) {
      const int n = *J.neg_len;
      double cs, ts, ws;
      pair_sums(J.neg_c, J.neg_t, n > 0 ? n - 1 : 0, ratio, cs, ts, ws);
      sh.status = round_to_int(ws) == 0 ? ST_ZERO_CYCLE : ST_NO_NEGATIVE_CYCLE;
    }
    __syncthreads();
    return sh.status;
  }

  // mooreBellmanFordSubroutine (ratio and upper-bound modes).
  // Returns status; on FOUND the new cycle is in neg_* and its ratio in sh.d0.
  __device__ int bellman_ford(double ratio, bool zeroSearch, bool ubSearch, int bid, int onlyV, int64_t& sweeps,
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
      for (int i = 0; i < MAXC; ++i) {
        const int v = vert(i);
        backup[i] = v >= 0 ? D[v] : PGC_INF;
      }
      sweeps++;
      if (!run_layers(ratio, ub)) return zero_or_none(ratio);

      if (ubSearch) {
        // path from the closing copy of the start vertex back to it (:370-397)
        if (threadIdx.x == 0) {
          const int start = sh.indexNode;
          const int64_t maxLength = (int64_t)L * K;
          int cnt = 0;
          int64_t firstV = start + (int64_t)npl * L * K;
          J.tmp_nodes[cnt++] = orig_index(npl, K, firstV);
          int64_t p = J.pred[firstV];
          J.tmp_nodes[cnt++] = orig_index(npl, K, p);
          for (int64_t kk = maxLength - 1; kk > 0 && cnt < J.cyc_cap; --kk) {
            if (p == -1) break;
            p = J.pred[p];
            J.tmp_nodes[cnt++] = orig_index(npl, K, p);
            if (p == start) break;
          }
          for (int i = 0, j = cnt - 1; i < j; ++i, --j) {
            const int tmp = J.tmp_nodes[i];
            J.tmp_nodes[i] = J.tmp_nodes[j];
            J.tmp_nodes[j] = tmp;
          }
          sh.i0 = cnt;
        }
        __syncthreads();
        const int cnt = sh.i0;
        if (!lookup_pairs(cnt)) err = ERR_EDGE_LOOKUP;
        copy_tmp_to_neg(cnt);
        if (threadIdx.x == 0) {
          double cs, ts, ws;
          pair_sums(J.neg_c, J.neg_t, cnt - 1, ratio, cs, ts, ws);
          const int w = round_to_int(ws);
          sh.status = w == 0 ? ST_ZERO_CYCLE : (w < 0 ? ST_FOUND_NEGATIVE_CYCLE : ST_NO_NEGATIVE_CYCLE);
        }
        __syncthreads();
        return sh.status;
      }

      if (!fold(bid, onlyV)) return zero_or_none(ratio);

      chase_wraps();
      __syncthreads();

      // candidates: branch vertices with a negative (or, searching zero
      // cycles, |D| < tol) layer-0 value
      const bool pickLast = (J.flags & FLAG_PICK_LAST_VERTEX) != 0;
      const int64_t capSteps = J.numNodes - 1;
      double bestKey = PGC_INF;
      int bestV = 0x7fffffff;
#pragma unroll
      for (int i = 0; i < MAXC; ++i) {
        const int v = vert(i);
        if (v < 0 || !in_branch(v, bid, onlyV)) continue;
        const double lb = D[v];
        if (!zeroSearch && lb >= 0) continue;
        if (zeroSearch && (lb <= -J.tol || lb >= J.tol)) continue;
        int vs;
        if (!walk_from(v, capSteps, vs)) continue;
        const double key = pickLast ? -(double)v : lb;
        if (key < bestKey || (key == bestKey && v < bestV)) {
          bestKey = key;
          bestV = v;
        }
      }
      block_min_key(sh, bestKey, bestV);
      if (bestV != 0x7fffffff) {
        // extract the cycle through v* (first repeated vertex of bestV's walk)
        if (threadIdx.x == 0) {
          int vs = -1;
          walk_from(bestV, capSteps, vs);
          int64_t m = 0;
          int x = vs, it = 0;
          do {
            m += J.wlen[x];
            x = J.wnext[x];
          } while (x != vs && ++it < 4 * npl + 8);
          sh.i0 = vs;
          sh.i1 = (m + 1 <= J.cyc_cap) ? (int)m : -1;
        }
        __syncthreads();
        const int vs = sh.i0;
        const int m = sh.i1;
        if (m < 0) {
          err = ERR_CYCLE_TOO_LONG;
          return ST_NO_NEGATIVE_CYCLE;
        }
        if (wave == 0) {
          int x = vs, offp = 0, guard = 0;
          do {
            const int len = J.wlen[x];
            for (int j = lane; j < len && offp + j < m; j += 64)
              J.tmp_nodes[m - 1 - (offp + j)] = J.chain[(int64_t)x * J.chain_stride + j];
            offp += len;
            x = J.wnext[x];
          } while (x != vs && x >= 0 && offp < m && ++guard < 4 * npl + 8);
          if (lane == 0) J.tmp_nodes[m] = vs + (int)J.numNodes;
        }
        __syncthreads();
        const int nn = m + 1;
        if (!lookup_pairs(nn)) {
          err = ERR_EDGE_LOOKUP;
          return ST_NO_NEGATIVE_CYCLE;
        }
        if (threadIdx.x == 0) {
          double cs, ts, ws;
          pair_sums(J.tmp_c, J.tmp_t, nn - 1, ratio, cs, ts, ws);
          sh.i2 = round_to_int(ws);
          sh.d0 = cs / ts;
        }
        __syncthreads();
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
      if (iter >= J.numNodes - 1) return ST_NO_NEGATIVE_CYCLE;
      if (sweeps >= J.max_sweeps) {
        err = ERR_SWEEP_CAP;
        return ST_NO_NEGATIVE_CYCLE;
      }
    }
  }

  // wraps and zero-layer hits of the current cycle (:606-634); thread 0
  __device__ void wraps_and_hits(int& numWraps, int& nHits) const {
    const int n = *J.neg_len;
    numWraps = 0;
    nHits = 0;
    bool entered = true, left = false;
    for (int i = 0; i < n; ++i) {
      const int v = J.neg_nodes[i];
      if (left && v < npl) {
        entered = true;
        left = false;
      }
      if (entered) {
        if (nHits < npl) J.hits[nHits] = v;
        nHits++;
        entered = false;
      }
      if (v >= npl && !left) {
        numWraps++;
        left = true;
      }
    }
  }

  // ------------------------------------------------------- modes
  __device__ void lawler() {
    int64_t sweeps = 0, bfc = 0, its = 0;
    int err = ERR_NONE;
    double ratio = J.ratio_in;
    double oldRatio = ratio;
    const double tol = J.tol;
    int i = 0;
    int lastStatus = ST_NO_NEGATIVE_CYCLE;
    while (true) {
      its++;
      bfc++;
      const int st = bellman_ford(ratio, false, false, J.branchId, -1, sweeps, err);
      lastStatus = st;
      if (err != ERR_NONE) break;
      if (st == ST_ZERO_CYCLE || st == ST_NO_NEGATIVE_CYCLE) break;
      ratio = sh.d0;
      if (fabs(oldRatio - ratio) < tol) break;
      if (ratio - tol > oldRatio) break;
      oldRatio = ratio;
      if (i > 1000) break;
      i++;
    }
    if (threadIdx.x == 0) {
      int nw_ = 0, nh = 0;
      wraps_and_hits(nw_, nh);
      JobResult& r = *J.res;
      r.status = lastStatus;
      r.error = err;
      r.ratio = ratio;
      r.ratio_ub = 0;
      r.num_wraps = nw_;
      r.num_in_branch = sh.numInBranch;
      r.n_hits = nh;
      r.neg_len = *J.neg_len;
      r.best_v = -1;
      r.sweeps = sweeps;
      r.bf_calls = bfc;
      r.lawler_iters = its;
      r.relax = 0;
    }
  }

  // findInitalUpperBound (:652-707)
  __device__ void upper_bound() {
    // start vertex: minimal float mean of the in-edge costs of its layer-1
    // copy (accumulated in ascending edge id), among vertices with > 3 in-edges
    double bestKey = PGC_INF;
    int bestV = 0x7fffffff;
    for (int v = threadIdx.x; v < npl; v += blockDim.x) {
      const int node_layer = 1;
      float mean = 0.0f;
      int n = 0;
      int lastE = -1;
      while (true) {
        int nextE = 0x7fffffff;
        double nextC = 0;
        for (int cls = 0; cls < 2; ++cls) {
          const SlotClass& S = cls == 0 ? J.intra : J.inter;
          const int ch = v >> 6, ln = v & 63;
          const int base = S.off[node_layer * nC + ch];
          const int width = (S.off[node_layer * nC + ch + 1] - base) >> 6;
          for (int j = 0; j < width; ++j) {
            const int idx = base + j * 64 + ln;
            if (S.src[idx] < 0) break;
            const int e = S.eid[idx];
            if (e > lastE && e < nextE) {
              nextE = e;
              nextC = (double)((const CT*)S.cost)[idx];
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
    block_min_key(sh, bestKey, bestV);
    const int start = (bestV == 0x7fffffff) ? 0 : bestV;
    int64_t sweeps = 0;
    int err = ERR_NONE;
    double ratio = 10.0;
    bellman_ford(ratio, true, true, 0, start, sweeps, err);
    if (threadIdx.x == 0) {
      const int n = *J.neg_len;
      double cs, ts, ws;
      pair_sums(J.neg_c, J.neg_t, n > 0 ? n - 1 : 0, ratio, cs, ts, ws);
      double cycleCost = (double)(float)ws;
      int guard = 0;
      while ((cycleCost > 0 || isnan(cycleCost)) && guard++ < 4096) {
        ratio += ratio;
        pair_sums(J.neg_c, J.neg_t, n > 0 ? n - 1 : 0, ratio, cs, ts, ws);
        cycleCost = (double)(float)ws;
      }
      const double r_ub = (double)(float)(cs / ts);
      JobResult& r = *J.res;
      r.status = ST_NO_NEGATIVE_CYCLE;
      r.error = err;
      r.ratio = r_ub;
      r.ratio_ub = r_ub;
      r.num_wraps = 0;
      r.num_in_branch = 1;
      r.n_hits = 0;
      r.neg_len = n;
      r.best_v = start;
      r.sweeps = sweeps;
      r.bf_calls = 1;
      r.lawler_iters = 0;
      r.relax = 0;
    }
  }

  // one sweep of the minimum-cost solver + lower bounds and paths
  __device__ void min_cost() {
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
      if (!run_layers(0.0, J.upper_bound)) status = ST_CANNOT_CONTAIN;
    }
    if (status == ST_CAN_CONTAIN) {
      const int lastRow = L % 2;
      const bool cpuRules = (J.flags & FLAG_MINCOST_CPU_RULES) != 0;
      const int stride = J.chain_stride;
      const int64_t maxLen = (int64_t)L * K;
#pragma unroll
      for (int i = 0; i < MAXC; ++i) {
        const int v = vert(i);
        if (v < 0) continue;
        if (J.branch[v] != bid) {
          J.lb_key[v] = PGC_INF;
          J.back[v] = -1;
          J.wlen[v] = 0;
          continue;
        }
        double key = D[lastRow * npl + v];
        if (!cpuRules) key = (double)(float)key;
        J.lb_key[v] = key;
        int back = -1, cnt = 0;
        if (!isinf((float)key)) {
          int64_t p = J.pred[(int64_t)v + (int64_t)npl * L * K];
          if (p != -1) {
            J.chain[(int64_t)v * stride + cnt++] = orig_index(npl, K, p);
            for (int64_t kk = maxLen; kk > 0; --kk) {
              const int64_t q = J.pred[p];
              if (q == -1) break;
              p = q;
              if (cnt < stride) J.chain[(int64_t)v * stride + cnt] = orig_index(npl, K, p);
              cnt++;
            }
            back = orig_index(npl, K, p);
          }
        }
        J.back[v] = back;
        J.wlen[v] = cnt < stride ? cnt : stride;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      JobResult& r = *J.res;
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
    }
  }
};

template <typename CT, bool HAS_TIME, bool MINCOST, int MAXC>
__global__ void __launch_bounds__(PGC_BLOCK) solve_kernel(const SolverJob* __restrict__ jobs) {
  extern __shared__ double lds[];
  __shared__ Shared sh;
  const SolverJob& J = jobs[blockIdx.x];
  double* D = lds;
  double* Dmin = lds + 2 * J.K * J.npl;
  Solver<CT, HAS_TIME, MINCOST, MAXC> s(J, D, Dmin, sh);
  if (MINCOST) {
    s.min_cost();
  } else if (J.mode == MODE_UPPER_BOUND) {
    s.upper_bound();
  } else {
    s.lawler();
  }
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

// ------------------------------------------------------------ launchers
namespace pgc {

size_t solver_lds_bytes(int npl, int K) { return (size_t)(2 * K + 1) * npl * sizeof(double); }

template <typename CT, bool HT, bool MC, int MAXC>
static hipError_t launch_one(const SolverJob* jobs, int njobs, size_t lds, hipStream_t st) {
  static bool attr_set = false;
  if (!attr_set) {
    hipError_t e = hipFuncSetAttribute((const void*)solve_kernel<CT, HT, MC, MAXC>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 1024);
    if (e != hipSuccess) return e;
    attr_set = true;
  }
  hipLaunchKernelGGL((solve_kernel<CT, HT, MC, MAXC>), dim3(njobs), dim3(PGC_BLOCK), lds, st, jobs);
  return hipGetLastError();
}

template <typename CT, bool HT, bool MC>
static hipError_t launch_maxc(int maxc, const SolverJob* jobs, int njobs, size_t lds, hipStream_t st) {
  switch (maxc) {
    case 1: return launch_one<CT, HT, MC, 1>(jobs, njobs, lds, st);
    case 2: return launch_one<CT, HT, MC, 2>(jobs, njobs, lds, st);
    case 3:
    case 4: return launch_one<CT, HT, MC, 4>(jobs, njobs, lds, st);
    default: return launch_one<CT, HT, MC, 8>(jobs, njobs, lds, st);
  }
}

// cost_kind: 0 float cost (q-units), 1 double cost
hipError_t launch_solver(const SolverJob* d_jobs, int njobs, int maxChunksPerWave, bool mincost, bool hasTime,
                         bool doubleCost, size_t lds, hipStream_t st) {
  if (lds > 160 * 1024 - 1024) return hipErrorInvalidValue;
  if (maxChunksPerWave > 8) return hipErrorInvalidValue;
  if (mincost) {
    if (doubleCost) return launch_maxc<double, false, true>(maxChunksPerWave, d_jobs, njobs, lds, st);
    return launch_maxc<float, false, true>(maxChunksPerWave, d_jobs, njobs, lds, st);
  }
  if (hasTime) return launch_maxc<float, true, false>(maxChunksPerWave, d_jobs, njobs, lds, st);
  return launch_maxc<float, false, false>(maxChunksPerWave, d_jobs, njobs, lds, st);
}

hipError_t launch_lookup(const SolverJob* d_job, const int32_t* d_nodes, int n, int32_t* d_eids, bool hasTime,
                         bool doubleCost, hipStream_t st) {
  if (n < 2) return hipSuccess;
  dim3 b(256), g((n + 255) / 256);
  if (doubleCost)
    hipLaunchKernelGGL((lookup_kernel<double, false>), g, b, 0, st, d_job, d_nodes, n, d_eids);
  else if (hasTime)
    hipLaunchKernelGGL((lookup_kernel<float, true>), g, b, 0, st, d_job, d_nodes, n, d_eids);
  else
    hipLaunchKernelGGL((lookup_kernel<float, false>), g, b, 0, st, d_job, d_nodes, n, d_eids);
  return hipGetLastError();
}

}  // namespace pgc
