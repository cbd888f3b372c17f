// Exact minimum cost/time-ratio once-wrapping cycle of a layered product
// graph with integer edge costs and unit edge times. TEST INFRASTRUCTURE ONLY:
// a checker of the oracle's and the device's Lawler results at sizes brute
// force cannot reach; not a restatement of the reference.
//
// The cycles are those the reference's solver searches
// (cost_time_ratio_solver.cpp:251-635): a cycle enters layer 0 at a start
// vertex s at depth 0, takes at most K-1 intralayer steps per layer (depth
// k-1 -> k) and one interlayer step per layer (any depth -> depth 0 of the
// next layer), and returns to s on the copy of layer 0. With the quantized
// integer costs q_e and unit times (the mean problem: no time column), the
// ratio of a cycle is C/T with C = sum q_e and T its number of edges.
//
// Dinkelbach's method in exact integer arithmetic: for a candidate ratio a/b
// (b > 0) the edge weights b*q_e - a are int64; a shortest path from (0, 0, s)
// to s on the copy of layer 0, per start s, over the depth-expanded DAG, is a
// cycle with C/T < a/b exactly when its weight is negative. The most negative
// cycle over all starts gives the next a/b; when no cycle is negative, a/b is
// the minimum ratio (a certificate: no once-wrapping cycle is below it). All
// starts run together, one DP column per start, so every relaxation is a
// contiguous min over the starts. A start with no negative cycle at a/b has
// none at any smaller ratio, and drops out.
#include <stdint.h>

#include <algorithm>
#include <limits>
#include <string>
#include <vector>

namespace {

constexpr int64_t kInf = std::numeric_limits<int64_t>::max() / 4;

struct Layered {
  int64_t npl = 0, L = 0;
  // per layer l < L: intralayer in-edges (target v in layer l) and interlayer
  // in-edges (target v in layer l+1, or in layer 0 for l = L-1), CSR by target
  std::vector<std::vector<int64_t>> intraOff, interOff;
  std::vector<std::vector<int32_t>> intraSrc, interSrc, intraE, interE;
};

bool build(const int32_t* src, const int32_t* trgt, int64_t E, int64_t npl, Layered& G, std::string& err) {
  int64_t mx = -1;
  for (int64_t e = 0; e < E; ++e) mx = std::max<int64_t>(mx, src[e]);
  const int64_t nNodes = mx + 1;
  if (npl <= 0 || nNodes % npl != 0) {
    err = "node count is not a multiple of npl";
    return false;
  }
  G.npl = npl;
  G.L = nNodes / npl;
  const int64_t L = G.L;
  std::vector<std::vector<std::vector<std::pair<int32_t, int32_t>>>> intra(L), inter(L);
  for (int64_t l = 0; l < L; ++l) {
    intra[l].resize(npl);
    inter[l].resize(npl);
  }
  for (int64_t e = 0; e < E; ++e) {
    const int64_t s = src[e], t = trgt[e];
    if (s < 0 || t < 0) continue;
    const int64_t ls = s / npl, lt = t / npl;
    const int32_t u = (int32_t)(s - ls * npl), v = (int32_t)(t - lt * npl);
    if (ls >= L || lt >= L) continue;
    if (ls == lt)
      intra[ls][v].push_back({u, (int32_t)e});
    else if (lt == (ls + 1) % L)
      inter[ls][v].push_back({u, (int32_t)e});
  }
  G.intraOff.assign(L, {});
  G.interOff.assign(L, {});
  G.intraSrc.assign(L, {});
  G.interSrc.assign(L, {});
  G.intraE.assign(L, {});
  G.interE.assign(L, {});
  for (int64_t l = 0; l < L; ++l) {
    for (int cls = 0; cls < 2; ++cls) {
      auto& lists = cls == 0 ? intra[l] : inter[l];
      auto& off = cls == 0 ? G.intraOff[l] : G.interOff[l];
      auto& sv = cls == 0 ? G.intraSrc[l] : G.interSrc[l];
      auto& ev = cls == 0 ? G.intraE[l] : G.interE[l];
      off.assign(npl + 1, 0);
      for (int64_t v = 0; v < npl; ++v) {
        for (auto& p : lists[v]) {
          sv.push_back(p.first);
          ev.push_back(p.second);
        }
        off[v + 1] = (int64_t)sv.size();
      }
    }
  }
  return true;
}

// One DP over the layers for the starts in `starts` (columns), weights
// w_e = b*q_e - a. out[i] = the shortest cycle weight of starts[i] (kInf or
// more if none: unreachable entries stay above kInf / 2, since no path adds
// more than L*K*|w| to them). PRED (one start): pred records, per layer,
// depth and vertex, the edge that set the value (-1: none); an interlayer
// entry is 8 * edge + the source's depth.
template <bool PRED>
void dp(const Layered& G, int K, const int64_t* q, int64_t a, int64_t b, const std::vector<int32_t>& starts,
        std::vector<int64_t>& out, std::vector<int32_t>* pred) {
  const int64_t npl = G.npl, L = G.L;
  const int64_t S = (int64_t)starts.size();
  std::vector<int64_t> cur((size_t)K * npl * S, kInf), nxt((size_t)npl * S, kInf);  // cur[k][v][i]
  auto at = [&](std::vector<int64_t>& d, int64_t k, int64_t v) { return d.data() + (k * npl + v) * S; };
  for (int64_t i = 0; i < S; ++i) at(cur, 0, starts[i])[i] = 0;
  if (PRED) pred->assign((size_t)(L + 1) * K * npl, -1);
  for (int64_t l = 0; l < L; ++l) {
    for (int k = 1; k < K; ++k) {  // intralayer steps, depth k-1 -> k
#pragma omp parallel for schedule(static) if (!PRED)
      for (int64_t v = 0; v < npl; ++v) {
        int64_t* dst = at(cur, k, v);
        for (int64_t j = G.intraOff[l][v]; j < G.intraOff[l][v + 1]; ++j) {
          const int64_t* sp = at(cur, k - 1, G.intraSrc[l][j]);
          const int32_t e = G.intraE[l][j];
          const int64_t w = b * q[e] - a;
          if (PRED) {
            if (sp[0] < kInf / 2 && sp[0] + w < dst[0]) {
              dst[0] = sp[0] + w;
              (*pred)[((size_t)l * K + k) * npl + v] = e;
            }
          } else {
            for (int64_t i = 0; i < S; ++i) dst[i] = std::min(dst[i], sp[i] + w);
          }
        }
      }
    }
    // interlayer step: any depth of layer l -> depth 0 of layer l+1 (layer 0's copy for l = L-1)
    std::fill(nxt.begin(), nxt.end(), kInf);
#pragma omp parallel for schedule(static) if (!PRED)
    for (int64_t v = 0; v < npl; ++v) {
      int64_t* dst = nxt.data() + v * S;
      for (int64_t j = G.interOff[l][v]; j < G.interOff[l][v + 1]; ++j) {
        const int32_t u = G.interSrc[l][j];
        const int32_t e = G.interE[l][j];
        const int64_t w = b * q[e] - a;
        for (int k = 0; k < K; ++k) {
          const int64_t* sp = at(cur, k, u);
          if (PRED) {
            if (sp[0] < kInf / 2 && sp[0] + w < dst[0]) {
              dst[0] = sp[0] + w;
              (*pred)[((size_t)(l + 1) * K + 0) * npl + v] = e * 8 + k;
            }
          } else {
            for (int64_t i = 0; i < S; ++i) dst[i] = std::min(dst[i], sp[i] + w);
          }
        }
      }
    }
    std::fill(cur.begin(), cur.end(), kInf);
    if (l + 1 < L) std::copy(nxt.begin(), nxt.end(), cur.begin());  // depth 0 of layer l+1
  }
  out.assign(S, kInf);
  for (int64_t i = 0; i < S; ++i) {
    const int64_t x = nxt[(int64_t)starts[i] * S + i];
    out[i] = x < kInf / 2 ? x : kInf;
  }
}

}  // namespace

extern "C" {

static std::string g_err;

const char* orc_exact_error(void) { return g_err.c_str(); }

// The minimum ratio a/b over the once-wrapping cycles (unit times), and one
// cycle attaining it: its start vertex and its edge ids in walk order.
// (a0, b0): a ratio not below the minimum (e.g. a known cycle's), b0 > 0.
// Returns 0, or -1 (no cycle / bad input / path buffer too small).
int orc_exact_ratio(const int32_t* src, const int32_t* trgt, int64_t E, const int64_t* q, int64_t npl, int K,
                    int64_t a0, int64_t b0, int64_t* a_out, int64_t* b_out, int32_t* start_out, int32_t* path,
                    int64_t path_cap, int64_t* path_len, int64_t* iters) {
  Layered G;
  if (K < 1 || b0 <= 0 || !build(src, trgt, E, npl, G, g_err)) {
    if (g_err.empty()) g_err = "bad arguments";
    return -1;
  }
  std::vector<int32_t> starts(npl);
  for (int64_t s = 0; s < npl; ++s) starts[s] = (int32_t)s;
  int64_t a = a0, b = b0;
  int64_t it = 0;
  int32_t bestS = -1;
  std::vector<int32_t> pred;
  std::vector<int64_t> out;
  while (true) {
    ++it;
    dp<false>(G, K, q, a, b, starts, out, nullptr);
    int64_t best = 0;
    int32_t who = -1;
    std::vector<int32_t> keep;
    for (size_t i = 0; i < starts.size(); ++i) {
      if (out[i] < 0) keep.push_back(starts[i]);
      if (out[i] < best) {
        best = out[i];
        who = starts[i];
      }
    }
    if (who < 0) break;  // no negative cycle: a/b is the minimum
    // the most negative cycle: its cost and length give the next ratio
    std::vector<int32_t> one{who};
    dp<true>(G, K, q, a, b, one, out, &pred);
    // walk back from (L, 0, who)
    std::vector<int32_t> edges;
    int64_t l = G.L, k = 0, v = who;
    while (l > 0 || k > 0) {
      const int32_t pe = pred[((size_t)l * K + k) * npl + v];
      if (pe < 0) {
        g_err = "broken predecessor chain";
        return -1;
      }
      if (k == 0) {
        const int32_t e = pe / 8;
        const int kk = pe % 8;
        edges.push_back(e);
        l -= 1;
        k = kk;
        v = src[e] - (int64_t)(src[e] / npl) * npl;
      } else {
        edges.push_back(pe);
        k -= 1;
        v = src[pe] - (int64_t)(src[pe] / npl) * npl;
      }
    }
    if (v != who) {
      g_err = "walk did not return to the start";
      return -1;
    }
    std::reverse(edges.begin(), edges.end());
    int64_t C = 0;
    for (int32_t e : edges) C += q[e];
    const int64_t T = (int64_t)edges.size();
    // strictly smaller ratio: C*b < a*T (the cycle's weight C*b - a*T < 0)
    if (!(C * b < a * T)) {
      g_err = "Dinkelbach step did not decrease the ratio";
      return -1;
    }
    a = C;
    b = T;
    bestS = who;
    if ((int64_t)edges.size() > path_cap) {
      g_err = "path buffer too small";
      return -1;
    }
    std::copy(edges.begin(), edges.end(), path);
    *path_len = T;
    starts = keep;
    if (starts.empty()) break;
  }
  *a_out = a;
  *b_out = b;
  *start_out = bestS;
  *iters = it;
  if (bestS < 0) *path_len = 0;  // (a0, b0) was already minimal: no cycle below it was found
  return 0;
}

}  // extern "C"
