// The layer sweep of the solver kernels: one Bellman-Ford pass over layers
// 0..L of the layered product graph (interlayer step, then the K-1
// intralayer depth steps of every layer), relaxing the distance rows in LDS.
//
// It is compiled as a separate (non-inlined) device function: the solver
// kernel around it keeps a large state (branch and bound, Lawler iteration,
// cycle extraction) live, and inside one function that state would take the
// registers the sweep needs for its slot prefetch and its LDS gathers. All
// pointers carry explicit address spaces (LDS for the distance rows and the
// barrier counters, global for the job), so the function emits no FLAT
// accesses.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pgc_internal.h"

namespace pgc {

#define PGC_LDS __attribute__((address_space(3)))

// Wave-uniform copies: the value of the first active lane, in scalar
// registers. The arguments of a device function arrive in vector registers,
// and values loaded through them would otherwise count as varying per lane:
// branches on them would run under exec masks and their address arithmetic on
// the vector ALU, with spills of the vector registers this takes.
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ int64_t uni(int64_t x) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)x);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)x >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ double uni(double x) { return __longlong_as_double(uni((int64_t)__double_as_longlong(x))); }
template <typename T>
__device__ __forceinline__ T* uni(T* p) {
  return (T*)(intptr_t)uni((int64_t)(intptr_t)p);
}

// Slot class of the job with wave-uniform fields.
__device__ __forceinline__ SlotClass uni_class(const PGC_GLOBAL SlotClass& c) {
  SlotClass r;
  r.src = uni(c.src);
  r.cost = uni(c.cost);
  r.time = uni(c.time);
  r.eid = uni(c.eid);
  r.w = uni(c.w);
  r.lstride = uni(c.lstride);
  r.src_inv = uni(c.src_inv);
  r.deg = uni(c.deg);
  r.grp = uni(c.grp);
  return r;
}

// Running minimum of the first-minimum scans: one v_min_f64 instead of a
// compare-and-select pair. The slot choice still uses the strict nv < best of
// the scan (ties keep the earlier slot). The values agree with the select
// except possibly in the sign of a zero minimum, which no comparison, sum or
// output distinguishes (distances are never NaN).
__device__ __forceinline__ double pgc_min(double a, double b) { return __builtin_fmin(a, b); }

// Depth of the winning source of an interlayer scan: the first kk < K whose
// rounded sum row(kk) + bw equals best (the reference's loop order), 0 if
// none. Up to four depths the rows are read together (one round trip instead
// of up to K dependent ones); the first match is then taken in depth order.
template <typename RowAt>
__device__ __forceinline__ int recover_depth(const RowAt& row, int K, double bw, double best) {
  if (K <= 4) {
    double d[4];
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) d[kk] = row(kk < K ? kk : K - 1);
    int ks = 0;
#pragma unroll
    for (int kk = 3; kk >= 0; --kk)
      if (kk < K && d[kk] + bw == best) ks = kk;
    return ks;
  }
  for (int kk = 0; kk < K; ++kk)
    if (row(kk) + bw == best) return kk;
  return 0;
}

struct SweepState {
  int complete;  // 0 when a layer 0<l<L had no improvement
  int p0, p1, odd;  // barrier-or state (see BarrierOr)
};

// ONLY_K1: the instance for jobs with K == 1 (interlayer blocks only). The K == 1
// loops and the general loop are separate functions, so that each gets its
// own register allocation. WE: interlayer slots read per node (< W when the
// class's largest in-degree leaves the last slots padding on every node; a
// padding slot never wins, so skipping it changes no result).
template <typename CT, bool HAS_TIME, bool MINCOST, int MAXC, int W, bool ONLY_K1 = false, int WE = W>
struct Sweep {
  static_assert(WE == W || ONLY_K1, "short interlayer slots: K == 1 loops only");
  const PGC_GLOBAL SolverJob* J;
  PGC_LDS double* D;     // [2K][ds] distance rows; entry npl of every row is +inf
  PGC_LDS double* Dmin;  // [ds] minimum over the depths of the previous layer
  PGC_LDS Shared* sh;
  PGC_GLOBAL int32_t* pred;
  PGC_LDS int32_t* predL;  // pred in LDS (small graphs), selected by predLds
  int predLds;
  // chain origins (K == 1, pred in LDS), or nullptr. Row l & 1 holds, for
  // every vertex x of layer l, where the predecessor chain from (l, x) ends:
  // the layer-0 vertex it reaches (>= 0), or -1 - m when the node of layer m
  // on it has no predecessor.
  PGC_LDS int32_t* orgL;
  int orgOn;  // orgL is in use (an explicit flag: null pointers do not survive the cast to LDS)
  SlotClass Sa, Sx;  // intralayer and interlayer slot classes, read once per sweep
  int npl, ds, K, L, nC, wave, lane, nw;
  int p0, p1, odd;

  __device__ __forceinline__ int chunk_of(int i) const { return wave + i * nw; }

  __device__ __forceinline__ int vert(int i) const {
    const int c = wave + i * nw;
    if (c >= nC) return -1;
    const int v = c * 64 + lane;
    return v < npl ? v : -1;
  }

  // Block-wide OR with one barrier (the scheme of BarrierOr).
  __device__ __forceinline__ bool bor(bool pred_) {
    PGC_LDS int* cnt = odd ? &sh->orCount[1] : &sh->orCount[0];
    const unsigned long long b = __ballot(pred_);
    if (b && lane == 0) __atomic_fetch_add(cnt, 1, __ATOMIC_RELAXED);
    __syncthreads();
    const int c = uni(*cnt);
    bool any;
    if (odd) {
      any = c != p1;
      p1 = c;
    } else {
      any = c != p0;
      p0 = c;
    }
    odd = !odd;
    return any;
  }

  // Slots of one owned chunk (pass i) of a layer block. The layout's slot
  // width is this kernel's W, so the W loads are unconditional and
  // independent (no branch between them: the compiler then counts them
  // exactly and a later wait covers only the loads it needs). Padding slots
  // point at the sentinel vertex npl (distance +inf) with cost 0.
  __device__ __forceinline__ void load_pass(const int32_t* src_, const void* cost_, const float* time_, int b,
                                            int (&s)[W], CT (&c)[W], float (&t)[W]) const {
    const PGC_GLOBAL int* __restrict__ src = gp(src_);
    const PGC_GLOBAL CT* __restrict__ cost = gp((const CT*)cost_);
    const PGC_GLOBAL float* __restrict__ tim = gp(time_);
#pragma unroll
    for (int j = 0; j < WE; ++j) {
      s[j] = src[b + j * 64];
      c[j] = cost[b + j * 64];
      if (HAS_TIME) t[j] = tim[b + j * 64];
    }
  }

  // c - lambda*t; rt = lambda*tconst (the same product for every edge when
  // the time is constant, so hoisting it is exact)
  __device__ __forceinline__ double weight(CT c, float t, double ratio, double rt) const {
    if (MINCOST) return (double)c;
    return __dsub_rn((double)c, HAS_TIME ? __dmul_rn(ratio, (double)t) : rt);
  }

  // Interlayer step of one owned vertex (depth 0 of layer `layer`). K1: the
  // job's K is 1 (no depth to recover, so the winner's cost is not tracked).
  template <bool K1>
  __device__ __forceinline__ bool inter_vertex(int i, int v, int layer, int cur, int prev, const int (&s)[W],
                                               const CT (&c)[W], const float (&t)[W], double ratio, double rt,
                                               double ub, double (&dmin_r)[MAXC]) {
    double best = PGC_INF;
    int bs = npl;
    CT bc = 0;
    float bt = 0;
    const PGC_LDS double* Dp = K1 ? D + prev * ds : Dmin;
    // all gathers first (independent LDS reads in flight together), then the
    // first-minimum scan in slot order
    double dv[W];
#pragma unroll
    for (int j = 0; j < WE; ++j) dv[j] = Dp[s[j]];
#pragma unroll
    for (int j = 0; j < WE; ++j) {
      const double nv = dv[j] + weight(c[j], t[j], ratio, rt);
      const bool better = nv < best;
      best = pgc_min(best, nv);
      bs = better ? s[j] : bs;
      if (!K1) {
        bc = better ? c[j] : bc;
        if (HAS_TIME) bt = better ? t[j] : bt;
      }
    }
    double res = PGC_INF;
    bool imp = false;
    if (best < PGC_INF && best < ub) {
      int ks = 0;
      if (!K1) {
        const double bw = weight(bc, bt, ratio, rt);
        ks = recover_depth([&](int kk) { return D[(prev + kk) * ds + bs]; }, K, bw, best);
      }
      const int64_t pi = (int64_t)layer * npl * K + v;
      const int pv = bs + (layer - 1) * npl * K + ks * npl;
      if (predLds)
        predL[pi] = pv;
      else
        pred[pi] = pv;
      res = best;
      imp = true;
    }
    D[cur * ds + v] = res;
    dmin_r[i] = res;
    if (K1 && orgOn) {
      // origin of (layer, v): through its new predecessor, or through the
      // one it keeps from an earlier sweep (K == 1: always a node of layer-1)
      const PGC_LDS int32_t* prevRow = orgL + ((layer - 1) & 1) * npl;
      int o;
      if (imp) {
        o = layer == 1 ? bs : prevRow[bs];
      } else {
        const int p = predL[(int64_t)layer * npl + v];
        o = p == -1 ? -1 - layer : (layer == 1 ? p : prevRow[p - (layer - 1) * npl]);
      }
      orgL[(layer & 1) * npl + v] = o;
    }
    return imp;
  }

  // Intralayer step (depth k) of one owned vertex.
  __device__ __forceinline__ bool intra_vertex(int i, int v, int layer, int cur, int k, const int (&s)[W],
                                               const CT (&c)[W], const float (&t)[W], double ratio, double rt,
                                               double ub, double (&dmin_r)[MAXC]) {
    double best = PGC_INF;
    int bs = npl;
    const PGC_LDS double* Dk = D + (cur + k - 1) * ds;
    double dv[W];
#pragma unroll
    for (int j = 0; j < W; ++j) dv[j] = Dk[s[j]];
#pragma unroll
    for (int j = 0; j < W; ++j) {
      const double nv = dv[j] + weight(c[j], t[j], ratio, rt);
      const bool better = nv < best;
      best = pgc_min(best, nv);
      bs = better ? s[j] : bs;
    }
    const double curv = (layer == 0) ? D[(cur + k) * ds + v] : PGC_INF;
    double res = curv;
    bool imp = false;
    if (best < curv && best < ub) {
      const int64_t pi = (int64_t)layer * npl * K + (int64_t)k * npl + v;
      const int pv = bs + layer * npl * K + (k - 1) * npl;
      if (predLds)
        predL[pi] = pv;
      else
        pred[pi] = pv;
      res = best;
      imp = true;
    }
    D[(cur + k) * ds + v] = res;
    dmin_r[i] = res < dmin_r[i] ? res : dmin_r[i];
    if (k == K - 1) Dmin[v] = dmin_r[i];
    return imp;
  }

  // Slot registers of one (class, layer) block of the owned chunks.
  struct Slots {
    int s[MAXC][W];
    CT c[MAXC][W];
    float t[MAXC][W];
  };

  // Issues the loads of the interlayer (inter) or intralayer block of a
  // layer: independent, back to back, from one load site for both classes.
  __device__ __forceinline__ void load_block(Slots& b, bool inter, int layer) const {
    // field-wise selects of wave-uniform values (no indexed struct access)
    const int32_t* src = inter ? Sx.src : Sa.src;
    const void* cost = inter ? Sx.cost : Sa.cost;
    const float* tim = inter ? Sx.time : Sa.time;
    const int base = layer * (inter ? Sx.lstride : Sa.lstride) + lane;
#pragma unroll
    for (int i = 0; i < MAXC; ++i) {
      const int ci = chunk_of(i);
      // passes past the last chunk read chunk 0 (in bounds) and are discarded
      load_pass(src, cost, tim, base + (ci < nC ? ci * W * 64 : 0), b.s[i], b.c[i], b.t[i]);
    }
  }

  // One relaxation step over the owned chunks with a loaded block.
  // INTER: 0 intralayer, 1 interlayer with K > 1, 2 interlayer with K == 1.
  template <int INTER>
  __device__ __forceinline__ bool compute(const Slots& b, int layer, int cur, int prev, int k, double ratio,
                                          double rt, double ub, double (&dmin_r)[MAXC]) {
    bool imp = false;
#pragma unroll
    for (int i = 0; i < MAXC; ++i) {
      const int v = vert(i);
      if (v >= 0) {
        if (INTER)
          imp |= inter_vertex<INTER == 2>(i, v, layer, cur, prev, b.s[i], b.c[i], b.t[i], ratio, rt, ub, dmin_r);
        else
          imp |= intra_vertex(i, v, layer, cur, k, b.s[i], b.c[i], b.t[i], ratio, rt, ub, dmin_r);
      }
    }
    return imp;
  }

  // Block b of a sweep: with K > 1 the blocks are intra(0), inter(1),
  // intra(1), ..., inter(L), intra(L); with K == 1 inter(1), ..., inter(L).
  __device__ __forceinline__ void block_of(int b, bool& inter, int& layer) const {
    if (K > 1) {
      inter = (b & 1) != 0;
      layer = (b + 1) >> 1;
    } else {
      inter = true;
      layer = b + 1;
    }
  }

  // The pass with K == 1 (interlayer blocks only) and prefetch: the two
  // register sets alternate roles (the loop is unrolled by two), so no block
  // is copied between them, and a block's wait covers only its own loads.
  // The prefetch is unconditional (past the last layer it reloads layer L,
  // in bounds, unused): with a branch around it the wait analysis would
  // merge the two paths and wait for the prefetch as well.
  // run_layers_k1 for layer-invariant interlayer sources: the sources of the
  // owned chunk are loaded once (from layer 1) and held in registers, and a
  // block loads only its costs and times.
  struct SlotsCT {
    CT c[MAXC][W];
    float t[MAXC][W];
  };
  __device__ __forceinline__ void load_ct(SlotsCT& b, int layer) const {
    const PGC_GLOBAL CT* __restrict__ cost = gp((const CT*)Sx.cost);
    const PGC_GLOBAL float* __restrict__ tim = gp(Sx.time);
#pragma unroll
    for (int i = 0; i < MAXC; ++i) {
      const int ci = chunk_of(i);
      const int base = layer * Sx.lstride + lane + (ci < nC ? ci * W * 64 : 0);
#pragma unroll
      for (int j = 0; j < WE; ++j) {
        b.c[i][j] = cost[base + j * 64];
        if (HAS_TIME) b.t[i][j] = tim[base + j * 64];
      }
    }
  }
  __device__ __forceinline__ bool compute_inv(const int (&s)[MAXC][W], const SlotsCT& b, int layer, double ratio,
                                              double rt, double ub, double (&dmin_r)[MAXC]) {
    bool imp = false;
    const int cur = layer & 1, prev = 1 - cur;
#pragma unroll
    for (int i = 0; i < MAXC; ++i) {
      const int v = vert(i);
      if (v >= 0) imp |= inter_vertex<true>(i, v, layer, cur, prev, s[i], b.c[i], b.t[i], ratio, rt, ub, dmin_r);
    }
    return imp;
  }
  __device__ __forceinline__ int run_layers_k1_inv(double ratio, double rt, double ub, double (&dmin_r)[MAXC]) {
    const int nb = L;
    int s[MAXC][W];
    {
      const PGC_GLOBAL int* __restrict__ src = gp(Sx.src);
#pragma unroll
      for (int i = 0; i < MAXC; ++i) {
        const int ci = chunk_of(i);
        const int base = Sx.lstride + lane + (ci < nC ? ci * W * 64 : 0);
#pragma unroll
        for (int j = 0; j < WE; ++j) s[i][j] = src[base + j * 64];
      }
    }
    SlotsCT A, B;
    load_ct(A, 1);
    for (int b = 0; b < nb; b += 2) {
      int layer = b + 1;
      load_ct(B, min(layer + 1, L));
      bool imp = compute_inv(s, A, layer, ratio, rt, ub, dmin_r);
      if (layer == L - 1) {
        if (!bor(imp)) return 0;
      } else {
        __syncthreads();
      }
      if (b + 1 >= nb) break;
      layer = b + 2;
      load_ct(A, min(layer + 1, L));
      imp = compute_inv(s, B, layer, ratio, rt, ub, dmin_r);
      if (layer == L - 1) {
        if (!bor(imp)) return 0;
      } else {
        __syncthreads();
      }
    }
    return 1;
  }

  // run_layers_k1_inv with implicit costs: a slot's q-cost is recomputed
  // from feat_diff instead of loaded. The interlayer edge (l-1, x0) -> (l, x1)
  // of a generated non-conjugate graph (vanilla cost, no time column, no
  // length term) has the energy 0.5 (FD[x0, ys] + FD[x1, yd]) with (ys, yd)
  // the Y edge of layer l-1 (deformationEnergy.cpp:66-69), and its slot cost
  // is that energy quantized with the graph's scale s, min(round(s e), 1e7),
  // stored as float (cost_time_ratio_solver.cpp:709-721, utils.hpp:196-209):
  // the same operations in the same order here give the same bits. Column ys
  // of the next layer is staged in LDS (stage[(l+1) & 1], entry npl = 0 for
  // the padding sentinel) while the current layer is relaxed; FD[x1, yd] of
  // the owned vertex is a register.
  __device__ __forceinline__ float iq(double a, double b, double s) const {
    const double e = __dmul_rn(0.5, __dadd_rn(a, b));
    double q = round(__dmul_rn(s, e));
    if (q > 10000000.0) q = 10000000.0;
    return (float)q;
  }
  __device__ __forceinline__ int run_layers_k1_implicit(PGC_LDS double* stage, double ratio, double rt, double ub,
                                                        double (&dmin_r)[MAXC]) {
    const PGC_GLOBAL double* __restrict__ col = gp(J->icol);
    const PGC_GLOBAL int32_t* __restrict__ ey = gp(J->iey);
    const double sc = uni(*gp(J->iscale));
    int s[MAXC][W];
    {
      const PGC_GLOBAL int* __restrict__ src = gp(Sx.src);
#pragma unroll
      for (int i = 0; i < MAXC; ++i) {
        const int ci = chunk_of(i);
        const int base = Sx.lstride + lane + (ci < nC ? ci * W * 64 : 0);
#pragma unroll
        for (int j = 0; j < WE; ++j) s[i][j] = src[base + j * 64];
      }
    }
    // owned vertices (a pass past the last chunk reads vertex 0, unused)
    int vo[MAXC];
#pragma unroll
    for (int i = 0; i < MAXC; ++i) vo[i] = max(vert(i), 0);
    // stage layer 1's source column and its targets' values
    double bt[MAXC];
    {
      const int ys = uni(ey[0]), yd = uni(ey[1]);
      const PGC_GLOBAL double* cs = col + (int64_t)ys * npl;
      const PGC_GLOBAL double* cd = col + (int64_t)yd * npl;
      for (int x = threadIdx.x; x <= npl; x += PGC_BLOCK) {
        stage[ds + x] = x < npl ? cs[x] : 0.0;
        stage[x] = 0.0;
      }
#pragma unroll
      for (int i = 0; i < MAXC; ++i) bt[i] = cd[vo[i]];
    }
    __syncthreads();
    for (int layer = 1; layer <= L; ++layer) {
      // next layer's column and target values (in flight during this layer)
      const int ln = min(layer + 1, L);
      const int ys = uni(ey[2 * (ln - 1)]), yd = uni(ey[2 * (ln - 1) + 1]);
      const PGC_GLOBAL double* cs = col + (int64_t)ys * npl;
      const PGC_GLOBAL double* cd = col + (int64_t)yd * npl;
      double nxt[4];  // npl <= 4 * PGC_BLOCK (checked by the host)
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int x = threadIdx.x + k * PGC_BLOCK;
        nxt[k] = x < npl ? cs[x] : 0.0;
      }
      double bn[MAXC];
#pragma unroll
      for (int i = 0; i < MAXC; ++i) bn[i] = cd[vo[i]];
      // this layer's costs from the staged column
      const PGC_LDS double* A = stage + (layer & 1) * ds;
      SlotsCT b;
#pragma unroll
      for (int i = 0; i < MAXC; ++i) {
        double a[W];
#pragma unroll
        for (int j = 0; j < WE; ++j) a[j] = A[s[i][j]];
#pragma unroll
        for (int j = 0; j < WE; ++j) b.c[i][j] = (CT)iq(a[j], bt[i], sc);
      }
      const bool imp = compute_inv(s, b, layer, ratio, rt, ub, dmin_r);
      // the next layer's buffer was last read before the previous barrier
      PGC_LDS double* An = stage + ((layer + 1) & 1) * ds;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int x = threadIdx.x + k * PGC_BLOCK;
        if (x < npl) An[x] = nxt[k];
      }
#pragma unroll
      for (int i = 0; i < MAXC; ++i) bt[i] = bn[i];
      if (layer == L - 1) {
        if (!bor(imp)) return 0;
      } else {
        __syncthreads();
      }
    }
    return 1;
  }

  __device__ __forceinline__ int run_layers_k1(Slots& A, Slots& B, double ratio, double rt, double ub,
                                               double (&dmin_r)[MAXC]) {
    const int nb = L;
    load_block(A, true, 1);
    for (int b = 0; b < nb; b += 2) {
      int layer = b + 1;
      load_block(B, true, min(layer + 1, L));
      bool imp = compute<2>(A, layer, (layer & 1), 1 - (layer & 1), 0, ratio, rt, ub, dmin_r);
      if (layer == L - 1) {
        if (!bor(imp)) return 0;
      } else {
        __syncthreads();
      }
      if (b + 1 >= nb) break;
      layer = b + 2;
      load_block(A, true, min(layer + 1, L));
      imp = compute<2>(B, layer, (layer & 1), 1 - (layer & 1), 0, ratio, rt, ub, dmin_r);
      if (layer == L - 1) {
        if (!bor(imp)) return 0;
      } else {
        __syncthreads();
      }
    }
    return 1;
  }

  // One pass over layers 0..L. Returns 0 when a layer 0<l<L had no improvement.
  //
  // A layer without improvement has only +inf rows, and so has every layer
  // after it: whether some layer 0<l<L had none is whether layer L-1 had
  // none. Only that layer's steps end in a block-wide OR; the others end in a
  // plain barrier (a pass that dies early runs on with +inf rows and writes
  // no predecessor, and its caller reads no row of it).
  //
  // The slots of a block do not depend on the distances, so when a wave owns
  // at most two chunks (two blocks fit in registers) the next block's loads
  // are issued before the current block is relaxed: a step then waits for
  // LDS and the barrier, not for global memory. One intralayer block serves
  // the K-1 intralayer steps of its layer.
  __device__ __forceinline__ int run_layers(double ratio, double ub) {
    constexpr bool PF = MAXC <= 2;
    const double rt = __dmul_rn(ratio, uni(J->time_const));
    double dmin_r[MAXC];
    const int nb = K > 1 ? 2 * L + 1 : L;
    Slots A, B;
    if constexpr (ONLY_K1 && PF) {
      if (nb <= 0) return 1;
      if (Sx.src_inv && !(uni((int)J->flags) & (int)FLAG_NO_CACHED_SOURCES)) return run_layers_k1_inv(ratio, rt, ub, dmin_r);
      return run_layers_k1(A, B, ratio, rt, ub, dmin_r);
    }
    bool inter;
    int layer;
    if (PF && nb > 0) {
      block_of(0, inter, layer);
      load_block(A, inter, layer);
      // wait here for the first block: inside the loop A then only ever
      // holds copies of B made after B's loads completed, and the relaxation
      // needs no wait on global memory (the compiler's wait analysis merges
      // all paths, so a pending first load would put a wait in every step)
      __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    }
    bool imp = false;
    for (int b = 0; b < nb; ++b) {
      block_of(b, inter, layer);
      const int cur = (layer & 1) * K;
      const int prev = K - cur;
      if (PF) {
        if (b + 1 < nb) {
          bool in2;
          int l2;
          block_of(b + 1, in2, l2);
          load_block(B, in2, l2);
        }
      } else {
        load_block(A, inter, layer);
      }
      const bool check = layer == L - 1 && layer > 0;
      if (inter) {
        if (K == 1) {
          imp = compute<2>(A, layer, cur, prev, 0, ratio, rt, ub, dmin_r);
          if (check) {
            if (!bor(imp)) return 0;
          } else {
            __syncthreads();
          }
        } else {
          imp = compute<1>(A, layer, cur, prev, 0, ratio, rt, ub, dmin_r);
          __syncthreads();
        }
      } else {
        if (layer == 0) {
          imp = false;
#pragma unroll
          for (int i = 0; i < MAXC; ++i) {
            const int v = vert(i);
            if (v >= 0) dmin_r[i] = D[v];
          }
        }
        for (int k = 1; k < K; ++k) {
          imp |= compute<0>(A, layer, cur, prev, k, ratio, rt, ub, dmin_r);
          if (k == K - 1 && check) {
            if (!bor(imp)) return 0;
          } else {
            __syncthreads();
          }
        }
      }
      if (PF) A = B;
    }
    return 1;
  }

};

template <typename CT, bool HAS_TIME, bool MINCOST, int MAXC, int W, bool ONLY_K1 = false, int WE = W>
__device__ __noinline__ SweepState sweep_layers(const SolverJob* Jp, double* Dg, double* Dming, Shared* shg,
                                                int32_t* predLg, int predLds, int p0, int p1, int odd, double ratio,
                                                double ub, int32_t* orgLg = nullptr, int orgOn = 0) {
  Sweep<CT, HAS_TIME, MINCOST, MAXC, W, ONLY_K1, WE> w;
  w.orgOn = ONLY_K1 ? uni(orgOn) : 0;
  w.orgL = (PGC_LDS int32_t*)uni(orgLg);
  w.J = (const PGC_GLOBAL SolverJob*)uni(Jp);
  w.D = (PGC_LDS double*)uni(Dg);
  w.Dmin = (PGC_LDS double*)uni(Dming);
  w.sh = (PGC_LDS Shared*)uni(shg);
  w.pred = (PGC_GLOBAL int32_t*)uni(w.J->pred);
  w.predL = (PGC_LDS int32_t*)uni(predLg);
  w.predLds = uni(predLds);
  w.Sa = uni_class(w.J->intra);
  w.Sx = uni_class(w.J->inter);
  w.npl = uni(w.J->npl);
  w.ds = w.npl + 1;
  w.K = uni(w.J->K);
  w.L = uni(w.J->L);
  w.nC = uni(w.J->nChunks);
  w.wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  w.lane = threadIdx.x & 63;
  w.nw = PGC_BLOCK / PGC_WAVE;
  w.p0 = uni(p0);
  w.p1 = uni(p1);
  w.odd = uni(odd);
  SweepState r;
  r.complete = w.run_layers(uni(ratio), uni(ub));
  r.p0 = w.p0;
  r.p1 = w.p1;
  r.odd = w.odd;
  return r;
}

// The K == 1 sweep with implicit costs (run_layers_k1_implicit), apart from
// sweep_layers so that the materialized-cost sweep keeps its own register
// allocation. stageLg: 2 (npl + 1) doubles of LDS after the origin rows.
template <typename CT, int MAXC, int W, int WE>
__device__ __noinline__ SweepState sweep_layers_k1_implicit(const SolverJob* Jp, double* Dg, double* Dming,
                                                            Shared* shg, int32_t* predLg, int p0, int p1, int odd,
                                                            double ratio, double ub, int32_t* orgLg, double* stageLg) {
  Sweep<CT, false, false, MAXC, W, true, WE> w;
  w.orgOn = 1;
  w.orgL = (PGC_LDS int32_t*)uni(orgLg);
  w.J = (const PGC_GLOBAL SolverJob*)uni(Jp);
  w.D = (PGC_LDS double*)uni(Dg);
  w.Dmin = (PGC_LDS double*)uni(Dming);
  w.sh = (PGC_LDS Shared*)uni(shg);
  w.pred = (PGC_GLOBAL int32_t*)uni(w.J->pred);
  w.predL = (PGC_LDS int32_t*)uni(predLg);
  w.predLds = 1;
  w.Sa = uni_class(w.J->intra);
  w.Sx = uni_class(w.J->inter);
  w.npl = uni(w.J->npl);
  w.ds = w.npl + 1;
  w.K = 1;
  w.L = uni(w.J->L);
  w.nC = uni(w.J->nChunks);
  w.wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  w.lane = threadIdx.x & 63;
  w.nw = PGC_BLOCK / PGC_WAVE;
  w.p0 = uni(p0);
  w.p1 = uni(p1);
  w.odd = uni(odd);
  SweepState r;
  const double ra = uni(ratio);
  double dmin_r[MAXC];
  r.complete = w.L > 0 ? w.run_layers_k1_implicit((PGC_LDS double*)uni(stageLg), ra,
                                                  __dmul_rn(ra, uni(w.J->time_const)), uni(ub), dmin_r)
                       : 1;
  r.p0 = w.p0;
  r.p1 = w.p1;
  r.odd = w.odd;
  return r;
}

// ------------------------------------------------------------------------
// The same sweep with one chunk at a time: every wave loops over its chunks at
// run time, so only one chunk's slots are live, and the per-vertex minimum
// over the depths of a layer is recomputed from the rows at its last
// intralayer step (the minimum is exact, so this equals the running minimum
// of the unrolled sweep). The rows live in LDS for graphs with several chunks
// per wave or wide slots, and in global memory (L2-resident) for graphs whose
// distance table does not fit the LDS.
template <bool LD>
struct DistPtr {
  typedef PGC_GLOBAL double* type;
};
template <>
struct DistPtr<true> {
  typedef PGC_LDS double* type;
};

// LD: the rows live in LDS (graphs with many chunks per wave, for which the
// unrolled sweep would hold too many registers), else in global memory.
// MC > 0: at most MC chunks per wave (rows in LDS); with layer-invariant
// sources in both classes, the sources of the owned chunks are loaded once
// per pass and held in registers, and a step loads only costs and times.
// WX, WI: interlayer and intralayer slots a step reads per node (below W
// when the class's largest in-degree leaves the last slots padding on every
// node; a padding slot never wins, so skipping it changes no result).
// GRID: the device-wide sweep (cooperative launch): wave and nw number the
// waves of the whole grid, the steps end in device-wide barriers and the
// reachability test in a device-wide OR.
template <typename CT, bool HAS_TIME, bool MINCOST, int W, bool LD = false, int MC = 0, int WX = W, int WI = W,
          bool GRID = false>
struct SweepG {
  static_assert(WX <= W && WI <= W, "short slots");
  static_assert(!(GRID && LD), "the device-wide sweep keeps its rows in global memory");
  typedef typename DistPtr<LD>::type DP;
  const PGC_GLOBAL SolverJob* J;
  DP D;
  DP Dmin;
  PGC_LDS Shared* sh;
  PGC_GLOBAL int32_t* pred;
  PGC_LDS int32_t* predL;  // pred in LDS (small graphs, LD only), selected by predLds
  int predLds;
  SlotClass Sa, Sx;
  int npl, ds, K, L, nC, wave, lane, nw;
  int p0, p1, odd;
  GridState* gs;  // GRID only
  uint64_t bgen;
  uint32_t nblk;
  bool alive = true;  // GRID: no barrier has timed out

  // the barrier between two steps (GRID: device-wide, without the L2
  // write-back: the rows that cross workgroups are stored through dput)
  __device__ __forceinline__ void sync() {
    if constexpr (GRID)
      alive = grid_sync<false>(gs, bgen, nblk, (int*)&sh->gridOk) && alive;
    else
      __syncthreads();
  }

  // distance rows and per-node minima. GRID: stored with device-scope
  // (write-through) stores, read with ordinary loads after the step
  // barrier's acquire
  __device__ __forceinline__ double dget(const DP p) const { return *p; }
  __device__ __forceinline__ void dput(DP p, double v) const {
    if constexpr (GRID)
      __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
      *p = v;
  }

  __device__ __forceinline__ bool bor(bool pred_) {
    if constexpr (GRID) {
      GridOr g;
      g.prev0 = (unsigned)p0;
      g.prev1 = (unsigned)p1;
      g.odd = odd != 0;
      const bool any = g(gs, bgen, nblk, (int*)&sh->gridFlag, (int*)&sh->gridOk, pred_, alive);
      p0 = (int)g.prev0;
      p1 = (int)g.prev1;
      odd = g.odd ? 1 : 0;
      return any && alive;
    }
    PGC_LDS int* cnt = odd ? &sh->orCount[1] : &sh->orCount[0];
    const unsigned long long b = __ballot(pred_);
    if (b && lane == 0) __atomic_fetch_add(cnt, 1, __ATOMIC_RELAXED);
    __syncthreads();
    const int c = uni(*cnt);
    bool any;
    if (odd) {
      any = c != p1;
      p1 = c;
    } else {
      any = c != p0;
      p0 = c;
    }
    odd = !odd;
    return any;
  }

  __device__ __forceinline__ double weight(CT c, float t, double ratio, double rt) const {
    if (MINCOST) return (double)c;
    return __dsub_rn((double)c, HAS_TIME ? __dmul_rn(ratio, (double)t) : rt);
  }

  template <int N>
  __device__ __forceinline__ void load(const SlotClass& S, int b, int (&s)[N], CT (&c)[N], float (&t)[N]) const {
    const PGC_GLOBAL int* __restrict__ src = gp(S.src);
    const PGC_GLOBAL CT* __restrict__ cost = gp((const CT*)S.cost);
    const PGC_GLOBAL float* __restrict__ tim = gp(S.time);
#pragma unroll
    for (int j = 0; j < N; ++j) {
      s[j] = src[b + j * 64];
      c[j] = cost[b + j * 64];
      if (HAS_TIME) t[j] = tim[b + j * 64];
    }
  }

  // First-minimum scan and write-back of one interlayer vertex v whose slots
  // and gathered source distances are loaded.
  template <int N>
  __device__ __forceinline__ bool inter_relax(int layer, int cur, int prev, int v, const int (&s)[N],
                                              const CT (&cc)[N], const float (&t)[N], const double (&dv)[N],
                                              double ratio, double rt, double ub) {
    double best = PGC_INF;
    int bs = npl;
    CT bc = 0;
    float bt = 0;
#pragma unroll
    for (int j = 0; j < N; ++j) {
      const double nv = dv[j] + weight(cc[j], t[j], ratio, rt);
      const bool better = nv < best;
      best = pgc_min(best, nv);
      bs = better ? s[j] : bs;
      bc = better ? cc[j] : bc;
      if (HAS_TIME) bt = better ? t[j] : bt;
    }
    return inter_finish(layer, cur, prev, v, best, bs, bc, bt, ratio, rt, ub);
  }

  // The write-back of an interlayer vertex after its scan: depth of the
  // winning source, predecessor, distance.
  __device__ __forceinline__ bool inter_finish(int layer, int cur, int prev, int v, double best, int bs, CT bc,
                                               float bt, double ratio, double rt, double ub) {
    double res = PGC_INF;
    bool imp = false;
    if (best < PGC_INF && best < ub) {
      int ks = 0;
      if (K > 1) {
        const double bw = weight(bc, bt, ratio, rt);
        ks = recover_depth([&](int kk) { return dget(D + (prev + kk) * ds + bs); }, K, bw, best);
      }
      const int64_t pi = (int64_t)layer * npl * K + v;
      const int pv = bs + (layer - 1) * npl * K + ks * npl;
      if (LD && predLds)
        predL[pi] = pv;
      else
        pred[pi] = pv;
      res = best;
      imp = true;
    }
    dput(D + cur * ds + v, res);
    return imp;
  }

  // The same for an intralayer vertex at depth k (and Dmin at the last depth).
  template <int N>
  __device__ __forceinline__ bool intra_relax(int layer, int cur, int k, int v, const int (&s)[N],
                                              const CT (&cc)[N], const float (&t)[N], const double (&dv)[N],
                                              double ratio, double rt, double ub) {
    double best = PGC_INF;
    int bs = npl;
#pragma unroll
    for (int j = 0; j < N; ++j) {
      const double nv = dv[j] + weight(cc[j], t[j], ratio, rt);
      const bool better = nv < best;
      best = pgc_min(best, nv);
      bs = better ? s[j] : bs;
    }
    return intra_finish(layer, cur, k, v, best, bs, ub);
  }

  __device__ __forceinline__ bool intra_finish(int layer, int cur, int k, int v, double best, int bs, double ub) {
    const double curv = (layer == 0) ? dget(D + (cur + k) * ds + v) : PGC_INF;
    double res = curv;
    bool imp = false;
    if (best < curv && best < ub) {
      const int64_t pi = (int64_t)layer * npl * K + (int64_t)k * npl + v;
      const int pv = bs + layer * npl * K + (k - 1) * npl;
      if (LD && predLds)
        predL[pi] = pv;
      else
        pred[pi] = pv;
      res = best;
      imp = true;
    }
    dput(D + (cur + k) * ds + v, res);
    if (k == K - 1) {
      double m = dget(D + cur * ds + v);
      for (int kk = 1; kk < K; ++kk) {
        const double x = dget(D + (cur + kk) * ds + v);
        m = x < m ? x : m;
      }
      dput(Dmin + v, m);
    }
    return imp;
  }

  // Chunks per loop iteration. One: with two chunks' loads and gathers in
  // flight together the global-table sweep (C2-size graphs) measured no
  // faster, and one chunk keeps the registers free.
  static constexpr int NG = 1;

  // One interlayer (k == 0, N = WX slots) or intralayer (depth k, N = WI)
  // step over the wave's chunks.
  template <int N>
  __device__ __forceinline__ bool step_n(int layer, int cur, int prev, int k, double ratio, double rt, double ub) {
    bool imp = false;
    const bool inter = k == 0;
    const SlotClass& S = inter ? Sx : Sa;
    const DP Dg = inter ? ((K == 1) ? D + prev * ds : Dmin) : D + (cur + k - 1) * ds;
    for (int c0 = wave; c0 < nC; c0 += NG * nw) {
      int s[NG][N];
      CT cc[NG][N];
      float t[NG][N];
#pragma unroll
      for (int g = 0; g < NG; ++g) {
        const int c = c0 + g * nw;
        load(S, layer * S.lstride + (c < nC ? c : c0) * W * 64 + lane, s[g], cc[g], t[g]);
      }
      double dv[NG][N];
#pragma unroll
      for (int g = 0; g < NG; ++g)
#pragma unroll
        for (int j = 0; j < N; ++j) dv[g][j] = dget(Dg + s[g][j]);
#pragma unroll
      for (int g = 0; g < NG; ++g) {
        const int c = c0 + g * nw;
        const int v = c * 64 + lane;
        if (c >= nC || v >= npl) continue;
        if (inter)
          imp |= inter_relax(layer, cur, prev, v, s[g], cc[g], t[g], dv[g], ratio, rt, ub);
        else
          imp |= intra_relax(layer, cur, k, v, s[g], cc[g], t[g], dv[g], ratio, rt, ub);
      }
    }
    return imp;
  }
  // Runtime slot width (W == 0: a node of the graph has more than 32 in-edges
  // of one class). A chunk's slots are scanned in groups of eight, as many as
  // its largest in-degree needs (SlotClass::grp), the running first minimum
  // carried across the groups in slot order.
  __device__ __forceinline__ bool step_rt(int layer, int cur, int prev, int k, double ratio, double rt, double ub) {
    bool imp = false;
    const bool inter = k == 0;
    const SlotClass& S = inter ? Sx : Sa;
    const DP Dg = inter ? ((K == 1) ? D + prev * ds : Dmin) : D + (cur + k - 1) * ds;
    const int w = S.w;
    for (int c = wave; c < nC; c += nw) {
      const int ng = uni(gp(S.grp)[c]);
      const int base = layer * S.lstride + c * w * 64 + lane;
      double best = PGC_INF;
      int bs = npl;
      CT bc = 0;
      float bt = 0;
      for (int g = 0; g < ng; ++g) {
        int sg[8];
        CT cg[8];
        float tg[8];
        load<8>(S, base + g * 8 * 64, sg, cg, tg);
        double dv[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) dv[j] = dget(Dg + sg[j]);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const double nv = dv[j] + weight(cg[j], tg[j], ratio, rt);
          const bool better = nv < best;
          best = pgc_min(best, nv);
          bs = better ? sg[j] : bs;
          bc = better ? cg[j] : bc;
          if (HAS_TIME) bt = better ? tg[j] : bt;
        }
      }
      const int v = c * 64 + lane;
      if (v >= npl) continue;
      imp |= inter ? inter_finish(layer, cur, prev, v, best, bs, bc, bt, ratio, rt, ub)
                   : intra_finish(layer, cur, k, v, best, bs, ub);
    }
    return imp;
  }

  __device__ __forceinline__ bool inter_step(int layer, int cur, int prev, double ratio, double rt, double ub) {
    if constexpr (W == 0) return step_rt(layer, cur, prev, 0, ratio, rt, ub);
    else return step_n<WX>(layer, cur, prev, 0, ratio, rt, ub);
  }
  __device__ __forceinline__ bool intra_step(int layer, int cur, int k, double ratio, double rt, double ub) {
    if constexpr (W == 0) return step_rt(layer, cur, K - cur, k, ratio, rt, ub);
    else return step_n<WI>(layer, cur, K - cur, k, ratio, rt, ub);
  }

  // --- steps with the sources held in registers (MC > 0)
  template <int N>
  __device__ __forceinline__ void load_ct(const SlotClass& S, int b, CT (&c)[N], float (&t)[N]) const {
    const PGC_GLOBAL CT* __restrict__ cost = gp((const CT*)S.cost);
    const PGC_GLOBAL float* __restrict__ tim = gp(S.time);
#pragma unroll
    for (int j = 0; j < N; ++j) {
      c[j] = cost[b + j * 64];
      if (HAS_TIME) t[j] = tim[b + j * 64];
    }
  }
  // The cached steps load each chunk's costs while the wave relaxes its
  // previous chunk (chunk i+1's loads are issued before chunk i's gathers),
  // so only the first chunk of a step waits for its costs. A load past the
  // last chunk rereads the current one (in bounds, unused).
  __device__ __forceinline__ bool inter_cached(int layer, int cur, int prev, const int (&sX)[MC > 0 ? MC : 1][WX],
                                               double ratio, double rt, double ub) {
    bool imp = false;
    const DP Dp = (K == 1) ? D + prev * ds : Dmin;
    CT ccs[MC > 0 ? MC : 1][WX];
    float ts[MC > 0 ? MC : 1][WX];
    load_ct(Sx, layer * Sx.lstride + (wave < nC ? wave : 0) * W * 64 + lane, ccs[0], ts[0]);
#pragma unroll
    for (int i = 0; i < MC; ++i) {
      const int c = wave + i * nw;
      if (c >= nC) break;
      if (i + 1 < MC) {
        const int c1 = c + nw;
        load_ct(Sx, layer * Sx.lstride + (c1 < nC ? c1 : c) * W * 64 + lane, ccs[i + 1 < MC ? i + 1 : i],
                ts[i + 1 < MC ? i + 1 : i]);
      }
      const CT(&cc)[WX] = ccs[i];
      const float(&t)[WX] = ts[i];
      const int v = c * 64 + lane;
      if (v >= npl) continue;
      const int(&s)[WX] = sX[i];
      double dv[WX];
#pragma unroll
      for (int j = 0; j < WX; ++j) dv[j] = dget(Dp + s[j]);
      double best = PGC_INF;
      int bs = npl;
      CT bc = 0;
      float bt = 0;
#pragma unroll
      for (int j = 0; j < WX; ++j) {
        const double nv = dv[j] + weight(cc[j], t[j], ratio, rt);
        const bool better = nv < best;
        best = pgc_min(best, nv);
        bs = better ? s[j] : bs;
        bc = better ? cc[j] : bc;
        if (HAS_TIME) bt = better ? t[j] : bt;
      }
      double res = PGC_INF;
      if (best < PGC_INF && best < ub) {
        int ks = 0;
        if (K > 1) {
          const double bw = weight(bc, bt, ratio, rt);
          ks = recover_depth([&](int kk) { return dget(D + (prev + kk) * ds + bs); }, K, bw, best);
        }
        const int64_t pi = (int64_t)layer * npl * K + v;
        const int pv = bs + (layer - 1) * npl * K + ks * npl;
        if (LD && predLds)
          predL[pi] = pv;
        else
          pred[pi] = pv;
        res = best;
        imp = true;
      }
      dput(D + cur * ds + v, res);
    }
    return imp;
  }
  __device__ __forceinline__ bool intra_cached(int layer, int cur, int k, const int (&sI)[MC > 0 ? MC : 1][WI],
                                               double ratio, double rt, double ub) {
    bool imp = false;
    const DP Dk = D + (cur + k - 1) * ds;
    CT ccs[MC > 0 ? MC : 1][WI];
    float ts[MC > 0 ? MC : 1][WI];
    load_ct(Sa, layer * Sa.lstride + (wave < nC ? wave : 0) * W * 64 + lane, ccs[0], ts[0]);
#pragma unroll
    for (int i = 0; i < MC; ++i) {
      const int c = wave + i * nw;
      if (c >= nC) break;
      if (i + 1 < MC) {
        const int c1 = c + nw;
        load_ct(Sa, layer * Sa.lstride + (c1 < nC ? c1 : c) * W * 64 + lane, ccs[i + 1 < MC ? i + 1 : i],
                ts[i + 1 < MC ? i + 1 : i]);
      }
      const CT(&cc)[WI] = ccs[i];
      const float(&t)[WI] = ts[i];
      const int v = c * 64 + lane;
      if (v >= npl) continue;
      const int(&s)[WI] = sI[i];
      double dv[WI];
#pragma unroll
      for (int j = 0; j < WI; ++j) dv[j] = dget(Dk + s[j]);
      double best = PGC_INF;
      int bs = npl;
#pragma unroll
      for (int j = 0; j < WI; ++j) {
        const double nv = dv[j] + weight(cc[j], t[j], ratio, rt);
        const bool better = nv < best;
        best = pgc_min(best, nv);
        bs = better ? s[j] : bs;
      }
      const double curv = (layer == 0) ? dget(D + (cur + k) * ds + v) : PGC_INF;
      double res = curv;
      if (best < curv && best < ub) {
        const int64_t pi = (int64_t)layer * npl * K + (int64_t)k * npl + v;
        const int pv = bs + layer * npl * K + (k - 1) * npl;
        if (LD && predLds)
          predL[pi] = pv;
        else
          pred[pi] = pv;
        res = best;
        imp = true;
      }
      dput(D + (cur + k) * ds + v, res);
      if (k == K - 1) {
        double m = dget(D + cur * ds + v);
        for (int kk = 1; kk < K; ++kk) {
          const double x = dget(D + (cur + kk) * ds + v);
          m = x < m ? x : m;
        }
        dput(Dmin + v, m);
      }
    }
    return imp;
  }
  // Device-wide sweep with K == 2 and one chunk per wave. Every step reads its
  // rows from memory shared by the whole grid, so each dependent read costs a
  // full memory round trip; this schedule has one per step. The interlayer
  // step gathers both depth rows of its sources and takes the first minimum in
  // (slot, depth) order in registers. That is the winner the minimum-over-
  // depths row and the depth recovery give: rounding is monotonic, so
  // min_k(D_k + w) = min_k(D_k) + w, and the first slot attaining the minimum
  // and the first depth of that slot attaining it are the same in both. With
  // no minimum row, the intralayer step writes only its depth row.
  // The interlayer gathers are the step's bound: scattered 8-byte reads, each
  // lane its own line through the CU's L1. Every node therefore also stores
  // its two depth values of layer l side by side (pairs[l & 1][v] = (D0, D1),
  // after the table's rows and minimum row), and the interlayer step gathers
  // each source's two depths with one 16-byte load: half the requests. The
  // rows are still written; the rest of the solve reads them.
  // The sources are held as 32-bit byte offsets into a row (a row has at most
  // 2^28 entries here), so every gather is a scalar row base plus a vector
  // offset and no 64-bit address stays live across the loop.
  // The empty asm marks the offset as rewritten at every use: without it the
  // compiler hoists a 64-bit address per (slot, row parity) out of the layer
  // loop and spills them.
  __device__ __forceinline__ double rget(const DP row, uint32_t off) const {
    asm volatile("" : "+v"(off));
    return *(const PGC_GLOBAL double*)((const PGC_GLOBAL char*)row + off);
  }
  // both depths of a node from the interleaved pairs (one 16-byte load);
  // off is the node's 8-byte row offset, doubled
  typedef double dpair __attribute__((ext_vector_type(2)));
  __device__ __forceinline__ dpair pget(const DP pairs, uint32_t off) const {
    off <<= 1;
    asm volatile("" : "+v"(off));
    return *(const PGC_GLOBAL dpair*)((const PGC_GLOBAL char*)pairs + off);
  }
  __device__ __forceinline__ int run_layers_grid_k2(double ratio, double rt, double ub) {
    const int c = wave;
    const bool own = c < nC;
    const int o = (own ? c : 0) * W * 64 + lane;
    const int v = c * 64 + lane;
    const bool act = own && v < npl;
    // the sources of both classes stay in registers (layer 1's interlayer
    // and layer 0's intralayer pattern), so every gather of a step is issued
    // at once, beside its cost loads
    uint32_t oX[WX], oI[WI];
    {
      const PGC_GLOBAL int* __restrict__ srcX = gp(Sx.src);
      const PGC_GLOBAL int* __restrict__ srcI = gp(Sa.src);
#pragma unroll
      for (int j = 0; j < WX; ++j) oX[j] = (uint32_t)srcX[Sx.lstride + o + j * 64] << 3;
#pragma unroll
      for (int j = 0; j < WI; ++j) oI[j] = (uint32_t)srcI[o + j * 64] << 3;
    }
    const DP P = D + ((2 * K + 1) * ds + 1) / 2 * 2;
    // the padding sentinel of both pair buffers (the rows' entry npl is +inf)
    if (blockIdx.x == 0 && threadIdx.x < 4) dput(P + (threadIdx.x >> 1) * 2 * ds + 2 * npl + (threadIdx.x & 1), PGC_INF);
    for (int layer = 0; layer <= L; ++layer) {
      const int cur = (layer & 1) * 2;
      const DP Pc = P + (layer & 1) * 2 * ds;
      bool imp = false;
      const bool check = layer == L - 1 && layer > 0;
      if (layer > 0) {
        if (own) {
          CT cc[WX];
          float t[WX];
          load_ct(Sx, layer * Sx.lstride + o, cc, t);
          const DP Pp = P + ((layer - 1) & 1) * 2 * ds;
          double d0[WX], d1[WX];
#pragma unroll
          for (int j = 0; j < WX; ++j) {
            const dpair pr = pget(Pp, oX[j]);
            d0[j] = pr.x;
            d1[j] = pr.y;
          }
          double best = PGC_INF;
          uint32_t bo = (uint32_t)npl << 3;
          int ks = 0;
#pragma unroll
          for (int j = 0; j < WX; ++j) {
            const double w = weight(cc[j], t[j], ratio, rt);
            const double a = d0[j] + w;
            const bool b0 = a < best;
            best = pgc_min(best, a);
            const double b = d1[j] + w;
            const bool b1 = b < best;
            best = pgc_min(best, b);
            bo = b0 || b1 ? oX[j] : bo;
            ks = b1 ? 1 : (b0 ? 0 : ks);
          }
          if (act) {
            double res = PGC_INF;
            if (best < PGC_INF && best < ub) {
              pred[(int64_t)layer * npl * 2 + v] = (int)(bo >> 3) + (layer - 1) * npl * 2 + ks * npl;
              res = best;
              imp = true;
            }
            dput(D + cur * ds + v, res);
            dput(Pc + 2 * v, res);
          }
        }
        sync();
      }
      if (layer < L) {
        if (own) {
          CT cc[WI];
          float t[WI];
          load_ct(Sa, layer * Sa.lstride + o, cc, t);
          const DP D0 = D + cur * ds;
          double dv[WI];
#pragma unroll
          for (int j = 0; j < WI; ++j) dv[j] = rget(D0, oI[j]);
          const double curv = (layer == 0 && act) ? dget(D + (cur + 1) * ds + v) : PGC_INF;
          double best = PGC_INF;
          uint32_t bo = (uint32_t)npl << 3;
#pragma unroll
          for (int j = 0; j < WI; ++j) {
            const double nv = dv[j] + weight(cc[j], t[j], ratio, rt);
            const bool better = nv < best;
            best = pgc_min(best, nv);
            bo = better ? oI[j] : bo;
          }
          if (act) {
            double res = curv;
            if (best < curv && best < ub) {
              pred[(int64_t)layer * npl * 2 + npl + v] = (int)(bo >> 3) + layer * npl * 2;
              res = best;
              imp = true;
            }
            dput(D + (cur + 1) * ds + v, res);
            dput(Pc + 2 * v + 1, res);
            // layer 0's depth-0 row was set before the sweep
            if (layer == 0) dput(Pc + 2 * v, dget(D + cur * ds + v));
          }
        }
      } else if (act) {
        // the closing layer has no intralayer edges: its depth-1 row is +inf
        dput(D + (cur + 1) * ds + v, PGC_INF);
      }
      if (check) {
        if (!bor(imp)) return 0;
      } else {
        sync();
      }
    }
    return 1;
  }

  __device__ __forceinline__ int run_layers_cached(double ratio, double rt, double ub) {
    constexpr int M = MC > 0 ? MC : 1;
    int sX[M][WX], sI[M][WI];
    {
      const PGC_GLOBAL int* __restrict__ srcX = gp(Sx.src);
      const PGC_GLOBAL int* __restrict__ srcI = gp(Sa.src);
#pragma unroll
      for (int i = 0; i < M; ++i) {
        const int c = wave + i * nw;
        const int o = (c < nC ? c : 0) * W * 64 + lane;
#pragma unroll
        for (int j = 0; j < WX; ++j) sX[i][j] = srcX[Sx.lstride + o + j * 64];  // interlayer pattern of layer 1
#pragma unroll
        for (int j = 0; j < WI; ++j) sI[i][j] = K > 1 ? srcI[o + j * 64] : 0;  // intralayer pattern of layer 0
      }
    }
    for (int layer = 0; layer <= L; ++layer) {
      const int cur = (layer & 1) * K;
      const int prev = K - cur;
      bool imp = false;
      const bool check = layer == L - 1 && layer > 0;
      if (layer > 0) {
        imp = inter_cached(layer, cur, prev, sX, ratio, rt, ub);
        if (K == 1) {
          if (check) {
            if (!bor(imp)) return 0;
          } else {
            sync();
          }
          continue;
        }
        sync();
      }
      for (int k = 1; k < K; ++k) {
        // the closing layer L has its own (empty) intralayer slots
        imp |= layer < L ? intra_cached(layer, cur, k, sI, ratio, rt, ub) : intra_step(layer, cur, k, ratio, rt, ub);
        if (k == K - 1 && check) {
          if (!bor(imp)) return 0;
        } else {
          sync();
        }
      }
    }
    return 1;
  }

  __device__ __forceinline__ int run_layers(double ratio, double ub) {
    const double rt = __dmul_rn(ratio, uni(J->time_const));
    if constexpr (MC > 0) {
      if (Sx.src_inv && (K == 1 || Sa.src_inv) && !(uni((int)J->flags) & (int)FLAG_NO_CACHED_SOURCES))
        return run_layers_cached(ratio, rt, ub);
    }
    for (int layer = 0; layer <= L; ++layer) {
      const int cur = (layer & 1) * K;
      const int prev = K - cur;
      bool imp = false;
      // only layer L-1 decides (see the LDS sweep)
      const bool check = layer == L - 1 && layer > 0;
      if (layer > 0) {
        imp = inter_step(layer, cur, prev, ratio, rt, ub);
        if (K == 1) {
          if (check) {
            if (!bor(imp)) return 0;
          } else {
            sync();
          }
          continue;
        }
        sync();
      }
      for (int k = 1; k < K; ++k) {
        imp |= intra_step(layer, cur, k, ratio, rt, ub);
        if (k == K - 1 && check) {
          if (!bor(imp)) return 0;
        } else {
          sync();
        }
      }
    }
    return 1;
  }
};

// K2: the device-wide K == 2 schedule (run_layers_grid_k2), a separate
// instance so that its registers are allocated apart from the general loops;
// the caller has checked that both slot classes have layer-invariant sources.
template <typename CT, bool HAS_TIME, bool MINCOST, int W, bool LD = false, int MC = 0, int WX = W, int WI = W,
          bool GRID = false, bool K2 = false>
__device__ __noinline__ SweepState sweep_layers_global(const SolverJob* Jp, double* Dg, double* Dming, Shared* shg,
                                                       int32_t* predLg, int predLds, int p0, int p1, int odd,
                                                       double ratio, double ub, GridState* gs = nullptr,
                                                       uint64_t* bgen = nullptr) {
  typedef SweepG<CT, HAS_TIME, MINCOST, W, LD, MC, WX, WI, GRID> SW;
  SW w;
  w.J = (const PGC_GLOBAL SolverJob*)uni(Jp);
  w.D = (typename SW::DP)uni(Dg);
  w.Dmin = (typename SW::DP)uni(Dming);
  w.sh = (PGC_LDS Shared*)uni(shg);
  w.pred = (PGC_GLOBAL int32_t*)uni(w.J->pred);
  w.predL = (PGC_LDS int32_t*)uni(predLg);
  w.predLds = LD ? uni(predLds) : 0;
  w.Sa = uni_class(w.J->intra);
  w.Sx = uni_class(w.J->inter);
  w.npl = uni(w.J->npl);
  w.ds = w.npl + 1;
  w.K = uni(w.J->K);
  w.L = uni(w.J->L);
  w.nC = uni(w.J->nChunks);
  w.lane = threadIdx.x & 63;
  if constexpr (GRID) {
    w.wave = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (PGC_BLOCK / PGC_WAVE) + (threadIdx.x >> 6)));
    w.nw = (int)gridDim.x * (PGC_BLOCK / PGC_WAVE);
    w.gs = uni(gs);
    w.bgen = uni((int64_t)*bgen);
    w.nblk = gridDim.x;
  } else {
    w.wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    w.nw = PGC_BLOCK / PGC_WAVE;
  }
  w.p0 = uni(p0);
  w.p1 = uni(p1);
  w.odd = uni(odd);
  SweepState r;
  if constexpr (K2 && GRID && MC == 1 && W != 0)
    r.complete = w.run_layers_grid_k2(uni(ratio), __dmul_rn(uni(ratio), uni(w.J->time_const)), uni(ub));
  else
    r.complete = w.run_layers(uni(ratio), uni(ub));
  r.p0 = w.p0;
  r.p1 = w.p1;
  r.odd = w.odd;
  if constexpr (GRID) *bgen = w.bgen;
  return r;
}

}  // namespace pgc
