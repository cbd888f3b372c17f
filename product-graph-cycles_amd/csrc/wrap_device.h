// Wrap map of the layer-0 vertices and the candidate walk of the negative-
// cycle search, for the solver variants with the distance table in LDS (at
// most MAXC chunks per wave).
//
// Both run once per Bellman-Ford sweep. They are separate (non-inlined)
// functions for the same reason as the sweep: inside the solver kernel the
// branch-and-bound state holds most registers, and these loops would run on
// spilled addresses and counters. Every argument is made wave-uniform, and
// the placement of the predecessor table and of the wrap map (LDS or global
// memory) is decided once per call, not per step.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sweep_device.h"

namespace pgc {

struct WrapGeom {
  int npl, nC, wave, lane, nw;
  __device__ __forceinline__ int vert(int i) const {
    const int c = wave + i * nw;
    if (c >= nC) return -1;
    const int v = c * 64 + lane;
    return v < npl ? v : -1;
  }
};

__device__ __forceinline__ WrapGeom wrap_geom(int npl, int nC) {
  WrapGeom g;
  g.npl = uni(npl);
  g.nC = uni(nC);
  g.wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  g.lane = threadIdx.x & 63;
  g.nw = PGC_BLOCK / PGC_WAVE;
  return g;
}

// Chase of every owned layer-0 vertex v: follow predecessors from v until a
// layer-0 node id (< npl) is reached. wnext[v] = that node, wlen[v] = steps;
// -1 when a chain ends (no predecessor), -2 past stride-1 steps.
//
// The chains of a wave advance in lockstep, one predecessor per iteration,
// with the outcome of every lane kept by selects (no per-lane branches): the
// loop ends when no lane of the wave is still walking. Lanes that are done
// read entry 0 (in bounds, unused).
template <int MAXC, typename PT, typename WT>
__device__ __forceinline__ void chase_run(const WrapGeom& g, PT P, WT Wn, WT Wl, int stride) {
  int pp[MAXC], pos[MAXC], rn[MAXC], rl[MAXC];
  bool done[MAXC];
#pragma unroll
  for (int i = 0; i < MAXC; ++i) {
    const int v = g.vert(i);
    pp[i] = P[v >= 0 ? v : 0];
    pos[i] = 0;
    rn[i] = -1;
    rl[i] = 0;
    done[i] = v < 0 || pp[i] == -1;
  }
  while (true) {
    bool walking = false;
#pragma unroll
    for (int i = 0; i < MAXC; ++i) {
      const bool act = !done[i];
      const bool hit = act && pp[i] < g.npl;
      const bool cap = act && !hit && pos[i] >= stride - 1;
      const bool adv = act && !hit && !cap;
      const int nxt = P[adv ? pp[i] : 0];
      pos[i] += (hit || adv) ? 1 : 0;
      const bool end = adv && nxt == -1;
      rn[i] = hit ? pp[i] : (cap ? -2 : (end ? -1 : rn[i]));
      rl[i] = (hit || cap || end) ? pos[i] : rl[i];
      pp[i] = adv ? nxt : pp[i];
      done[i] = done[i] || hit || cap || end;
      walking |= !done[i];
    }
    if (!__any(walking)) break;
  }
#pragma unroll
  for (int i = 0; i < MAXC; ++i) {
    const int v = g.vert(i);
    if (v >= 0) {
      Wn[v] = rn[i];
      Wl[v] = rl[i];
    }
  }
}

template <int MAXC>
__device__ __noinline__ void chase_wraps_fn(int32_t* pred, int predLds, int32_t* wn, int32_t* wl, int wrapLds, int npl,
                                            int nC, int stride) {
  const WrapGeom g = wrap_geom(npl, nC);
  pred = uni(pred);
  wn = uni(wn);
  wl = uni(wl);
  stride = uni(stride);
  if (uni(predLds)) {
    const PGC_LDS int32_t* P = (const PGC_LDS int32_t*)pred;
    if (uni(wrapLds))
      chase_run<MAXC>(g, P, (PGC_LDS int32_t*)wn, (PGC_LDS int32_t*)wl, stride);
    else
      chase_run<MAXC>(g, P, (PGC_GLOBAL int32_t*)wn, (PGC_GLOBAL int32_t*)wl, stride);
  } else {
    const PGC_GLOBAL int32_t* P = (const PGC_GLOBAL int32_t*)pred;
    if (uni(wrapLds))
      chase_run<MAXC>(g, P, (PGC_LDS int32_t*)wn, (PGC_LDS int32_t*)wl, stride);
    else
      chase_run<MAXC>(g, P, (PGC_GLOBAL int32_t*)wn, (PGC_GLOBAL int32_t*)wl, stride);
  }
}

// Walk of the reference (cost_time_ratio_solver.cpp:443-489) on the wrap map:
// the first repeated layer-0 vertex of the walk from v (Floyd's tortoise and
// hare over wnext), accepted when the cycle's step count is within capSteps.
template <typename WT>
__device__ __forceinline__ bool walk_wrap(WT Wn, WT Ln, int npl, int v, int64_t capSteps, int& vstar) {
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

struct Candidate {
  double key;
  int v;
};

// This thread's best candidate: owned branch vertices with a negative (or,
// searching zero cycles, |D| < tol) layer-0 value whose walk closes within
// the step cap; key -v (pick last) or D, ties to the lower vertex.
template <int MAXC, typename WT>
__device__ __forceinline__ Candidate candidates_run(const WrapGeom& g, const PGC_LDS double* D, uint32_t inb,
                                                    int zeroSearch, double tol, int pickLast, int64_t capSteps, WT Wn,
                                                    WT Wl) {
  Candidate c;
  c.key = PGC_INF;
  c.v = 0x7fffffff;
#pragma unroll
  for (int i = 0; i < MAXC; ++i) {
    const int v = g.vert(i);
    if (v < 0 || ((inb >> i) & 1u) == 0) continue;
    const double lb = D[v];
    if (!zeroSearch && lb >= 0) continue;
    if (zeroSearch && (lb <= -tol || lb >= tol)) continue;
    int vs;
    if (!walk_wrap(Wn, Wl, g.npl, v, capSteps, vs)) continue;
    const double key = pickLast ? -(double)v : lb;
    if (key < c.key || (key == c.key && v < c.v)) {
      c.key = key;
      c.v = v;
    }
  }
  return c;
}

// inb: bit i set when the thread's owned vertex i is in the current branch.
template <int MAXC>
__device__ __noinline__ Candidate candidates_fn(double* D, uint32_t inb, int zeroSearch, double tol, int pickLast,
                                                int64_t capSteps, int32_t* wn, int32_t* wl, int wrapLds, int npl,
                                                int nC) {
  const WrapGeom g = wrap_geom(npl, nC);
  const PGC_LDS double* Dl = (const PGC_LDS double*)uni(D);
  zeroSearch = uni(zeroSearch);
  tol = uni(tol);
  pickLast = uni(pickLast);
  capSteps = uni(capSteps);
  wn = uni(wn);
  wl = uni(wl);
  if (uni(wrapLds))
    return candidates_run<MAXC>(g, Dl, inb, zeroSearch, tol, pickLast, capSteps, (const PGC_LDS int32_t*)wn,
                                (const PGC_LDS int32_t*)wl);
  return candidates_run<MAXC>(g, Dl, inb, zeroSearch, tol, pickLast, capSteps, (const PGC_GLOBAL int32_t*)wn,
                              (const PGC_GLOBAL int32_t*)wl);
}

}  // namespace pgc
