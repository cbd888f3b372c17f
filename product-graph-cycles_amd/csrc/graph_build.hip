// On-device construction of the layered product graph and its per-edge
// cost/time, q-unit quantization, and the sliced-ELL in-edge layout used by
// the solver kernels.
//
// Reference counterparts:
//   ProductSpace::computeCombinations        product_spaces/product_space.cpp:19-132
//   DeformationEnergy::computeEnergy         energy/deformationEnergy.cpp:11-103
//   robustFloatToIntConversion + median      cost_time_ratio_solver.cpp:709-721, helper/utils.hpp:196-209
//   buildingInGraph (+ VECSET flattening)    cost_time_ratio_solver.cpp:42-66, ctr_bnb_subroutine.cuh:232-287
// All kernels are HBM-streaming integer/fp64 work: one thread per product
// edge (or per slot), coalesced row-major writes, no LDS staging needed
// except for the 256-bin radix-select histograms.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <math.h>
#include <stdint.h>

#include "fastdiv.h"
#include "pgc_internal.h"

namespace pgc {

// ------------------------------------------------------- product space (a1)
// Edge e of layer i = e / B, r = e % B with B = 2|EX| + nVX (|EX| + nVX when
// intralayer edges are pruned): r < 2|EX| -> X edge j = r/2, intralayer when r
// is even; otherwise the degenerate edge of vertex r - 2|EX|.
__global__ void k_product_space(const int32_t* __restrict__ EX, int64_t nEX, const int32_t* __restrict__ EY,
                                int64_t nEY, int64_t nVX, int prune, int32_t* __restrict__ rows,
                                int32_t* __restrict__ src, int32_t* __restrict__ trgt, int64_t E) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  const int64_t per_x = prune ? 1 : 2;
  const int64_t B = per_x * nEX + nVX;
  const int64_t i = e / B;
  const int64_t r = e - i * B;
  const int64_t t = (i + 1) % nEY;
  const int32_t y0 = EY[2 * i], y1 = EY[2 * i + 1];
  int32_t r0, r1, x0, x1;
  int64_t s, d;
  if (r < per_x * nEX) {
    const int64_t j = r / per_x;
    const bool intra = !prune && (r % 2 == 0);
    x0 = EX[2 * j];
    x1 = EX[2 * j + 1];
    r0 = y0;
    r1 = intra ? y0 : y1;
    s = i * nVX + x0;
    d = (intra ? i : t) * nVX + x1;
  } else {
    const int64_t x = r - per_x * nEX;
    x0 = x1 = (int32_t)x;
    r0 = y0;
    r1 = y1;
    s = i * nVX + x;
    d = t * nVX + x;
  }
  int4 row;
  row.x = r0;
  row.y = r1;
  row.z = x0;
  row.w = x1;
  reinterpret_cast<int4*>(rows)[e] = row;
  src[e] = (int32_t)s;
  trgt[e] = (int32_t)d;
}

// ------------------------------------------------------------- energy (a3)
// Length of a product edge (the line-integral factor) and its feature sum.
__device__ __forceinline__ double edge_length(const double* __restrict__ VX, const double* __restrict__ VY,
                                              int64_t y0, int64_t y1, int64_t x0, int64_t x1) {
  double len = 0;
  for (int j = 0; j < 3; ++j) {
    const double a = VY[3 * y0 + j] - VY[3 * y1 + j];
    len = len + __dmul_rn(a, a);
    const double b = VX[3 * x0 + j] - VX[3 * x1 + j];
    len = len + __dmul_rn(b, b);
  }
  return sqrt(len);
}

// FD is |VX| x |VY| row-major (fp64); out is column-major E x ncols.
__global__ void k_energy(const double* __restrict__ VX, const double* __restrict__ VY,
                         const int32_t* __restrict__ rows, int64_t E, const double* __restrict__ FD, int64_t nVY,
                         int lineIntegral, int lengthMode, double* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  const int4 row = reinterpret_cast<const int4*>(rows)[e];
  const int64_t y0 = row.x, y1 = row.y, x0 = row.z, x1 = row.w;
  const double len = lineIntegral ? edge_length(VX, VY, y0, y1, x0, x1) : 1.0;
  const double fd = FD[x0 * nVY + y0] + FD[x1 * nVY + y1];
  if (!lengthMode) {
    out[e] = __dmul_rn(__dmul_rn(0.5, fd), len);
  } else {
    out[e] = fd;
    out[E + e] = len;
  }
}

// k_energy for a list of graphs (blockIdx.y = graph), each with its own
// feat_diff; a float32 feat_diff is widened exactly.
// Row (y0, y1, x0, x1) of edge e of a non-conjugate product space, as
// k_product_space writes it: layer i = e / B with B edges per layer (per X
// edge an intralayer and an interlayer edge, or only the interlayer one when
// pruned, then one stay edge per X vertex).
__device__ __forceinline__ int4 decode_row(const EItem& it, int64_t e) {
  const int64_t per_x = it.prune ? 1 : 2;
  const int64_t B = per_x * it.nEX + it.nVX;
  const int64_t i = (e < 0xffffffffll && B < 0xffffffffll) ? (int64_t)((uint32_t)e / (uint32_t)B) : e / B;
  const int64_t r = e - i * B;
  const int32_t y0 = it.ey[2 * i], y1 = it.ey[2 * i + 1];
  int4 row;
  row.x = y0;
  if (r < per_x * it.nEX) {
    const int64_t j = r / per_x;
    const bool intra = !it.prune && (r % 2 == 0);
    row.y = intra ? y0 : y1;
    row.z = it.ex[2 * j];
    row.w = it.ex[2 * j + 1];
  } else {
    row.y = y1;
    row.z = row.w = (int32_t)(r - per_x * it.nEX);
  }
  return row;
}

__global__ void k_benergy(const EItem* __restrict__ items) {
  const EItem it = items[blockIdx.y];
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < it.E; e += (int64_t)gridDim.x * blockDim.x) {
    const int4 row = it.ex ? decode_row(it, e) : reinterpret_cast<const int4*>(it.rows)[e];
    const int64_t y0 = row.x, y1 = row.y, x0 = row.z, x1 = row.w;
    const double len = it.line ? edge_length(it.vx, it.vy, y0, y1, x0, x1) : 1.0;
    const int64_t a = x0 * it.nVY + y0, b = x1 * it.nVY + y1;
    const double fd = it.f32 ? (double)((const float*)it.fd)[a] + (double)((const float*)it.fd)[b]
                             : ((const double*)it.fd)[a] + ((const double*)it.fd)[b];
    if (!it.lengthMode) {
      it.energy[e] = __dmul_rn(__dmul_rn(0.5, fd), len);
    } else {
      it.energy[e] = fd;
      it.energy[it.E + e] = len;
    }
  }
}

// --------------------------------- cycle ratio in raw units and its gradient
// lambda = sum_e c_e / sum_e t_e over a cycle's edges (walk order), with c_e
// the raw energy of k_energy and t_e its time column (1 per edge without
// one); d lambda / d FD[x,y] = sum over edge endpoints (x,y) of dc_e/dFD / T.

__device__ __forceinline__ void cycle_edge(const CycleItem& it, int i, int64_t& a, int64_t& b, double& fd,
                                           double& len) {
  const int4 row = reinterpret_cast<const int4*>(it.rows)[it.eids[i]];
  a = (int64_t)row.z * it.nVY + row.x;
  b = (int64_t)row.w * it.nVY + row.y;
  len = it.line ? edge_length(it.vx, it.vy, row.x, row.y, row.z, row.w) : 1.0;
  fd = it.fd ? it.fd[a] + it.fd[b] : 0.0;
}

// one workgroup per item; sequential sums in walk order (deterministic)
__global__ void k_cycle_value(const CycleItem* __restrict__ items, int n, double* __restrict__ values) {
  const CycleItem it = items[blockIdx.x];
  if (threadIdx.x != 0) return;
  double cs = 0, ts = 0;
  for (int i = 0; i < it.m; ++i) {
    int64_t a, b;
    double fd, len;
    cycle_edge(it, i, a, b, fd, len);
    if (!it.length) {
      cs += __dmul_rn(__dmul_rn(0.5, fd), len);
      ts += 1.0;
    } else {
      cs += fd;
      ts += len;
    }
  }
  values[blockIdx.x] = cs / ts;
}

__global__ void k_cycle_grad(const CycleItem* __restrict__ items, int n) {
  const CycleItem it = items[blockIdx.x];
  __shared__ double T;
  if (threadIdx.x == 0) {
    double ts = 0;
    for (int i = 0; i < it.m; ++i) {
      int64_t a, b;
      double fd, len;
      cycle_edge(it, i, a, b, fd, len);
      ts += it.length ? len : 1.0;
    }
    T = ts;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < it.m; i += blockDim.x) {
    int64_t a, b;
    double fd, len;
    cycle_edge(it, i, a, b, fd, len);
    const double coef = (it.length ? 1.0 : 0.5 * len) / T * it.gout;
    atomicAdd(&it.dfd[a], coef);
    atomicAdd(&it.dfd[b], coef);
  }
}

// ------------------------------------------- exact median by radix select
// Order-preserving map of doubles to unsigned integers.
__device__ __forceinline__ uint64_t dkey(double x) {
  const uint64_t u = (uint64_t)__double_as_longlong(x);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double dkey_inv(uint64_t k) {
  const uint64_t u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return __longlong_as_double((long long)u);
}

struct SelectState {
  uint64_t prefix[2];
  uint64_t mask[2];
  int64_t k[2];
  uint32_t hist[2][256];
  double value[2];
};

// Single-column select: eight passes of 8-bit digits, most significant
// first. Pass p histograms digit p of the keys that match each rank's prefix
// (digits 0..p-1); every block of pass p first picks digit p-1 itself from
// the previous pass's histograms (a 256-bin scan in LDS), so no launch sits
// between two passes; block 0 records the state for the next pass (two
// buffers: pass p reads buffer (p-1) & 1 and writes buffer p & 1).
struct Select1State {
  uint64_t prefix[2][2];  // [buffer][rank]
  uint64_t mask[2][2];
  int64_t k[2][2];        // rank inside the keys that match the prefix
  uint32_t hist[8][2][256];  // [pass][rank][bin]
  uint32_t rcnt[512];  // keys workgroup b gathered into its region of the gather buffer
  uint32_t over;       // some region overflowed: the later passes read every key
};

__global__ void k_select_init(Select1State* st, int64_t k0, int64_t k1) {
  for (int i = threadIdx.x; i < 8 * 2 * 256; i += blockDim.x) (&st->hist[0][0][0])[i] = 0;
  if (threadIdx.x == 0) {
    st->over = 0;
    st->prefix[0][0] = st->prefix[0][1] = 0;
    st->mask[0][0] = st->mask[0][1] = 0;
    st->k[0][0] = k0;
    st->k[0][1] = k1;
  }
}

// The digit at `shift` of each rank from histograms h (keys matching the
// prefix): the first bin b with k < (keys in bins <= b); k becomes the rank
// inside that bin (bin 255 and k minus all the keys when no bin qualifies,
// which a consistent histogram never gives). All 256 threads of the block
// take part (thread = bin) and all get the new state.
__device__ void pick_digit(const uint32_t (*h)[256], int shift, uint64_t (&pre)[2], uint64_t (&msk)[2],
                           int64_t (&k)[2]) {
  __shared__ int64_t sc[2][256];
  __shared__ int64_t rk[2];
  __shared__ uint64_t rp[2], rm[2];
  const int b = threadIdx.x;
  const int64_t c0 = h[0][b], c1 = h[1][b];
  sc[0][b] = c0;
  sc[1][b] = c1;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {
    const int64_t a0 = b >= o ? sc[0][b - o] : 0, a1 = b >= o ? sc[1][b - o] : 0;
    __syncthreads();
    sc[0][b] += a0;
    sc[1][b] += a1;
    __syncthreads();
  }
  for (int t = 0; t < 2; ++t) {
    const int64_t incl = sc[t][b], excl = incl - (t ? c1 : c0);
    const bool hit = excl <= k[t] && k[t] < incl;
    const bool none = b == 255 && k[t] >= incl;
    if (hit || none) {
      rk[t] = k[t] - (hit ? excl : incl);
      rp[t] = pre[t] | (((uint64_t)b) << shift);
      rm[t] = msk[t] | (((uint64_t)255) << shift);
    }
  }
  __syncthreads();
  for (int t = 0; t < 2; ++t) {
    k[t] = rk[t];
    pre[t] = rp[t];
    msk[t] = rm[t];
  }
}

// Histograms of digit `pass` of the keys that match each rank's prefix. In
// the first passes most keys share their high digits: when every active lane
// of a wave has the same digit, the wave adds its two match counts with one
// LDS atomic each instead of one per lane (the same totals). A thread loads
// kSelUnroll keys before it counts any, so their latencies overlap.
// With a gather buffer (cand, ccap keys), pass kSelGather also gathers the
// keys that match its prefix (either rank: the only keys any later pass
// counts), each workgroup into its own region of ccap / gridDim keys (an LDS
// counter, no global atomics), and in the later passes each workgroup reads
// its region instead of its share of all n keys, unless a region overflowed.
// The grid is the same for every pass.
constexpr int kSelUnroll = 8;
constexpr int kSelGather = 2;
__global__ void __launch_bounds__(256) k_select_hist(const double* __restrict__ x, int64_t n, Select1State* st,
                                                     int pass, double* __restrict__ cand, int64_t ccap) {
  __shared__ uint32_t h[2][256];
  h[0][threadIdx.x] = 0;
  h[1][threadIdx.x] = 0;
  uint64_t pre[2], msk[2];
  int64_t k[2];
  const int rb = pass == 0 ? 0 : (pass - 1) & 1;  // the state the previous pass used
  for (int t = 0; t < 2; ++t) {
    pre[t] = st->prefix[rb][t];
    msk[t] = st->mask[rb][t];
    k[t] = st->k[rb][t];
  }
  if (pass > 0) {
    pick_digit(st->hist[pass - 1], 64 - 8 * pass, pre, msk, k);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      for (int t = 0; t < 2; ++t) {
        st->prefix[pass & 1][t] = pre[t];
        st->mask[pass & 1][t] = msk[t];
        st->k[pass & 1][t] = k[t];
      }
    }
  }
  __syncthreads();
  const int shift = 56 - 8 * pass;
  const int lane = threadIdx.x & 63;
  // both ranks on one prefix (every pass for an odd count, the early ones
  // for an even one): count the first histogram and copy it
  const bool same = pre[0] == pre[1] && msk[0] == msk[1];
  const bool gather = cand && pass == kSelGather;
  const int64_t rs = cand ? ccap / gridDim.x : 0;  // region of this workgroup
  __shared__ uint32_t lc;
  int64_t start = (int64_t)blockIdx.x * blockDim.x * kSelUnroll;
  int64_t step = (int64_t)gridDim.x * blockDim.x * kSelUnroll;
  if (cand && pass > kSelGather && __hip_atomic_load(&st->over, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
    x = cand + (int64_t)blockIdx.x * rs;
    n = st->rcnt[blockIdx.x];
    start = 0;
    step = (int64_t)blockDim.x * kSelUnroll;
  }
  if (threadIdx.x == 0) lc = 0;
  __syncthreads();
  for (int64_t base = start + threadIdx.x; base < n; base += step) {
    double v[kSelUnroll];
#pragma unroll
    for (int u = 0; u < kSelUnroll; ++u) {
      const int64_t i = base + (int64_t)u * blockDim.x;
      v[u] = i < n ? x[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kSelUnroll; ++u) {
      if (base + (int64_t)u * blockDim.x >= n) continue;
      const uint64_t kk = dkey(v[u]);
      const uint32_t b = (uint32_t)(kk >> shift) & 255u;
      const bool in0 = (kk & msk[0]) == pre[0], in1 = !same && (kk & msk[1]) == pre[1];
      const uint64_t act = __ballot(1);
      const uint32_t bf = (uint32_t)__builtin_amdgcn_readfirstlane((int)b);
      if (__ballot(b == bf) == act) {
        const uint64_t v0 = __ballot(in0), v1 = __ballot(in1);
        if (lane == __ffsll((unsigned long long)act) - 1) {
          if (v0) atomicAdd(&h[0][bf], (uint32_t)__popcll(v0));
          if (v1) atomicAdd(&h[1][bf], (uint32_t)__popcll(v1));
        }
      } else {
        if (in0) atomicAdd(&h[0][b], 1u);
        if (in1) atomicAdd(&h[1][b], 1u);
      }
      if (gather) {
        const bool keep = in0 || ((kk & msk[1]) == pre[1]);
        const uint64_t kb = __ballot(keep);
        if (kb) {
          const int first = __ffsll((unsigned long long)act) - 1;
          uint32_t o = 0;
          if (lane == first) o = atomicAdd(&lc, (uint32_t)__popcll(kb));
          o = (uint32_t)__shfl((int)o, first);
          const int64_t at = (int64_t)o + __popcll(kb & ((1ull << lane) - 1ull));
          if (keep && at < rs) cand[(int64_t)blockIdx.x * rs + at] = v[u];
        }
      }
    }
  }
  __syncthreads();
  if (gather && threadIdx.x == 0) {
    st->rcnt[blockIdx.x] = (uint32_t)min<int64_t>(lc, rs);
    if (lc > rs) atomicOr(&st->over, 1u);
  }
  if (same) h[1][threadIdx.x] = h[0][threadIdx.x];
  if (h[0][threadIdx.x]) atomicAdd(&st->hist[pass][0][threadIdx.x], h[0][threadIdx.x]);
  if (h[1][threadIdx.x]) atomicAdd(&st->hist[pass][1][threadIdx.x], h[1][threadIdx.x]);
}

// col <- min(round(s*col), 100*scaling) with s = scaling / (2 max(median, 0.001));
// scale_out (if not null) <- s
__global__ void __launch_bounds__(256) k_quantize(double* __restrict__ col, int64_t n, const Select1State* st, int even,
                                                  double scaling, double* __restrict__ scale_out) {
  // the last digit, from the eighth pass's histograms
  uint64_t pre[2], msk[2];
  int64_t k[2];
  for (int t = 0; t < 2; ++t) {
    pre[t] = st->prefix[1][t];
    msk[t] = st->mask[1][t];
    k[t] = st->k[1][t];
  }
  pick_digit(st->hist[7], 0, pre, msk, k);
  const double v0 = dkey_inv(pre[0]), v1 = dkey_inv(pre[1]);
  const double med = even ? (v0 + v1) / 2.0 : v0;
  const double m = med > 0.001 ? med : 0.001;
  const double s = scaling / (2.0 * m);
  const double maxVal = 100 * scaling;
  if (scale_out && blockIdx.x == 0 && threadIdx.x == 0) *scale_out = s;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double q = round(__dmul_rn(s, col[i]));
    if (q > maxVal) q = maxVal;
    col[i] = q;
  }
}

// ---- batched forms: one launch over many columns (grid.y = column), the
// same arithmetic per column as the single-column kernels above.
//
// Batched median by a sampled bracket. Per column: sort kMedianSample
// strided keys (all keys when the column is that short: then the median is
// read off the sorted sample), take the keys kMedianMargin sample positions
// below and above the median's expected position as a bracket [lo, hi], count
// the keys below lo and gather the keys inside; when the counts show the two
// median ranks inside the bracket, radix-select them among the gathered keys,
// else (or when the gathered keys overflow the column's capacity) run the
// full radix select on the column. Either way the result is the exact
// k-th smallest key, as the full select computes it.
struct BracketState {
  uint64_t lo, hi;
  unsigned long long nlo, cnt;
  int32_t done, miss;
};

__device__ __forceinline__ void median_ranks(int64_t n, int64_t& k0, int64_t& k1) {
  k0 = (n % 2 == 0) ? (n - 2) / 2 : n / 2;
  k1 = (n % 2 == 0) ? k0 + 1 : k0;
}

__global__ void __launch_bounds__(1024) k_bsample(const QItem* __restrict__ items, SelectState* __restrict__ states,
                                                 BracketState* __restrict__ brs) {
  __shared__ uint64_t key[kMedianSample];
  const QItem it = items[blockIdx.x];
  SelectState* st = states + blockIdx.x;
  BracketState* b = brs + blockIdx.x;
  const int64_t n = it.n;
  const bool all = n <= kMedianSample;
  for (int j = threadIdx.x; j < kMedianSample; j += blockDim.x) {
    const bool in = all ? j < n : true;
    const int64_t i = all ? j : (int64_t)j * n / kMedianSample;
    key[j] = in ? dkey(it.col[i]) : ~0ull;
  }
  __syncthreads();
  // bitonic sort, ascending
  for (int size = 2; size <= kMedianSample; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = threadIdx.x; t < kMedianSample / 2; t += blockDim.x) {
        const int i = 2 * t - (t & (stride - 1));
        const int j = i + stride;
        const bool up = (i & size) == 0;
        const uint64_t a = key[i], c = key[j];
        if ((a > c) == up) {
          key[i] = c;
          key[j] = a;
        }
      }
      __syncthreads();
    }
  }
  for (int i = threadIdx.x; i < 512; i += blockDim.x) (&st->hist[0][0])[i] = 0;
  if (threadIdx.x == 0) {
    int64_t k0, k1;
    median_ranks(n, k0, k1);
    st->prefix[0] = st->prefix[1] = 0;
    st->mask[0] = st->mask[1] = 0;
    st->k[0] = k0;
    st->k[1] = k1;
    b->nlo = 0;
    b->cnt = 0;
    b->miss = 0;
    b->lo = 0;
    b->hi = ~0ull;
    if (all) {
      b->done = 1;
      st->value[0] = n > 0 ? dkey_inv(key[k0]) : 0.0;
      st->value[1] = n > 0 ? dkey_inv(key[k1]) : 0.0;
    } else {
      b->done = 0;
      const int64_t lo = k0 * kMedianSample / n - kMedianMargin;
      const int64_t hi = k1 * kMedianSample / n + kMedianMargin;
      b->lo = lo <= 0 ? 0ull : key[lo];
      b->hi = hi >= kMedianSample - 1 ? ~0ull : key[hi];
    }
  }
}

// Keys below the bracket counted, keys inside gathered. A block takes tiles
// of 4096 keys (16 per thread, coalesced); a block-wide prefix sum of the
// threads' counts places them, so a tile costs one atomic on the column's
// counter (not one per wave: 55 million contended atomics for the batch).
__global__ void __launch_bounds__(256) k_bbracket(const QItem* __restrict__ items, BracketState* __restrict__ brs,
                                                  double* __restrict__ cand) {
  constexpr int PER = 16;
  constexpr int TILE = 256 * PER;
  __shared__ int wtot[4];
  __shared__ unsigned long long tbase;
  const QItem it = items[blockIdx.y];
  BracketState* b = brs + blockIdx.y;
  if (b->done) return;
  const uint64_t lo = b->lo, hi = b->hi;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t below = 0;
  for (int64_t t0 = (int64_t)blockIdx.x * TILE; t0 < it.n; t0 += (int64_t)gridDim.x * TILE) {
    double x[PER];
    uint32_t inm = 0;
    int cnt = 0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int64_t i = t0 + q * 256 + threadIdx.x;
      const bool valid = i < it.n;
      x[q] = it.col[valid ? i : 0];
      const uint64_t k = dkey(x[q]);
      below += (valid && k < lo) ? 1u : 0u;
      const bool in = valid && k >= lo && k <= hi;
      inm |= (in ? 1u : 0u) << q;
      cnt += in ? 1 : 0;
    }
    // inclusive scan of the counts in the wave, wave totals through LDS
    int inc = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(inc, o);
      if (lane >= o) inc += y;
    }
    if (lane == 63) wtot[w] = inc;
    __syncthreads();
    int wbase = 0, total = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int t = wtot[k];
      wbase += k < w ? t : 0;
      total += t;
    }
    if (threadIdx.x == 0) tbase = total ? atomicAdd(&b->cnt, (unsigned long long)total) : 0ull;
    __syncthreads();
    int64_t pos = (int64_t)tbase + wbase + inc - cnt;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      if ((inm >> q) & 1u) {
        if (pos < it.cand_cap) cand[it.cand_off + pos] = x[q];
        ++pos;
      }
    }
    __syncthreads();  // wtot and tbase are rewritten by the next tile
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) below += __shfl_xor(below, o);
  if (lane == 0 && below) atomicAdd(&b->nlo, (unsigned long long)below);
}

// ranks inside the bracket, or the full select when the bracket missed
__global__ void k_bcand_init(const QItem* __restrict__ items, SelectState* __restrict__ states,
                             BracketState* __restrict__ brs) {
  if (threadIdx.x != 0) return;
  const QItem it = items[blockIdx.x];
  SelectState* st = states + blockIdx.x;
  BracketState* b = brs + blockIdx.x;
  if (b->done) return;
  int64_t k0, k1;
  median_ranks(it.n, k0, k1);
  const int64_t nlo = (int64_t)b->nlo, cnt = (int64_t)b->cnt;
  if (cnt > it.cand_cap || k0 < nlo || k1 >= nlo + cnt) {
    b->miss = 1;
  } else {
    st->k[0] = k0 - nlo;
    st->k[1] = k1 - nlo;
  }
}

// histogram pass over the gathered keys (CAND) or over the whole column of a
// column whose bracket missed (!CAND); other columns return at once
template <bool CAND>
__global__ void k_bselect_hist_br(const QItem* __restrict__ items, SelectState* __restrict__ states,
                                  const BracketState* __restrict__ brs, const double* __restrict__ cand, int shift) {
  const BracketState* b = brs + blockIdx.y;
  if (b->done || (CAND ? b->miss != 0 : b->miss == 0)) return;
  __shared__ uint32_t h[2][256];
  const QItem it = items[blockIdx.y];
  SelectState* st = states + blockIdx.y;
  const double* col = CAND ? cand + it.cand_off : it.col;
  const int64_t n = CAND ? (int64_t)b->cnt : it.n;
  h[0][threadIdx.x] = 0;
  h[1][threadIdx.x] = 0;
  __syncthreads();
  const uint64_t p0 = st->prefix[0], m0 = st->mask[0], p1 = st->prefix[1], m1 = st->mask[1];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t kk = dkey(col[i]);
    const uint32_t bk = (uint32_t)(kk >> shift) & 255u;
    if ((kk & m0) == p0) atomicAdd(&h[0][bk], 1u);
    if ((kk & m1) == p1) atomicAdd(&h[1][bk], 1u);
  }
  __syncthreads();
  if (h[0][threadIdx.x]) atomicAdd(&st->hist[0][threadIdx.x], h[0][threadIdx.x]);
  if (h[1][threadIdx.x]) atomicAdd(&st->hist[1][threadIdx.x], h[1][threadIdx.x]);
}

template <bool CAND>
__global__ void k_bselect_step_br(SelectState* __restrict__ states, const BracketState* __restrict__ brs, int shift) {
  const BracketState* b = brs + blockIdx.x;
  if (b->done || (CAND ? b->miss != 0 : b->miss == 0)) return;
  SelectState* st = states + blockIdx.x;
  if (threadIdx.x < 2) {
    const int t = threadIdx.x;
    int64_t cum = 0;
    int bk = 0;
    for (; bk < 256; ++bk) {
      const int64_t c = st->hist[t][bk];
      if (st->k[t] < cum + c) break;
      cum += c;
    }
    if (bk > 255) bk = 255;
    st->k[t] -= cum;
    st->prefix[t] |= ((uint64_t)bk) << shift;
    st->mask[t] |= ((uint64_t)255) << shift;
    if (shift == 0) st->value[t] = dkey_inv(st->prefix[t]);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 512; i += blockDim.x) (&st->hist[0][0])[i] = 0;
}

__global__ void k_bquantize(const QItem* __restrict__ items, const SelectState* __restrict__ states, double scaling) {
  const QItem it = items[blockIdx.y];
  const SelectState* st = states + blockIdx.y;
  const double med = (it.n % 2 == 0) ? (st->value[0] + st->value[1]) / 2.0 : st->value[0];
  const double m = med > 0.001 ? med : 0.001;
  const double s = scaling / (2.0 * m);
  const double maxVal = 100 * scaling;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < it.n; i += (int64_t)gridDim.x * blockDim.x) {
    double q = round(__dmul_rn(s, it.col[i]));
    if (q > maxVal) q = maxVal;
    it.col[i] = q;
  }
}

__global__ void k_bfill_costs(const FItem* __restrict__ items) {
  const FItem it = items[blockIdx.y];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < it.nslots;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = it.eid[i];  // -1: a padding slot (see k_fill_costs)
    if (e < 0) {
      it.cost[i] = 0.f;
      if (it.time) it.time[i] = 0.f;
      continue;
    }
    it.cost[i] = (float)it.energy[e];
    if (it.time) it.time[i] = (float)it.energy[it.E + e];
  }
}

// ------------------------------------------------- sliced-ELL in-edge layout
// class 0: source on the target's layer (intralayer), class 1: source on the
// previous layer (interlayer). Targets on layer 0 whose source id is > npl
// move to the closing copy numNodes + target (the reference's dummy nodes).
// Node ids are int32 and numNodes = max(src) + 1 <= 2^31, so every id here,
// the closing copies included (< numNodes + npl <= 2^32), fits 32 unsigned
// bits and the layer of an id is one FastDiv by npl. Returns false for an
// edge that is neither (or names a negative id): it is dropped.
__device__ __forceinline__ bool classify(int32_t s_, int32_t t_, uint32_t npl, uint32_t numNodes, const FastDiv& fd,
                                         uint32_t& node, uint32_t& layer, int& cls) {
  if (s_ < 0 || t_ < 0) return false;
  const uint32_t s = (uint32_t)s_;
  uint32_t t = (uint32_t)t_;
  if (t < npl && s > npl) t += numNodes;
  node = t;
  layer = fd.div(t);
  const uint32_t ls = fd.div(s);
  cls = (ls == layer) ? 0 : ((ls + 1 == layer) ? 1 : -1);
  return cls >= 0;
}

__global__ void k_count(const int32_t* __restrict__ src, const int32_t* __restrict__ trgt, int64_t E, uint32_t npl,
                        uint32_t numNodes, FastDiv fd, int32_t* __restrict__ cnt, int32_t* __restrict__ dropped) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  uint32_t node, layer;
  int cls;
  const int64_t nTot = (int64_t)numNodes + npl;
  if (!classify(src[e], trgt[e], npl, numNodes, fd, node, layer, cls) || node >= nTot) {
    atomicAdd(dropped, 1);
    return;
  }
  atomicAdd(&cnt[cls * nTot + node], 1);
}

// per class: slots per lane (the maximum in-degree over all nodes) in
// out[0..1] and the number of edges (the sum of in-degrees) in out[2..3]
__global__ void k_class_width(const int32_t* __restrict__ cnt, int64_t nTot, int32_t* __restrict__ out) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int w0 = 0, w1 = 0, n0 = 0, n1 = 0;
  for (int64_t i = g; i < nTot; i += (int64_t)gridDim.x * blockDim.x) {
    const int a = cnt[i], b = cnt[nTot + i];
    w0 = max(w0, a);
    w1 = max(w1, b);
    n0 += a;
    n1 += b;
  }
  for (int o = 32; o > 0; o >>= 1) {
    w0 = max(w0, __shfl_xor(w0, o));
    w1 = max(w1, __shfl_xor(w1, o));
    n0 += __shfl_xor(n0, o);
    n1 += __shfl_xor(n1, o);
  }
  if ((threadIdx.x & 63) == 0) {
    if (w0) atomicMax(&out[0], w0);
    if (w1) atomicMax(&out[1], w1);
    if (n0) atomicAdd(&out[2], n0);
    if (n1) atomicAdd(&out[3], n1);
  }
}

__global__ void k_fill_slots(const int32_t* __restrict__ src, const int32_t* __restrict__ trgt, int64_t E,
                             uint32_t npl, uint32_t numNodes, FastDiv fd, int nChunks, int32_t* __restrict__ fill,
                             const int32_t* __restrict__ width, int32_t* __restrict__ s0, int32_t* __restrict__ s1,
                             int32_t* __restrict__ e0, int32_t* __restrict__ e1) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  uint32_t node, layer;
  int cls;
  const int32_t s = src[e];
  const int64_t nTot = (int64_t)numNodes + npl;
  if (!classify(s, trgt[e], npl, numNodes, fd, node, layer, cls) || node >= nTot) return;
  const int pos = atomicAdd(&fill[cls * nTot + node], 1);
  const uint32_t v = node - layer * npl;
  const int c = (int)(v >> 6), lane = (int)(v & 63);
  const int64_t w = width[cls];
  const int64_t slot = ((int64_t)layer * nChunks + c) * w * 64 + (int64_t)pos * 64 + lane;
  const int32_t sl = (int32_t)((uint32_t)s - (cls == 0 ? layer : layer - 1) * npl);
  (cls == 0 ? s0 : s1)[slot] = sl;
  (cls == 0 ? e0 : e1)[slot] = (int32_t)e;
}

// order the slots of every node by (source, edge id); nTot = nLayers * npl
// < 2^32 (the count table's node ids)
__global__ void k_sort_slots(const int32_t* __restrict__ cnt, uint32_t npl, int64_t nTot, FastDiv fd, int nChunks,
                             const int32_t* __restrict__ width, int32_t* __restrict__ s0, int32_t* __restrict__ s1,
                             int32_t* __restrict__ e0, int32_t* __restrict__ e1) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= 2 * nTot) return;
  const int n = cnt[g];
  if (n <= 1) return;
  const int cls = g >= nTot ? 1 : 0;
  const uint32_t node = (uint32_t)(g - cls * nTot);
  const uint32_t layer = fd.div(node);
  const uint32_t v = node - layer * npl;
  const int c = (int)(v >> 6), lane = (int)(v & 63);
  int32_t* S = cls == 0 ? s0 : s1;
  int32_t* Ee = cls == 0 ? e0 : e1;
  const int64_t base = ((int64_t)layer * nChunks + c) * (int64_t)width[cls] * 64 + lane;
  for (int i = 1; i < n; ++i) {
    const int32_t ks = S[base + (int64_t)i * 64], ke = Ee[base + (int64_t)i * 64];
    int j = i - 1;
    while (j >= 0) {
      const int32_t js = S[base + (int64_t)j * 64], je = Ee[base + (int64_t)j * 64];
      if (js < ks || (js == ks && je < ke)) break;
      S[base + (int64_t)(j + 1) * 64] = js;
      Ee[base + (int64_t)(j + 1) * 64] = je;
      --j;
    }
    S[base + (int64_t)(j + 1) * 64] = ks;
    Ee[base + (int64_t)(j + 1) * 64] = ke;
  }
}

// groups of eight slots per chunk (per class): the largest in-degree of the
// chunk's nodes over all layers, rounded up to eight, divided by eight
__global__ void k_chunk_groups(const int32_t* __restrict__ cnt, uint32_t npl, int64_t nTot, FastDiv fd,
                               int nChunks, int32_t* __restrict__ grp) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= 2 * nTot) return;
  const int n = cnt[g];
  if (n <= 0) return;
  const int cls = g >= nTot ? 1 : 0;
  const uint32_t node = (uint32_t)(g - cls * nTot);
  const uint32_t layer = fd.div(node);
  const uint32_t v = node - layer * npl;
  atomicMax(&grp[cls * nChunks + (int)(v >> 6)], (n + 7) >> 3);
}

// padding slots (src < 0) point at the sentinel vertex npl of the solver's
// distance rows (always +inf), so the solver needs no per-slot validity test
__global__ void k_pad_src(int32_t* __restrict__ s, int64_t n, int32_t npl) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && s[i] < 0) s[i] = npl;
}

// Clears *flag when a source slot of layer l > first differs from the same
// slot of layer `first` (then the class's sources are not layer-invariant).
__global__ void k_src_invariant(const int32_t* __restrict__ s, int64_t lstride, int64_t first, int64_t nLayers,
                                int32_t* __restrict__ flag) {
  const int64_t n = (nLayers - first - 1) * lstride;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t o = i % lstride;
    if (s[(first + 1) * lstride + i] != s[first * lstride + o]) *flag = 0;
  }
}

// ---- in-edge layout of a generated non-conjugate product graph, in closed
// form. Every layer repeats the X graph: node (l, x1) has the intralayer
// in-edges (l, x0) -> (l, x1) of the X edges j = (x0, x1), edge id l*B + 2j,
// on layers 0..L-1 (none on the closing layer), and the interlayer in-edges
// (l-1, x0) -> (l, x1), edge id (l-1)*B + 2j+1 (j when pruned), plus the stay
// edge (l-1, x1) -> (l, x1), edge id (l-1)*B + per_x*|EX| + x1, on layers
// 1..L (layer L holds the closing copies). The slots of one layer (the
// template) are built from the X graph, ordered by (source, edge id) as the
// generic build orders them, and copied to every layer with the layer's edge
// id base.
//
// in-degrees per class of the X graph: threads [0, nEX) take the X edges,
// [nEX, nEX + nVX) the stay edges.
__global__ void k_xcount(const int32_t* __restrict__ EX, int64_t nEX, int64_t nVX, int prune,
                         int32_t* __restrict__ cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nEX) {
    const int32_t x1 = EX[2 * i + 1];
    if (!prune) atomicAdd(&cnt[x1], 1);
    atomicAdd(&cnt[nVX + x1], 1);
  } else if (i < nEX + nVX) {
    atomicAdd(&cnt[nVX + (i - nEX)], 1);
  }
}

// template slots (source vertex, edge id inside its layer) of the one layer;
// the order inside a node is fixed afterwards by k_sort_slots
__global__ void k_xfill(const int32_t* __restrict__ EX, int64_t nEX, int64_t nVX, int prune, int nChunks,
                        const int32_t* __restrict__ width, int32_t* __restrict__ fill, int32_t* __restrict__ s0,
                        int32_t* __restrict__ s1, int32_t* __restrict__ r0, int32_t* __restrict__ r1) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per_x = prune ? 1 : 2;
  int32_t x0, x1;
  int64_t rI, rX;
  if (i < nEX) {
    x0 = EX[2 * i];
    x1 = EX[2 * i + 1];
    rI = 2 * i;
    rX = per_x * i + (prune ? 0 : 1);
  } else if (i < nEX + nVX) {
    x0 = x1 = (int32_t)(i - nEX);
    rI = -1;
    rX = per_x * nEX + x0;
  } else {
    return;
  }
  const int c = x1 >> 6, lane = x1 & 63;
  if (rI >= 0 && !prune) {
    const int pos = atomicAdd(&fill[x1], 1);
    const int64_t slot = ((int64_t)c * width[0] + pos) * 64 + lane;
    s0[slot] = x0;
    r0[slot] = (int32_t)rI;
  }
  const int pos = atomicAdd(&fill[nVX + x1], 1);
  const int64_t slot = ((int64_t)c * width[1] + pos) * 64 + lane;
  s1[slot] = x0;
  r1[slot] = (int32_t)rX;
}

// every layer of one class from the template: src (padding -> the sentinel
// npl) and edge id base + r (padding -> -1). Intralayer class (cls 0): layers
// 0..L-1, base l*B; interlayer class: layers 1..L, base (l-1)*B. Layers
// without in-edges of the class, and the kSlotPad tail, are all padding.
__global__ void k_xreplicate(const int32_t* __restrict__ ts, const int32_t* __restrict__ tr, int64_t lstride,
                             int64_t nLayers, int64_t total, int cls, int64_t B, int32_t npl,
                             int32_t* __restrict__ s, int32_t* __restrict__ e) {
  const int64_t L = nLayers - 1;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t l = i / lstride;
    const int64_t o = i - l * lstride;
    const bool active = l < nLayers && (cls == 0 ? l < L : l >= 1);
    const int32_t r = active ? tr[o] : -1;
    s[i] = r >= 0 ? ts[o] : npl;
    e[i] = r >= 0 ? (int32_t)((cls == 0 ? l : l - 1) * B + r) : -1;
  }
}

// per-slot cost (float q-units or raw double) and time gathered by edge id
template <typename CT>
__global__ void k_fill_costs(const int32_t* __restrict__ eid, const int32_t* __restrict__ s, int64_t nSlots,
                             const double* __restrict__ energy, int64_t E, int ncols, CT* __restrict__ cost,
                             float* __restrict__ time, int wmask, int deg) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nSlots) return;
  // slots at or past the class's largest in-degree are padding on every node;
  // the caller zeroed them when it allocated the buffers
  if ((int)((i >> 6) & wmask) >= deg) return;
  // padding slots are the ones without an edge (their source is the
  // sentinel npl, never negative once k_pad_src has run)
  const int64_t e = eid[i];
  if (e < 0) {
    cost[i] = (CT)0;
    if (time) time[i] = 0.f;
    return;
  }
  cost[i] = (CT)energy[e];
  if (time) time[i] = (float)energy[E + e];
}

// rows of the given edges (output matching)
__global__ void k_gather_rows(const int32_t* __restrict__ rows, int cols, const int32_t* __restrict__ eids,
                              int64_t n, int32_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * cols) return;
  const int64_t r = i / cols, c = i - r * cols;
  out[i] = rows[(int64_t)eids[r] * cols + c];
}

// out[j] = base[j * stride], j < n (one column of a column-strided table)
__global__ void k_gather_strided(const int32_t* __restrict__ base, int64_t stride, int64_t n,
                                 int32_t* __restrict__ out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) out[j] = base[j * stride];
}

// layer-0 intralayer in-edges (branch graph of the min-cost branching)
__global__ void k_int_max(const int32_t* __restrict__ x, int64_t n, int32_t* __restrict__ out) {
  int m = -2147483647 - 1;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    m = max(m, x[i]);
  for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

// ------------------------------------------------------------ launchers
static inline unsigned nb(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

hipError_t launch_product_space(const int32_t* EX, int64_t nEX, const int32_t* EY, int64_t nEY, int64_t nVX,
                                int prune, int32_t* rows, int32_t* src, int32_t* trgt, int64_t E, hipStream_t st) {
  hipLaunchKernelGGL(k_product_space, dim3(nb(E, 256)), dim3(256), 0, st, EX, nEX, EY, nEY, nVX, prune, rows, src,
                     trgt, E);
  return hipGetLastError();
}

hipError_t launch_energy(const double* VX, const double* VY, const int32_t* rows, int64_t E, const double* FD,
                         int64_t nVY, int lineIntegral, int lengthMode, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_energy, dim3(nb(E, 256)), dim3(256), 0, st, VX, VY, rows, E, FD, nVY, lineIntegral,
                     lengthMode, out);
  return hipGetLastError();
}

// out[y * nVX + x] = FD[x * nVY + y]: feat_diff by columns (one Y vertex's
// values contiguous), through 32 x 32 tiles in LDS
__global__ void k_transpose_fd(const double* __restrict__ FD, int64_t nVX, int64_t nVY, double* __restrict__ out) {
  __shared__ double tile[32][33];
  const int64_t x0 = (int64_t)blockIdx.x * 32, y0 = (int64_t)blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 256 threads: 8 rows per pass
  for (int r = ty; r < 32; r += 8) {
    const int64_t x = x0 + r, y = y0 + tx;
    if (x < nVX && y < nVY) tile[r][tx] = FD[x * nVY + y];
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int64_t y = y0 + r, x = x0 + tx;
    if (x < nVX && y < nVY) out[y * nVX + x] = tile[tx][r];
  }
}

hipError_t launch_transpose_fd(const double* FD, int64_t nVX, int64_t nVY, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_transpose_fd, dim3(nb(nVX, 32), nb(nVY, 32)), dim3(256), 0, st, FD, nVX, nVY, out);
  return hipGetLastError();
}

// cand: a buffer of ccap keys for the keys the later radix passes read, or
// null (every pass reads all n keys)
hipError_t launch_quantize(double* col, int64_t n, void* selectState, double* scale_out, double* cand, int64_t ccap,
                           hipStream_t st) {
  Select1State* S = (Select1State*)selectState;
  const int64_t k1 = (n % 2 == 0) ? (n - 2) / 2 : n / 2;
  const int64_t k2 = (n % 2 == 0) ? k1 + 1 : k1;
  hipLaunchKernelGGL(k_select_init, dim3(1), dim3(256), 0, st, S, k1, k2);
  // at most 512 workgroups: each adds its partial histograms with global
  // atomics on the same 512 counters
  const unsigned gh = (unsigned)std::min<int64_t>(512, std::max<int64_t>(1, (n + 256 * kSelUnroll - 1) / (256 * kSelUnroll)));
  for (int pass = 0; pass < 8; ++pass)
    hipLaunchKernelGGL(k_select_hist, dim3(gh), dim3(256), 0, st, col, n, S, pass, cand, cand ? ccap : 0);
  const unsigned g = (unsigned)std::min<int64_t>(2048, std::max<int64_t>(1, (n + 255) / 256));
  hipLaunchKernelGGL(k_quantize, dim3(g), dim3(256), 0, st, col, n, S, (int)(n % 2 == 0), 100000.0, scale_out);
  return hipGetLastError();
}

size_t select_state_bytes() { return sizeof(Select1State); }

size_t bselect_state_bytes() { return sizeof(SelectState); }

size_t bracket_state_bytes() { return sizeof(BracketState); }

// Quantization of n columns (the median of each, then q-units): one full pass
// over the keys for the bracket, eight radix passes over the gathered keys,
// eight full passes only for columns whose bracket missed (the others return
// at once), then the quantization pass.
hipError_t launch_quantize_batch(const QItem* d_items, int n, void* states, void* brackets, double* cand,
                                 hipStream_t st) {
  if (n <= 0) return hipSuccess;
  SelectState* S = (SelectState*)states;
  BracketState* B = (BracketState*)brackets;
  hipLaunchKernelGGL(k_bsample, dim3(n), dim3(1024), 0, st, d_items, S, B);
  hipLaunchKernelGGL(k_bbracket, dim3(32, n), dim3(256), 0, st, d_items, B, cand);
  hipLaunchKernelGGL(k_bcand_init, dim3(n), dim3(64), 0, st, d_items, S, B);
  for (int shift = 56; shift >= 0; shift -= 8) {
    hipLaunchKernelGGL(k_bselect_hist_br<true>, dim3(8, n), dim3(256), 0, st, d_items, S, B, cand, shift);
    hipLaunchKernelGGL(k_bselect_step_br<true>, dim3(n), dim3(256), 0, st, S, B, shift);
  }
  for (int shift = 56; shift >= 0; shift -= 8) {
    hipLaunchKernelGGL(k_bselect_hist_br<false>, dim3(32, n), dim3(256), 0, st, d_items, S, B, cand, shift);
    hipLaunchKernelGGL(k_bselect_step_br<false>, dim3(n), dim3(256), 0, st, S, B, shift);
  }
  hipLaunchKernelGGL(k_bquantize, dim3(32, n), dim3(256), 0, st, d_items, S, 100000.0);
  return hipGetLastError();
}

hipError_t launch_energy_batch(const EItem* d_items, int n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_benergy, dim3(64, n), dim3(256), 0, st, d_items);
  return hipGetLastError();
}

// grid.y = item: fills of many device arrays (the predecessor tables of a
// batch's jobs) in one launch instead of one memset call each
__global__ void k_bfill_i32(const FillItem* __restrict__ items, int32_t value) {
  const FillItem it = items[blockIdx.y];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < it.n; i += (int64_t)gridDim.x * blockDim.x)
    it.p[i] = value;
}

hipError_t launch_fill_i32_batch(const FillItem* d_items, int n, int32_t value, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_bfill_i32, dim3(64, n), dim3(256), 0, st, d_items, value);
  return hipGetLastError();
}

hipError_t launch_fill_costs_batch(const FItem* d_items, int n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_bfill_costs, dim3(64, n), dim3(256), 0, st, d_items);
  return hipGetLastError();
}

hipError_t launch_layout_count(const int32_t* src, const int32_t* trgt, int64_t E, int64_t npl, int64_t numNodes,
                               int32_t* cnt, int32_t* dropped, hipStream_t st) {
  if (numNodes + npl > ((int64_t)1 << 32) || numNodes > ((int64_t)1 << 31)) return hipErrorInvalidValue;
  FastDiv fd;
  fd.init((uint32_t)npl);
  hipLaunchKernelGGL(k_count, dim3(nb(E, 256)), dim3(256), 0, st, src, trgt, E, (uint32_t)npl, (uint32_t)numNodes, fd,
                     cnt, dropped);
  return hipGetLastError();
}

__global__ void k_fill_i32(int32_t* __restrict__ p, int64_t n, int32_t value) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = value;
}

// p[0, n) <- value (a kernel, not a memset node: callers may capture it)
hipError_t launch_fill_i32(int32_t* p, int64_t n, int32_t value, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)std::min<int64_t>(nb(n, 256), 4096)), dim3(256), 0, st, p, n, value);
  return hipGetLastError();
}

hipError_t launch_layout_width(const int32_t* cnt, int64_t nTot, int32_t* width, hipStream_t st) {
  hipError_t e = launch_fill_i32(width, 4, 0, st);
  if (e != hipSuccess) return e;
  // few blocks: every wave ends with four atomics on the same four words
  const unsigned g = (unsigned)std::min<int64_t>(32, std::max<int64_t>(1, (nTot + 255) / 256));
  hipLaunchKernelGGL(k_class_width, dim3(g), dim3(256), 0, st, cnt, nTot, width);
  return hipGetLastError();
}

hipError_t launch_layout_fill(const int32_t* src, const int32_t* trgt, int64_t E, int64_t npl, int64_t numNodes,
                              int nChunks, int32_t* fill, const int32_t* cnt, const int32_t* width, int32_t* s0,
                              int32_t* s1, int32_t* e0, int32_t* e1, hipStream_t st) {
  if (numNodes + npl > ((int64_t)1 << 32) || numNodes > ((int64_t)1 << 31)) return hipErrorInvalidValue;
  FastDiv fd;
  fd.init((uint32_t)npl);
  hipLaunchKernelGGL(k_fill_slots, dim3(nb(E, 256)), dim3(256), 0, st, src, trgt, E, (uint32_t)npl,
                     (uint32_t)numNodes, fd, nChunks, fill, width, s0, s1, e0, e1);
  const int64_t nTot = (numNodes / npl + 1) * npl;
  hipLaunchKernelGGL(k_sort_slots, dim3(nb(2 * nTot, 256)), dim3(256), 0, st, cnt, (uint32_t)npl, nTot, fd, nChunks,
                     width, s0, s1, e0, e1);
  return hipGetLastError();
}

hipError_t launch_xlayout_count(const int32_t* EX, int64_t nEX, int64_t nVX, int prune, int32_t* cnt,
                                hipStream_t st) {
  hipLaunchKernelGGL(k_xcount, dim3(nb(nEX + nVX, 256)), dim3(256), 0, st, EX, nEX, nVX, prune, cnt);
  return hipGetLastError();
}

// template slots of one layer, ordered by (source, edge id) inside each node
hipError_t launch_xlayout_fill(const int32_t* EX, int64_t nEX, int64_t nVX, int prune, int nChunks,
                               const int32_t* width, const int32_t* cnt, int32_t* fill, int32_t* s0, int32_t* s1,
                               int32_t* r0, int32_t* r1, hipStream_t st) {
  hipLaunchKernelGGL(k_xfill, dim3(nb(nEX + nVX, 256)), dim3(256), 0, st, EX, nEX, nVX, prune, nChunks, width, fill,
                     s0, s1, r0, r1);
  FastDiv fd;
  fd.init((uint32_t)nVX);
  hipLaunchKernelGGL(k_sort_slots, dim3(nb(2 * nVX, 256)), dim3(256), 0, st, cnt, (uint32_t)nVX, nVX, fd, nChunks,
                     width, s0, s1, r0, r1);
  return hipGetLastError();
}

hipError_t launch_xlayout_replicate(const int32_t* ts, const int32_t* tr, int64_t lstride, int64_t nLayers,
                                    int64_t total, int cls, int64_t B, int32_t npl, int32_t* s, int32_t* e,
                                    hipStream_t st) {
  if (total <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_xreplicate, dim3((unsigned)std::min<int64_t>(nb(total, 256), 8192)), dim3(256), 0, st, ts, tr,
                     lstride, nLayers, total, cls, B, npl, s, e);
  return hipGetLastError();
}

hipError_t launch_chunk_groups(const int32_t* cnt, int64_t npl, int64_t nTot, int nChunks, int32_t* grp,
                               hipStream_t st) {
  FastDiv fd;
  fd.init((uint32_t)npl);
  hipLaunchKernelGGL(k_chunk_groups, dim3(nb(2 * nTot, 256)), dim3(256), 0, st, cnt, (uint32_t)npl, nTot, fd, nChunks,
                     grp);
  return hipGetLastError();
}

hipError_t launch_src_invariant(const int32_t* s, int64_t lstride, int64_t first, int64_t nLayers, int32_t* flag,
                                hipStream_t st) {
  const int64_t n = (nLayers - first - 1) * lstride;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_src_invariant, dim3((unsigned)std::min<int64_t>(nb(n, 256), 4096)), dim3(256), 0, st, s,
                     lstride, first, nLayers, flag);
  return hipGetLastError();
}

hipError_t launch_pad_src(int32_t* s, int64_t n, int32_t npl, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_pad_src, dim3(nb(n, 256)), dim3(256), 0, st, s, n, npl);
  return hipGetLastError();
}

// W: the layout's slot width when it is a power of two (chunk rows of W x 64
// slots), with deg the class's largest in-degree: the fill skips the slots
// j >= deg of every node, which the caller zeroed; W = 0 fills every slot.
hipError_t launch_fill_costs(const int32_t* eid, const int32_t* s, int64_t nSlots, const double* energy, int64_t E,
                             int ncols, bool asDouble, void* cost, float* time, int W, int deg, hipStream_t st) {
  if (nSlots == 0) return hipSuccess;
  const bool skip = W > 0 && (W & (W - 1)) == 0 && deg >= 0 && deg < W;
  const int wmask = skip ? W - 1 : 0, d = skip ? deg : 1;
  if (asDouble)
    hipLaunchKernelGGL(k_fill_costs<double>, dim3(nb(nSlots, 256)), dim3(256), 0, st, eid, s, nSlots, energy, E,
                       ncols, (double*)cost, time, wmask, d);
  else
    hipLaunchKernelGGL(k_fill_costs<float>, dim3(nb(nSlots, 256)), dim3(256), 0, st, eid, s, nSlots, energy, E,
                       ncols, (float*)cost, time, wmask, d);
  return hipGetLastError();
}

hipError_t launch_gather_rows(const int32_t* rows, int cols, const int32_t* eids, int64_t n, int32_t* out,
                              hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_rows, dim3(nb(n * cols, 256)), dim3(256), 0, st, rows, cols, eids, n, out);
  return hipGetLastError();
}

hipError_t launch_gather_strided(const int32_t* base, int64_t stride, int64_t n, int32_t* out, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_strided, dim3(nb(n, 256)), dim3(256), 0, st, base, stride, n, out);
  return hipGetLastError();
}

hipError_t launch_int_max(const int32_t* x, int64_t n, int32_t* out, hipStream_t st) {
  const unsigned g = (unsigned)std::min<int64_t>(1024, std::max<int64_t>(1, (n + 255) / 256));
  hipLaunchKernelGGL(k_int_max, dim3(g), dim3(256), 0, st, x, n, out);
  return hipGetLastError();
}

hipError_t launch_cycle_value(const CycleItem* items, int n, double* values, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_cycle_value, dim3(n), dim3(64), 0, st, items, n, values);
  return hipGetLastError();
}

hipError_t launch_cycle_grad(const CycleItem* items, int n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_cycle_grad, dim3(n), dim3(256), 0, st, items, n);
  return hipGetLastError();
}

}  // namespace pgc
