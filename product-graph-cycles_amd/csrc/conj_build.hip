// Conjugate product graph construction on the device (gfx950): product
// space, node rotations and the higher-order energy.
//
//   product space   pgg/product_spaces/conj_product_space.cpp:70-168
//   rotations       pgg/energy/higherOrderDeformationEnergy.cpp:16-53, 95-134
//                   (USE_MATLAB_CONVENTION: one rotation per product node)
//   energy          pgg/energy/higherOrderDeformationEnergy.cpp:209-339
//
// Nodes are product edges, 3|EX| per layer: [0,|EX|) non-degenerate,
// [|EX|,2|EX|) degenerate in 2D, then degenerate in 3D ([|EX|,2|EX|) when the
// intralayer edges are pruned). Edge ids: per layer, the intralayer block (two
// edges per (j, it) pair) then the interlayer block (three, or two when
// pruned, per pair, then two per j); pairs (j, it) run over the out-edges it
// of EX(j,1) in ascending edge id, skipping the edge back to EX(j,0) (the
// reference iterates a hash set; (source, target) node pairs are unique, so
// no solve depends on this order).
//
// A node's rotation is the rotation R in SO(3) maximising tr(R^T X^T Y) for
// the row-stacked (normal, edge, edge x normal) frames X (2D) and Y (3D) -- the
// reference's Kabsch solution of Eigen's JacobiSVD -- computed as the dominant
// eigenvector of Horn's 4x4 quaternion matrix by cyclic Jacobi rotations, in
// fp64 with every operation rounded separately.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "pgc_internal.h"

namespace pgc {

namespace {

// The product space in two passes. Every layer has the same edge pattern
// (layer i's edges are [i perLayer, (i + 1) perLayer)), so the first pass
// lays out one layer's edges as descriptors (X edge j, out-edge it, kind), one
// thread per X edge j in the order above; the second writes every edge of
// every layer from its descriptor and its layer's Y edges, one thread per
// edge, so the rows, sources and targets are written contiguously.
// Kinds: 0, 1 intralayer (per (j, it) pair); 2, 3, 4 interlayer (per pair;
// 3 only unpruned); 5, 6 the two interlayer edges closing j's block.
__global__ void k_conj_desc(const int32_t* __restrict__ EX, int64_t nEX, int prune,
                            const int32_t* __restrict__ outOff, const int32_t* __restrict__ outList,
                            const int64_t* __restrict__ pre, int64_t intraTotal, int2* __restrict__ desc) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nEX) return;
  const int s3 = EX[2 * j], m3 = EX[2 * j + 1];
  int64_t a = 2 * pre[j];                                      // intralayer slots of j
  int64_t b = intraTotal + (prune ? 2 : 3) * pre[j] + 2 * j;  // interlayer slots of j
  const int jj = (int)j;
  for (int k = outOff[m3]; k < outOff[m3 + 1]; ++k) {
    const int it = outList[k];
    if (EX[2 * it + 1] == s3) continue;
    if (!prune) {
      desc[a++] = make_int2(jj | (0 << 28), it);
      desc[a++] = make_int2(jj | (1 << 28), it);
    }
    desc[b++] = make_int2(jj | (2 << 28), it);
    if (!prune) desc[b++] = make_int2(jj | (3 << 28), it);
    desc[b++] = make_int2(jj | (4 << 28), it);
  }
  desc[b++] = make_int2(jj | (5 << 28), -1);
  desc[b++] = make_int2(jj | (6 << 28), -1);
}

__global__ void k_conj_product_space(const int32_t* __restrict__ EX, int64_t nEX, const int32_t* __restrict__ EY,
                                     int64_t nEY, int prune, const int2* __restrict__ desc, int64_t perLayer,
                                     int32_t* __restrict__ rows, int32_t* __restrict__ src,
                                     int32_t* __restrict__ trgt) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nEY * perLayer) return;
  // edge ids are below 2^31: 32-bit division
  const int64_t i = (uint32_t)e / (uint32_t)perLayer, q = e - i * perLayer;
  const int64_t nx = i + 1 == nEY ? 0 : i + 1;
  const int64_t npl = 3 * nEX;
  const int64_t cur = i * npl, nxt = nx * npl;
  const int64_t off2d = nEX, off3d = prune ? nEX : 2 * nEX;
  const int s2 = EY[2 * i], m2 = EY[2 * i + 1], t2 = EY[2 * nx + 1];
  const int2 d = desc[q];
  const int64_t j = d.x & ((1 << 28) - 1);
  const int kind = d.x >> 28;
  const int it = d.y;
  const int s3 = EX[2 * j], m3 = EX[2 * j + 1];
  const int t3 = it >= 0 ? EX[2 * it + 1] : m3;
  int r[6];
  int64_t sv, tv;
  switch (kind) {
    case 0: r[0] = s2; r[1] = s2; r[2] = s2; r[3] = s3; r[4] = m3; r[5] = t3; sv = cur + off2d + j; tv = cur + off2d + it; break;
    case 1: r[0] = s2; r[1] = s2; r[2] = m2; r[3] = s3; r[4] = m3; r[5] = t3; sv = cur + off2d + j; tv = cur + it; break;
    case 2: r[0] = s2; r[1] = m2; r[2] = t2; r[3] = s3; r[4] = m3; r[5] = t3; sv = cur + j; tv = nxt + it; break;
    case 3: r[0] = s2; r[1] = m2; r[2] = m2; r[3] = s3; r[4] = m3; r[5] = t3; sv = cur + j; tv = nxt + off2d + it; break;
    case 4: r[0] = s2; r[1] = m2; r[2] = t2; r[3] = m3; r[4] = m3; r[5] = t3; sv = cur + off3d + j; tv = nxt + it; break;
    case 5: r[0] = s2; r[1] = m2; r[2] = t2; r[3] = s3; r[4] = m3; r[5] = m3; sv = cur + j; tv = nxt + off3d + j; break;
    default: r[0] = s2; r[1] = m2; r[2] = t2; r[3] = m3; r[4] = m3; r[5] = m3; sv = cur + off3d + j; tv = nxt + off3d + j; break;
  }
  int2* rw = (int2*)(rows + 6 * e);  // 8-byte aligned: 24 e bytes
  rw[0] = make_int2(r[0], r[1]);
  rw[1] = make_int2(r[2], r[3]);
  rw[2] = make_int2(r[4], r[5]);
  src[e] = (int32_t)sv;
  trgt[e] = (int32_t)tv;
}

struct V3 {
  double x, y, z;
};
__device__ __forceinline__ V3 vsub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 vcross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ V3 vnormalized(V3 a) {
  const double n2 = a.x * a.x + a.y * a.y + a.z * a.z;
  if (!(n2 > 0)) return a;
  const double n = sqrt(n2);
  return {a.x / n, a.y / n, a.z / n};
}
__device__ __forceinline__ V3 row3(const double* M, int64_t r) { return {M[3 * r], M[3 * r + 1], M[3 * r + 2]}; }

// Cyclic Jacobi rotations of a symmetric 4x4 matrix; v receives the eigenvectors.
__device__ void jacobi4(double a[4][4], double v[4][4]) {
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) v[i][j] = i == j ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 50; ++sweep) {
    double off = 0;
    for (int p = 0; p < 4; ++p)
      for (int q = p + 1; q < 4; ++q) off += a[p][q] * a[p][q];
    if (off < 1e-300) break;
    for (int p = 0; p < 4; ++p) {
      for (int q = p + 1; q < 4; ++q) {
        if (a[p][q] == 0.0) continue;
        const double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0);
        const double s = t * c;
        for (int k = 0; k < 4; ++k) {
          const double akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - s * akq;
          a[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 4; ++k) {
          const double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - s * aqk;
          a[q][k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 4; ++k) {
          const double vkp = v[k][p], vkq = v[k][q];
          v[k][p] = c * vkp - s * vkq;
          v[k][q] = s * vkp + c * vkq;
        }
      }
    }
  }
}

__device__ void rotation_quat(V3 v2, V3 n2, V3 c2, V3 v3, V3 n3, V3 c3, double q[4]) {
  const V3 X[3] = {vnormalized(n2), vnormalized(v2), vnormalized(c2)};
  const V3 Y[3] = {vnormalized(n3), vnormalized(v3), vnormalized(c3)};
  const double Xm[3][3] = {{X[0].x, X[0].y, X[0].z}, {X[1].x, X[1].y, X[1].z}, {X[2].x, X[2].y, X[2].z}};
  const double Ym[3][3] = {{Y[0].x, Y[0].y, Y[0].z}, {Y[1].x, Y[1].y, Y[1].z}, {Y[2].x, Y[2].y, Y[2].z}};
  double A[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += Xm[k][i] * Ym[k][j];
      A[i][j] = s;
    }
  // Horn's matrix for max tr(R M), M = A^T (S_ab = M[a][b] = A[b][a])
  const double Sxx = A[0][0], Syy = A[1][1], Szz = A[2][2];
  const double Sxy = A[1][0], Syx = A[0][1], Sxz = A[2][0], Szx = A[0][2], Syz = A[2][1], Szy = A[1][2];
  double N[4][4] = {{Sxx + Syy + Szz, Syz - Szy, Szx - Sxz, Sxy - Syx},
                    {Syz - Szy, Sxx - Syy - Szz, Sxy + Syx, Szx + Sxz},
                    {Szx - Sxz, Sxy + Syx, -Sxx + Syy - Szz, Syz + Szy},
                    {Sxy - Syx, Szx + Sxz, Syz + Szy, -Sxx - Syy + Szz}};
  double V[4][4];
  jacobi4(N, V);
  // the eigenvector of the largest eigenvalue (the first of equal ones),
  // selected with constant indices so that V stays in registers
  double bv = N[0][0], w = V[0][0], x = V[1][0], y = V[2][0], z = V[3][0];
#pragma unroll
  for (int i = 1; i < 4; ++i) {
    if (N[i][i] > bv) {
      bv = N[i][i];
      w = V[0][i];
      x = V[1][i];
      y = V[2][i];
      z = V[3][i];
    }
  }
  const double n = sqrt(w * w + x * x + y * y + z * z);
  const double sg = w < 0 ? -1.0 : 1.0;
  q[0] = sg * w / n;
  q[1] = sg * x / n;
  q[2] = sg * y / n;
  q[3] = sg * z / n;
}

// A node's rotation depends only on its Y edge and its X edge: node j of
// layer i uses Y edge i (i - 1 for the 2D-degenerate nodes [|EX|, 2|EX|)
// when unpruned) and X edge j mod |EX|. So one rotation per (Y edge, X edge)
// pair is computed, quat[ie * |EX| + ex], a third of the nodes, and the energy
// looks a node's rotation up with quat_index.
// (node ids are below 2^31; the layer is one 32-bit division, the rest
// compares)
__device__ __forceinline__ int64_t quat_index(int32_t node, int32_t nEX, int32_t nEY, int prune) {
  const uint32_t npl = 3u * (uint32_t)nEX;
  const uint32_t i = (uint32_t)node / npl, j = (uint32_t)node - i * npl;
  const uint32_t n = (uint32_t)nEX;
  const bool prev2d = j >= n && j < 2 * n && !prune;
  const uint32_t ie = prev2d ? (i == 0 ? (uint32_t)nEY - 1 : i - 1) : i;
  const uint32_t ex = j >= 2 * n ? j - 2 * n : (j >= n ? j - n : j);
  return (int64_t)ie * nEX + ex;
}

__global__ void __launch_bounds__(128) k_conj_quat(const double* __restrict__ VX, const double* __restrict__ NX,
                            const double* __restrict__ VY, const double* __restrict__ NY,
                            const int32_t* __restrict__ EX, int64_t nEX, const int32_t* __restrict__ EY,
                            int64_t nEY, double* __restrict__ quat) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= nEX * nEY) return;
  const int64_t ie = u / nEX, ex = u - ie * nEX;
  const V3 e2 = vsub(row3(VY, EY[2 * ie + 1]), row3(VY, EY[2 * ie]));
  const V3 n2 = row3(NY, EY[2 * ie]);
  const V3 c2 = vcross(e2, n2);
  const V3 e3 = vsub(row3(VX, EX[2 * ex + 1]), row3(VX, EX[2 * ex]));
  const V3 na = row3(NX, EX[2 * ex]), nb = row3(NX, EX[2 * ex + 1]);
  const V3 n3 = {(na.x + nb.x) / 2, (na.y + nb.y) / 2, (na.z + nb.z) / 2};
  const V3 c3 = vcross(e3, n3);
  double q[4];
  rotation_quat(e2, n2, c2, e3, n3, c3, q);
  for (int k = 0; k < 4; ++k) quat[4 * u + k] = q[k];
}

__device__ __forceinline__ double rho(double x, double alpha, double c, double pwr) {
  const double xDivCSqr = pow((x / c), pwr);
  const double absAlphaMinusTwo = fabs(alpha - 2);
  return absAlphaMinusTwo / alpha * (pow(xDivCSqr / absAlphaMinusTwo + 1, 0.5 * alpha) - 1);
}

__global__ void k_conj_energy(const double* __restrict__ VX, const double* __restrict__ VY, int64_t nVY,
                              const int32_t* __restrict__ rows, const int32_t* __restrict__ SRC,
                              const int32_t* __restrict__ TRGT, int64_t E, const double* __restrict__ FD,
                              const double* __restrict__ quat, int64_t nEX, int64_t nEY, int prune, int kind,
                              int lineIntegral, double alpha, double c, double pwr, double maxSO3Val,
                              double maxFeatDiff, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= E) return;
  const int32_t* r = rows + 6 * i;
  const int64_t y0 = r[0], y1 = r[1], y2 = r[2], x0 = r[3], x1 = r[4], x2 = r[5];
  double so3 = 0.0;
  if (quat) {
    const double* a = quat + 4 * quat_index(SRC[i], (int32_t)nEX, (int32_t)nEY, prune);
    const double* b = quat + 4 * quat_index(TRGT[i], (int32_t)nEX, (int32_t)nEY, prune);
    double ip = fabs(a[0] * b[0] + (a[1] * b[1] + a[2] * b[2] + a[3] * b[3]));
    if (ip > 1) ip = 1;
    so3 = isnan(ip) ? 0.0 : 2 * acos(ip);
  }
  const double fd = FD[x2 * nVY + y2];
  double len = 0;
  if (lineIntegral) {
    for (int j = 0; j < 3; ++j) {
      const double dy = VY[3 * y0 + j] - VY[3 * y1 + j];
      len += __dmul_rn(dy, dy);
      const double dx = VX[3 * x0 + j] - VX[3 * x1 + j];
      len += __dmul_rn(dx, dx);
    }
    len = sqrt(len);
  }
  const double pen = (y2 == y1 || x2 == x1) ? 0.1 : 1;
  const double twoPi = 2 * M_PI;
  double c0 = 0, c1 = 0;
  switch (kind) {
    case CONJ_VV: c0 = rho(so3, alpha, c, pwr) + fd; break;
    case CONJ_VL: c0 = rho(so3, alpha, c, pwr) + fd; c1 = len; break;
    case CONJ_LV: c0 = len; c1 = rho(so3, alpha, c, pwr) + fd; break;
    case CONJ_VP: c0 = rho(so3, alpha, c, pwr) + fd; c1 = pen; break;
    case CONJ_PV: c0 = pen; c1 = rho(so3, alpha, c, pwr) + fd; break;
    case CONJ_FSP: c0 = fd; c1 = fmax(0.0, twoPi - so3); break;
    case CONJ_FSR: c0 = fd; c1 = fmax(0.0, maxSO3Val - rho(so3, alpha, c, pwr)); break;
    case CONJ_SRF: c0 = rho(so3, alpha, c, pwr); c1 = fmax(0.0, maxFeatDiff - fd); break;
    default: c0 = so3; c1 = fmax(0.0, maxFeatDiff - fd); break;  // CONJ_SPF
  }
  out[i] = c0;
  if (kind != CONJ_VV) out[E + i] = c1;
}

__global__ void k_max_f64(const double* __restrict__ x, int64_t n, double* __restrict__ out) {
  __shared__ double red[256];
  double m = -__builtin_huge_val();
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) m = fmax(m, x[i]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

hipError_t launch_conj_product_space(const int32_t* EX, int64_t nEX, const int32_t* EY, int64_t nEY, int prune,
                                     const int32_t* outOff, const int32_t* outList, const int64_t* pre,
                                     int64_t perLayer, int64_t intraTotal, int2* desc, int32_t* rows, int32_t* src,
                                     int32_t* trgt, hipStream_t st) {
  const int64_t n = nEY * perLayer;
  if (n == 0 || nEX == 0) return hipSuccess;
  hipLaunchKernelGGL(k_conj_desc, dim3(nblk(nEX, 256)), dim3(256), 0, st, EX, nEX, prune, outOff, outList, pre,
                     intraTotal, desc);
  hipLaunchKernelGGL(k_conj_product_space, dim3(nblk(n, 256)), dim3(256), 0, st, EX, nEX, EY, nEY, prune, desc,
                     perLayer, rows, src, trgt);
  return hipGetLastError();
}

hipError_t launch_conj_quat(const double* VX, const double* NX, const double* VY, const double* NY,
                            const int32_t* EX, int64_t nEX, const int32_t* EY, int64_t nEY, int prune, double* quat,
                            hipStream_t st) {
  (void)prune;
  const int64_t n = nEX * nEY;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_conj_quat, dim3(nblk(n, 128)), dim3(128), 0, st, VX, NX, VY, NY, EX, nEX, EY, nEY,
                     quat);
  return hipGetLastError();
}

hipError_t launch_conj_energy(const double* VX, const double* VY, int64_t nVY, const int32_t* rows,
                              const int32_t* SRC, const int32_t* TRGT, int64_t E, const double* FD,
                              const double* quat, int64_t nEX, int64_t nEY, int prune, int kind, int lineIntegral,
                              double alpha, double c, double pwr, double maxSO3Val, double maxFeatDiff, double* out,
                              hipStream_t st) {
  if (E == 0) return hipSuccess;
  hipLaunchKernelGGL(k_conj_energy, dim3(nblk(E, 256)), dim3(256), 0, st, VX, VY, nVY, rows, SRC, TRGT, E, FD,
                     quat, nEX, nEY, prune, kind, lineIntegral, alpha, c, pwr, maxSO3Val, maxFeatDiff, out);
  return hipGetLastError();
}

hipError_t launch_max_f64(const double* x, int64_t n, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_max_f64, dim3(1), dim3(256), 0, st, x, n, out);
  return hipGetLastError();
}

}  // namespace pgc
