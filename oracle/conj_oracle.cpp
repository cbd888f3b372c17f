// TEST INFRASTRUCTURE ONLY (oracle).
//
// CPU restatement of the reference's conjugate product graph and its
// higher-order energy, the parity checker of the HIP construction kernels.
//
// Reference files restated here (paths relative to the reference root):
//   product_graph_generators/product_spaces/conj_product_space.cpp:16-228
//       (orc_conj_product_space, orc_conj_branch_graph)
//   product_graph_generators/energy/higherOrderDeformationEnergy.cpp:16-420
//       with USE_MATLAB_CONVENTION (higherOrderDeformationEnergy.hpp:14):
//       one rotation per product node (orc_conj_quaternions, orc_conj_energy)
//
// Deterministic choices:
//   * the out-edges of an X vertex are visited in ascending edge id (the
//     reference iterates a tsl::robin_set<int>, conj_product_space.cpp:109,132);
//     (source, target) node pairs are unique, so the solvers' results do not
//     depend on this order, only the edge ids do;
//   * the rotation of a node is the rotation R in SO(3) maximising
//     tr(R^T X^T Y) (the reference's Kabsch solution U diag(1,1,det(UV^T)) V^T
//     from Eigen's JacobiSVD, :36-51), computed as the dominant eigenvector of
//     Horn's 4x4 quaternion matrix with cyclic Jacobi rotations. The angle
//     between two nodes, 2 acos(|<q0, q1>|), is sign-free. Checked against a
//     numpy SVD Kabsch solution in tests/test_oracle_conj.py.
//
// Mode handling (stringToMode, :393-410, keeps the reference's inverted
// "penalisedegenerate" test): (vanilla, nomode) aborts in the reference
// (:216-218 with :332-335); here it is the one-column mean problem
// rho(so3) + fd. (length, vanilla) and (penalisedegenerate, vanilla) write
// column 1 of a one-column matrix in the reference (:299-313); here they give
// two columns. Other unsupported pairs return -1.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

namespace {

enum Mode { NOMODE, VANILLA, FEATURE, LENGTH, SO3_PLAIN, SO3_ROBUST, PEN_DEG };

Mode to_mode(const char* s) {
  const std::string m(s ? s : "");
  if (m == "nomode") return NOMODE;
  if (m == "vanilla") return VANILLA;
  if (m == "feature") return FEATURE;
  if (m == "lengthnormalisation") return LENGTH;
  if (m == "plainso3") return SO3_PLAIN;
  if (m == "robustso3") return SO3_ROBUST;
  if (m != "penalisedegenerate") return PEN_DEG;
  return VANILLA;
}

std::vector<std::vector<int>> out_edges(const int32_t* EX, long nEX, long nVX) {
  std::vector<std::vector<int>> out(nVX);
  for (long e = 0; e < nEX; ++e) out[EX[2 * e]].push_back((int)e);
  return out;  // ascending edge id by construction
}

struct V3 {
  double x, y, z;
};
V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
V3 normalized(V3 a) {
  const double n2 = a.x * a.x + a.y * a.y + a.z * a.z;
  if (!(n2 > 0)) return a;
  const double n = std::sqrt(n2);
  return {a.x / n, a.y / n, a.z / n};
}
V3 row3(const double* M, long r) { return {M[3 * r], M[3 * r + 1], M[3 * r + 2]}; }

// Dominant eigenvector of a symmetric 4x4 matrix by cyclic Jacobi rotations.
void jacobi4(double a[4][4], double v[4][4]) {
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
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0);
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

// Unit quaternion (w, x, y, z) of the rotation maximising tr(R^T A),
// A = X^T Y with X, Y the row-stacked (normal, edge, edge x normal) frames.
void rotation_quat(V3 v2, V3 n2, V3 c2, V3 v3, V3 n3, V3 c3, double q[4]) {
  const V3 X[3] = {normalized(n2), normalized(v2), normalized(c2)};
  const V3 Y[3] = {normalized(n3), normalized(v3), normalized(c3)};
  double A[3][3];
  // A[i][j] = sum_k X[k][i] * Y[k][j]
  const double Xm[3][3] = {{X[0].x, X[0].y, X[0].z}, {X[1].x, X[1].y, X[1].z}, {X[2].x, X[2].y, X[2].z}};
  const double Ym[3][3] = {{Y[0].x, Y[0].y, Y[0].z}, {Y[1].x, Y[1].y, Y[1].z}, {Y[2].x, Y[2].y, Y[2].z}};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += Xm[k][i] * Ym[k][j];
      A[i][j] = s;
    }
  // Horn's matrix for max tr(R M) with M = A^T (S_ab = M[a][b] = A[b][a])
  const double Sxx = A[0][0], Syy = A[1][1], Szz = A[2][2];
  const double Sxy = A[1][0], Syx = A[0][1], Sxz = A[2][0], Szx = A[0][2], Syz = A[2][1], Szy = A[1][2];
  double N[4][4] = {{Sxx + Syy + Szz, Syz - Szy, Szx - Sxz, Sxy - Syx},
                    {Syz - Szy, Sxx - Syy - Szz, Sxy + Syx, Szx + Sxz},
                    {Szx - Sxz, Sxy + Syx, -Sxx + Syy - Szz, Syz + Szy},
                    {Sxy - Syx, Szx + Sxz, Syz + Szy, -Sxx - Syy + Szz}};
  double V[4][4];
  jacobi4(N, V);
  int best = 0;
  for (int i = 1; i < 4; ++i)
    if (N[i][i] > N[best][best]) best = i;
  double w = V[0][best], x = V[1][best], y = V[2][best], z = V[3][best];
  const double n = std::sqrt(w * w + x * x + y * y + z * z);
  const double sg = w < 0 ? -1.0 : 1.0;
  q[0] = sg * w / n;
  q[1] = sg * x / n;
  q[2] = sg * y / n;
  q[3] = sg * z / n;
}

// precomputeAllQuaternionds with USE_MATLAB_CONVENTION (:95-134)
void node_quat(const double* VX, const double* NX, const double* VY, const double* NY, const int32_t* EX, long nEX,
               const int32_t* EY, long nEY, int prune, long node, double q[4]) {
  const long npl = 3 * nEX;
  const long i = node / npl, j = node % npl;
  auto frame2d = [&](long ie, V3& e, V3& n, V3& c) {
    e = sub(row3(VY, EY[2 * ie + 1]), row3(VY, EY[2 * ie]));
    n = row3(NY, EY[2 * ie]);
    c = cross(e, n);
  };
  V3 e2, n2, c2;
  frame2d(i, e2, n2, c2);
  const long ex = j - (j / nEX) * nEX;
  const V3 e3 = sub(row3(VX, EX[2 * ex + 1]), row3(VX, EX[2 * ex]));
  const V3 na = row3(NX, EX[2 * ex]), nb = row3(NX, EX[2 * ex + 1]);
  const V3 n3 = {(na.x + nb.x) / 2, (na.y + nb.y) / 2, (na.z + nb.z) / 2};
  const V3 c3 = cross(e3, n3);
  if (j >= nEX && j < 2 * nEX && !prune) {
    V3 e2p, n2p, c2p;
    frame2d((i + nEY - 1) % nEY, e2p, n2p, c2p);
    rotation_quat(e2p, n2p, c2p, e3, n3, c3, q);
  } else {
    rotation_quat(e2, n2, c2, e3, n3, c3, q);
  }
}

double rho(double x, double alpha, double c, double pwr) {
  const double xDivCSqr = std::pow((x / c), pwr);
  const double absAlphaMinusTwo = std::abs(alpha - 2);
  return absAlphaMinusTwo / alpha * (std::pow(xDivCSqr / absAlphaMinusTwo + 1, 0.5 * alpha) - 1);
}

}  // namespace

extern "C" {

// Conjugate product space (conj_product_space.cpp:70-168). Rows
// [src2d, mid2d, trgt2d, src3d, mid3d, trgt3d] row-major; returns the edge
// count (call with rows == NULL to count).
long orc_conj_product_space(const int32_t* EX, long nEX, const int32_t* EY, long nEY, int prune, int32_t* rows,
                            int32_t* src, int32_t* trgt) {
  long nVX = 0;
  for (long i = 0; i < 2 * nEX; ++i) nVX = std::max<long>(nVX, EX[i] + 1);
  const auto out = out_edges(EX, nEX, nVX);
  const long npl = 3 * nEX;
  const long off2d = nEX, off3d = prune ? nEX : 2 * nEX;
  long n = 0;
  auto emit = [&](int a, int b, int c, int d, int e, int f, long s, long t) {
    if (rows) {
      int32_t* r = rows + 6 * n;
      r[0] = a;
      r[1] = b;
      r[2] = c;
      r[3] = d;
      r[4] = e;
      r[5] = f;
      src[n] = (int32_t)s;
      trgt[n] = (int32_t)t;
    }
    n++;
  };
  for (long i = 0; i < nEY; ++i) {
    const long nx = (i + 1) % nEY;
    const long cur = i * npl, nxt = nx * npl;
    const int s2 = EY[2 * i], m2 = EY[2 * i + 1], t2 = EY[2 * nx + 1];
    if (!prune) {
      for (long j = 0; j < nEX; ++j) {
        const int s3 = EX[2 * j], m3 = EX[2 * j + 1];
        for (int it : out[m3]) {
          const int t3 = EX[2 * it + 1];
          if (t3 == s3) continue;
          emit(s2, s2, s2, s3, m3, t3, cur + off2d + j, cur + off2d + it);
          emit(s2, s2, m2, s3, m3, t3, cur + off2d + j, cur + it);
        }
      }
    }
    for (long j = 0; j < nEX; ++j) {
      const int s3 = EX[2 * j], m3 = EX[2 * j + 1];
      for (int it : out[m3]) {
        const int t3 = EX[2 * it + 1];
        if (t3 == s3) continue;
        emit(s2, m2, t2, s3, m3, t3, cur + j, nxt + it);
        if (!prune) emit(s2, m2, m2, s3, m3, t3, cur + j, nxt + off2d + it);
        emit(s2, m2, t2, m3, m3, t3, cur + off3d + j, nxt + it);
      }
      emit(s2, m2, t2, s3, m3, m3, cur + j, nxt + off3d + j);
      emit(s2, m2, t2, m3, m3, m3, cur + off3d + j, nxt + off3d + j);
    }
  }
  return n;
}

// branchGraph of the conjugate product space (:170-225) as CSR over the
// 3*nEX layer-0 nodes, values sorted per node. Returns the value count
// (call with vals == NULL to count; off must hold 3*nEX+1 entries).
long orc_conj_branch_graph(const int32_t* EX, long nEX, int prune, int64_t* off, int64_t* vals) {
  long nVX = 0;
  for (long i = 0; i < 2 * nEX; ++i) nVX = std::max<long>(nVX, EX[i] + 1);
  const auto out = out_edges(EX, nEX, nVX);
  const long npl = 3 * nEX;
  std::vector<std::vector<int64_t>> bg(npl);
  for (long j = 0; j < nEX; ++j) {
    bg[j].push_back(j + nEX);
    if (!prune) {
      bg[j + nEX].push_back(j + 2 * nEX);
      bg[j + 2 * nEX].push_back(j);
    } else {
      bg[j + nEX].push_back(j);
    }
    // both loops of the reference read EX(j, 0) (:189, :206)
    for (int rep = 0; rep < 2; ++rep) {
      for (int it : out[EX[2 * j]]) {
        bg[j].push_back(it);
        bg[j + nEX].push_back(it + nEX);
        if (!prune) bg[j + 2 * nEX].push_back(it + 2 * nEX);
      }
    }
  }
  long n = 0;
  if (off) off[0] = 0;
  for (long v = 0; v < npl; ++v) {
    std::sort(bg[v].begin(), bg[v].end());
    bg[v].erase(std::unique(bg[v].begin(), bg[v].end()), bg[v].end());
    for (int64_t x : bg[v]) {
      if (vals) vals[n] = x;
      n++;
    }
    if (off) off[v + 1] = n;
  }
  return n;
}

// Rotation quaternions of the product nodes [0, 3*nEX*nEY) (w, x, y, z).
void orc_conj_quaternions(const double* VX, const double* NX, const double* VY, const double* NY, const int32_t* EX,
                          long nEX, const int32_t* EY, long nEY, int prune, double* q) {
  const long nn = 3 * nEX * nEY;
#pragma omp parallel for
  for (long n = 0; n < nn; ++n) node_quat(VX, NX, VY, NY, EX, nEX, EY, nEY, prune, n, q + 4 * n);
}

// Conjugate energy (higherOrderDeformationEnergy.cpp:209-339). rows: E x 6
// row-major; FD |VX| x |VY| row-major; out column-major E x ncols. Returns
// ncols, or -1 for an unsupported (cost, time) pair.
int orc_conj_energy(const double* VX, const double* NX, const double* VY, long nVY, const double* NY,
                    const int32_t* EX, long nEX, const int32_t* EY, long nEY, const int32_t* rows,
                    const int32_t* SRC, const int32_t* TRGT, long E, const double* FD, long nVX, int reg,
                    const char* cost_mode, const char* time_mode, double alpha, double c, double pwr, int prune,
                    double* out) {
  const Mode cm = to_mode(cost_mode), tm = to_mode(time_mode);
  enum { VV, VL, LV, VP, PV, FSP, FSR, SRF, SPF } kind;
  if (cm == VANILLA && (tm == VANILLA || tm == NOMODE)) kind = VV;
  else if (cm == VANILLA && tm == LENGTH) kind = VL;
  else if (cm == LENGTH && tm == VANILLA) kind = LV;
  else if (cm == VANILLA && tm == PEN_DEG) kind = VP;
  else if (cm == PEN_DEG && tm == VANILLA) kind = PV;
  else if (cm == FEATURE && tm == SO3_PLAIN) kind = FSP;
  else if (cm == FEATURE && tm == SO3_ROBUST) kind = FSR;
  else if (cm == SO3_ROBUST && tm == FEATURE) kind = SRF;
  else if (cm == SO3_PLAIN && tm == FEATURE) kind = SPF;
  else return -1;
  const int ncols = kind == VV ? 1 : 2;
  const bool lineIntegral = cm == LENGTH || tm == LENGTH;
  const double twoPi = 2 * M_PI;
  const float maxSO3Val = (float)rho(twoPi, alpha, c, pwr);
  float maxFeatDiff = 0;
  if (tm == FEATURE) {
    double m = -INFINITY;
    for (long i = 0; i < nVX * nVY; ++i) m = std::max(m, FD[i]);
    maxFeatDiff = (float)m;
  }
  std::vector<double> quat;
  if (reg) {
    quat.resize(4 * 3 * nEX * nEY);
    orc_conj_quaternions(VX, NX, VY, NY, EX, nEX, EY, nEY, prune, quat.data());
  }
#pragma omp parallel for
  for (long i = 0; i < E; ++i) {
    const int32_t* r = rows + 6 * i;
    const long y0 = r[0], y1 = r[1], y2 = r[2], x0 = r[3], x1 = r[4], x2 = r[5];
    double so3 = 0.0;
    if (reg) {
      const double* a = quat.data() + 4 * (long)SRC[i];
      const double* b = quat.data() + 4 * (long)TRGT[i];
      double ip = std::fabs(a[0] * b[0] + (a[1] * b[1] + a[2] * b[2] + a[3] * b[3]));
      if (ip > 1) ip = 1;
      so3 = std::isnan(ip) ? 0.0 : 2 * std::acos(ip);
    }
    const double fd = FD[x2 * nVY + y2];
    double len = 0;
    if (lineIntegral) {
      for (int j = 0; j < 3; ++j) {
        len += std::pow(VY[3 * y0 + j] - VY[3 * y1 + j], 2);
        len += std::pow(VX[3 * x0 + j] - VX[3 * x1 + j], 2);
      }
      len = std::sqrt(len);
    }
    const double pen = (y2 == y1 || x2 == x1) ? 0.1 : 1;
    double c0 = 0, c1 = 0;
    switch (kind) {
      case VV: c0 = rho(so3, alpha, c, pwr) + fd; break;
      case VL: c0 = rho(so3, alpha, c, pwr) + fd; c1 = len; break;
      case LV: c0 = len; c1 = rho(so3, alpha, c, pwr) + fd; break;
      case VP: c0 = rho(so3, alpha, c, pwr) + fd; c1 = pen; break;
      case PV: c0 = pen; c1 = rho(so3, alpha, c, pwr) + fd; break;
      case FSP: c0 = fd; c1 = std::max(0.0, twoPi - so3); break;
      case FSR: c0 = fd; c1 = std::max(0.0, maxSO3Val - rho(so3, alpha, c, pwr)); break;
      case SRF: c0 = rho(so3, alpha, c, pwr); c1 = std::max(0.0, maxFeatDiff - fd); break;
      case SPF: c0 = so3; c1 = std::max(0.0, maxFeatDiff - fd); break;
    }
    out[i] = c0;
    if (ncols > 1) out[E + i] = c1;
  }
  return ncols;
}

}  // extern "C"
