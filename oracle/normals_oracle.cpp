// CPU restatement of the reference's normal helpers. TEST INFRASTRUCTURE ONLY:
// loaded by tests/ through oracle/pgc_oracle.py, never by the product package.
//
//   orc_normals2d  helper/utils.cpp:93-157 (utils::normals2d), step by step in
//                  plain C++ with the reference's scalar types: the bounding
//                  box values and the multiplier are float, everything else
//                  double; rowwise normalisation divides by the row norm with
//                  no zero guard (Eigen 3.4's VectorwiseOp::normalized), so a
//                  zero row becomes NaN.
//   orc_normals3d  helper/utils.cpp:86-90 (utils::normals3d), which calls
//                  igl::per_vertex_normals(V, F, N) of libigl b1bd5b1 (not in
//                  the reference tree; a FetchContent dependency,
//                  cmake/libigl.cmake:9). Restated from libigl's published
//                  per_vertex_normals / per_face_normals / doublearea: default
//                  weighting = face double area; face normal = (v1-v0) x (v2-v0)
//                  normalised, (0,0,0) for a degenerate face; corners
//                  accumulated face by face; rows normalised as above.
//
// Pins: the closed forms of tests/test_normals.py (circle, ellipse, the
// coordinate-plane permutation, polyhedra, a corner whose area weighting
// differs from uniform and angle weighting). The library itself is absent, so
// the 3D restatement has no vectors from it: its parity is pinned by those
// closed forms only.
#include <cmath>
#include <cstdint>
#include <vector>

namespace {

inline void normalize_rows(double* M, long n) {
  for (long i = 0; i < n; ++i) {
    double* r = M + 3 * i;
    const double nrm = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    r[0] /= nrm;
    r[1] /= nrm;
    r[2] /= nrm;
  }
}

}  // namespace

extern "C" {

// V [nV x 3] row-major, E [nE x 2] row-major. out [nV x 3] row-major.
// Returns 0, or 1 (V.rows() != E.rows()) / 2 (no coordinate column sums to
// |.| < 1e-4): the reference prints an error and returns a 1x1 matrix; out is
// left untouched. *zero_dim receives the dimension the reference dropped.
int orc_normals2d(const double* Vin, long nV, const int32_t* E, long nE, double* out, int* zero_dim) {
  if (nV != nE) return 1;
  int zd = -1;
  for (int i = 0; i < 3; ++i) {
    double s = 0.0;
    for (long r = 0; r < nV; ++r) s += Vin[3 * r + i];
    if (std::fabs(s) < 1e-4) {
      zd = i;
      break;
    }
  }
  *zero_dim = zd;
  if (zd == -1) return 2;
  // resorting: the dropped dimension becomes column 2
  int perm[3] = {0, 1, 2};
  if (zd == 1) {
    perm[0] = 0; perm[1] = 2; perm[2] = 1;
  } else if (zd == 0) {
    perm[0] = 1; perm[1] = 2; perm[2] = 0;
  }
  std::vector<double> V(3 * nV);
  for (long r = 0; r < nV; ++r)
    for (int c = 0; c < 3; ++c) V[3 * r + c] = Vin[3 * r + perm[c]];

  // edge vectors, lengths, and edge normals (edge vector times R^T with
  // R = [[0,1,0],[-1,0,0],[0,0,1]]: (x, y, z) -> (y, -x, z))
  std::vector<double> ev(3 * nE), el(nE), en(3 * nE);
  for (long e = 0; e < nE; ++e) {
    const long a = E[2 * e], b = E[2 * e + 1];
    for (int c = 0; c < 3; ++c) ev[3 * e + c] = V[3 * a + c] - V[3 * b + c];
    const double x = ev[3 * e], y = ev[3 * e + 1], z = ev[3 * e + 2];
    el[e] = std::sqrt(x * x + y * y + z * z);
    // edgeVectors * R^T: row (x, y, z) times R^T = (x*R00 + y*R01 + z*R02, ...)
    en[3 * e + 0] = x * 0.0 + y * 1.0 + z * 0.0;
    en[3 * e + 1] = x * -1.0 + y * 0.0 + z * 0.0;
    en[3 * e + 2] = x * 0.0 + y * 0.0 + z * 1.0;
  }
  normalize_rows(en.data(), nE);

  // outward check on the bounding box, in float as the reference declares it
  double mxX = V[0], mnX = V[0], mxY = V[1], mnY = V[1];
  for (long r = 1; r < nV; ++r) {
    mxX = std::fmax(mxX, V[3 * r]);
    mnX = std::fmin(mnX, V[3 * r]);
    mxY = std::fmax(mxY, V[3 * r + 1]);
    mnY = std::fmin(mnY, V[3 * r + 1]);
  }
  const float maxXVal = (float)mxX, minXVal = (float)mnX;
  const float maxYVal = (float)mxY, minYVal = (float)mnY;
  const float dx = std::fabs(maxXVal - minXVal), dy = std::fabs(maxYVal - minYVal);
  const float multiplier = (float)(0.1 * (double)(dx < dy ? dx : dy));
  double mx = -INFINITY;
  for (long r = 0; r < nV; ++r) {
    // (V + multiplier * edgeN).col(0): row r of edgeN (edges and vertices
    // have the same count)
    const double x = V[3 * r] + (double)multiplier * en[3 * r];
    mx = r == 0 ? x : std::fmax(mx, x);
  }
  const float maxXValmult = (float)mx;
  if (maxXValmult < maxXVal)
    for (double& x : en) x = -x;

  // vertex normals: edge-length-weighted sum of the two adjacent edge normals
  // (row i - 1 and row i mod n, for i = 1..n)
  for (long i = 1; i < nV + 1; ++i) {
    const long nx = i % nV;
    for (int c = 0; c < 3; ++c) out[3 * nx + c] = el[i - 1] * en[3 * (i - 1) + c] + el[nx] * en[3 * nx + c];
  }
  normalize_rows(out, nV);
  return 0;
}

// V [nV x 3], F [nF x 3] row-major; out [nV x 3] row-major.
void orc_normals3d(const double* V, long nV, const int32_t* F, long nF, double* out) {
  for (long i = 0; i < 3 * nV; ++i) out[i] = 0.0;
  for (long f = 0; f < nF; ++f) {
    const double* p0 = V + 3 * (long)F[3 * f];
    const double* p1 = V + 3 * (long)F[3 * f + 1];
    const double* p2 = V + 3 * (long)F[3 * f + 2];
    // face normal: (v1 - v0) x (v2 - v0), normalised; zero for a degenerate face
    const double a[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    const double b[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
    double n[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    const double r = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (r == 0.0) {
      n[0] = n[1] = n[2] = 0.0;
    } else {
      n[0] /= r;
      n[1] /= r;
      n[2] /= r;
    }
    // double area: root of the summed squares of the double areas projected
    // on the planes (x,y), (y,z), (z,x), each from corner 2:
    // r = v0 - v2, s = v1 - v2, proj = r_d * s_e - r_e * s_d
    double w = 0.0;
    for (int d = 0; d < 3; ++d) {
      const int e = (d + 1) % 3;
      const double rd = p0[d] - p2[d], re = p0[e] - p2[e];
      const double sd = p1[d] - p2[d], se = p1[e] - p2[e];
      const double pa = rd * se - re * sd;
      w += pa * pa;
    }
    w = std::sqrt(w);
    for (int c = 0; c < 3; ++c) {
      double* o = out + 3 * (long)F[3 * f + c];
      o[0] += w * n[0];
      o[1] += w * n[1];
      o[2] += w * n[2];
    }
  }
  normalize_rows(out, nV);
}

}  // extern "C"
