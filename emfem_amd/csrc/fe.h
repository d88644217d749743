// Lowest-order Nedelec (Whitney) element on a tetrahedron, shared by the device
// kernels and the host-side right-hand-side builders.
//
// Reference: FEValues (src/em_fe.cc:284-413), the element loop of assemble_matrix
// (src/em_fwd.cc:343-388) and the quadrature rules (src/em_fe.cc:6-154).
//
// The reference integrates the 6x6 matrices with the 4-point degree-2 rule.  Here
// the same integrals are evaluated in closed form: curl N is constant per tet and
// the mass integrand is a quadratic in the barycentrics, integrated exactly by
// int(lambda_p lambda_q) = V (1 + delta_pq) / 20.  The 4-point rule reproduces
// these to within the 15 significant digits of its abscissae.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define EMF_HD __host__ __device__
#else
#define EMF_HD
#endif

namespace emfem {

constexpr double kPi = 3.14159265358979323846;
constexpr double kMu = 4 * kPi * 1E-7;        // src/em_defs.h:16-17
constexpr double kDtoR = kPi / 180.0;          // src/em_defs.h:21
constexpr double kGeomEps = 1.0E-6;            // src/em_defs.h:23

// Canonical local edge order: local edge i joins local vertices kLocalEdge[i].
// Orientation follows the global vertex ids (lower id first, src/em_mesh.cc:448-455).
EMF_HD inline void local_edge(int i, int &p, int &q) {
  const int a[6] = {0, 0, 0, 1, 1, 2};
  const int b[6] = {1, 2, 3, 2, 3, 3};
  p = a[i];
  q = b[i];
}

struct TetGeom {
  double g[4][3];  // gradients of the barycentric coordinates
  double det;      // det([1; x_v]) = 6 V (signed, as the reference uses it)
};

EMF_HD inline void cross3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

EMF_HD inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Barycentric gradients from the inverse of [x1-x0, x2-x0, x3-x0]
// (equivalent to the 4x4 inverse of FEValues::calculate_coefs, src/em_fe.cc:398-413).
EMF_HD inline void tet_geometry(const double x[4][3], TetGeom &G) {
  double c1[3], c2[3], c3[3], r1[3], r2[3], r3[3];
  for (int k = 0; k < 3; ++k) {
    c1[k] = x[1][k] - x[0][k];
    c2[k] = x[2][k] - x[0][k];
    c3[k] = x[3][k] - x[0][k];
  }
  cross3(c2, c3, r1);
  cross3(c3, c1, r2);
  cross3(c1, c2, r3);
  G.det = dot3(c1, r1);
  double inv = 1.0 / G.det;
  for (int k = 0; k < 3; ++k) {
    G.g[1][k] = r1[k] * inv;
    G.g[2][k] = r2[k] * inv;
    G.g[3][k] = r3[k] * inv;
    G.g[0][k] = -(G.g[1][k] + G.g[2][k] + G.g[3][k]);
  }
}

// Barycentric coordinates of p.
EMF_HD inline void barycentric(const double x[4][3], const TetGeom &G, const double *p, double *lam) {
  double d[3] = {p[0] - x[0][0], p[1] - x[0][1], p[2] - x[0][2]};
  lam[1] = dot3(G.g[1], d);
  lam[2] = dot3(G.g[2], d);
  lam[3] = dot3(G.g[3], d);
  lam[0] = 1.0 - lam[1] - lam[2] - lam[3];
}

// Local vertices (a, b) of local edge i, oriented by global vertex id.
EMF_HD inline void edge_ends(const int *vid, int i, int &a, int &b) {
  int p, q;
  local_edge(i, p, q);
  if (vid[p] < vid[q]) {
    a = p;
    b = q;
  } else {
    a = q;
    b = p;
  }
}

// Symmetric conductivity stored as (xx, yy, zz, xy, xz, yz).
EMF_HD inline void sig_apply(const double *s, const double *v, double *o) {
  o[0] = s[0] * v[0] + s[3] * v[1] + s[4] * v[2];
  o[1] = s[3] * v[0] + s[1] * v[1] + s[5] * v[2];
  o[2] = s[4] * v[0] + s[5] * v[1] + s[2] * v[2];
}

// Whitney value N_i(p) = lambda_a grad lambda_b - lambda_b grad lambda_a (src/em_fe.cc:304-308).
EMF_HD inline void whitney_value(const TetGeom &G, const double *lam, int a, int b, double *v) {
  for (int k = 0; k < 3; ++k) v[k] = lam[a] * G.g[b][k] - lam[b] * G.g[a][k];
}

// curl N_i = 2 grad lambda_a x grad lambda_b (src/em_fe.cc:310-313).
EMF_HD inline void whitney_curl(const TetGeom &G, int a, int b, double *c) {
  cross3(G.g[a], G.g[b], c);
  c[0] *= 2.0;
  c[1] *= 2.0;
  c[2] *= 2.0;
}

// Raw (pre-Dirichlet) element matrices, stored as 6x6 row-major:
//   C_ij = (1/mu) int curl N_i . curl N_j,  M_ij = omega int N_i . sigma N_j
// (src/em_fwd.cc:358-375).
EMF_HD inline void element_matrices(const TetGeom &G, const int *vid, const double *sig, double omega,
                                    double C[6][6], double M[6][6]) {
  int ea[6], eb[6];
  double cr[6][3];
  double S[4][4];
  for (int i = 0; i < 6; ++i) {
    edge_ends(vid, i, ea[i], eb[i]);
    cross3(G.g[ea[i]], G.g[eb[i]], cr[i]);
  }
  for (int p = 0; p < 4; ++p) {
    double sg[3];
    sig_apply(sig, G.g[p], sg);
    for (int q = p; q < 4; ++q) {
      S[p][q] = dot3(G.g[q], sg);
      S[q][p] = S[p][q];
    }
  }
  const double vol = G.det / 6.0;
  const double cc = 4.0 * vol / kMu;
  const double mm = omega * vol / 20.0;
  for (int i = 0; i < 6; ++i) {
    for (int j = 0; j <= i; ++j) {
      const int a = ea[i], b = eb[i], c = ea[j], d = eb[j];
      double cv = cc * dot3(cr[i], cr[j]);
      double mv = mm * ((a == c ? 2.0 : 1.0) * S[b][d] - (a == d ? 2.0 : 1.0) * S[b][c] -
                        (b == c ? 2.0 : 1.0) * S[a][d] + (b == d ? 2.0 : 1.0) * S[a][c]);
      C[i][j] = C[j][i] = cv;
      M[i][j] = M[j][i] = mv;
    }
  }
}

// The reference's Dirichlet rule (src/em_fwd.cc:377-388): for a boundary edge i,
// C keeps only its diagonal in row/column i and M loses row/column i entirely.
EMF_HD inline void apply_dirichlet(const bool *bdr, double C[6][6], double M[6][6]) {
  for (int i = 0; i < 6; ++i) {
    if (!bdr[i]) continue;
    for (int j = 0; j < 6; ++j) {
      if (j != i) C[i][j] = C[j][i] = 0.0;
      M[i][j] = M[j][i] = 0.0;
    }
  }
}

}  // namespace emfem
