// Element assembly and the goal-oriented estimator on the device.
//
// Assembly (src/em_fwd.cc:318-400): one lane per tet; the 6x6 curl-curl and mass
// matrices (21 + 21 unique values) stay in registers and are added into the CSR
// values (C, M interleaved per nonzero) by a read-modify-write scatter.  Tets are
// launched one colour at a time; tets of one colour share no edge, so no two lanes
// ever touch the same CSR row: the scatter needs no atomics and the summation order
// is fixed, which makes the result bitwise reproducible.
//
// Estimator (src/em_fwd.cc:963-1180): one lane per tet computes its volume residual
// and its four face jumps.  An interior face is evaluated from the side of the lower
// tet index (the side the reference visits first, src/em_fwd.cc:1148-1159), so both
// tets sharing it compute the identical value without any exchange.
#include <hip/hip_runtime.h>

#include "common.h"
#include "fe.h"
#include "device.h"
#include "kernels.h"

namespace emfem {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ int csr_find(const int *__restrict__ ci, int b, int e, int c) {
  // rows hold ~16 sorted columns: binary search
  while (b < e) {
    int m = (b + e) >> 1;
    if (ci[m] < c) b = m + 1;
    else e = m;
  }
  return b;
}

__device__ __forceinline__ void load_tet(int t, const int *__restrict__ tets, const double *__restrict__ pts, int vid[4],
                                         double x[4][3]) {
  const int4 v = reinterpret_cast<const int4 *>(tets)[t];
  vid[0] = v.x;
  vid[1] = v.y;
  vid[2] = v.z;
  vid[3] = v.w;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) x[i][k] = pts[(size_t)vid[i] * 3 + k];
}

__global__ __launch_bounds__(kBlock) void k_assemble(int ntets, const int *__restrict__ tet_ids,
                                                     const int *__restrict__ tets, const int *__restrict__ tet2edge,
                                                     const double *__restrict__ pts, const int *__restrict__ attr,
                                                     const double *__restrict__ sig6, const uint8_t *__restrict__ edge_bdr,
                                                     double omega, const int *__restrict__ rp,
                                                     const int *__restrict__ ci, double2 *K) {
  const int i = xcd_block() * kBlock + threadIdx.x;
  if (i >= ntets) return;
  const int t = tet_ids[i];
  int vid[4];
  double x[4][3];
  load_tet(t, tets, pts, vid, x);
  int edge[6];
  bool bdr[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    edge[k] = tet2edge[(size_t)t * 6 + k];
    bdr[k] = edge_bdr[edge[k]] != 0;
  }
  double sig[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) sig[k] = sig6[attr[t] * 6 + k];
  TetGeom G;
  tet_geometry(x, G);
  double C[6][6], M[6][6];
  element_matrices(G, vid, sig, omega, C, M);
  apply_dirichlet(bdr, C, M);
#pragma unroll
  for (int a = 0; a < 6; ++a) {
    const int r = edge[a];
    const int b = rp[r], e = rp[r + 1];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const int p = csr_find(ci, b, e, edge[c]);
      double2 v = K[p];
      v.x += C[a][c];
      v.y += M[a][c];
      K[p] = v;
    }
  }
}

// ---------------------------------------------------------------------------------
// estimator
// ---------------------------------------------------------------------------------

struct TetData {
  int vid[4];
  double x[4][3];
  int edge[6];
  TetGeom G;
  double sig[6];
};

__device__ __forceinline__ void load_full(int t, const int *__restrict__ tets, const int *__restrict__ tet2edge,
                                          const double *__restrict__ pts, const int *__restrict__ attr,
                                          const double *__restrict__ sig6, TetData &D) {
  load_tet(t, tets, pts, D.vid, D.x);
#pragma unroll
  for (int k = 0; k < 6; ++k) D.edge[k] = tet2edge[(size_t)t * 6 + k];
#pragma unroll
  for (int k = 0; k < 6; ++k) D.sig[k] = sig6[attr[t] * 6 + k];
  tet_geometry(D.x, D.G);
}

__device__ __forceinline__ double dist(const double *a, const double *b) {
  const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  return sqrt(dx * dx + dy * dy + dz * dz);
}

// circumradius of the tet (TetAccessor::diameter, src/em_mesh.h:232-254)
__device__ double tet_circumradius(const TetData &D) {
  const double a = dist(D.x[1], D.x[2]), b = dist(D.x[0], D.x[2]), c = dist(D.x[0], D.x[1]);
  const double aa = dist(D.x[0], D.x[3]), bb = dist(D.x[1], D.x[3]), cc = dist(D.x[2], D.x[3]);
  const double la = a * aa, lb = b * bb, lc = c * cc, s = 0.5 * (la + lb + lc);
  return sqrt(s * (s - la) * (s - lb) * (s - lc)) / D.G.det;
}

// circumradius of face f (TetAccessor::face_diameter, src/em_mesh.h:256-274)
__device__ double face_circumradius(const TetData &D, int f) {
  const double *v0 = D.x[(f + 1) & 3], *v1 = D.x[(f + 2) & 3], *v2 = D.x[(f + 3) & 3];
  const double a = dist(v1, v2), b = dist(v0, v2), c = dist(v0, v1);
  const double area = 0.25 * sqrt((a + b + c) * (b + c - a) * (c + a - b) * (a + b - c));
  return a * b * c / (4.0 * area);
}

// Field value (complex, as re/im triples) of the solution at p inside D.
__device__ void field_value(const TetData &D, const double *lam, const double *sol, int ld, int col, double er[3],
                            double ei[3]) {
  er[0] = er[1] = er[2] = ei[0] = ei[1] = ei[2] = 0.0;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    int a, b;
    edge_ends(D.vid, i, a, b);
    double N[3];
    whitney_value(D.G, lam, a, b, N);
    const double vr = sol[(size_t)D.edge[i] * ld + col], vi = sol[(size_t)D.edge[i] * ld + col + 1];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      er[k] += vr * N[k];
      ei[k] += vi * N[k];
    }
  }
}

__device__ void field_curl(const TetData &D, const double *sol, int ld, int col, double cr[3], double ci_[3]) {
  cr[0] = cr[1] = cr[2] = ci_[0] = ci_[1] = ci_[2] = 0.0;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    int a, b;
    edge_ends(D.vid, i, a, b);
    double Cn[3];
    whitney_curl(D.G, a, b, Cn);
    const double vr = sol[(size_t)D.edge[i] * ld + col], vi = sol[(size_t)D.edge[i] * ld + col + 1];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      cr[k] += vr * Cn[k];
      ci_[k] += vi * Cn[k];
    }
  }
}

constexpr int kMaxSol = 8;

// NS solutions at compile time keep the per-solution sums in registers; the two sides
// of a face are loaded into their own structs (no run-time choice between structs,
// which would force them into scratch memory).
template <int NS>
__global__ __launch_bounds__(128) void k_estimate(int nt, const int *__restrict__ tets, const int *__restrict__ tet2edge,
                                                  const int *__restrict__ neighbor, const int *__restrict__ non,
                                                  const double *__restrict__ pts, const int *__restrict__ attr,
                                                  const double *__restrict__ sig6, double omega,
                                                  const int *__restrict__ slots, const double *__restrict__ sol, int ld,
                                                  double *errors) {
  const int t = xcd_block() * 128 + threadIdx.x;
  if (t >= nt) return;
  double err[NS];
  {
    TetData D;
    load_full(t, tets, tet2edge, pts, attr, sig6, D);
    // volume term: h^2 sum_q |i w sigma E(q)|^2 jxw, 4-point rule (src/em_fe.cc:225-250)
    const double qa = 0.585410196624969, qb = 0.138196601125011;
    const double jxw = D.G.det / 24.0;
    const double h = tet_circumradius(D);
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int col = 2 * slots[s];
      double res = 0.0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        double lam[4] = {qb, qb, qb, qb};
        // reference point q has natural coordinate (1 - p0 - p1 - p2) on vertex 0
        lam[0] = q == 3 ? 1.0 - 3.0 * qb : 1.0 - qa - 2.0 * qb;
        if (q < 3) lam[q + 1] = qa;
        double er[3], ei[3], sr[3], si[3];
        field_value(D, lam, sol, ld, col, er, ei);
        sig_apply(D.sig, er, sr);
        sig_apply(D.sig, ei, si);
        res += omega * omega * (dot3(sr, sr) + dot3(si, si)) * jxw;
      }
      err[s] = res * h * h;
    }
  }
  // face terms
  for (int f = 0; f < 4; ++f) {
    const int n = neighbor[t * 4 + f];
    if (n < 0) continue;
    // evaluate from the lower-index side: P is the side whose face rule and normal
    // are used
    const bool own = t < n;
    TetData P, Q;
    load_full(own ? t : n, tets, tet2edge, pts, attr, sig6, P);
    load_full(own ? n : t, tets, tet2edge, pts, attr, sig6, Q);
    const int pf = own ? f : non[t * 4 + f];
    double nrm[3];
    {
      const double *g = P.G.g[pf];
      const double gl = sqrt(dot3(g, g));
      nrm[0] = -g[0] / gl;
      nrm[1] = -g[1] / gl;
      nrm[2] = -g[2] / gl;
    }
    // 3-point face rule (src/em_fe.cc:134-208): weights |grad lambda_f| det 6 / 6
    const double fj = sqrt(dot3(P.G.g[pf], P.G.g[pf])) * fabs(P.G.det);
    const double hf = face_circumradius(P, pf);
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int col = 2 * slots[s];
      double cpr[3], cpi[3], cqr[3], cqi[3];
      field_curl(P, sol, ld, col, cpr, cpi);
      field_curl(Q, sol, ld, col, cqr, cqi);
      double dr[3] = {cpr[0] - cqr[0], cpr[1] - cqr[1], cpr[2] - cqr[2]};
      double di[3] = {cpi[0] - cqi[0], cpi[1] - cqi[1], cpi[2] - cqi[2]};
      double xr[3], xi[3];
      cross3(dr, nrm, xr);
      cross3(di, nrm, xi);
      const double cross2 = dot3(xr, xr) + dot3(xi, xi);
      double dotj = 0.0;
      for (int q = 0; q < 3; ++q) {
        double L[4];
        L[pf] = 0.0;
        for (int j = 0; j < 3; ++j) L[(pf + j + 1) & 3] = (j == q) ? 2.0 / 3.0 : 1.0 / 6.0;
        double p[3];
        for (int k = 0; k < 3; ++k) p[k] = L[0] * P.x[0][k] + L[1] * P.x[1][k] + L[2] * P.x[2][k] + L[3] * P.x[3][k];
        double lp[4], lq[4];
        barycentric(P.x, P.G, p, lp);
        barycentric(Q.x, Q.G, p, lq);
        double er[3], ei[3], enr[3], eni[3], sr[3], si[3], snr[3], sni[3];
        field_value(P, lp, sol, ld, col, er, ei);
        field_value(Q, lq, sol, ld, col, enr, eni);
        sig_apply(P.sig, er, sr);
        sig_apply(P.sig, ei, si);
        sig_apply(Q.sig, enr, snr);
        sig_apply(Q.sig, eni, sni);
        const double jr = (sr[0] - snr[0]) * nrm[0] + (sr[1] - snr[1]) * nrm[1] + (sr[2] - snr[2]) * nrm[2];
        const double ji = (si[0] - sni[0]) * nrm[0] + (si[1] - sni[1]) * nrm[1] + (si[2] - sni[2]) * nrm[2];
        dotj += (jr * jr + ji * ji) * fj;
      }
      // curl jump is constant over the face; three points each weighted fj
      err[s] += (dotj + 3.0 * cross2 * fj) * hf / 2.0;
    }
  }
#pragma unroll
  for (int s = 0; s < NS; ++s) errors[(size_t)s * nt + t] = err[s];
}

}  // namespace

void assemble_colour(hipStream_t s, int ntets, const int *tet_ids, const int *tets, const int *tet2edge,
                     const double *pts, const int *attr, const double *sig6, const uint8_t *edge_bdr, double omega,
                     const int *rp, const int *ci, double2 *K) {
  if (ntets == 0) return;
  hipLaunchKernelGGL(k_assemble, dim3((ntets + kBlock - 1) / kBlock), dim3(kBlock), 0, s, ntets, tet_ids, tets,
                     tet2edge, pts, attr, sig6, edge_bdr, omega, rp, ci, K);
  EMF_HIP(hipGetLastError());
}

void estimate(hipStream_t s, int nt, const int *tets, const int *tet2edge, const int *neighbor, const int *non,
              const double *pts, const int *attr, const double *sig6, double omega, int ns, const int *slots,
              const double *sol, int ld, double *errors) {
  EMF_REQUIRE(ns <= kMaxSol, kErrArg, "estimator supports at most 8 solutions per call");
  if (nt == 0) return;
  switch (ns) {
#define EMF_EST(K)                                                                                                 \
  case K:                                                                                                          \
    hipLaunchKernelGGL((k_estimate<K>), dim3((nt + 127) / 128), dim3(128), 0, s, nt, tets, tet2edge, neighbor, non, \
                       pts, attr, sig6, omega, slots, sol, ld, errors);                                            \
    break;
    EMF_EST(1) EMF_EST(2) EMF_EST(3) EMF_EST(4) EMF_EST(5) EMF_EST(6) EMF_EST(7) EMF_EST(8)
#undef EMF_EST
    default: EMF_REQUIRE(false, kErrArg, "estimator supports 1 to 8 solutions per call");
  }
  EMF_HIP(hipGetLastError());
}

}  // namespace emfem
