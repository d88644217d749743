/*
 * CPU FP64 oracle for the EMFEM forward-modelling hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates the reference's algorithms in
 * plain C99 so that the HIP implementation can be checked against it.  It is
 * imported only by tests/, by __graft_entry__.smoke() and by bench.py's CPU
 * baseline leg; the product library never links or calls it.
 *
 * Every function cites the reference file:line it follows (paths relative to
 * the reference checkout).  Dense 4x4 algebra that the reference delegates to
 * Eigen is restated here directly (Gauss-Jordan inverse, cofactor determinant).
 *
 * Topology conventions (shared by every function here):
 *   tets[t*4+v]       tet -> vertex ids (TetGen tetrahedronlist)
 *   edgelist[e*2+j]   edge -> vertex ids, sorted ascending (src/em_mesh.cc:448-455)
 *   tet2edge[t*6+i]   tet -> edge ids (TetGen tet2edgelist); any local order is accepted
 *   neighbor[t*4+f]   tet across the face opposite vertex f, -1 on the boundary
 *   non[t*4+f]        that neighbour's face index pointing back (src/em_mesh.cc:463-481)
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_PI 3.14159265358979323846
#define OR_MU (4 * OR_PI * 1E-7)
#define OR_DTOR (OR_PI / 180.0)
#define OR_EPS 1.0E-6
#define OR_MAXLAY 256

typedef double complex zdouble;

/* ------------------------------------------------------------------------- */
/* Topology                                                                  */
/* ------------------------------------------------------------------------- */

/* Local edge order used when the oracle derives edges itself. */
static const int LOCAL_EDGE[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};

static int cmp_u64(const void *a, const void *b) {
  uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

static int find_u64(const uint64_t *keys, int n, uint64_t k) {
  int lo = 0, hi = n - 1;
  while (lo <= hi) {
    int mid = (lo + hi) / 2;
    if (keys[mid] == k) return mid;
    if (keys[mid] < k) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

/* Edges of the mesh: unique vertex pairs numbered in lexicographic order of
 * (lower id, higher id); each stored ascending as the reference orients them
 * (src/em_mesh.cc:448-455).  TetGen's own numbering is not reproducible offline,
 * so this rule is the documented replacement.  Returns the edge count. */
int or_build_edges(int nt, const int *tets, int *edgelist, int *tet2edge) {
  int t, i, n = 0;
  uint64_t *keys = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)nt * 6);
  for (t = 0; t < nt; ++t)
    for (i = 0; i < 6; ++i) {
      uint64_t a = (uint64_t)tets[t * 4 + LOCAL_EDGE[i][0]], b = (uint64_t)tets[t * 4 + LOCAL_EDGE[i][1]];
      if (a > b) { uint64_t s = a; a = b; b = s; }
      keys[(size_t)t * 6 + i] = (a << 32) | b;
    }
  uint64_t *uniq = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)nt * 6);
  memcpy(uniq, keys, sizeof(uint64_t) * (size_t)nt * 6);
  qsort(uniq, (size_t)nt * 6, sizeof(uint64_t), cmp_u64);
  for (i = 0; i < nt * 6; ++i)
    if (i == 0 || uniq[i] != uniq[i - 1]) uniq[n++] = uniq[i];
  for (i = 0; i < n; ++i) {
    edgelist[i * 2 + 0] = (int)(uniq[i] >> 32);
    edgelist[i * 2 + 1] = (int)(uniq[i] & 0xffffffffu);
  }
  for (t = 0; t < nt; ++t)
    for (i = 0; i < 6; ++i) tet2edge[t * 6 + i] = find_u64(uniq, n, keys[(size_t)t * 6 + i]);
  free(keys);
  free(uniq);
  return n;
}

typedef struct { int a, b, c, t, f; } face_rec;

static int cmp_face(const void *x, const void *y) {
  const face_rec *p = (const face_rec *)x, *q = (const face_rec *)y;
  if (p->a != q->a) return p->a < q->a ? -1 : 1;
  if (p->b != q->b) return p->b < q->b ? -1 : 1;
  if (p->c != q->c) return p->c < q->c ? -1 : 1;
  if (p->t != q->t) return p->t < q->t ? -1 : 1;
  return 0;
}

static void sort3(int *v) {
  int s;
  if (v[0] > v[1]) { s = v[0]; v[0] = v[1]; v[1] = s; }
  if (v[1] > v[2]) { s = v[1]; v[1] = v[2]; v[2] = s; }
  if (v[0] > v[1]) { s = v[0]; v[0] = v[1]; v[1] = s; }
}

/* Face neighbours: neighbor[t*4+f] is the tet sharing the face opposite vertex f
 * (TetGen neighborlist convention), and non[] the reverse face index
 * (src/em_mesh.cc:463-481). */
void or_neighbors(int nt, const int *tets, int *neighbor, int *non) {
  int t, f, i;
  face_rec *fr = (face_rec *)malloc(sizeof(face_rec) * (size_t)nt * 4);
  for (t = 0; t < nt; ++t)
    for (f = 0; f < 4; ++f) {
      int v[3];
      for (i = 0; i < 3; ++i) v[i] = tets[t * 4 + (f + i + 1) % 4];
      sort3(v);
      face_rec r = {v[0], v[1], v[2], t, f};
      fr[t * 4 + f] = r;
      neighbor[t * 4 + f] = -1;
      non[t * 4 + f] = -1;
    }
  qsort(fr, (size_t)nt * 4, sizeof(face_rec), cmp_face);
  for (i = 0; i + 1 < nt * 4; ++i) {
    face_rec *p = &fr[i], *q = &fr[i + 1];
    if (p->a == q->a && p->b == q->b && p->c == q->c) {
      neighbor[p->t * 4 + p->f] = q->t;
      neighbor[q->t * 4 + q->f] = p->t;
      non[p->t * 4 + p->f] = q->f;
      non[q->t * 4 + q->f] = p->f;
    }
  }
  free(fr);
}

/* Edge endpoint -> tet-local vertex (TetAccessor::edge_end_point, src/em_mesh.h:156-172). */
static int edge_end_point(const int *tv, const int *edgelist, int edge, int ep) {
  int v;
  for (v = 0; v < 4; ++v)
    if (edgelist[edge * 2 + ep] == tv[v]) return v;
  return -1;
}

/* Boundary markers: every edge and vertex of a face without a second tet
 * (src/em_mesh.cc:483-502). */
void or_boundary_markers(int nt, const int *tets, const int *neighbor, const int *tet2edge,
                         const int *edgelist, unsigned char *edge_bdr, unsigned char *vertex_bdr) {
  int t, f, e, i;
  for (t = 0; t < nt; ++t)
    for (f = 0; f < 4; ++f) {
      if (neighbor[t * 4 + f] >= 0) continue;
      for (i = 0; i < 3; ++i) vertex_bdr[tets[t * 4 + (f + i + 1) % 4]] = 1;
      /* the face's edges are the tet edges that do not touch vertex f */
      for (e = 0; e < 6; ++e) {
        int g = tet2edge[t * 6 + e];
        int a = edge_end_point(&tets[t * 4], edgelist, g, 0);
        int b = edge_end_point(&tets[t * 4], edgelist, g, 1);
        if (a != f && b != f) edge_bdr[g] = 1;
      }
    }
}

/* CSR sparsity of the edge matrices: row = edge, columns = sorted unique edges of
 * every tet that contains the row edge (src/em_fwd.cc:191-246).  Returns nnz, or
 * -1 when col_cap is too small. */
static int cmp_int(const void *a, const void *b) {
  int x = *(const int *)a, y = *(const int *)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

int or_make_pattern(int ne, int nt, const int *tet2edge, int *rowptr, int *col, int col_cap) {
  int t, i, j, e;
  int *cnt = (int *)calloc((size_t)ne + 1, sizeof(int));
  for (t = 0; t < nt; ++t)
    for (i = 0; i < 6; ++i) cnt[tet2edge[t * 6 + i] + 1] += 6;
  for (e = 0; e < ne; ++e) cnt[e + 1] += cnt[e];
  int *buf = (int *)malloc(sizeof(int) * (size_t)cnt[ne]);
  int *fill = (int *)malloc(sizeof(int) * (size_t)ne);
  for (e = 0; e < ne; ++e) fill[e] = cnt[e];
  for (t = 0; t < nt; ++t)
    for (i = 0; i < 6; ++i) {
      int r = tet2edge[t * 6 + i];
      for (j = 0; j < 6; ++j) buf[fill[r]++] = tet2edge[t * 6 + j];
    }
  int nnz = 0;
  for (e = 0; e < ne; ++e) {
    int *row = buf + cnt[e], n = cnt[e + 1] - cnt[e], m = 0;
    qsort(row, (size_t)n, sizeof(int), cmp_int);
    rowptr[e] = nnz;
    for (i = 0; i < n; ++i)
      if (i == 0 || row[i] != row[i - 1]) {
        if (nnz >= col_cap) { free(cnt); free(buf); free(fill); return -1; }
        col[nnz++] = row[i];
        ++m;
      }
    (void)m;
  }
  rowptr[ne] = nnz;
  free(cnt);
  free(buf);
  free(fill);
  return nnz;
}

static int csr_find(const int *rowptr, const int *col, int r, int c) {
  int lo = rowptr[r], hi = rowptr[r + 1] - 1;
  while (lo <= hi) {
    int mid = (lo + hi) / 2;
    if (col[mid] == c) return mid;
    if (col[mid] < c) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

/* ------------------------------------------------------------------------- */
/* Conductivity tensor (src/em_fwd.cc:14-72)                                  */
/* ------------------------------------------------------------------------- */

static void mat3_mul(const double *a, const double *b, double *c) {
  int i, j, k;
  double r[9];
  for (i = 0; i < 3; ++i)
    for (j = 0; j < 3; ++j) {
      r[i * 3 + j] = 0;
      for (k = 0; k < 3; ++k) r[i * 3 + j] += a[i * 3 + k] * b[k * 3 + j];
    }
  memcpy(c, r, sizeof(r));
}

static void mat3_t(const double *a, double *b) {
  int i, j;
  for (i = 0; i < 3; ++i)
    for (j = 0; j < 3; ++j) b[j * 3 + i] = a[i * 3 + j];
}

/* sigma (row-major 3x3) for one resistivity row; returns -1 for an unknown form. */
int or_sigma(int aniso_form, const double *r, double *sigma) {
  memset(sigma, 0, sizeof(double) * 9);
  switch (aniso_form) {
  case 0:
    sigma[0] = sigma[4] = sigma[8] = 1.0 / r[0];
    return 0;
  case 1:
    sigma[0] = sigma[4] = 1.0 / r[0];
    sigma[8] = 1.0 / r[1];
    return 0;
  case 2:
    sigma[0] = 1.0 / r[0]; sigma[4] = 1.0 / r[1]; sigma[8] = 1.0 / r[2];
    return 0;
  case 3: {
    double xr[9] = {0}, yr[9] = {0}, zr[9] = {0}, t[9], tt[9];
    sigma[0] = 1.0 / r[0]; sigma[4] = 1.0 / r[1]; sigma[8] = 1.0 / r[2];
    xr[0] = 1.0;
    xr[4] = xr[8] = cos(r[3] * OR_DTOR);
    xr[5] = sin(r[3] * OR_DTOR);
    xr[7] = -sin(r[3] * OR_DTOR);
    yr[4] = 1.0;
    yr[0] = yr[8] = cos(r[4] * OR_DTOR);
    yr[2] = -sin(r[4] * OR_DTOR);
    yr[6] = sin(r[4] * OR_DTOR);
    zr[8] = 1.0;
    zr[0] = zr[4] = cos(r[5] * OR_DTOR);
    zr[1] = sin(r[5] * OR_DTOR);
    zr[3] = -sin(r[5] * OR_DTOR);
    /* sigma = Rz Ry Rx S Rx^T Ry^T Rz^T */
    mat3_mul(zr, yr, t);
    mat3_mul(t, xr, t);
    mat3_mul(t, sigma, tt);
    mat3_t(xr, t); mat3_mul(tt, t, tt);
    mat3_t(yr, t); mat3_mul(tt, t, tt);
    mat3_t(zr, t); mat3_mul(tt, t, tt);
    memcpy(sigma, tt, sizeof(tt));
    return 0;
  }
  default:
    return -1;
  }
}

static int rho_cols(int form) {
  static const int c[4] = {1, 2, 3, 6};
  return (form >= 0 && form < 4) ? c[form] : 0;
}

/* ------------------------------------------------------------------------- */
/* FEValues (src/em_fe.cc:284-413) and quadrature (src/em_fe.cc:134-282)      */
/* ------------------------------------------------------------------------- */

typedef struct {
  double x[4][3];          /* vertex coordinates */
  double coef[4][4];       /* inverse of [1; x_v]: coef[n] = (a, b, c, d) of lambda_n */
  double det;              /* det([1; x_v]) = 6 V */
  int ev[6][2];            /* edge -> local vertices (src/em_fe.cc:366-378) */
  int fe[4][3];            /* face -> local edges (src/em_fe.cc:380-395) */
  int edge[6];             /* global edge ids */
} fe_values;

static double det4(double m[4][4]) {
  /* cofactor expansion along the first row */
  double d = 0;
  int c;
  for (c = 0; c < 4; ++c) {
    double s[3][3];
    int i, j, k;
    for (i = 1; i < 4; ++i)
      for (j = 0, k = 0; j < 4; ++j)
        if (j != c) s[i - 1][k++] = m[i][j];
    double m3 = s[0][0] * (s[1][1] * s[2][2] - s[1][2] * s[2][1]) -
                s[0][1] * (s[1][0] * s[2][2] - s[1][2] * s[2][0]) +
                s[0][2] * (s[1][0] * s[2][1] - s[1][1] * s[2][0]);
    d += ((c & 1) ? -1.0 : 1.0) * m[0][c] * m3;
  }
  return d;
}

static void inv4(double m[4][4], double out[4][4]) {
  double a[4][8];
  int i, j, k;
  for (i = 0; i < 4; ++i)
    for (j = 0; j < 4; ++j) {
      a[i][j] = m[i][j];
      a[i][j + 4] = (i == j) ? 1.0 : 0.0;
    }
  for (k = 0; k < 4; ++k) {
    int p = k;
    for (i = k + 1; i < 4; ++i)
      if (fabs(a[i][k]) > fabs(a[p][k])) p = i;
    if (p != k)
      for (j = 0; j < 8; ++j) { double s = a[k][j]; a[k][j] = a[p][j]; a[p][j] = s; }
    double piv = a[k][k];
    for (j = 0; j < 8; ++j) a[k][j] /= piv;
    for (i = 0; i < 4; ++i)
      if (i != k) {
        double f = a[i][k];
        for (j = 0; j < 8; ++j) a[i][j] -= f * a[k][j];
      }
  }
  for (i = 0; i < 4; ++i)
    for (j = 0; j < 4; ++j) out[i][j] = a[i][j + 4];
}

static void fe_reinit(fe_values *fv, const double *pts, const int *tets, const int *tet2edge,
                      const int *edgelist, int t) {
  int i, j, k;
  double jac[4][4], inv[4][4];
  const int *tv = &tets[t * 4];
  for (i = 0; i < 4; ++i)
    for (j = 0; j < 3; ++j) fv->x[i][j] = pts[tv[i] * 3 + j];
  for (i = 0; i < 4; ++i) {
    jac[0][i] = 1.0;
    for (j = 0; j < 3; ++j) jac[j + 1][i] = fv->x[i][j];
  }
  fv->det = det4(jac);
  inv4(jac, inv);
  /* a_ = col 0, b_ = col 1, ... of the inverse (src/em_fe.cc:398-413) */
  for (i = 0; i < 4; ++i)
    for (j = 0; j < 4; ++j) fv->coef[i][j] = inv[i][j];
  for (i = 0; i < 6; ++i) {
    fv->edge[i] = tet2edge[t * 6 + i];
    for (j = 0; j < 2; ++j) fv->ev[i][j] = edge_end_point(tv, edgelist, fv->edge[i], j);
  }
  for (i = 0; i < 4; ++i)
    for (j = 0; j < 3; ++j) {
      int va = (i + j + 1) % 4, vb = (i + (j + 1) % 3 + 1) % 4;
      for (k = 0; k < 6; ++k)
        if ((va == fv->ev[k][0] && vb == fv->ev[k][1]) || (va == fv->ev[k][1] && vb == fv->ev[k][0]))
          fv->fe[i][j] = k;
    }
}

static double value_q(const fe_values *fv, int n, const double *p) {
  return fv->coef[n][0] + p[0] * fv->coef[n][1] + p[1] * fv->coef[n][2] + p[2] * fv->coef[n][3];
}

static void grad_q(const fe_values *fv, int n, double *g) {
  g[0] = fv->coef[n][1]; g[1] = fv->coef[n][2]; g[2] = fv->coef[n][3];
}

/* N_e = lambda_a grad lambda_b - lambda_b grad lambda_a (src/em_fe.cc:304-308) */
static void value_nedelec(const fe_values *fv, int e, const double *p, double *v) {
  double ga[3], gb[3];
  double la = value_q(fv, fv->ev[e][0], p), lb = value_q(fv, fv->ev[e][1], p);
  grad_q(fv, fv->ev[e][0], ga);
  grad_q(fv, fv->ev[e][1], gb);
  int i;
  for (i = 0; i < 3; ++i) v[i] = la * gb[i] - lb * ga[i];
}

/* curl N_e = 2 grad lambda_a x grad lambda_b (src/em_fe.cc:310-313) */
static void curl_nedelec(const fe_values *fv, int e, double *c) {
  double ga[3], gb[3];
  grad_q(fv, fv->ev[e][0], ga);
  grad_q(fv, fv->ev[e][1], gb);
  c[0] = 2.0 * (ga[1] * gb[2] - ga[2] * gb[1]);
  c[1] = 2.0 * (ga[2] * gb[0] - ga[0] * gb[2]);
  c[2] = 2.0 * (ga[0] * gb[1] - ga[1] * gb[0]);
}

/* outward normal -grad lambda_f / |grad lambda_f| (src/em_fe.cc:335-338) */
static void face_normal(const fe_values *fv, int f, double *n) {
  double g[3];
  grad_q(fv, f, g);
  double s = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
  n[0] = -g[0] / s; n[1] = -g[1] / s; n[2] = -g[2] / s;
}

/* 4-point degree-2 cell rule (src/em_fe.cc:210-256) mapped to the tet
 * (src/em_fe.cc:258-282): qp[q] = sum_v L_v x_v with L = (1-p0-p1-p2, p0, p1, p2). */
static void cell_quadrature(const fe_values *fv, double qp[4][3], double jxw[4]) {
  const double a = 0.585410196624969, b = 0.138196601125011;
  const double ref[4][3] = {{a, b, b}, {b, a, b}, {b, b, a}, {b, b, b}};
  int q, i, j;
  for (q = 0; q < 4; ++q) {
    double L[4] = {1 - ref[q][0] - ref[q][1] - ref[q][2], ref[q][0], ref[q][1], ref[q][2]};
    for (j = 0; j < 3; ++j) {
      qp[q][j] = 0;
      for (i = 0; i < 4; ++i) qp[q][j] += fv->x[i][j] * L[i];
    }
    jxw[q] = fv->det * (1.0 / 24.0);
  }
}

/* 3-point face rule on face f (src/em_fe.cc:134-208): natural coordinate f is 0 and
 * coordinate (f+j+1)%4 = p_j; jxw = |grad lambda_f| * det * 6 * (1/6). */
static void face_quadrature(const fe_values *fv, int f, double qp[3][3], double jxw[3]) {
  const double ref[3][3] = {{2.0 / 3, 1.0 / 6, 1.0 / 6}, {1.0 / 6, 2.0 / 3, 1.0 / 6}, {1.0 / 6, 1.0 / 6, 2.0 / 3}};
  double g[3];
  int q, i, j;
  grad_q(fv, f, g);
  double s = sqrt(pow(g[0] * fv->det * 6, 2) + pow(g[1] * fv->det * 6, 2) + pow(g[2] * fv->det * 6, 2));
  for (q = 0; q < 3; ++q) {
    double L[4];
    L[f] = 0.0;
    for (j = 0; j < 3; ++j) L[(f + j + 1) % 4] = ref[q][j];
    for (j = 0; j < 3; ++j) {
      qp[q][j] = 0;
      for (i = 0; i < 4; ++i) qp[q][j] += fv->x[i][j] * L[i];
    }
    jxw[q] = s * (1.0 / 6.0);
  }
}

/* Raw 6x6 element matrices C (1/mu curl.curl) and M (omega N.sigma N), lower
 * triangle computed and mirrored (src/em_fwd.cc:354-375). */
static void element_matrices(const fe_values *fv, const double *sigma, double omega, double C[6][6],
                             double M[6][6]) {
  double qp[4][3], jxw[4];
  int i, j, q, k;
  cell_quadrature(fv, qp, jxw);
  for (i = 0; i < 6; ++i)
    for (j = 0; j < 6; ++j) C[i][j] = M[i][j] = 0.0;
  for (i = 0; i < 6; ++i)
    for (j = 0; j <= i; ++j) {
      for (q = 0; q < 4; ++q) {
        double ci[3], cj[3], vi[3], vj[3], sv[3];
        curl_nedelec(fv, i, ci);
        curl_nedelec(fv, j, cj);
        C[i][j] += (1.0 / OR_MU) * (ci[0] * cj[0] + ci[1] * cj[1] + ci[2] * cj[2]) * jxw[q];
        value_nedelec(fv, i, qp[q], vi);
        value_nedelec(fv, j, qp[q], vj);
        for (k = 0; k < 3; ++k) sv[k] = sigma[k * 3 + 0] * vj[0] + sigma[k * 3 + 1] * vj[1] + sigma[k * 3 + 2] * vj[2];
        M[i][j] += omega * (vi[0] * sv[0] + vi[1] * sv[1] + vi[2] * sv[2]) * jxw[q];
      }
      if (i != j) {
        C[j][i] = C[i][j];
        M[j][i] = M[i][j];
      }
    }
}

/* Per-tet element matrices exported for tests (C then M, 36 each, row-major). */
int or_element_matrices(const double *pts, const int *tets, const int *tet2edge, const int *edgelist,
                        int t, const double *sigma, double omega, double *C, double *M) {
  fe_values fv;
  double c[6][6], m[6][6];
  fe_reinit(&fv, pts, tets, tet2edge, edgelist, t);
  element_matrices(&fv, sigma, omega, c, m);
  memcpy(C, c, sizeof(c));
  memcpy(M, m, sizeof(m));
  return 0;
}

/* Barycentric gradients and the signed volume (for tests). */
void or_tet_geometry(const double *pts, const int *tets, int t, double *coef16, double *det) {
  fe_values fv;
  int dummy_t2e[6] = {0, 0, 0, 0, 0, 0}, dummy_el[2] = {tets[t * 4], tets[t * 4 + 1]};
  (void)dummy_t2e; (void)dummy_el;
  int i, j;
  double jac[4][4], inv[4][4];
  for (i = 0; i < 4; ++i)
    for (j = 0; j < 3; ++j) fv.x[i][j] = pts[tets[t * 4 + i] * 3 + j];
  for (i = 0; i < 4; ++i) {
    jac[0][i] = 1.0;
    for (j = 0; j < 3; ++j) jac[j + 1][i] = fv.x[i][j];
  }
  *det = det4(jac);
  inv4(jac, inv);
  for (i = 0; i < 4; ++i)
    for (j = 0; j < 4; ++j) coef16[i * 4 + j] = inv[i][j];
}

/* ------------------------------------------------------------------------- */
/* Assembly (src/em_fwd.cc:318-400)                                          */
/* ------------------------------------------------------------------------- */

/* Assembles C and M on the given pattern with the reference's Dirichlet rule:
 * for a boundary edge i of a tet, zero C(i,j) and C(j,i) for j != i and the
 * whole M row/column i (src/em_fwd.cc:377-388).  bdr_cell[t] is set for every
 * tet with a boundary edge.  Returns -1 on an unknown anisotropy form. */
int or_assemble_matrix(int nt, const double *pts, const int *tets, const int *tet2edge, const int *edgelist,
                       const unsigned char *edge_bdr, const int *attr, int aniso_form, const double *rho,
                       double freq, const int *rowptr, const int *col, double *Cv, double *Mv,
                       unsigned char *bdr_cell) {
  int t, i, j;
  double omega = 2 * OR_PI * freq;
  int nc = rho_cols(aniso_form);
  int nnz = rowptr[0];
  (void)nnz;
  for (t = 0; t < nt; ++t) {
    fe_values fv;
    double sigma[9], C[6][6], M[6][6];
    if (or_sigma(aniso_form, &rho[attr[t] * nc], sigma) != 0) return -1;
    fe_reinit(&fv, pts, tets, tet2edge, edgelist, t);
    element_matrices(&fv, sigma, omega, C, M);
    bdr_cell[t] = 0;
    for (i = 0; i < 6; ++i)
      if (edge_bdr[fv.edge[i]]) {
        for (j = 0; j < 6; ++j) {
          if (j != i) C[i][j] = C[j][i] = 0.0;
          M[i][j] = M[j][i] = 0.0;
        }
        bdr_cell[t] = 1;
      }
    for (i = 0; i < 6; ++i)
      for (j = 0; j < 6; ++j) {
        int p = csr_find(rowptr, col, fv.edge[i], fv.edge[j]);
        Cv[p] += C[i][j];
        Mv[p] += M[i][j];
      }
  }
  return 0;
}

/* or_assemble_matrix over nthreads OpenMP threads (the CPU baseline's timing only):
 * element matrices in parallel, scatter-adds as atomic updates, so the sums agree with
 * the serial loop to rounding but not bit for bit.  C, M must be zeroed. */
int or_assemble_matrix_mt(int nt, const double *pts, const int *tets, const int *tet2edge, const int *edgelist,
                          const unsigned char *edge_bdr, const int *attr, int aniso_form, const double *rho,
                          double freq, const int *rowptr, const int *col, double *Cv, double *Mv,
                          unsigned char *bdr_cell, int nthreads) {
  double omega = 2 * OR_PI * freq;
  int nc = rho_cols(aniso_form), bad = 0, t;
#pragma omp parallel for schedule(static) num_threads(nthreads) reduction(| : bad)
  for (t = 0; t < nt; ++t) {
    fe_values fv;
    double sigma[9], C[6][6], M[6][6];
    int i, j;
    if (or_sigma(aniso_form, &rho[attr[t] * nc], sigma) != 0) { bad = 1; continue; }
    fe_reinit(&fv, pts, tets, tet2edge, edgelist, t);
    element_matrices(&fv, sigma, omega, C, M);
    bdr_cell[t] = 0;
    for (i = 0; i < 6; ++i)
      if (edge_bdr[fv.edge[i]]) {
        for (j = 0; j < 6; ++j) {
          if (j != i) C[i][j] = C[j][i] = 0.0;
          M[i][j] = M[j][i] = 0.0;
        }
        bdr_cell[t] = 1;
      }
    for (i = 0; i < 6; ++i)
      for (j = 0; j < 6; ++j) {
        int p = csr_find(rowptr, col, fv.edge[i], fv.edge[j]);
#pragma omp atomic
        Cv[p] += C[i][j];
#pragma omp atomic
        Mv[p] += M[i][j];
      }
  }
  return bad ? -1 : 0;
}

/* ------------------------------------------------------------------------- */
/* Point location (src/em_mesh.cc:504-566, src/em_mesh.h:208-230)             */
/* ------------------------------------------------------------------------- */

/* orient3d(a,b,c,d) = det[a-d; b-d; c-d] evaluated in double.  The reference uses
 * TetGen's robust predicate; with whole-metre coordinates every product here is
 * an integer below 2**53, so this evaluation is exact for the synthetic meshes. */
static double orient3d(const double *a, const double *b, const double *c, const double *d) {
  double adx = a[0] - d[0], bdx = b[0] - d[0], cdx = c[0] - d[0];
  double ady = a[1] - d[1], bdy = b[1] - d[1], cdy = c[1] - d[1];
  double adz = a[2] - d[2], bdz = b[2] - d[2], cdz = c[2] - d[2];
  return adx * (bdy * cdz - bdz * cdy) + bdx * (cdy * adz - cdz * ady) + cdx * (ady * bdz - adz * bdy);
}

int or_contains_point(const double *pts, const int *tets, int t, const double *p) {
  const double *v0 = &pts[tets[t * 4 + 0] * 3], *v1 = &pts[tets[t * 4 + 1] * 3];
  const double *v2 = &pts[tets[t * 4 + 2] * 3], *v3 = &pts[tets[t * 4 + 3] * 3];
  double d[4];
  d[0] = orient3d(v0, v1, v3, p);
  d[1] = orient3d(v1, v2, v3, p);
  d[2] = orient3d(v1, v0, v2, p);
  d[3] = orient3d(v0, v3, v2, p);
  return d[0] >= 0.0 && d[1] >= 0.0 && d[2] >= 0.0 && d[3] >= 0.0;
}

/* vertex_to_tet: the last tet (in index order) listing each vertex (src/em_mesh.cc:510-515). */
void or_vertex_to_tet(int nt, const int *tets, int *v2t) {
  int t, v;
  for (t = 0; t < nt; ++t)
    for (v = 0; v < 4; ++v) v2t[tets[t * 4 + v]] = t;
}

/* Nearest vertex by brute force (the reference asks a KD-tree; ties go to the lowest id). */
int or_nearest_vertex(int nv, const double *pts, const double *p) {
  int v, best = 0;
  double bd = 1e300;
  for (v = 0; v < nv; ++v) {
    double dx = pts[v * 3] - p[0], dy = pts[v * 3 + 1] - p[1], dz = pts[v * 3 + 2] - p[2];
    double d = dx * dx + dy * dy + dz * dz; /* squared: the kd-tree's L2 metric, src/em_mesh.h:60-61 */
    if (d < bd) { bd = d; best = v; }
  }
  return best;
}

/* Breadth-first search from vertex_to_tet[nearest vertex] in neighbour order; the
 * first tet that contains p wins (src/em_mesh.cc:527-566).  Returns -1 if none. */
int or_find_cell(int nv, const double *pts, int nt, const int *tets, const int *neighbor, const int *v2t,
                 const double *p) {
  int head = 0, tail = 0, found = -1, i;
  int *queue = (int *)malloc(sizeof(int) * ((size_t)nt * 4 + 1));
  unsigned char *touched = (unsigned char *)calloc((size_t)nt, 1);
  queue[tail++] = v2t[or_nearest_vertex(nv, pts, p)];
  while (head < tail) {
    int t = queue[head++];
    if (touched[t]) continue;
    touched[t] = 1;
    if (or_contains_point(pts, tets, t, p)) { found = t; break; }
    for (i = 0; i < 4; ++i) {
      int n = neighbor[t * 4 + i];
      if (n >= 0 && !touched[n]) queue[tail++] = n;
    }
  }
  free(queue);
  free(touched);
  return found;
}

/* ------------------------------------------------------------------------- */
/* 1-D plane wave (src/em_pw.h:6-139)                                        */
/* ------------------------------------------------------------------------- */

typedef struct {
  int nlay;
  double omega;
  double ztop[OR_MAXLAY], lsig[OR_MAXLAY];
  zdouble a[OR_MAXLAY + 1], b[OR_MAXLAY + 1], c[OR_MAXLAY + 1], d[OR_MAXLAY + 1];
} or_pw;

static void pw_init(or_pw *pw, int nlay, const double *ztop, const double *lsig, double freq) {
  int i;
  zdouble kk[OR_MAXLAY + 1], Rp[OR_MAXLAY + 1], expmgh[OR_MAXLAY + 1];
  const zdouble II = I;
  pw->nlay = nlay;
  for (i = 0; i < nlay; ++i) { pw->ztop[i] = ztop[i]; pw->lsig[i] = lsig[i]; }
  pw->omega = 2 * OR_PI * freq;
  for (i = 1; i <= nlay; ++i) kk[i] = csqrt(-II * pw->omega * OR_MU * lsig[i - 1]);
  kk[0] = csqrt(-II * pw->omega * OR_MU * lsig[0]);
  for (i = 0; i <= nlay; ++i) expmgh[i] = 0;
  for (i = 1; i <= nlay - 1; ++i) expmgh[i] = cexp(-kk[i] * (ztop[i] - ztop[i - 1]));
  for (i = 0; i <= nlay; ++i) Rp[i] = 0;
  for (i = nlay - 1; i >= 1; --i) {
    zdouble s = kk[i] + kk[i + 1];
    zdouble gmogp = (-II * pw->omega * OR_MU * (lsig[i - 1] - lsig[i])) / (s * s);
    zdouble rjexp = Rp[i + 1] * expmgh[i + 1];
    Rp[i] = (gmogp + rjexp) / (1.0 + gmogp * rjexp) * expmgh[i];
  }
  for (i = 0; i <= nlay; ++i) pw->a[i] = pw->b[i] = pw->c[i] = pw->d[i] = 0;
  pw->a[0] = 1.0;
  pw->c[0] = 1.0;
  for (i = 1; i <= nlay; ++i) {
    zdouble Atop = pw->a[i - 1] + pw->b[i - 1] * expmgh[i - 1];
    pw->b[i] = Atop / (1.0 + Rp[i] * expmgh[i]);
    pw->a[i] = pw->b[i] * Rp[i];
    Atop = pw->c[i - 1] + pw->d[i - 1] * expmgh[i - 1];
    pw->d[i] = Atop / (1.0 - Rp[i] * expmgh[i]);
    pw->c[i] = -pw->d[i] * Rp[i];
  }
}

static void pw_field(const or_pw *pw, double z, zdouble eh[4]) {
  int i, ilay = 0;
  const zdouble II = I;
  for (i = pw->nlay - 1; i >= 0; --i)
    if (z >= pw->ztop[i]) { ilay = i; break; }
  zdouble kk = csqrt(-II * pw->omega * OR_MU * pw->lsig[ilay]);
  zdouble expp = 0, expm, exppdz = 0, expmdz;
  expm = cexp(-kk * (z - pw->ztop[ilay]));
  expmdz = -kk * expm;
  if (ilay != pw->nlay - 1) {
    expp = cexp(kk * (z - pw->ztop[ilay + 1]));
    exppdz = kk * expp;
  }
  eh[0] = pw->a[ilay + 1] * expp + pw->b[ilay + 1] * expm;
  eh[1] = (pw->a[ilay + 1] * exppdz + pw->b[ilay + 1] * expmdz) / (II * OR_MU * pw->omega);
  eh[2] = (pw->c[ilay + 1] * exppdz + pw->d[ilay + 1] * expmdz) / pw->lsig[ilay];
  eh[3] = pw->c[ilay + 1] * expp + pw->d[ilay + 1] * expm;
  for (i = 0; i < 4; ++i)
    if (cabs(eh[i]) < 1E-80) eh[i] = 0;
}

/* E of the 1-D solution: XY (mode -1) -> conj(eh[0]) in x; YX (mode -2) -> conj(eh[2]) in y
 * (src/em_pw.h:108-119).  H likewise (src/em_pw.h:121-132). */
void or_pw_eval(int nlay, const double *ztop, const double *lsig, double freq, double z, int mode,
                double *e_ri, double *h_ri) {
  or_pw pw;
  zdouble eh[4], e[3] = {0, 0, 0}, h[3] = {0, 0, 0};
  int i;
  pw_init(&pw, nlay, ztop, lsig, freq);
  pw_field(&pw, z, eh);
  if (mode == -1) { e[0] = conj(eh[0]); h[1] = conj(eh[1]); }
  else if (mode == -2) { e[1] = conj(eh[2]); h[0] = conj(eh[3]); }
  for (i = 0; i < 3; ++i) {
    e_ri[i * 2] = creal(e[i]); e_ri[i * 2 + 1] = cimag(e[i]);
    h_ri[i * 2] = creal(h[i]); h_ri[i * 2 + 1] = cimag(h[i]);
  }
}

/* ------------------------------------------------------------------------- */
/* Background model from the 4 top corners (src/em_fwd.cc:74-159)             */
/* ------------------------------------------------------------------------- */

typedef struct {
  double corners[4][3];
  int nlay[4];
  double ztop[4][OR_MAXLAY], lsig[4][OR_MAXLAY];
} or_background;

/* Walks down the vertical line through p, recording (z, mean diagonal sigma) at every
 * vertex, then merges equal neighbours into layers (src/em_fwd.cc:74-133). */
static int layered_model(int nv, const double *pts, int nt, const int *tets, const int *neighbor, const int *v2t,
                         const int *attr, int aniso_form, const double *rho, const double *p, const double *br,
                         int *nlay, double *ztop, double *lsig) {
  double fp[3] = {p[0], p[1], p[2]};
  int nc = rho_cols(aniso_form);
  int cap = 1 << 16, n = 0, i;
  double *z = (double *)malloc(sizeof(double) * cap), *sg = (double *)malloc(sizeof(double) * cap);
  while (fp[2] < br[2]) {
    int v = or_nearest_vertex(nv, pts, fp);
    double q[3] = {fp[0], fp[1], fp[2] + OR_EPS};
    int t = or_find_cell(nv, pts, nt, tets, neighbor, v2t, q);
    double sigma[9];
    if (t < 0 || n >= cap) { free(z); free(sg); return -1; }
    or_sigma(aniso_form, &rho[attr[t] * nc], sigma);
    z[n] = fp[2];
    sg[n] = (sigma[0] + sigma[4] + sigma[8]) / 3.0;
    ++n;
    for (i = 0; i < 4; ++i) {
      int vi = tets[t * 4 + i];
      const double *x = &pts[vi * 3];
      if (vi != v && fabs(x[0] - p[0]) < OR_EPS && fabs(x[1] - p[1]) < OR_EPS) {
        fp[0] = x[0]; fp[1] = x[1]; fp[2] = x[2];
        break;
      }
    }
    if (i == 4) { free(z); free(sg); return -1; }
  }
  int m = 1;
  ztop[0] = z[0];
  lsig[0] = sg[0];
  for (i = 0; i < n - 1; ++i)
    if (sg[i] != sg[i + 1]) {
      if (m >= OR_MAXLAY) { free(z); free(sg); return -1; }
      ztop[m] = z[i + 1];
      lsig[m] = sg[i + 1];
      ++m;
    }
  *nlay = m;
  free(z);
  free(sg);
  return 0;
}

/* update_background_model (src/em_fwd.cc:135-159).  Output layout:
 * corners[12], nlay[4], ztop[4*OR_MAXLAY], lsig[4*OR_MAXLAY]. */
int or_background_model(int nv, const double *pts, int nt, const int *tets, const int *neighbor, const int *attr,
                        int aniso_form, const double *rho, double *corners, int *nlay, double *ztop, double *lsig) {
  int *v2t = (int *)malloc(sizeof(int) * (size_t)nv);
  double p[3] = {-1.0E+15, -1.0E+15, -1.0E+15};
  double tl[3], br[3];
  int k, rc = 0;
  or_vertex_to_tet(nt, tets, v2t);
  int vtl = or_nearest_vertex(nv, pts, p);
  p[0] = p[1] = p[2] = 1.0E+15;
  int vbr = or_nearest_vertex(nv, pts, p);
  for (k = 0; k < 3; ++k) { tl[k] = pts[vtl * 3 + k]; br[k] = pts[vbr * 3 + k]; }
  double c[4][3] = {{tl[0], tl[1], tl[2]}, {br[0], tl[1], tl[2]}, {br[0], br[1], tl[2]}, {tl[0], br[1], tl[2]}};
  for (k = 0; k < 4 && rc == 0; ++k) {
    memcpy(&corners[k * 3], c[k], sizeof(double) * 3);
    rc = layered_model(nv, pts, nt, tets, neighbor, v2t, attr, aniso_form, rho, c[k], br, &nlay[k],
                       &ztop[k * OR_MAXLAY], &lsig[k * OR_MAXLAY]);
  }
  free(v2t);
  return rc;
}

int or_max_layers(void) { return OR_MAXLAY; }

/* ------------------------------------------------------------------------- */
/* MT right-hand side (src/em_fwd.cc:402-580)                                 */
/* ------------------------------------------------------------------------- */

/* Weights of the four corner columns at the edge midpoint ec (src/em_fwd.cc:438-471): a
 * cosine taper between the two corners of a side face (x = x_min, y = y_min, x = x_max,
 * y = y_max, tested in that order), bilinear in (x, y) elsewhere (top and bottom faces).
 * corners[4][3] = (x_min, y_min), (x_max, y_min), (x_max, y_max), (x_min, y_max). */
void or_blend_weights(const double *corners, const double *ec, double *coeffs) {
  const double(*tc)[3] = (const double(*)[3])corners;
  if (fabs(ec[0] - tc[0][0]) < OR_EPS) {
    double s = (1.0 - cos(OR_PI * (ec[1] - tc[0][1]) / (tc[2][1] - tc[0][1]))) / 2.0;
    coeffs[0] = 1.0 - s; coeffs[1] = 0.0; coeffs[2] = 0.0; coeffs[3] = s;
  } else if (fabs(ec[1] - tc[0][1]) < OR_EPS) {
    double s = (1.0 - cos(OR_PI * (ec[0] - tc[0][0]) / (tc[2][0] - tc[0][0]))) / 2.0;
    coeffs[0] = 1.0 - s; coeffs[1] = s; coeffs[2] = 0.0; coeffs[3] = 0.0;
  } else if (fabs(ec[0] - tc[2][0]) < OR_EPS) {
    double s = (1.0 - cos(OR_PI * (ec[1] - tc[0][1]) / (tc[2][1] - tc[0][1]))) / 2.0;
    coeffs[0] = 0.0; coeffs[1] = 1.0 - s; coeffs[2] = s; coeffs[3] = 0.0;
  } else if (fabs(ec[1] - tc[2][1]) < OR_EPS) {
    double s = (1.0 - cos(OR_PI * (ec[0] - tc[0][0]) / (tc[2][0] - tc[0][0]))) / 2.0;
    coeffs[0] = 0.0; coeffs[1] = 0.0; coeffs[2] = s; coeffs[3] = 1.0 - s;
  } else {
    double area = (tc[2][0] - tc[0][0]) * (tc[2][1] - tc[0][1]);
    coeffs[0] = ((tc[2][0] - ec[0]) * (tc[2][1] - ec[1])) / area;
    coeffs[1] = ((ec[0] - tc[3][0]) * (tc[3][1] - ec[1])) / area;
    coeffs[2] = ((ec[0] - tc[0][0]) * (ec[1] - tc[0][1])) / area;
    coeffs[3] = ((tc[1][0] - ec[0]) * (ec[1] - tc[1][1])) / area;
  }
}

/* Boundary values: 1-D fields at the 4 corners blended over the edge midpoint, projected on
 * the edge (src/em_fwd.cc:402-486).  bval is (re, im) per edge; only boundary edges are set.
 * The first bdr cell (ascending tet index) that holds an edge defines its value. */
int or_boundary_values(int nt, const double *pts, const int *tets, const int *tet2edge, const int *edgelist,
                       const unsigned char *edge_bdr, const double *corners, const int *nlay, const double *ztop,
                       const double *lsig, double freq, int mode, int ne, double *bval, unsigned char *have) {
  or_pw pw[4];
  int k, t, i;
  for (k = 0; k < 4; ++k) pw_init(&pw[k], nlay[k], &ztop[k * OR_MAXLAY], &lsig[k * OR_MAXLAY], freq);
  memset(have, 0, (size_t)ne);
  for (t = 0; t < nt; ++t) {
    int any = 0;
    for (i = 0; i < 6; ++i) any |= edge_bdr[tet2edge[t * 6 + i]];
    if (!any) continue;
    fe_values fv;
    fe_reinit(&fv, pts, tets, tet2edge, edgelist, t);
    for (i = 0; i < 6; ++i) {
      int g = fv.edge[i];
      if (!edge_bdr[g]) continue;
      double ec[3], ed[3], coeffs[4], N[3];
      const double *xa = fv.x[fv.ev[i][0]], *xb = fv.x[fv.ev[i][1]];
      for (k = 0; k < 3; ++k) { ec[k] = (xa[k] + xb[k]) / 2.0; ed[k] = xb[k] - xa[k]; }
      or_blend_weights(corners, ec, coeffs);
      zdouble ep[3] = {0, 0, 0};
      for (k = 0; k < 4; ++k) {
        zdouble eh[4], e[3] = {0, 0, 0};
        pw_field(&pw[k], ec[2], eh);
        if (mode == -1) e[0] = conj(eh[0]);
        else if (mode == -2) e[1] = conj(eh[2]);
        ep[0] += coeffs[k] * e[0]; ep[1] += coeffs[k] * e[1]; ep[2] += coeffs[k] * e[2];
      }
      value_nedelec(&fv, i, ec, N);
      zdouble dv = (ed[0] * ep[0] + ed[1] * ep[1] + ed[2] * ep[2]) / (ed[0] * N[0] + ed[1] * N[1] + ed[2] * N[2]);
      if (!have[g]) {
        if (cabs(dv) > 1E-13) { bval[g * 2] = creal(dv); bval[g * 2 + 1] = cimag(dv); }
        else { bval[g * 2] = 0.0; bval[g * 2 + 1] = 0.0; }
        have[g] = 1;
      }
    }
  }
  return 0;
}

/* Dirichlet lifting into (s_re, s_im) from the raw element matrices of the bdr cells
 * (src/em_fwd.cc:488-580).  s must be zeroed by the caller. */
int or_assemble_rhs_mt(int nt, const double *pts, const int *tets, const int *tet2edge, const int *edgelist,
                       const unsigned char *edge_bdr, const int *attr, int aniso_form, const double *rho,
                       double freq, const double *bval, double *s_re, double *s_im) {
  int t, i, j;
  double omega = 2 * OR_PI * freq;
  int nc = rho_cols(aniso_form);
  for (t = 0; t < nt; ++t) {
    int any = 0;
    for (i = 0; i < 6; ++i) any |= edge_bdr[tet2edge[t * 6 + i]];
    if (!any) continue;
    fe_values fv;
    double sigma[9], C[6][6], M[6][6], b_re[6] = {0}, b_im[6] = {0};
    if (or_sigma(aniso_form, &rho[attr[t] * nc], sigma) != 0) return -1;
    fe_reinit(&fv, pts, tets, tet2edge, edgelist, t);
    element_matrices(&fv, sigma, omega, C, M);
    for (i = 0; i < 6; ++i) {
      if (!edge_bdr[fv.edge[i]]) continue;
      double bv_re = bval[fv.edge[i] * 2], bv_im = bval[fv.edge[i] * 2 + 1];
      b_re[i] = bv_re * C[i][i];
      b_im[i] = -bv_im * C[i][i];
      for (j = 0; j < 6; ++j)
        if (!edge_bdr[fv.edge[j]]) {
          b_re[j] -= C[j][i] * bv_re;
          b_im[j] += M[j][i] * bv_re;
          b_re[j] += M[j][i] * bv_im;
          b_im[j] += C[j][i] * bv_im;
        }
    }
    for (i = 0; i < 6; ++i) {
      s_re[fv.edge[i]] += b_re[i];
      s_im[fv.edge[i]] += b_im[i];
    }
  }
  return 0;
}

/* ------------------------------------------------------------------------- */
/* CSEM and dual right-hand sides (src/em_fwd.cc:582-717)                     */
/* ------------------------------------------------------------------------- */

/* Splits a line source into point dipoles (src/em_fwd.cc:582-614).  out rows are
 * (cx, cy, cz, dx, dy, dz, current, weight); returns the dipole count. */
int or_divide_line_source(const double *tx, int n_div, double *out) {
  double dir[3], c[3] = {tx[0], tx[1], tx[2]};
  double current = tx[5], length = tx[6];
  int i, k;
  dir[0] = cos(tx[3] * OR_DTOR) * cos(tx[4] * OR_DTOR);
  dir[1] = sin(tx[3] * OR_DTOR) * cos(tx[4] * OR_DTOR);
  dir[2] = sin(tx[4] * OR_DTOR);
  if (fabs(tx[6]) <= 0.1) {
    for (k = 0; k < 3; ++k) { out[k] = c[k]; out[3 + k] = dir[k]; }
    out[6] = current;
    out[7] = 1.0;
    return 1;
  }
  double A[3], B[3];
  for (k = 0; k < 3; ++k) A[k] = c[k] - dir[k] * length / 2;
  for (i = 0; i < n_div; ++i) {
    for (k = 0; k < 3; ++k) B[k] = A[k] + dir[k] * length / n_div;
    for (k = 0; k < 3; ++k) { out[i * 8 + k] = (A[k] + B[k]) / 2; out[i * 8 + 3 + k] = dir[k]; }
    out[i * 8 + 6] = current;
    out[i * 8 + 7] = 1.0 / n_div;
    for (k = 0; k < 3; ++k) A[k] = B[k];
  }
  return n_div;
}

/* b_im[i] += omega I (N_i(c).d) w on the tet holding each dipole, b_re = 0
 * (src/em_fwd.cc:616-671).  Returns -1 if a dipole is outside the mesh. */
int or_assemble_rhs_csem(int nv, const double *pts, int nt, const int *tets, const int *tet2edge,
                         const int *edgelist, const int *neighbor, double freq, const double *tx, int n_div,
                         double *s_re, double *s_im) {
  double omega = 2 * OR_PI * freq;
  double *dip = (double *)malloc(sizeof(double) * 8 * (size_t)(n_div > 1 ? n_div : 1));
  int *v2t = (int *)malloc(sizeof(int) * (size_t)nv);
  int nd = or_divide_line_source(tx, n_div, dip), j, i;
  or_vertex_to_tet(nt, tets, v2t);
  for (j = 0; j < nd; ++j) {
    double *d = &dip[j * 8];
    int t = or_find_cell(nv, pts, nt, tets, neighbor, v2t, d);
    if (t < 0) { free(dip); free(v2t); return -1; }
    fe_values fv;
    fe_reinit(&fv, pts, tets, tet2edge, edgelist, t);
    for (i = 0; i < 6; ++i) {
      double N[3];
      value_nedelec(&fv, i, d, N);
      s_re[fv.edge[i]] += 0.0;
      s_im[fv.edge[i]] += omega * d[6] * (N[0] * d[3] + N[1] * d[4] + N[2] * d[5]) * d[7];
    }
  }
  free(dip);
  free(v2t);
  return 0;
}

/* b_re[j] = b_im[j] = sum over receivers of the component sum of N_j(rx)
 * (src/em_fwd.cc:673-717). */
int or_assemble_rhs_dual(int nv, const double *pts, int nt, const int *tets, const int *tet2edge,
                         const int *edgelist, const int *neighbor, int nrx, const double *rx, double *s_re,
                         double *s_im) {
  int *v2t = (int *)malloc(sizeof(int) * (size_t)nv);
  int r, i;
  or_vertex_to_tet(nt, tets, v2t);
  for (r = 0; r < nrx; ++r) {
    int t = or_find_cell(nv, pts, nt, tets, neighbor, v2t, &rx[r * 3]);
    if (t < 0) { free(v2t); return -1; }
    fe_values fv;
    fe_reinit(&fv, pts, tets, tet2edge, edgelist, t);
    for (i = 0; i < 6; ++i) {
      double N[3];
      value_nedelec(&fv, i, &rx[r * 3], N);
      s_re[fv.edge[i]] += N[0] + N[1] + N[2];
      s_im[fv.edge[i]] += N[0] + N[1] + N[2];
    }
  }
  free(v2t);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Receivers (src/em_fwd.cc:719-871)                                         */
/* ------------------------------------------------------------------------- */

/* E = sum e_i N_i(p), H = -sum e_i curl N_i / (i omega mu) (src/em_fwd.cc:719-744).
 * e is interleaved (re, im) per edge; es/hs are (re, im) x 3. */
void or_interpolate_fields(const double *pts, const int *tets, const int *tet2edge, const int *edgelist, int t,
                           const double *p, double freq, const double *e, double *es, double *hs) {
  fe_values fv;
  zdouble E[3] = {0, 0, 0}, H[3] = {0, 0, 0};
  const zdouble II = I;
  int i, k;
  fe_reinit(&fv, pts, tets, tet2edge, edgelist, t);
  for (i = 0; i < 6; ++i) {
    double N[3], Cn[3];
    zdouble v = e[fv.edge[i] * 2] + II * e[fv.edge[i] * 2 + 1];
    value_nedelec(&fv, i, p, N);
    curl_nedelec(&fv, i, Cn);
    for (k = 0; k < 3; ++k) {
      E[k] += v * N[k];
      H[k] += -1.0 / (II * 2.0 * OR_PI * freq * OR_MU) * v * Cn[k];
    }
  }
  for (k = 0; k < 3; ++k) {
    es[k * 2] = creal(E[k]); es[k * 2 + 1] = cimag(E[k]);
    hs[k * 2] = creal(H[k]); hs[k * 2 + 1] = cimag(H[k]);
  }
}

/* MT transfer functions from the two polarisations (src/em_fwd.cc:746-802).
 * Returns (re, im) of the requested code, or -1 for a code the reference does not
 * compute.  The apparent-resistivity/phase codes 212/222/312-342 fall through to an
 * error in the reference; they are answered here as log(rho_a) + i*phi with
 * rho_a = |Z|^2/(omega mu) (parity unpinned: no reference value exists). */
int or_mt_response(int otype, double freq, const double *exy, const double *hxy, const double *eyx,
                   const double *hyx, double *out) {
  const zdouble II = I;
  zdouble e_xy[3], h_xy[3], e_yx[3], h_yx[3], r;
  int k;
  for (k = 0; k < 3; ++k) {
    e_xy[k] = exy[k * 2] + II * exy[k * 2 + 1];
    h_xy[k] = hxy[k * 2] + II * hxy[k * 2 + 1];
    e_yx[k] = eyx[k * 2] + II * eyx[k * 2 + 1];
    h_yx[k] = hyx[k * 2] + II * hyx[k * 2 + 1];
  }
  zdouble det = h_xy[0] * h_yx[1] - h_yx[0] * h_xy[1];
  zdouble zxy = (e_yx[0] * h_xy[0] - e_xy[0] * h_yx[0]) / det;
  zdouble zyx = (e_xy[1] * h_yx[1] - e_yx[1] * h_xy[1]) / det;
  zdouble zxx = (e_xy[0] * h_yx[1] - e_yx[0] * h_xy[1]) / det;
  zdouble zyy = (e_yx[1] * h_xy[0] - e_xy[1] * h_yx[0]) / det;
  double omega = 2 * OR_PI * freq;
  switch (otype) {
  case 311: r = zxx; break;
  case 321: r = zxy; break;
  case 331: r = zyx; break;
  case 341: r = zyy; break;
  case 351: r = (h_xy[2] * h_yx[1] - h_yx[2] * h_xy[1]) / det; break;
  case 361: r = (h_yx[2] * h_xy[0] - h_xy[2] * h_yx[0]) / det; break;
  case 212: r = log(cabs(zxy) * cabs(zxy) / (omega * OR_MU)) + II * carg(zxy); break;
  case 222: r = log(cabs(zyx) * cabs(zyx) / (omega * OR_MU)) + II * carg(zyx); break;
  case 312: r = clog(zxx); break;
  case 322: r = clog(zxy); break;
  case 332: r = clog(zyx); break;
  case 342: r = clog(zyy); break;
  default: return -1;
  }
  out[0] = creal(r);
  out[1] = cimag(r);
  return 0;
}

/* CSEM outputs: E/H components (RI) or their complex log (AP) (src/em_fwd.cc:804-871). */
int or_csem_response(int otype, const double *es, const double *hs, double *out) {
  const zdouble II = I;
  int comp = otype / 10 % 10 - 1;   /* 1..6 -> Ex..Hz */
  int form = otype % 10, ap = form == 2;   /* 1 = RI, 2 = AP (src/em_defs.h:49-61) */
  zdouble v;
  if (otype < 111 || otype > 162 || comp < 0 || comp > 5 || (form != 1 && form != 2)) return -1;
  if (comp < 3) v = es[comp * 2] + II * es[comp * 2 + 1];
  else v = hs[(comp - 3) * 2] + II * hs[(comp - 3) * 2 + 1];
  if (ap) v = clog(v);
  out[0] = creal(v);
  out[1] = cimag(v);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Goal-oriented error estimator (src/em_fwd.cc:963-1180, src/em_mesh.h:232-274) */
/* ------------------------------------------------------------------------- */

static double dist3(const double *a, const double *b) {
  return sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]));
}

/* tet circumradius (TetAccessor::diameter, src/em_mesh.h:232-254) */
double or_tet_diameter(const double *pts, const int *tets, int t) {
  const double *v[4];
  int i;
  for (i = 0; i < 4; ++i) v[i] = &pts[tets[t * 4 + i] * 3];
  double a = dist3(v[1], v[2]), b = dist3(v[0], v[2]), c = dist3(v[0], v[1]);
  double aa = dist3(v[0], v[3]), bb = dist3(v[1], v[3]), cc = dist3(v[2], v[3]);
  double la = a * aa, lb = b * bb, lc = c * cc, s = 0.5 * (la + lb + lc);
  double jac[4][4];
  int j;
  for (i = 0; i < 4; ++i) {
    jac[0][i] = 1;
    for (j = 0; j < 3; ++j) jac[j + 1][i] = v[i][j];
  }
  double vol = det4(jac) / 6.0;
  return sqrt(s * (s - la) * (s - lb) * (s - lc)) / (6.0 * vol);
}

/* triangle circumradius of face f (TetAccessor::face_diameter, src/em_mesh.h:256-274) */
double or_face_diameter(const double *pts, const int *tets, int t, int f) {
  const double *v[3];
  int i;
  for (i = 0; i < 3; ++i) v[i] = &pts[tets[t * 4 + (f + i + 1) % 4] * 3];
  double a = dist3(v[1], v[2]), b = dist3(v[0], v[2]), c = dist3(v[0], v[1]);
  double area = 0.25 * sqrt((a + b + c) * (b + c - a) * (c + a - b) * (a + b - c));
  return a * b * c / (4.0 * area);
}

static void fn_value(const fe_values *fv, const double *p, const double *sol, zdouble *e) {
  int i, k;
  e[0] = e[1] = e[2] = 0;
  for (i = 0; i < 6; ++i) {
    double N[3];
    zdouble v = sol[fv->edge[i] * 2] + I * sol[fv->edge[i] * 2 + 1];
    value_nedelec(fv, i, p, N);
    for (k = 0; k < 3; ++k) e[k] += v * N[k];
  }
}

static void fn_curl(const fe_values *fv, const double *sol, zdouble *ce) {
  int i, k;
  ce[0] = ce[1] = ce[2] = 0;
  for (i = 0; i < 6; ++i) {
    double Cn[3];
    zdouble v = sol[fv->edge[i] * 2] + I * sol[fv->edge[i] * 2 + 1];
    curl_nedelec(fv, i, Cn);
    for (k = 0; k < 3; ++k) ce[k] += v * Cn[k];
  }
}

static void sig_mul(const double *s, const zdouble *e, zdouble *o) {
  int k;
  for (k = 0; k < 3; ++k) o[k] = s[k * 3] * e[0] + s[k * 3 + 1] * e[1] + s[k * 3 + 2] * e[2];
}

/* Per-tet squared indicators for ns solutions (sols[s] = interleaved re/im per edge):
 * errors[s*nt + t] = h_K^2 sum_q |i omega sigma E|^2 jxw + sum_f face_err
 * (integrate_over_cell / integrate_over_face / estimate_error,
 * src/em_fwd.cc:1005-1166).  face_err is computed once per interior face from the
 * lower-index tet (the first one the reference visits) and copied to both sides. */
int or_estimate_error(int nt, const double *pts, const int *tets, const int *tet2edge, const int *edgelist,
                      const int *neighbor, const int *non, const int *attr, int aniso_form, const double *rho,
                      double freq, int ns, const double *const *sols, double *errors) {
  double omega = 2 * OR_PI * freq;
  int nc = rho_cols(aniso_form);
  int t, f, s, q, k;
  double *face_err = (double *)malloc(sizeof(double) * (size_t)ns * nt * 4);
  for (k = 0; k < ns * nt * 4; ++k) face_err[k] = -1.0;
  for (k = 0; k < ns * nt; ++k) errors[k] = 0.0;
  for (t = 0; t < nt; ++t) {
    fe_values fv;
    double sigma[9], qp[4][3], jxw[4];
    if (or_sigma(aniso_form, &rho[attr[t] * nc], sigma) != 0) { free(face_err); return -1; }
    fe_reinit(&fv, pts, tets, tet2edge, edgelist, t);
    cell_quadrature(&fv, qp, jxw);
    double h = or_tet_diameter(pts, tets, t);
    for (s = 0; s < ns; ++s) {
      double res = 0.0;
      for (q = 0; q < 4; ++q) {
        zdouble e[3], se[3];
        fn_value(&fv, qp[q], sols[s], e);
        sig_mul(sigma, e, se);
        double nrm = 0;
        for (k = 0; k < 3; ++k) { zdouble v = I * omega * se[k]; nrm += creal(v) * creal(v) + cimag(v) * cimag(v); }
        res += nrm * jxw[q];
      }
      errors[s * nt + t] += res * h * h;
    }
    for (f = 0; f < 4; ++f) {
      int n = neighbor[t * 4 + f];
      if (n < 0)
        for (s = 0; s < ns; ++s) face_err[(s * nt + t) * 4 + f] = 0.0;
      if (face_err[(0 * nt + t) * 4 + f] >= 0.0) continue;
      int nn = non[t * 4 + f];
      fe_values fvn;
      double sign[9], fqp[3][3], fjxw[3], fn[3];
      or_sigma(aniso_form, &rho[attr[n] * nc], sign);
      fe_reinit(&fvn, pts, tets, tet2edge, edgelist, n);
      face_normal(&fv, f, fn);
      face_quadrature(&fv, f, fqp, fjxw);
      double hf = or_face_diameter(pts, tets, t, f);
      for (s = 0; s < ns; ++s) {
        double dotj = 0.0, crossh = 0.0;
        for (q = 0; q < 3; ++q) {
          zdouble e[3], en[3], se[3], sen[3], ce[3], cen[3], d, cr[3];
          fn_value(&fv, fqp[q], sols[s], e);
          fn_value(&fvn, fqp[q], sols[s], en);
          sig_mul(sigma, e, se);
          sig_mul(sign, en, sen);
          d = (se[0] - sen[0]) * fn[0] + (se[1] - sen[1]) * fn[1] + (se[2] - sen[2]) * fn[2];
          dotj += (creal(d) * creal(d) + cimag(d) * cimag(d)) * fjxw[q];
          fn_curl(&fv, sols[s], ce);
          fn_curl(&fvn, sols[s], cen);
          zdouble dc[3] = {ce[0] - cen[0], ce[1] - cen[1], ce[2] - cen[2]};
          cr[0] = dc[1] * fn[2] - dc[2] * fn[1];
          cr[1] = dc[2] * fn[0] - dc[0] * fn[2];
          cr[2] = dc[0] * fn[1] - dc[1] * fn[0];
          double c2 = 0;
          for (k = 0; k < 3; ++k) c2 += creal(cr[k]) * creal(cr[k]) + cimag(cr[k]) * cimag(cr[k]);
          crossh += c2 * fjxw[q];
        }
        double v = (dotj + crossh) * hf / 2.0;
        face_err[(s * nt + t) * 4 + f] = v;
        face_err[(s * nt + n) * 4 + nn] = v;
      }
    }
    for (s = 0; s < ns; ++s)
      for (f = 0; f < 4; ++f) errors[s * nt + t] += face_err[(s * nt + t) * 4 + f];
  }
  free(face_err);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Marking (src/em_fwd.cc:1182-1240)                                         */
/* ------------------------------------------------------------------------- */

/* or_estimate_error over nthreads OpenMP threads (the CPU baseline's timing only), with
 * the serial loop's arithmetic: each interior face is evaluated from its lower-index
 * tet, which writes both sides' slots (no other thread writes them), then every tet
 * sums its cell and face terms.  Results equal the serial ones bit for bit. */
int or_estimate_error_mt(int nt, const double *pts, const int *tets, const int *tet2edge, const int *edgelist,
                         const int *neighbor, const int *non, const int *attr, int aniso_form, const double *rho,
                         double freq, int ns, const double *const *sols, double *errors, int nthreads) {
  double omega = 2 * OR_PI * freq;
  int nc = rho_cols(aniso_form), bad = 0, t;
  double *face_err = (double *)calloc((size_t)ns * nt * 4, sizeof(double));
  double *cell_err = (double *)calloc((size_t)ns * nt, sizeof(double));
#pragma omp parallel for schedule(static) num_threads(nthreads) reduction(| : bad)
  for (t = 0; t < nt; ++t) {
    fe_values fv;
    double sigma[9], qp[4][3], jxw[4];
    int f, s, q, k;
    if (or_sigma(aniso_form, &rho[attr[t] * nc], sigma) != 0) { bad = 1; continue; }
    fe_reinit(&fv, pts, tets, tet2edge, edgelist, t);
    cell_quadrature(&fv, qp, jxw);
    double h = or_tet_diameter(pts, tets, t);
    for (s = 0; s < ns; ++s) {
      double res = 0.0;
      for (q = 0; q < 4; ++q) {
        zdouble e[3], se[3];
        fn_value(&fv, qp[q], sols[s], e);
        sig_mul(sigma, e, se);
        double nrm = 0;
        for (k = 0; k < 3; ++k) { zdouble v = I * omega * se[k]; nrm += creal(v) * creal(v) + cimag(v) * cimag(v); }
        res += nrm * jxw[q];
      }
      cell_err[s * nt + t] = res * h * h;
    }
    for (f = 0; f < 4; ++f) {
      int n = neighbor[t * 4 + f];
      if (n < t) continue;  /* boundary faces stay 0; shared faces belong to the lower index */
      int nn = non[t * 4 + f];
      fe_values fvn;
      double sign[9], fqp[3][3], fjxw[3], fn[3];
      or_sigma(aniso_form, &rho[attr[n] * nc], sign);
      fe_reinit(&fvn, pts, tets, tet2edge, edgelist, n);
      face_normal(&fv, f, fn);
      face_quadrature(&fv, f, fqp, fjxw);
      double hf = or_face_diameter(pts, tets, t, f);
      for (s = 0; s < ns; ++s) {
        double dotj = 0.0, crossh = 0.0;
        for (q = 0; q < 3; ++q) {
          zdouble e[3], en[3], se[3], sen[3], ce[3], cen[3], d, cr[3];
          fn_value(&fv, fqp[q], sols[s], e);
          fn_value(&fvn, fqp[q], sols[s], en);
          sig_mul(sigma, e, se);
          sig_mul(sign, en, sen);
          d = (se[0] - sen[0]) * fn[0] + (se[1] - sen[1]) * fn[1] + (se[2] - sen[2]) * fn[2];
          dotj += (creal(d) * creal(d) + cimag(d) * cimag(d)) * fjxw[q];
          fn_curl(&fv, sols[s], ce);
          fn_curl(&fvn, sols[s], cen);
          zdouble dc[3] = {ce[0] - cen[0], ce[1] - cen[1], ce[2] - cen[2]};
          cr[0] = dc[1] * fn[2] - dc[2] * fn[1];
          cr[1] = dc[2] * fn[0] - dc[0] * fn[2];
          cr[2] = dc[0] * fn[1] - dc[1] * fn[0];
          double c2 = 0;
          for (k = 0; k < 3; ++k) c2 += creal(cr[k]) * creal(cr[k]) + cimag(cr[k]) * cimag(cr[k]);
          crossh += c2 * fjxw[q];
        }
        double v = (dotj + crossh) * hf / 2.0;
        face_err[(s * nt + t) * 4 + f] = v;
        face_err[(s * nt + n) * 4 + nn] = v;
      }
    }
  }
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (t = 0; t < nt; ++t) {
    int s, f;
    for (s = 0; s < ns; ++s) {
      double v = 0.0 + cell_err[s * nt + t];
      for (f = 0; f < 4; ++f) v += face_err[(s * nt + t) * 4 + f];
      errors[s * nt + t] = v;
    }
  }
  free(face_err);
  free(cell_err);
  return bad ? -1 : 0;
}

static int cmp_desc(const void *a, const void *b) {
  double x = *(const double *)a, y = *(const double *)b;
  return x > y ? -1 : (x < y ? 1 : 0);
}

/* Doerfler marking: accumulate sorted errors until fraction*total, threshold = next
 * value, mark error > threshold (src/em_fwd.cc:1182-1215).  Flags are OR-ed in. */
void or_refine_fixed_fraction(int n, const double *error, double fraction, unsigned char *flag) {
  double total = 0, sum = 0, thr = 0;
  int i;
  double *s = (double *)malloc(sizeof(double) * (size_t)n);
  for (i = 0; i < n; ++i) { total += error[i]; s[i] = error[i]; }
  qsort(s, (size_t)n, sizeof(double), cmp_desc);
  for (i = 0; i < n; ++i) {
    if (sum < fraction * total) sum += s[i];
    else { thr = s[i]; break; }
  }
  for (i = 0; i < n; ++i)
    if (error[i] > thr) flag[i] = 1;
  free(s);
}

/* Fixed number: threshold = the (n*fraction)-th largest value (src/em_fwd.cc:1217-1240). */
void or_refine_fixed_number(int n, const double *error, double fraction, unsigned char *flag) {
  int k = (int)(n * fraction), i;
  double *s = (double *)malloc(sizeof(double) * (size_t)n);
  for (i = 0; i < n; ++i) s[i] = error[i];
  qsort(s, (size_t)n, sizeof(double), cmp_desc);
  double thr = s[k < n ? k : n - 1];
  for (i = 0; i < n; ++i)
    if (error[i] > thr) flag[i] = 1;
  free(s);
}
