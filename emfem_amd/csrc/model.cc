// Conductivity table, 1-D plane-wave boundary data and host right-hand-side builders.
// See model.h for the reference mapping.
#include "model.h"

#include <cmath>
#include <string>

#include "common.h"
#include "fe.h"

namespace emfem {

int rho_columns(int aniso_form) {
  switch (aniso_form) {
    case 0: return 1;
    case 1: return 2;
    case 2: return 3;
    case 3: return 6;
    default: return 0;
  }
}

// sigma = R diag(1/rho) R^T with R = Rz Ry Rx, angles in degrees (src/em_fwd.cc:39-65).
void make_sigma_table(int form, int n_rho, const double *rho, std::vector<double> &s6) {
  const int nc = rho_columns(form);
  EMF_REQUIRE(nc > 0, kErrUser, "Unsupported anisotropy form " + std::to_string(form) + ".");
  s6.assign((size_t)n_rho * 6, 0.0);
  for (int i = 0; i < n_rho; ++i) {
    const double *r = rho + (size_t)i * nc;
    double p[3];
    if (form == 0) p[0] = p[1] = p[2] = 1.0 / r[0];
    else if (form == 1) { p[0] = p[1] = 1.0 / r[0]; p[2] = 1.0 / r[1]; }
    else { p[0] = 1.0 / r[0]; p[1] = 1.0 / r[1]; p[2] = 1.0 / r[2]; }
    double R[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    if (form == 3) {
      const double cx = std::cos(r[3] * kDtoR), sx = std::sin(r[3] * kDtoR);
      const double cy = std::cos(r[4] * kDtoR), sy = std::sin(r[4] * kDtoR);
      const double cz = std::cos(r[5] * kDtoR), sz = std::sin(r[5] * kDtoR);
      const double Rx[3][3] = {{1, 0, 0}, {0, cx, sx}, {0, -sx, cx}};
      const double Ry[3][3] = {{cy, 0, -sy}, {0, 1, 0}, {sy, 0, cy}};
      const double Rz[3][3] = {{cz, sz, 0}, {-sz, cz, 0}, {0, 0, 1}};
      double ZY[3][3];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          ZY[a][b] = 0;
          for (int k = 0; k < 3; ++k) ZY[a][b] += Rz[a][k] * Ry[k][b];
        }
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          R[a][b] = 0;
          for (int k = 0; k < 3; ++k) R[a][b] += ZY[a][k] * Rx[k][b];
        }
    }
    auto s = [&](int a, int b) { return R[a][0] * p[0] * R[b][0] + R[a][1] * p[1] * R[b][1] + R[a][2] * p[2] * R[b][2]; };
    double *o = &s6[(size_t)i * 6];
    o[0] = s(0, 0);
    o[1] = s(1, 1);
    o[2] = s(2, 2);
    o[3] = s(0, 1);
    o[4] = s(0, 2);
    o[5] = s(1, 2);
  }
}

// ---------------------------------------------------------------------------------
// Plane wave (src/em_pw.h:8-132)
// ---------------------------------------------------------------------------------

void PlaneWave::init(const Column &c, double freq) {
  n_ = (int)c.ztop.size();
  ztop_ = c.ztop;
  lsig_ = c.lsig;
  omega_ = 2 * kPi * freq;
  const cplx iwm(0.0, omega_ * kMu);
  // wavenumber of layer l (1-based; index 0 repeats the top layer)
  std::vector<cplx> k(n_ + 1), q(n_ + 1, cplx(0)), R(n_ + 1, cplx(0));
  for (int l = 1; l <= n_; ++l) k[l] = std::sqrt(-iwm * lsig_[l - 1]);
  k[0] = k[1];
  for (int l = 1; l < n_; ++l) q[l] = std::exp(-k[l] * (ztop_[l] - ztop_[l - 1]));
  // reflection coefficients from the bottom half-space upward
  for (int l = n_ - 1; l >= 1; --l) {
    cplx ks = k[l] + k[l + 1];
    cplx gamma = (-iwm * (lsig_[l - 1] - lsig_[l])) / (ks * ks);
    cplx below = R[l + 1] * q[l + 1];
    R[l] = (gamma + below) / (1.0 + gamma * below) * q[l];
  }
  a_.assign(n_ + 1, cplx(0));
  b_.assign(n_ + 1, cplx(0));
  c_.assign(n_ + 1, cplx(0));
  d_.assign(n_ + 1, cplx(0));
  a_[0] = 1.0;
  c_[0] = 1.0;
  for (int l = 1; l <= n_; ++l) {
    cplx top_e = a_[l - 1] + b_[l - 1] * q[l - 1];
    b_[l] = top_e / (1.0 + R[l] * q[l]);
    a_[l] = b_[l] * R[l];
    cplx top_h = c_[l - 1] + d_[l - 1] * q[l - 1];
    d_[l] = top_h / (1.0 - R[l] * q[l]);
    c_[l] = -d_[l] * R[l];
  }
}

void PlaneWave::field(double z, cplx eh[4]) const {
  int L = 0;
  for (int i = n_ - 1; i >= 0; --i)
    if (z >= ztop_[i]) {
      L = i;
      break;
    }
  const cplx iwm(0.0, omega_ * kMu);
  cplx k = std::sqrt(-iwm * lsig_[L]);
  cplx em = std::exp(-k * (z - ztop_[L]));
  cplx ep(0.0), dep(0.0);
  cplx dem = -k * em;
  if (L != n_ - 1) {
    ep = std::exp(k * (z - ztop_[L + 1]));
    dep = k * ep;
  }
  eh[0] = a_[L + 1] * ep + b_[L + 1] * em;
  eh[1] = (a_[L + 1] * dep + b_[L + 1] * dem) / iwm;
  eh[2] = (c_[L + 1] * dep + d_[L + 1] * dem) / lsig_[L];
  eh[3] = c_[L + 1] * ep + d_[L + 1] * em;
  for (int i = 0; i < 4; ++i)
    if (std::abs(eh[i]) < 1E-80) eh[i] = 0.0;
}

void PlaneWave::efield(double z, int mode, cplx e[3]) const {
  cplx eh[4];
  e[0] = e[1] = e[2] = 0.0;
  field(z, eh);
  if (mode == -1) e[0] = std::conj(eh[0]);
  else if (mode == -2) e[1] = std::conj(eh[2]);
}

// ---------------------------------------------------------------------------------
// Background columns (src/em_fwd.cc:74-159)
// ---------------------------------------------------------------------------------

static void corner_column(const Topology &T, const std::vector<double> &s6, const double *p, double zbottom,
                          Column &col) {
  double fp[3] = {p[0], p[1], p[2]};
  std::vector<double> z, sg;
  while (fp[2] < zbottom) {
    int v = nearest_vertex(T, fp);
    double probe[3] = {fp[0], fp[1], fp[2] + kGeomEps};
    int t = find_cell(T, probe);
    EMF_REQUIRE(t >= 0, kErrUser, "background column left the mesh");
    const double *s = &s6[(size_t)T.attr[t] * 6];
    z.push_back(fp[2]);
    sg.push_back((s[0] + s[1] + s[2]) / 3.0);
    int next = -1;
    for (int i = 0; i < 4; ++i) {
      int vi = T.tets[t * 4 + i];
      const double *x = &T.pts[(size_t)vi * 3];
      if (vi != v && std::abs(x[0] - p[0]) < kGeomEps && std::abs(x[1] - p[1]) < kGeomEps) {
        next = vi;
        break;
      }
    }
    EMF_REQUIRE(next >= 0, kErrUser, "background column is not a chain of vertical mesh edges");
    for (int k = 0; k < 3; ++k) fp[k] = T.pts[(size_t)next * 3 + k];
  }
  col.ztop.clear();
  col.lsig.clear();
  for (size_t i = 0; i < z.size(); ++i)
    if (i == 0 || sg[i] != sg[i - 1]) {
      col.ztop.push_back(z[i]);
      col.lsig.push_back(sg[i]);
    }
}

void update_background(const Topology &T, const std::vector<double> &s6, Background &bg) {
  EMF_REQUIRE(T.nt > 0, kErrState, "no mesh");
  const double far_lo[3] = {-1.0E+15, -1.0E+15, -1.0E+15}, far_hi[3] = {1.0E+15, 1.0E+15, 1.0E+15};
  const double *tl = &T.pts[(size_t)nearest_vertex(T, far_lo) * 3];
  const double *br = &T.pts[(size_t)nearest_vertex(T, far_hi) * 3];
  const double c[4][3] = {{tl[0], tl[1], tl[2]}, {br[0], tl[1], tl[2]}, {br[0], br[1], tl[2]}, {tl[0], br[1], tl[2]}};
  for (int k = 0; k < 4; ++k) {
    for (int j = 0; j < 3; ++j) bg.corners[k][j] = c[k][j];
    corner_column(T, s6, c[k], br[2], bg.col[k]);
  }
  bg.ready = true;
}

// ---------------------------------------------------------------------------------
// Right-hand sides
// ---------------------------------------------------------------------------------

namespace {

struct TetLocal {
  double x[4][3];
  int vid[4];
  int edge[6];
  TetGeom G;
};

void load_tet(const Topology &T, int t, TetLocal &L) {
  for (int v = 0; v < 4; ++v) {
    L.vid[v] = T.tets[t * 4 + v];
    for (int k = 0; k < 3; ++k) L.x[v][k] = T.pts[(size_t)L.vid[v] * 3 + k];
  }
  for (int i = 0; i < 6; ++i) L.edge[i] = T.tet2edge[(size_t)t * 6 + i];
  tet_geometry(L.x, L.G);
}

// Blend weights of the 4 corner solutions at the edge midpoint (src/em_fwd.cc:438-464):
// a cosine taper along the side faces, bilinear elsewhere.
void corner_weights(const double (*tc)[3], const double *ec, double w[4]) {
  auto taper = [](double s) { return (1.0 - std::cos(kPi * s)) / 2.0; };
  const double sx = (ec[0] - tc[0][0]) / (tc[2][0] - tc[0][0]);
  const double sy = (ec[1] - tc[0][1]) / (tc[2][1] - tc[0][1]);
  if (std::abs(ec[0] - tc[0][0]) < kGeomEps) {
    double s = taper(sy);
    w[0] = 1.0 - s; w[1] = 0.0; w[2] = 0.0; w[3] = s;
  } else if (std::abs(ec[1] - tc[0][1]) < kGeomEps) {
    double s = taper(sx);
    w[0] = 1.0 - s; w[1] = s; w[2] = 0.0; w[3] = 0.0;
  } else if (std::abs(ec[0] - tc[2][0]) < kGeomEps) {
    double s = taper(sy);
    w[0] = 0.0; w[1] = 1.0 - s; w[2] = s; w[3] = 0.0;
  } else if (std::abs(ec[1] - tc[2][1]) < kGeomEps) {
    double s = taper(sx);
    w[0] = 0.0; w[1] = 0.0; w[2] = s; w[3] = 1.0 - s;
  } else {
    const double area = (tc[2][0] - tc[0][0]) * (tc[2][1] - tc[0][1]);
    w[0] = ((tc[2][0] - ec[0]) * (tc[2][1] - ec[1])) / area;
    w[1] = ((ec[0] - tc[3][0]) * (tc[3][1] - ec[1])) / area;
    w[2] = ((ec[0] - tc[0][0]) * (ec[1] - tc[0][1])) / area;
    w[3] = ((tc[1][0] - ec[0]) * (ec[1] - tc[1][1])) / area;
  }
}

}  // namespace

void rhs_mt(const Topology &T, const std::vector<double> &s6, const Background &bg, double freq, int mode,
            std::vector<double> &s_re, std::vector<double> &s_im, std::vector<cplx> *bval_out) {
  EMF_REQUIRE(bg.ready, kErrState, "background model not set (call emfem_update_background_model)");
  PlaneWave pw[4];
  for (int k = 0; k < 4; ++k) pw[k].init(bg.col[k], freq);
  std::vector<cplx> bval(T.ne, cplx(0));
  // Dirichlet values: the first bdr cell holding an edge defines it (src/em_fwd.cc:424-482).
  // The owner (cell, local edge) of every boundary edge is fixed by one sequential
  // pass; the values are then computed in parallel.
  const int nb = (int)T.bdr_cells.size();
  std::vector<int> own_cell, own_local;
  {
    std::vector<uint8_t> have(T.ne, 0);
    for (int q = 0; q < nb; ++q) {
      const int t = T.bdr_cells[q];
      for (int i = 0; i < 6; ++i) {
        const int g = T.tet2edge[(size_t)t * 6 + i];
        if (!T.edge_bdr[g] || have[g]) continue;
        have[g] = 1;
        own_cell.push_back(t);
        own_local.push_back(i);
      }
    }
  }
  parallel_ranges((int64_t)own_cell.size(), [&](int64_t qb, int64_t qe, int) {
    for (int64_t q = qb; q < qe; ++q) {
      TetLocal L;
      load_tet(T, own_cell[q], L);
      const int i = own_local[q];
      const int g = L.edge[i];
      int a, b;
      edge_ends(L.vid, i, a, b);
      double ec[3], ed[3];
      for (int k = 0; k < 3; ++k) {
        ec[k] = (L.x[a][k] + L.x[b][k]) / 2.0;
        ed[k] = L.x[b][k] - L.x[a][k];
      }
      double w[4];
      corner_weights(bg.corners, ec, w);
      cplx ep[3] = {0.0, 0.0, 0.0};
      for (int k = 0; k < 4; ++k) {
        cplx e[3];
        pw[k].efield(ec[2], mode, e);
        for (int j = 0; j < 3; ++j) ep[j] += w[k] * e[j];
      }
      double lam[4], N[3];
      barycentric(L.x, L.G, ec, lam);
      whitney_value(L.G, lam, a, b, N);
      cplx v = (ed[0] * ep[0] + ed[1] * ep[1] + ed[2] * ep[2]) / dot3(ed, N);
      bval[g] = std::abs(v) > 1E-13 ? v : cplx(0.0);
    }
  });
  // Lift the Dirichlet data with the raw element matrices (src/em_fwd.cc:516-572):
  // per-cell contributions in parallel, summed in cell order.
  s_re.assign(T.ne, 0.0);
  s_im.assign(T.ne, 0.0);
  const double omega = 2 * kPi * freq;
  std::vector<double> contrib((size_t)nb * 12);
  parallel_ranges(nb, [&](int64_t qb, int64_t qe, int) {
    for (int64_t q = qb; q < qe; ++q) {
      const int t = T.bdr_cells[q];
      TetLocal L;
      load_tet(T, t, L);
      double C[6][6], M[6][6];
      element_matrices(L.G, L.vid, &s6[(size_t)T.attr[t] * 6], omega, C, M);
      double *br = &contrib[(size_t)q * 12], *bi = br + 6;
      for (int i = 0; i < 6; ++i) br[i] = bi[i] = 0.0;
      for (int i = 0; i < 6; ++i) {
        if (!T.edge_bdr[L.edge[i]]) continue;
        const double vr = bval[L.edge[i]].real(), vi = bval[L.edge[i]].imag();
        br[i] = vr * C[i][i];
        bi[i] = -vi * C[i][i];
        for (int j = 0; j < 6; ++j) {
          if (T.edge_bdr[L.edge[j]]) continue;
          br[j] += -C[j][i] * vr + M[j][i] * vi;
          bi[j] += M[j][i] * vr + C[j][i] * vi;
        }
      }
    }
  });
  for (int q = 0; q < nb; ++q) {
    const int t = T.bdr_cells[q];
    const double *br = &contrib[(size_t)q * 12], *bi = br + 6;
    for (int i = 0; i < 6; ++i) {
      const int g = T.tet2edge[(size_t)t * 6 + i];
      s_re[g] += br[i];
      s_im[g] += bi[i];
    }
  }
  if (bval_out) bval_out->swap(bval);
}

Locate host_locate(const Topology &T) {
  return [&T](const std::vector<double> &pts) {
    std::vector<int> t(pts.size() / 3);
    for (size_t i = 0; i < t.size(); ++i) t[i] = find_cell(T, &pts[i * 3]);
    return t;
  };
}

void rhs_csem(const Topology &T, double freq, const double *tx, int n_div, const Locate &loc,
              std::vector<double> &s_re, std::vector<double> &s_im) {
  s_re.assign(T.ne, 0.0);
  s_im.assign(T.ne, 0.0);
  const double omega = 2 * kPi * freq;
  double dir[3] = {std::cos(tx[3] * kDtoR) * std::cos(tx[4] * kDtoR), std::sin(tx[3] * kDtoR) * std::cos(tx[4] * kDtoR),
                   std::sin(tx[4] * kDtoR)};
  const double current = tx[5], length = tx[6];
  // point dipoles (src/em_fwd.cc:582-614)
  std::vector<double> pos, wt;
  if (std::abs(length) <= 0.1) {
    pos = {tx[0], tx[1], tx[2]};
    wt = {1.0};
  } else {
    EMF_REQUIRE(n_div > 0, kErrArg, "n_tx_divisions must be positive");
    double A[3];
    for (int k = 0; k < 3; ++k) A[k] = tx[k] - dir[k] * length / 2;
    for (int i = 0; i < n_div; ++i) {
      double B[3];
      for (int k = 0; k < 3; ++k) B[k] = A[k] + dir[k] * length / n_div;
      for (int k = 0; k < 3; ++k) pos.push_back((A[k] + B[k]) / 2);
      wt.push_back(1.0 / n_div);
      for (int k = 0; k < 3; ++k) A[k] = B[k];
    }
  }
  const std::vector<int> cells = loc(pos);
  for (size_t j = 0; j < wt.size(); ++j) {
    const double *c = &pos[j * 3];
    const int t = cells[j];
    EMF_REQUIRE(t >= 0, kErrUser, "Transmitter dipole " + std::to_string(j) + " is not found in mesh.");
    TetLocal L;
    load_tet(T, t, L);
    double lam[4];
    barycentric(L.x, L.G, c, lam);
    for (int i = 0; i < 6; ++i) {
      int a, b;
      double N[3];
      edge_ends(L.vid, i, a, b);
      whitney_value(L.G, lam, a, b, N);
      s_im[L.edge[i]] += omega * current * dot3(N, dir) * wt[j];
    }
  }
}

void rhs_dual(const Topology &T, const std::vector<double> &rx, const Locate &loc, std::vector<double> &s_re,
              std::vector<double> &s_im) {
  s_re.assign(T.ne, 0.0);
  s_im.assign(T.ne, 0.0);
  const std::vector<int> cells = loc(rx);
  for (size_t r = 0; r < rx.size() / 3; ++r) {
    const double *p = &rx[r * 3];
    const int t = cells[r];
    EMF_REQUIRE(t >= 0, kErrUser, "Receiver " + std::to_string(r) + " is not found in mesh.");
    TetLocal L;
    load_tet(T, t, L);
    double lam[4];
    barycentric(L.x, L.G, p, lam);
    for (int i = 0; i < 6; ++i) {
      int a, b;
      double N[3];
      edge_ends(L.vid, i, a, b);
      whitney_value(L.G, lam, a, b, N);
      const double v = N[0] + N[1] + N[2];
      s_re[L.edge[i]] += v;
      s_im[L.edge[i]] += v;
    }
  }
}

void interpolate(const Topology &T, int t, const double *p, double freq, const double *e_ri, cplx E[3], cplx H[3]) {
  TetLocal L;
  load_tet(T, t, L);
  double lam[4];
  barycentric(L.x, L.G, p, lam);
  const cplx hscale = -1.0 / (cplx(0.0, 1.0) * 2.0 * kPi * freq * kMu);
  for (int k = 0; k < 3; ++k) E[k] = H[k] = 0.0;
  for (int i = 0; i < 6; ++i) {
    int a, b;
    double N[3], Cn[3];
    edge_ends(L.vid, i, a, b);
    whitney_value(L.G, lam, a, b, N);
    whitney_curl(L.G, a, b, Cn);
    const cplx v(e_ri[(size_t)L.edge[i] * 2], e_ri[(size_t)L.edge[i] * 2 + 1]);
    for (int k = 0; k < 3; ++k) {
      E[k] += v * N[k];
      H[k] += hscale * v * Cn[k];
    }
  }
}

bool mt_response(int otype, double freq, const cplx exy[3], const cplx hxy[3], const cplx eyx[3], const cplx hyx[3],
                 cplx &out) {
  const cplx det = hxy[0] * hyx[1] - hyx[0] * hxy[1];
  const cplx zxx = (exy[0] * hyx[1] - eyx[0] * hxy[1]) / det;
  const cplx zxy = (eyx[0] * hxy[0] - exy[0] * hyx[0]) / det;
  const cplx zyx = (exy[1] * hyx[1] - eyx[1] * hxy[1]) / det;
  const cplx zyy = (eyx[1] * hxy[0] - exy[1] * hyx[0]) / det;
  const double wm = 2 * kPi * freq * kMu;
  auto rho_phase = [&](cplx z) { return cplx(std::log(std::norm(z) / wm), std::arg(z)); };
  switch (otype) {
    case 311: out = zxx; return true;
    case 321: out = zxy; return true;
    case 331: out = zyx; return true;
    case 341: out = zyy; return true;
    case 351: out = (hxy[2] * hyx[1] - hyx[2] * hxy[1]) / det; return true;
    case 361: out = (hyx[2] * hxy[0] - hxy[2] * hyx[0]) / det; return true;
    case 212: out = rho_phase(zxy); return true;
    case 222: out = rho_phase(zyx); return true;
    case 312: out = std::log(zxx); return true;
    case 322: out = std::log(zxy); return true;
    case 332: out = std::log(zyx); return true;
    case 342: out = std::log(zyy); return true;
    default: return false;
  }
}

bool csem_response(int otype, const cplx E[3], const cplx H[3], cplx &out) {
  if (otype < 111 || otype > 162) return false;
  const int comp = otype / 10 % 10 - 1, form = otype % 10;
  if (comp < 0 || comp > 5 || (form != 1 && form != 2)) return false;
  cplx v = comp < 3 ? E[comp] : H[comp - 3];
  out = form == 2 ? std::log(v) : v;
  return true;
}

}  // namespace emfem
