// Forward driver, input/output formats and mesh refinement entry points.
//
// Reference: em_forward / forward_mt / forward_csem / distribute_tasks
// (src/em_fwd.cc:1273-1471), refine_rx_area / refine_tx_area / refine_uniform /
// refine_mesh (src/em_fwd.cc:873-961,1242-1271), read_mesh / read_rho / read_emd /
// save_rsp (src/em_io.cc:9-399) and parse_string (src/em_utils.cc:34-58).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>
#include <set>
#include <sstream>
#include <tuple>

#include "../../include/emfem_amd.h"
#include "context.h"
#include "refine.h"

using namespace emfem;
using namespace emfem_ctx;

extern "C" {
int emfem_set_mesh(emfem_context *c, int n_points, const double *points, int n_tets, const int *tets,
                   const int *attributes);
}

namespace {

constexpr double kRtoD = 180.0 / kPi;

// Reads a whole file with '#' comments removed (parse_string, src/em_utils.cc:34-58).
std::stringstream read_clean(const std::string &path, const char *what) {
  std::ifstream f(path);
  EMF_REQUIRE(f.good(), kErrUser, std::string("Unable to open file ") + path + ".");
  std::stringstream out;
  std::string line;
  while (std::getline(f, line)) {
    auto p = line.find('#');
    if (p != std::string::npos) line.erase(p);
    out << line << "\n";
  }
  (void)what;
  return out;
}

template <typename T>
T take(std::stringstream &ss, const char *what) {
  T v;
  ss >> v;
  EMF_REQUIRE(!ss.fail(), kErrUser, std::string("malformed input while reading ") + what);
  return v;
}

bool is_ap(int otype) {
  switch (otype) {
    case 212: case 222: case 112: case 122: case 132: case 142: case 152: case 162: return true;
    default: return false;
  }
}

RawMesh current_mesh(const emfem_context *c) {
  RawMesh m;
  m.pts = c->T.pts;
  m.tets = c->T.tets;
  m.attr = c->T.attr;
  return m;
}

void install_mesh(emfem_context *c, const RawMesh &m) {
  const int rc = emfem_set_mesh(c, m.nv(), m.pts.data(), m.nt(), m.tets.data(), m.attr.data());
  if (rc != EMFEM_OK) throw Error(rc, c->err);
}

double tet_volume(const Topology &T, int t) {
  const double *p[4];
  for (int k = 0; k < 4; ++k) p[k] = &T.pts[(size_t)T.tets[t * 4 + k] * 3];
  double a[3], b[3], d[3];
  for (int k = 0; k < 3; ++k) {
    a[k] = p[1][k] - p[0][k];
    b[k] = p[2][k] - p[0][k];
    d[k] = p[3][k] - p[0][k];
  }
  return (a[0] * (b[1] * d[2] - b[2] * d[1]) - a[1] * (b[0] * d[2] - b[2] * d[0]) + a[2] * (b[0] * d[1] - b[1] * d[0])) /
         6.0;
}

// Refines tets around the given points until each holds volume <= L^3
// (refine_rx_area / refine_tx_area, src/em_fwd.cc:885-961).
void refine_points(emfem_context *c, const std::vector<double> &pts, double L, const char *what) {
  if (L <= 0.0) return;
  for (int pass = 0; pass < 64; ++pass) {
    std::vector<uint8_t> flag(c->T.nt, 0);
    bool any = false;
    const std::vector<int> cells = locator(c)(pts);
    for (size_t i = 0; i < pts.size() / 3; ++i) {
      const int t = cells[i];
      EMF_REQUIRE(t >= 0, kErrUser, std::string(what) + " " + std::to_string(i) + " is not found in mesh.");
      if (tet_volume(c->T, t) > L * L * L) {
        flag[t] = 1;
        any = true;
      }
    }
    if (!any) return;
    RawMesh m = current_mesh(c);
    refine_bisect(m, flag);
    install_mesh(c, m);
  }
}

void restore_coarse(emfem_context *c) {
  RawMesh m;
  m.pts = c->coarse_pts;
  m.tets = c->coarse_tets;
  m.attr = c->coarse_attr;
  install_mesh(c, m);
}

// Dipole positions of a transmitter (divide_line_source, src/em_fwd.cc:582-614).
std::vector<double> dipole_points(const double *tx, int n_div) {
  const double dir[3] = {std::cos(tx[3] * kDtoR) * std::cos(tx[4] * kDtoR),
                         std::sin(tx[3] * kDtoR) * std::cos(tx[4] * kDtoR), std::sin(tx[4] * kDtoR)};
  std::vector<double> out;
  if (std::abs(tx[6]) <= 0.1) return {tx[0], tx[1], tx[2]};
  double A[3];
  for (int k = 0; k < 3; ++k) A[k] = tx[k] - dir[k] * tx[6] / 2;
  for (int i = 0; i < n_div; ++i) {
    double B[3];
    for (int k = 0; k < 3; ++k) B[k] = A[k] + dir[k] * tx[6] / n_div;
    for (int k = 0; k < 3; ++k) out.push_back((A[k] + B[k]) / 2);
    for (int k = 0; k < 3; ++k) A[k] = B[k];
  }
  return out;
}

void check(emfem_context *c, int rc) {
  if (rc != EMFEM_OK) throw Error(rc, c->err);
}

// save_mesh (src/em_io.cc:289-314) + Mesh::save_vtk (src/em_mesh.cc:184-248): legacy
// ASCII VTK with the subdomain (0: one process), the attribute, every resistivity
// column of the cell's row and the error indicators.  The data blocks follow the
// reference's std::map, i.e. they are ordered by name.
void write_vtk(const emfem_context *c, const std::string &path, int tidx, const double *err0, const double *err1) {
  static const char *kNames[4][6] = {{"rho"},
                                     {"rho_h", "rho_v"},
                                     {"rho_x", "rho_y", "rho_z"},
                                     {"rho_x", "rho_y", "rho_z", "eta_x", "eta_y", "eta_z"}};
  static const int kCols[4] = {1, 2, 3, 6};
  const Topology &T = c->T;
  EMF_REQUIRE(c->aniso_form >= 0 && c->aniso_form < 4 && c->n_rho > 0, kErrState, "no resistivity table");
  EMF_REQUIRE(err0 != nullptr && (tidx >= 0 || err1 != nullptr), kErrArg, "missing error indicators");
  const int nc = kCols[c->aniso_form];
  std::map<std::string, std::vector<double>> data;
  for (int i = 0; i < nc; ++i) {
    auto &d = data[kNames[c->aniso_form][i]];
    d.resize(T.nt);
    for (int t = 0; t < T.nt; ++t) d[t] = c->rho[(size_t)T.attr[t] * nc + i];
  }
  if (tidx < 0) {
    data["error_xy"].assign(err0, err0 + T.nt);
    data["error_yx"].assign(err1, err1 + T.nt);
  } else {
    data["error"].assign(err0, err0 + T.nt);
  }
  FILE *fp = std::fopen(path.c_str(), "w");
  EMF_REQUIRE(fp != nullptr, kErrUser, "Unable to open file " + path + ".");
  std::fprintf(fp, "# vtk DataFile Version 2.0\nUnstructured Grid\nASCII\nDATASET UNSTRUCTURED_GRID\n");
  std::fprintf(fp, "POINTS %d double\n", T.nv);
  for (int v = 0; v < T.nv; ++v)
    std::fprintf(fp, "%.17g %.17g %.17g\n", T.pts[(size_t)v * 3], T.pts[(size_t)v * 3 + 1], T.pts[(size_t)v * 3 + 2]);
  std::fprintf(fp, "\nCELLS %d %d\n", T.nt, T.nt * 5);
  for (int t = 0; t < T.nt; ++t) {
    std::fprintf(fp, "4");
    for (int k = 0; k < 4; ++k) std::fprintf(fp, " %4d", T.tets[(size_t)t * 4 + k]);
    std::fprintf(fp, "\n");
  }
  std::fprintf(fp, "\nCELL_TYPES %d\n", T.nt);
  for (int t = 0; t < T.nt; ++t) std::fprintf(fp, "10\n");
  std::fprintf(fp, "\nCELL_DATA %d\n", T.nt);
  auto block = [&](const char *name, auto value) {
    std::fprintf(fp, "SCALARS %s double 1\nLOOKUP_TABLE default\n", name);
    for (int t = 0; t < T.nt; ++t) std::fprintf(fp, "%.15g\n", value(t));
    std::fprintf(fp, "\n");
  };
  block("subdomain", [](int) { return 0.0; });
  block("tet_attr", [&](int t) { return (double)T.attr[t]; });
  for (const auto &kv : data) block(kv.first.c_str(), [&](int t) { return kv.second[t]; });
  std::fclose(fp);
}

// Per-cycle VTK output of the drivers (src/em_fwd.cc:1361-1363,1420-1422).  The
// reference requires -oprefix; here an empty prefix turns the output off.
bool saving_mesh(const emfem_context *c) { return c->opt("save_mesh") == "true" && !c->opt("oprefix").empty(); }

std::string vtk_name(const emfem_context *c, int fidx, int tidx, int cycle) {
  char buf[64];
  if (tidx < 0) std::snprintf(buf, sizeof(buf), "-%02d-%02d.vtk", fidx, cycle);
  else std::snprintf(buf, sizeof(buf), "-%02d-%03d-%02d.vtk", fidx, tidx, cycle);
  return c->opt("oprefix") + buf;
}

// Marks and refines by the squared indicators (refine_mesh, src/em_fwd.cc:1242-1271).
void refine_by_error(emfem_context *c, const std::vector<std::vector<double>> &errors) {
  std::vector<uint8_t> flag(c->T.nt, 0);
  const double frac = c->optd("refine_fraction");
  const bool fixed_fraction = c->opt("refine_strategy") == "fixed_fraction";
  for (const auto &e : errors) {
    if (fixed_fraction) check(c, emfem_refine_fixed_fraction((int)e.size(), e.data(), frac, flag.data()));
    else check(c, emfem_refine_fixed_number((int)e.size(), e.data(), frac, flag.data()));
  }
  RawMesh m = current_mesh(c);
  refine_bisect(m, flag);
  install_mesh(c, m);
}

void solve_slots(emfem_context *c, int first, int n, const std::vector<double> &rtols);

// The dual problem (point sources at receivers on the air-earth interface) drives
// gradient fields ~ 1/(omega sigma_air) in the air; it is solved with the cycle
// precision of -dual_precision ("same": that of the primal solves).
void solve_dual(emfem_context *c, int slot) {
  const std::string keep = c->opts["inner_precision"];
  if (c->opt("dual_precision") != "same") c->opts["inner_precision"] = c->opt("dual_precision");
  try {
    solve_slots(c, slot, 1, {c->optd("dual_rtol")});
  } catch (...) {
    c->opts["inner_precision"] = keep;
    throw;
  }
  c->opts["inner_precision"] = keep;
}

// Primal slots [first, first + n) at -e_rtol and the dual at slot first + n at
// -dual_rtol.  With the dual in the primal's cycle precision they are one block of
// n + 1 right-hand sides (one set of cycles for all) when the block keeps the
// many-column SpMMs on 16-byte rows (n + 1 <= 7 or n + 1 even); otherwise two solves.
void solve_primal_dual(emfem_context *c, int first, int n) {
  const std::string dp = c->opt("dual_precision");
  std::vector<double> tol(n, c->optd("e_rtol"));
  if ((dp == "same" || dp == c->opt("inner_precision")) && (n + 1 <= 7 || (n + 1) % 2 == 0)) {
    tol.push_back(c->optd("dual_rtol"));
    solve_slots(c, first, n + 1, tol);
    return;
  }
  solve_slots(c, first, n, tol);
  solve_dual(c, first + n);
}

void solve_slots(emfem_context *c, int first, int n, const std::vector<double> &rtols) {
  if (n > 1 && c->opt("forward_block_rhs") == "false") {
    // the reference's sequence: one solve per right-hand side (src/em_fwd.cc:1345-1355)
    for (int k = 0; k < n; ++k) solve_slots(c, first + k, 1, {rtols[k]});
    return;
  }
  std::vector<int> its(n);
  std::vector<double> rr(n);
  const int rc = emfem_solve_linear_system(c, first, n, c->opti("K_max_it"), rtols[0], rtols.data(), its.data(), rr.data());
  if (rc == EMFEM_ERR_NOT_CONVERGED) {
    // the reference continues with the partial solution (src/em_la.cc:51-52); count it
    c->stats["not_converged"] += 1;
    return;
  }
  check(c, rc);
  c->stats["forward_outer_its"] += *std::max_element(its.begin(), its.end());
}

// Wall time of one phase of the forward driver (every call it makes returns with its
// work done), accumulated into the stat fwd_<name>_ms.
template <typename F>
void phase(emfem_context *c, const char *name, F &&f) {
  const double t0 = now_ms();
  f();
  c->stats[std::string("fwd_") + name + "_ms"] += now_ms() - t0;
}

void forward_mt(emfem_context *c, int fidx) {
  restore_coarse(c);
  refine_points(c, c->rx, c->optd("max_rx_edge_length"), "Receiver");
  for (int i = 0; i < c->opti("n_uniform_refinements"); ++i) {
    RawMesh m = current_mesh(c);
    refine_uniform(m);
    install_mesh(c, m);
  }
  const int max_cycles = c->opti("max_adaptive_refinements");
  for (int cycle = 0; cycle <= max_cycles; ++cycle) {
    const bool last = cycle == max_cycles || c->T.ne >= c->opti("max_dofs");
    // learned indicator in intermediate cycles: no dual solve, no estimator
    const bool sur = c->opt("surrogate_cycles") == "true" && c->sur && !c->sur_rejected;
    phase(c, "system", [&]() { check(c, emfem_create_linear_system(c, 3)); });
    phase(c, "assemble", [&]() {
      check(c, emfem_assemble_matrix(c, fidx));
      check(c, emfem_assemble_rhs_mt(c, fidx, EMFEM_XY_POLAR, 0));
      check(c, emfem_assemble_rhs_mt(c, fidx, EMFEM_YX_POLAR, 1));
    });
    phase(c, "pc", [&]() { check(c, emfem_create_pc(c)); });
    const double e_rtol = c->optd("e_rtol");
    if (sur && !last) {
      // primal indicators exactly; the dual solve's indicator from the network
      phase(c, "solve", [&]() { solve_slots(c, 0, 2, {e_rtol, e_rtol}); });
      check(c, emfem_destroy_pc(c));
      const int nt = c->T.nt;
      const int slots[2] = {0, 1};
      std::vector<double> err((size_t)2 * nt), dual;
      phase(c, "estimate", [&]() {
        check(c, emfem_estimate_error(c, fidx, 2, slots, err.data()));
        dual = surrogate_indicator(c, fidx, 0, 1);
      });
      std::vector<double> exy(nt), eyx(nt);
      for (int t = 0; t < nt; ++t) {
        exy[t] = std::sqrt(err[t] * dual[t]);
        eyx[t] = std::sqrt(err[(size_t)nt + t] * dual[t]);
      }
      c->stats["surrogate_cycles_used"] += 1;
      if (saving_mesh(c)) write_vtk(c, vtk_name(c, fidx, -3, cycle), -3, exy.data(), eyx.data());
      phase(c, "system", [&]() { check(c, emfem_destroy_linear_system(c)); });
      phase(c, "refine", [&]() { refine_by_error(c, {exy, eyx}); });
      continue;
    }
    phase(c, "assemble", [&]() { check(c, emfem_assemble_rhs_dual(c, 2)); });
    phase(c, "solve", [&]() { solve_primal_dual(c, 0, 2); });
    check(c, emfem_destroy_pc(c));
    const int nt = c->T.nt;
    const int slots[3] = {0, 1, 2};
    std::vector<double> err((size_t)3 * nt);
    phase(c, "estimate", [&]() { check(c, emfem_estimate_error(c, fidx, 3, slots, err.data())); });
    std::vector<double> exy(nt), eyx(nt);
    for (int t = 0; t < nt; ++t) {
      exy[t] = std::sqrt(err[t] * err[(size_t)2 * nt + t]);
      eyx[t] = std::sqrt(err[(size_t)nt + t] * err[(size_t)2 * nt + t]);
    }
    if (saving_mesh(c)) write_vtk(c, vtk_name(c, fidx, -3, cycle), -3, exy.data(), eyx.data());
    if (sur && last) {
      // the plain estimator judges the surrogate on the final mesh: the indicators
      // with the predicted and with the solved dual
      const std::vector<double> dual = surrogate_indicator(c, fidx, 0, 1);
      std::vector<double> plain(nt), pred(nt);
      for (int t = 0; t < nt; ++t) {
        plain[t] = exy[t] + eyx[t];
        pred[t] = std::sqrt(err[t] * dual[t]) + std::sqrt(err[(size_t)nt + t] * dual[t]);
      }
      const double corr = rank_correlation(pred, plain);
      c->stats["surrogate_rank_corr"] = corr;
      if (corr < c->optd("surrogate_min_corr")) c->sur_rejected = true;
      c->stats["surrogate_rejected"] = c->sur_rejected ? 1.0 : 0.0;
    }
    if (last) {
      std::vector<double> rsp(c->obs4.size() / 2, NAN);
      phase(c, "responses", [&]() {
        check(c, emfem_calculate_response_mt(c, fidx, 0, 1, (int)c->obs4.size() / 4, c->obs4.data(), rsp.data()));
      });
      for (size_t o = 0; o < c->obs4.size() / 4; ++o)
        if (c->obs4[o * 4 + 1] == fidx && c->obs4[o * 4 + 2] == EMFEM_MT_TIDX)
          c->rsp[o] = cplx(rsp[o * 2], rsp[o * 2 + 1]);
      check(c, emfem_destroy_linear_system(c));
      break;
    }
    phase(c, "system", [&]() { check(c, emfem_destroy_linear_system(c)); });
    phase(c, "refine", [&]() { refine_by_error(c, {exy, eyx}); });
  }
}

// One frequency's CSEM transmitters.  When they share one mesh (no transmitter-area
// or adaptive refinement) they are solved as one block of right-hand sides; otherwise
// each transmitter gets its own refined mesh as in the reference.
void forward_csem(emfem_context *c, int fidx, const std::vector<int> &tidxs) {
  const int n_div = c->opti("n_tx_divisions");
  const bool shared = c->opti("max_adaptive_refinements") == 0 && c->optd("max_tx_edge_length") <= 0.0;
  std::vector<std::vector<int>> groups;
  if (shared) groups.push_back(tidxs);
  else
    for (int t : tidxs) groups.push_back({t});
  for (const auto &grp : groups) {
    restore_coarse(c);
    if (!shared) refine_points(c, dipole_points(&c->tx[(size_t)grp[0] * 7], n_div), c->optd("max_tx_edge_length"),
                               "Transmitter");
    refine_points(c, c->rx, c->optd("max_rx_edge_length"), "Receiver");
    for (int i = 0; i < c->opti("n_uniform_refinements"); ++i) {
      RawMesh m = current_mesh(c);
      refine_uniform(m);
      install_mesh(c, m);
    }
    const int max_cycles = c->opti("max_adaptive_refinements");
    const int ng = (int)grp.size();
    for (int cycle = 0; cycle <= max_cycles; ++cycle) {
      check(c, emfem_create_linear_system(c, ng + 1));
      check(c, emfem_assemble_matrix(c, fidx));
      check(c, emfem_create_pc(c));
      for (int k = 0; k < ng; ++k) check(c, emfem_assemble_rhs_csem(c, fidx, grp[k], k));
      check(c, emfem_assemble_rhs_dual(c, ng));
      solve_primal_dual(c, 0, ng);
      check(c, emfem_destroy_pc(c));
      const bool last = cycle == max_cycles || c->T.ne >= c->opti("max_dofs");
      if (saving_mesh(c)) {
        const int nt = c->T.nt;
        std::vector<double> err((size_t)2 * nt), ind(nt);
        for (int k = 0; k < ng; ++k) {
          const int slots[2] = {k, ng};
          check(c, emfem_estimate_error(c, fidx, 2, slots, err.data()));
          for (int t = 0; t < nt; ++t) ind[t] = std::sqrt(err[t] * err[(size_t)nt + t]);
          write_vtk(c, vtk_name(c, fidx, grp[k], cycle), grp[k], ind.data(), nullptr);
        }
      }
      if (last) {
        for (int k = 0; k < ng; ++k) {
          std::vector<double> rsp(c->obs4.size() / 2, NAN);
          check(c, emfem_calculate_response_csem(c, fidx, grp[k], k, (int)c->obs4.size() / 4, c->obs4.data(),
                                                 rsp.data()));
          for (size_t o = 0; o < c->obs4.size() / 4; ++o)
            if (c->obs4[o * 4 + 1] == fidx && c->obs4[o * 4 + 2] == grp[k]) c->rsp[o] = cplx(rsp[o * 2], rsp[o * 2 + 1]);
        }
        check(c, emfem_destroy_linear_system(c));
        break;
      }
      const int nt = c->T.nt;
      const int slots[2] = {0, ng};
      std::vector<double> err((size_t)2 * nt), ind(nt);
      check(c, emfem_estimate_error(c, fidx, 2, slots, err.data()));
      for (int t = 0; t < nt; ++t) ind[t] = std::sqrt(err[t] * err[(size_t)nt + t]);
      check(c, emfem_destroy_linear_system(c));
      refine_by_error(c, {ind});
    }
  }
}

}  // namespace

extern "C" {

int emfem_read_mesh(emfem_context *c, const char *path) {
  return guarded(c, [&]() {
    EMF_REQUIRE(path != nullptr, kErrArg, "null path");
    std::stringstream ss = read_clean(path, "mesh");
    const int np = take<int>(ss, "number of points");
    EMF_REQUIRE(np > 0, kErrUser, "mesh has no points");
    std::vector<double> pts((size_t)np * 3);
    for (auto &v : pts) v = take<double>(ss, "points");
    const int nt = take<int>(ss, "number of tets");
    EMF_REQUIRE(nt > 0, kErrUser, "mesh has no tets");
    std::vector<int> tets((size_t)nt * 4), attr(nt);
    for (int t = 0; t < nt; ++t) {
      for (int k = 0; k < 4; ++k) tets[(size_t)t * 4 + k] = take<int>(ss, "tets");
      attr[t] = (int)take<double>(ss, "tet attribute");
    }
    // read_rho (src/em_io.cc:9-86): n_rho form, then per row the values and two bounds
    const int n_rho = take<int>(ss, "number of resistivities");
    const int form = take<int>(ss, "anisotropy form");
    const int ncol = rho_columns(form);
    EMF_REQUIRE(ncol > 0, kErrUser, "Unsupported anisotropy form " + std::to_string(form) + ".");
    std::vector<double> rho((size_t)n_rho * ncol);
    for (int i = 0; i < n_rho; ++i) {
      for (int k = 0; k < ncol; ++k) rho[(size_t)i * ncol + k] = take<double>(ss, "resistivity");
      take<double>(ss, "lower bound");
      take<double>(ss, "upper bound");
    }
    check(c, emfem_set_mesh(c, np, pts.data(), nt, tets.data(), attr.data()));
    check(c, emfem_set_rho(c, n_rho, form, rho.data()));
    c->coarse_pts = pts;
    c->coarse_tets = tets;
    c->coarse_attr = attr;
  });
}

int emfem_set_observations(emfem_context *c, int n, const int *obs4, const double *values) {
  return guarded(c, [&]() {
    EMF_REQUIRE(n >= 0 && (n == 0 || obs4), kErrArg, "bad observation arrays");
    c->obs4.assign(obs4, obs4 + (size_t)n * 4);
    c->obs.assign(n, cplx(0));
    c->obserr.assign(n, cplx(0));
    c->rsp.assign(n, cplx(0));
    if (!values) return;
    for (int i = 0; i < n; ++i) {
      double re = values[i * 4], im = values[i * 4 + 1], er = values[i * 4 + 2], ei = values[i * 4 + 3];
      if (is_ap(obs4[i * 4])) {
        er /= 100.0;
        ei *= kDtoR;
        c->obs[i] = cplx(std::log(re), im * kDtoR);
      } else {
        c->obs[i] = cplx(re, im);
      }
      c->obserr[i] = cplx(er * er, ei * ei);
    }
  });
}

int emfem_read_emd(emfem_context *c, const char *path) {
  return guarded(c, [&]() {
    EMF_REQUIRE(path != nullptr, kErrArg, "null path");
    std::stringstream ss = read_clean(path, "emd");
    const int nf = take<int>(ss, "number of frequencies");
    std::vector<double> f(nf);
    for (auto &v : f) v = take<double>(ss, "frequencies");
    const int ntx = take<int>(ss, "number of transmitters");
    std::vector<double> tx((size_t)ntx * 7);
    for (auto &v : tx) v = take<double>(ss, "transmitters");
    const int nrx = take<int>(ss, "number of receivers");
    std::vector<double> rx((size_t)nrx * 3);
    for (auto &v : rx) v = take<double>(ss, "receivers");
    const int no = take<int>(ss, "number of observations");
    std::vector<int> obs4((size_t)no * 4);
    std::vector<double> val((size_t)no * 4);
    for (int i = 0; i < no; ++i) {
      for (int k = 0; k < 4; ++k) obs4[(size_t)i * 4 + k] = take<int>(ss, "observation index");
      for (int k = 0; k < 4; ++k) val[(size_t)i * 4 + k] = take<double>(ss, "observation value");
    }
    check(c, emfem_set_frequencies(c, nf, f.data()));
    check(c, emfem_set_transmitters(c, ntx, tx.data()));
    check(c, emfem_set_receivers(c, nrx, rx.data()));
    check(c, emfem_set_observations(c, no, obs4.data(), val.data()));
  });
}

int emfem_save_rsp(emfem_context *c, const char *path) {
  return guarded(c, [&]() {
    FILE *fp = std::fopen(path, "w");
    EMF_REQUIRE(fp != nullptr, kErrUser, std::string("Unable to open file ") + path + ".");
    std::fprintf(fp, "# frequencies\n%d\n", (int)c->freqs.size());
    for (double f : c->freqs) std::fprintf(fp, "%.4E\n", f);
    std::fprintf(fp, "# transmiters\n%d\n", (int)c->tx.size() / 7);
    for (size_t i = 0; i < c->tx.size() / 7; ++i) {
      for (int j = 0; j < 7; ++j) std::fprintf(fp, "% .4E ", c->tx[i * 7 + j]);
      std::fprintf(fp, "\n");
    }
    std::fprintf(fp, "# recievers\n%d\n", (int)c->rx.size() / 3);
    for (size_t i = 0; i < c->rx.size() / 3; ++i)
      std::fprintf(fp, "% .4E % .4E % .4E\n", c->rx[i * 3], c->rx[i * 3 + 1], c->rx[i * 3 + 2]);
    std::fprintf(fp, "# observations\n%d\n", (int)c->rsp.size());
    std::fprintf(fp, "#%6s %6s %5s %7s %13s %13s %13s %13s\n", "otype", "fidx", "tidx", "ridx", "amp", "phase", "amp`",
                 "phase`");
    for (size_t i = 0; i < c->rsp.size(); ++i) {
      const int *o = &c->obs4[i * 4];
      if (is_ap(o[0])) {
        const cplx ob(std::exp(c->obs[i].real()), c->obs[i].imag() * kRtoD);
        const cplx rs(std::exp(c->rsp[i].real()), c->rsp[i].imag() * kRtoD);
        std::fprintf(fp, "% 7d % 6d % 5d % 7d % 13.6E % 13.3f % 13.6E % 13.3f\n", o[0], o[1], o[2], o[3], ob.real(),
                     ob.imag(), rs.real(), rs.imag());
      } else {
        std::fprintf(fp, "% 7d % 6d % 5d % 7d % 13.6E % 13.6E % 13.6E % 13.6E\n", o[0], o[1], o[2], o[3],
                     c->obs[i].real(), c->obs[i].imag(), c->rsp[i].real(), c->rsp[i].imag());
      }
    }
    std::fclose(fp);
  });
}

int emfem_save_mesh(emfem_context *c, const char *path, int tidx, const double *error0, const double *error1) {
  return guarded(c, [&]() {
    require_mesh(c);
    EMF_REQUIRE(path != nullptr, kErrArg, "null path");
    write_vtk(c, path, tidx, error0, error1);
  });
}

int emfem_get_responses(const emfem_context *c, double *rsp) {
  if (!c || !rsp) return EMFEM_ERR_ARG;
  for (size_t i = 0; i < c->rsp.size(); ++i) {
    rsp[i * 2] = c->rsp[i].real();
    rsp[i * 2 + 1] = c->rsp[i].imag();
  }
  return EMFEM_OK;
}

int emfem_forward(emfem_context *c) {
  return guarded(c, [&]() {
    require_mesh(c);
    require_device(c);
    EMF_REQUIRE(c->n_rho > 0, kErrState, "no resistivity table");
    if (c->coarse_tets.empty()) {
      c->coarse_pts = c->T.pts;
      c->coarse_tets = c->T.tets;
      c->coarse_attr = c->T.attr;
    }
    // update_background_model on the coarse mesh (src/em_fwd.cc:1456)
    restore_coarse(c);
    update_background(c->T, c->s6, c->bg);
    const Background bg = c->bg;
    // distribute_tasks with one group: CSEM tasks then MT tasks, each in (fidx, tidx)
    // order (src/em_fwd.cc:1273-1318)
    std::set<std::pair<int, int>> tasks;
    for (size_t o = 0; o < c->obs4.size() / 4; ++o) tasks.insert({c->obs4[o * 4 + 1], c->obs4[o * 4 + 2]});
    std::map<int, std::vector<int>> csem;
    std::vector<int> mt;
    for (auto &t : tasks) {
      EMF_REQUIRE(t.first >= 0 && t.first < (int)c->freqs.size(), kErrUser, "observation frequency index out of range");
      if (t.second >= 0) {
        EMF_REQUIRE((size_t)t.second * 7 < c->tx.size(), kErrUser, "observation transmitter index out of range");
        csem[t.first].push_back(t.second);
      } else {
        mt.push_back(t.first);
      }
    }
    c->stats["not_converged"] = 0;
    c->stats["forward_outer_its"] = 0;
    for (const char *k : {"system", "assemble", "pc", "solve", "estimate", "refine", "responses"})
      c->stats[std::string("fwd_") + k + "_ms"] = 0.0;
    const double t0 = now_ms();
    for (auto &kv : csem) {
      forward_csem(c, kv.first, kv.second);
      c->bg = bg;
    }
    for (int f : mt) {
      forward_mt(c, f);
      c->bg = bg;
    }
    c->stats["forward_ms"] = now_ms() - t0;
  });
}

int emfem_refine_mesh(emfem_context *c, const unsigned char *flags) {
  return guarded(c, [&]() {
    require_mesh(c);
    EMF_REQUIRE(flags != nullptr, kErrArg, "null flags");
    RawMesh m = current_mesh(c);
    refine_bisect(m, std::vector<uint8_t>(flags, flags + m.nt()));
    install_mesh(c, m);
  });
}

int emfem_refine_uniform(emfem_context *c) {
  return guarded(c, [&]() {
    require_mesh(c);
    RawMesh m = current_mesh(c);
    refine_uniform(m);
    install_mesh(c, m);
  });
}

int emfem_get_mesh(const emfem_context *c, double *points, int *tets, int *attributes) {
  if (!c) return EMFEM_ERR_ARG;
  if (!c->have_mesh) return EMFEM_ERR_STATE;
  if (points) std::copy(c->T.pts.begin(), c->T.pts.end(), points);
  if (tets) std::copy(c->T.tets.begin(), c->T.tets.end(), tets);
  if (attributes) std::copy(c->T.attr.begin(), c->T.attr.end(), attributes);
  return EMFEM_OK;
}

}  // extern "C"
