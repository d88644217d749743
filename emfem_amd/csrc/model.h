// Conductivity table, 1-D plane-wave boundary data and host right-hand-side builders.
//
// Reference: get_cell_attribute (src/em_fwd.cc:14-72), PlaneWaveFunction
// (src/em_pw.h:6-139), generate_layered_model / update_background_model
// (src/em_fwd.cc:74-159), calculate_boundary_values / assemble_rhs_mt
// (src/em_fwd.cc:402-580), divide_line_source / assemble_rhs_csem
// (src/em_fwd.cc:582-671), assemble_rhs_dual (src/em_fwd.cc:673-717) and the
// receiver formulas (src/em_fwd.cc:719-871).
#pragma once

#include <complex>
#include <functional>
#include <vector>

#include "topology.h"

namespace emfem {

using cplx = std::complex<double>;

// Per-attribute symmetric conductivity (xx, yy, zz, xy, xz, yz) from the
// resistivity table; throws kErrUser for an unsupported form.
void make_sigma_table(int aniso_form, int n_rho, const double *rho, std::vector<double> &s6);

int rho_columns(int aniso_form);

// Layered 1-D model under one top corner: layer i starts at ztop[i].
struct Column {
  std::vector<double> ztop, lsig;
};

struct Background {
  double corners[4][3] = {};
  Column col[4];
  bool ready = false;
};

// update_background_model: walks the 4 vertical corner columns of the mesh.
void update_background(const Topology &T, const std::vector<double> &s6, Background &bg);

// 1-D layered plane wave normalised at the top of the model.
class PlaneWave {
 public:
  void init(const Column &c, double freq);
  // eh = (E_x, H_y, E_y, H_x) of the e^{-i w t} solution at depth z.
  void field(double z, cplx eh[4]) const;
  // E in the e^{+i w t} convention for polarisation mode (-1 XY, -2 YX).
  void efield(double z, int mode, cplx e[3]) const;

 private:
  int n_ = 0;
  double omega_ = 0;
  std::vector<double> ztop_, lsig_;
  std::vector<cplx> a_, b_, c_, d_;
};

// MT right-hand side (s_re, s_im) on the host.  bval receives the Dirichlet values
// of every boundary edge (NaN-free; interior entries stay 0).
void rhs_mt(const Topology &T, const std::vector<double> &s6, const Background &bg, double freq, int mode,
            std::vector<double> &s_re, std::vector<double> &s_im, std::vector<cplx> *bval = nullptr);

// Point location for a batch of points [n][3]: containing tet per point, -1 when
// outside (find_cell_around_point semantics).
using Locate = std::function<std::vector<int>(const std::vector<double> &pts)>;

// Locate by the host breadth-first search, one point at a time.
Locate host_locate(const Topology &T);

// CSEM right-hand side for one transmitter record (7 numbers).
void rhs_csem(const Topology &T, double freq, const double *tx, int n_div, const Locate &loc,
              std::vector<double> &s_re, std::vector<double> &s_im);

// Dual right-hand side summed over the receivers.
void rhs_dual(const Topology &T, const std::vector<double> &rx, const Locate &loc, std::vector<double> &s_re,
              std::vector<double> &s_im);

// E and H at p inside tet t for the complex edge solution e (interleaved re, im).
void interpolate(const Topology &T, int t, const double *p, double freq, const double *e_ri, cplx E[3],
                 cplx H[3]);

// MT transfer function / CSEM field for an observation code; returns false for a
// code with no definition.
bool mt_response(int otype, double freq, const cplx exy[3], const cplx hxy[3], const cplx eyx[3],
                 const cplx hyx[3], cplx &out);
bool csem_response(int otype, const cplx E[3], const cplx H[3], cplx &out);

}  // namespace emfem
