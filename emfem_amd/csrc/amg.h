// Auxiliary-space Maxwell preconditioner: host set-up.
//
// The reference hands B = C + M, the discrete gradient G (src/em_la.cc:230-295) and
// the vertex coordinates to HYPRE's AMS (src/em_la.cc:143-146), whose internals are
// not part of the reference.  This is a from-scratch auxiliary-space
// preconditioner in the same family: Chebyshev smoothing on the edge space plus
// corrections in the nodal gradient space G and the nodal vector space Pi, each
// solved approximately by one V-cycle of aggregation AMG.  Boundary vertices are left out
// of the auxiliary spaces because the Dirichlet edges they span are decoupled in B
// (src/em_fwd.cc:377-388).  The set-up (aggregation, prolongators, Galerkin
// products, dense coarsest inverse) runs here on the host in fp64; the cycles run on
// the device.
#pragma once

#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "topology.h"

namespace emfem {

// Allocator whose resize leaves new elements uninitialised (default-initialised): the
// large host arrays of the set-up are then first written, and their pages first
// touched, by the parallel loops that fill them instead of by a single-threaded fill
// with zeros.  Explicit values (assign, resize with a value) behave as usual.
template <typename T>
struct NoInitAlloc : std::allocator<T> {
  template <typename U>
  struct rebind {
    using other = NoInitAlloc<U>;
  };
  NoInitAlloc() = default;
  template <typename U>
  NoInitAlloc(const NoInitAlloc<U> &) {}
  template <typename U>
  void construct(U *p) {
    ::new (static_cast<void *>(p)) U;
  }
  template <typename U, typename... A>
  void construct(U *p, A &&...a) {
    ::new (static_cast<void *>(p)) U(std::forward<A>(a)...);
  }
};
template <typename T>
using HostVec = std::vector<T, NoInitAlloc<T>>;

struct HostCsr {
  int n = 0, m = 0;  // rows, columns
  HostVec<int> rp, ci;
  HostVec<double> v;
  int64_t nnz() const { return (int64_t)ci.size(); }
};

HostCsr csr_transpose(const HostCsr &A);
HostCsr csr_multiply(const HostCsr &A, const HostCsr &B);
HostCsr csr_triple(const HostCsr &R, const HostCsr &A, const HostCsr &P);

struct AmgLevelHost {
  HostCsr A, P, R;           // R = P^T
  std::vector<double> l1inv; // 1 / sum_j |a_ij|
};

struct AmgHost {
  std::vector<AmgLevelHost> levels;  // levels[0].A is the fine matrix
  int ncoarse = 0;
  // dense inverse of the coarsest matrix, transposed: row j holds column j (ncoarse x
  // ncoarse); coarse_span[2 i], coarse_span[2 i + 1]: first and one past the last row of
  // row i's connected component (the inverse is zero outside it)
  std::vector<double> coarse_inv_t;
  std::vector<int> coarse_span;
};

struct AmgParams {
  double theta = 0.08;     // strength threshold |a_ij| >= theta sqrt(a_ii a_jj)
  int max_coarse = 256;    // dense solve at or below this size
  int max_levels = 20;
  int smooth_levels = 0;   // levels with a Jacobi-smoothed prolongator (then plain aggregation)
  bool additive = true;    // auxiliary cycle 0-(G+Pi)-0 instead of 0-G-Pi-G-0
  std::map<std::string, double> *timing = nullptr;  // optional phase times (ms)
};

// A is taken by value: callers that no longer need it move it in.
void build_amg(HostCsr A, const AmgParams &p, AmgHost &H);

// Mesh-only parts of the auxiliary spaces: they depend on the topology and vertex
// coordinates, not on B, so one set serves every frequency on a mesh.
struct AmsAux {
  // E x V0 gradient restricted to interior vertices and its transpose
  HostCsr G0, G0T;
  // nodal-vector components Pi_d (E x V0) and their transposes
  HostCsr P[3], PT[3];
  // Auxiliary blocks side by side (build_pi_csr): E x (nblk V0), column k * V0 + v for
  // block k of interior vertex v.  Blocks are the three nodal-vector components Pi_d
  // (multiplicative cycle) or G0 followed by Pi_x, Pi_y, Pi_z (additive cycle).
  // Corrections of these blocks see the same residual, so their auxiliary problems
  // form one block-diagonal matrix diag(P_k^T B P_k) with one AMG hierarchy whose
  // aggregates never cross blocks.
  int nblk = 3;
  bool additive = false;
  int nv0 = 0;
  // pattern of every P_k^T B P_k (interior vertices sharing a tet, columns ascending)
  // and of their block-diagonal stack (values left empty)
  HostCsr Apat, APpat;
};

std::shared_ptr<const AmsAux> build_ams_aux(const Topology &T, bool additive);

// The blocks of X side by side as one CSR matrix Pi and its transpose (the
// interpolation of the CSR transfers, -ams_transfer csr; the structured transfers read
// G0 and Pi_d directly).
void build_pi_csr(const AmsAux &X, HostCsr &Pi, HostCsr &PiT);

// vals[k] = values of P_k^T B P_k on X.Apat, k = G0, Pi_x, Pi_y, Pi_z (the first form
// writes into caller arrays of Apat's nonzero count).
void galerkin_blocks(const AmsAux &X, const HostCsr &B, double *const vals[4]);
void galerkin_blocks(const AmsAux &X, const HostCsr &B, std::vector<double> (&vals)[4]);

struct AmsHost {
  std::shared_ptr<const AmsAux> aux;
  // storage of an earlier set-up's block-diagonal Galerkin stack, reused by build_ams,
  // and the auxiliary spaces it was built for (same spaces: its pattern is kept)
  HostCsr recycle;
  std::shared_ptr<const AmsAux> recycle_aux;
  AmgHost amgG;   // AMG of G0^T B G0 (multiplicative cycle only)
  AmgHost amgPi;  // AMG of diag(P_k^T B P_k)
  std::vector<double> l1inv;  // edge-space l1 row norms of B, inverted
  std::map<std::string, double> timing;  // set-up phase times (ms)
};

// B given on the edge pattern (fp64 values).  `aux` from an earlier call on the same
// mesh (same additive flag) is reused; otherwise it is built.
void build_ams(const Topology &T, const HostCsr &B, const AmgParams &p, AmsHost &H,
               std::shared_ptr<const AmsAux> aux = nullptr);

// The same with the four Galerkin value arrays (as galerkin_blocks writes them) filled
// by `fill` instead of from a host copy of B; H.l1inv is left empty (the device forms
// the edge-space l1 scaling itself).
using GalerkinFill = std::function<void(const AmsAux &, double *const[4])>;
void build_ams(const Topology &T, const GalerkinFill &fill, const AmgParams &p, AmsHost &H,
               std::shared_ptr<const AmsAux> aux = nullptr);

}  // namespace emfem
