// Device solvers: auxiliary-space preconditioner cycles, the inner flexible CG on
// B = C + M and the outer FGMRES on the reference's real-equivalent operator.
//
// Reference: pc_apply_b (src/em_la.cc:208-228; B_ksp = FCG, rtol 1e-2, max 100,
// unpreconditioned norm, src/em_la.cc:134-142) and solve_linear_system with the
// FGMRES A_ksp (src/em_la.cc:35-63,175-188).  The inner solve runs in precision T
// (fp32 in the mixed path, fp64 in the plain path); the outer FGMRES, its
// Gram-Schmidt and the true residual are fp64.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "amg.h"
#include "common.h"
#include "kernels.h"

namespace emfem {

struct SolveStats {
  int fell_back = 0;         // 1 if the solve switched to the fp64 cycle
  int outer_its = 0;         // max over right-hand sides
  long inner_its = 0;        // FCG iterations summed over preconditioner applications
  long inner_col_its = 0;    // FCG iterations summed over the columns still active in them
  long pc_applies = 0;       // preconditioner applications (each on 2k columns)
  long vcycles = 0;          // AMS cycles
  std::vector<int> its;      // per right-hand side
  std::vector<double> relres;
  std::vector<uint8_t> converged;  // the reference's stopping test was met
  int floor_accepts = 0;           // right-hand sides accepted at the fp64 residual floor
  std::vector<double> done_ms;     // per right-hand side: time from the start until it was accepted
};

// ms per launch of the complex operator on nrhs interleaved pairs (x -> y), from HIP
// events around `reps` back-to-back launches on stream s.
double time_operator(hipStream_t s, int ne, int nnz, const int *rp, const int *ci, const double2 *K, int nrhs,
                     const double *x, double *y, int reps);

struct PcOptions {
  int cheb_degree = 4;
  double cheb_ratio = 60.0;
  bool cheb_fourth = false;  // fourth-kind Chebyshev (no ratio) instead of first kind
  double inner_rtol = 1e-2;
  int inner_max_it = 100;
  AmgParams amg;
  bool use_graph = true;  // replay each auxiliary-space cycle as one captured hipGraph
  // smoother SpMM row mapping: 0 = tuned, 1 = lanes per row only (tuned; lanes on
  // columns at 16 or more columns), 2 = row blocks when possible (for the fp64 FCG
  // product too), 3 = LDS x-blocks
  // when possible, 4 = persistent pipelined row blocks when possible, 10 + n = n
  // lanes per row
  int spmm_rows = 0;
  // auxiliary-space transfers by the CSR matrices Pi / Pi^T instead of the structured
  // kernels (the same operator)
  bool aux_csr = false;
  // row-block SpMMs of B read 16-bit columns (compress_columns) when the pattern
  // allows; the results are the same bit for bit
  bool col16 = true;
};

// Preconditioner for B on nc real columns: flexible CG in fp64 with the
// auxiliary-space cycle in fp32 or fp64 (both built from one host set-up; the
// precision not asked for at creation is set up on first use).
class InnerPc {
 public:
  virtual ~InnerPc() = default;
  // x = approx B^-1 b by FCG + AMS; b, x are fp64 [E][nc] compact.
  virtual void apply(hipStream_t s, int nc, const double *b, double *x, SolveStats &st) = 0;
  // rows split into independent segments (a multi-model batch): the FCG runs per
  // (segment, column); the cycle itself acts on the block-diagonal operator as is
  virtual void set_segments(const Seg &sg) = 0;
  virtual void set_fp32(bool fp32, hipStream_t s) = 0;
  virtual bool fp32() const = 0;
  // replay the cycles as captured hipGraphs (off: plain launches, timed per SpMM)
  virtual void set_use_graph(bool g) = 0;
  // ms per launch of the B SpMM of the Chebyshev smoother (cycle precision) and of
  // the fp64 FCG on nc columns, from HIP events around `reps` back-to-back launches
  virtual double time_smoother(hipStream_t s, int nc, int reps) = 0;
  // in-cycle timing of the smoother SpMM with HIP events around each launch (graph
  // replays are not timed): on = start; off = stop and return ms per launch
  virtual double smoother_clock(bool on, hipStream_t s) = 0;
  virtual double time_fcg(hipStream_t s, int nc, int reps) = 0;
  // row mapping the smoother SpMM uses on nc columns: lanes per row, -1 = row blocks,
  // 0 = many-column kernel
  virtual int smoother_rows(hipStream_t s, int nc) = 0;
  // bytes of column index per nonzero the smoother SpMM (fcg false) or the fp64 FCG
  // product (fcg true) reads on nc columns (4, or 2 + 4 / 64 on 16-bit columns)
  virtual double index_bytes(hipStream_t s, int nc, bool fcg) = 0;
  // whether B's row-block SpMMs read the 16-bit column form
  virtual bool col16() const = 0;
  // Takes the set-up H of the next frequency of the same linear system (same pattern,
  // same auxiliary spaces) and the new (C, M) values K, keeping the device state that
  // depends only on the mesh; false (the preconditioner is then unusable) when H or
  // the options o differ in what that state was built for.
  virtual bool refresh(const AmsHost &H, const double2 *K, const PcOptions &o, hipStream_t s) = 0;
};

// The 16-bit column form of B's pattern (compress_columns) when its owner built it
// once for the pattern: built = true, off / base null when the pattern does not fit.
// built = false: each preconditioner builds its own (PcOptions::col16).
struct Col16 {
  bool built = false;
  const uint16_t *off = nullptr;
  const int *base = nullptr;
};

std::unique_ptr<InnerPc> make_inner_pc(bool fp32, const AmsHost &H, int ne, int nnz, const int *d_rp,
                                       const int *d_ci, const double2 *d_K, const PcOptions &o, hipStream_t s,
                                       const Col16 &c16 = Col16());

struct OuterOptions {
  int restart = 30;
  int max_it = 100;
  // switch the whole solve to the fp64 cycle when an fp32-preconditioned right-hand
  // side stops converging (less than halving its residual over `stall_window`
  // iterations)
  bool fp64_fallback = true;
  int stall_window = 15;
  // A right-hand side whose implicit residual met rtol while its true residual did not
  // is restarted from the true residual until that stops halving; it is then accepted
  // only if the true residual is within floor_factor * rtol * ||b||, and otherwise
  // reported as not converged.
  double floor_factor = 10.0;
  // warm start: x holds the initial guess on entry (otherwise the solve starts from 0)
  bool x_given = false;
  // > 0: return once at most this many right-hand sides are still unconverged (the
  // Arnoldi cycle in progress is closed when that many are still active, and the
  // following restart's true-residual check returns); 0 runs to completion
  int stop_remaining = 0;
  // Outer preconditioner.  false: the reference's block-diagonal diag(B, B)
  // (pc_apply_b, src/em_la.cc:208-228), one inner solve on the 2k real columns.
  // true: PRESB (Axelsson, Neytcheva and Ahmad, "A comparison of iterative methods to
  // solve complex valued linear algebraic systems", Numer. Algorithms 66, 2014) for
  // the block form [[C, -M], [M, C]] = S A, S = diag(I, -I): z = P^-1 S v with
  // h = B^-1 (f + g), y = B^-1 (g - M h), z = (h - y, y) for S v = (f, g) -- two
  // inner solves with the same B on k columns each, in sequence.  The spectrum of the
  // preconditioned operator lies in [1/2, 1], which halves the outer iterations of
  // these systems (measured with exact inner solves: marine CSEM 22 -> 11, adaptive MT
  // 10-16 -> 5-9).
  bool presb = false;
};

// Device workspace of the outer FGMRES, kept by the caller across solves so that
// repeated solves allocate nothing: the Krylov basis V (min(restart, max_it) + 1
// blocks), the preconditioned directions Z (min(restart, max_it) blocks), one work
// block and the reduction scratch.  Buffers only grow.
struct FgmresWork {
  DevBuf<double> V, Z, w, dots, coef, scal, work;
  DevBuf<double> pu, ph, py, px;  // PRESB: k-column right-hand sides and solutions, a 2k-column block
  // total device bytes held
  size_t bytes() const {
    return V.bytes() + Z.bytes() + w.bytes() + dots.bytes() + coef.bytes() + scal.bytes() + work.bytes() +
           pu.bytes() + ph.bytes() + py.bytes() + px.bytes();
  }
};

// FGMRES on nrhs interleaved right-hand sides b [E][2 nrhs] -> x [E][2 nrhs].  With
// segments (a block-diagonal multi-model batch) every (segment, right-hand side) is
// an independent system with its own Arnoldi process and stopping test; rtol and the
// per-RHS statistics are then indexed seg * nrhs + k.
// ws (may be null: a workspace local to the call) holds the device buffers.
void fgmres(hipStream_t s, int ne, int nnz, const int *d_rp, const int *d_ci, const double2 *d_K, InnerPc &pc,
            int nrhs, const double *b, double *x, const std::vector<double> &rtol, const OuterOptions &o,
            SolveStats &st, const Seg &sg = Seg(), FgmresWork *ws = nullptr);

// Grows d to at least n elements (contents are not kept when it grows).
template <typename T>
void grow(DevBuf<T> &d, size_t n) {
  if (d.size() < n) d.resize(n);
}

}  // namespace emfem
