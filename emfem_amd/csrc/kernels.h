// Launch wrappers of the device kernels (implemented in *.hip).
//
// Block vectors are row-major [rows][ld] with `nc` used columns.  Complex
// right-hand sides are stored as interleaved (re, im) column pairs, so a block of
// k complex vectors has nc = 2k real columns; the real-equivalent vector of the
// reference's VecNest {re, im} (src/em_la.cc:42-47) is one such column pair.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace emfem {

// CSR matrix view (device pointers).
template <typename T>
struct Csr {
  int nrows = 0, ncols = 0, nnz = 0;
  const int *rp = nullptr;
  const int *ci = nullptr;
  const T *v = nullptr;
  int lpr = 0;  // lanes per row of the narrow SpMM kernel; 0 = chosen from the row length
  int role = 0; // 1 = edge-space smoother operator (a separate kernel instance in traces)
  // lpr = -1: LDS-staged row blocks (-3: the same, persistent and software-pipelined);
  // blk_row [2 (nblk + 1)] are the first rows of the blocks and then their first
  // nonzeros (row_blocks), for the column count and precision the table was built for
  const int *blk_row = nullptr;
  int nblk = 0;
  // 16-bit column form read by the row blocks when set (see compress_columns): the
  // column of nonzero p is ci16_base[p / 64] + ci16[p]
  const uint16_t *ci16 = nullptr;
  const int *ci16_base = nullptr;
  // lpr = -2 (16 or more columns): LDS x-blocks of 64-column slabs.
  // Block b takes the rows [xb_row[b], xb_row[b+1]); its distinct columns
  // xb_ucol[xb_uptr[b] .. xb_uptr[b+1]) (at most xb_umax, ascending) are staged in LDS
  // and xb_lci [nnz] holds each nonzero's slot among them.
  const int *xb_row = nullptr, *xb_uptr = nullptr, *xb_ucol = nullptr;
  const int *xb_lci = nullptr;
  int xb_nblk = 0, xb_umax = 0;
  // lpr = -2, -4, -8 (fewer than 16 columns): narrow x-blocks with -lpr lanes per row;
  // tables as above with at most kXnCols distinct columns per block
  const int *xn_row = nullptr, *xn_uptr = nullptr, *xn_ucol = nullptr, *xn_lci = nullptr;
  int xn_nblk = 0;
};

// Host tables of the LDS x-block SpMM for a pattern (see Csr::xb_*): consecutive rows
// while their distinct columns number at most umax and the rows at most rmax.  Empty
// when a single row has more than umax distinct columns.
struct XBlocks {
  std::vector<int> row, uptr, ucol;
  std::vector<int> lci;
  int umax = 0;
  bool empty() const { return row.size() < 2; }
};
XBlocks x_blocks(const std::vector<int> &rp, const std::vector<int> &ci, int ncols, int umax, int rmax);
constexpr int kXBlockCols = 64;  // distinct columns per x-block
constexpr int kXBlockRows = 64;
constexpr int kXnCols = 256;  // distinct columns per narrow x-block
constexpr int kXnRows = 64;

// Threads per workgroup of the row-block SpMM: one wavefront for fp32 blocks of 3 to
// 6 columns (seven or nine nonzeros per thread, all of whose loads are in flight
// together), else 256.
template <typename T, int NC>
constexpr int stream_block() {
  return sizeof(T) == 4 && NC >= 3 && NC <= 6 ? 64 : 256;
}

// Nonzeros staged per workgroup by the row-block SpMM: for the one-wavefront blocks
// 576 at 3 columns (nine nonzeros per thread; at 1.54M edges 49.8 us per launch
// against 51.6 for 448, 50.7 for 512 and 50.2 for 640) and 448 at 4 to 6 (seven),
// 768 for fp64 blocks of 3 columns (80.6 us against 84.6 us for 1024), else 16 KB of
// products in fp32, 24 KB in fp64 (a multiple of the 256-thread block, at most 2048).
template <typename T, int NC>
constexpr int stream_tile() {
  constexpr int bytes = sizeof(T) == 8 ? 24576 : 16384;
  constexpr int t = (bytes / (NC * (int)sizeof(T))) / 256 * 256;
  return stream_block<T, NC>() == 64 ? (NC == 3 ? 9 * 64 : 7 * 64)
                                     : (sizeof(T) == 8 && NC == 3 ? 768 : (t > 2048 ? 2048 : t));
}
int stream_tile_for(int nc, bool fp64);

// Row blocks for the row-block SpMM: consecutive rows while their nonzeros fit in
// `tile`; returns the first row of every block and nrows (nblk + 1 entries), followed
// by the first nonzero of every block and nnz (nblk + 1 entries); empty when a row is
// longer than the tile.
std::vector<int> row_blocks(const std::vector<int> &rp, int tile);

// 16-bit column form of a CSR pattern: the nonzeros in aligned groups of kCol16Group,
// each group's smallest column in base[ceil(nnz / kCol16Group)] and every column as its
// offset from that in off[nnz].  The row-block SpMM then streams 2 + 4 / 64 bytes of
// index per nonzero instead of 4.  Returns false (base and off unusable) when some
// group spans more than 65536 columns; the pattern then keeps its 32-bit columns.
// flag is one device int of scratch.  Synchronises s.
constexpr int kCol16Group = 64;
bool compress_columns(hipStream_t s, int nnz, const int *ci, int *base, uint16_t *off, int *flag);

// y = P xc + zs .* z for a P with exactly one entry per row (column pci[row], value
// pv[row]): the plain-aggregation prolongation of the AMG cycle as one elementwise
// pass; vectors [rows][nc], xc [*][nc].
template <typename T>
void prolong_single(hipStream_t s, int rows, int nc, const int *pci, const T *pv, const T *xc, const T *zs,
                    const T *z, T *y);

// out[p] = v[p] * d[ci[p]]: the values of A diag(d) on A's pattern.
template <typename T>
void scale_columns(hipStream_t s, int nnz, const int *ci, const T *v, const T *d, T *out);

// A Chebyshev step fused into an SpMM epilogue (off when x is null): the SpMM's value
// v is not stored; y(row, j) receives x + c1 (x - d) + c2 v at (row, j) instead (in T
// with the coefficients rounded to T; d null = 0).  x is normally the SpMM's input
// block and d the output block (each read only at its own row); d and x are
// [rows][ld].
// Independently of the step, zcol (device, indexed by the column j of [rows][ld])
// scales the SpMM's z term per column: beta * zcol[j] * z(row, j).
template <typename T>
struct ChebFuse {
  const T *d = nullptr, *x = nullptr;
  int ld = 0;
  double c1 = 0.0, c2 = 0.0;
  const double *zcol = nullptr;
  __host__ __device__ bool on() const { return x != nullptr; }
};

// Picks the fastest row mapping of the narrow SpMM for A on nc columns by timing
// each with HIP events on stream s: lanes per row 2, 4, 8, 16, and -1 (LDS-staged
// row blocks) when A carries a block table for nc; for 16 or more columns the slab
// width of the wide kernel (16 .. 256) and, when A carries x-block tables, -2 and -3
// (LDS x-blocks).  x and y are scratch blocks of [A.ncols][nc] and [A.nrows][nc].
template <typename TA, typename TX, typename T>
int tune_lpr(hipStream_t s, const Csr<TA> &A, int nc, const TX *x, T *y, const T *rs = nullptr,
             const ChebFuse<T> &cf = ChebFuse<T>());

// y = rs .* (alpha * (A x) + gamma * w) + beta * z   (rows of A; rs may be null = 1,
// w / z may be null when their coefficient is 0).  y may alias z (same row only) but
// must not alias x.
// Matrix values TA, input TX, output and accumulation T; instantiated for
// (f32,f32,f32), (f64,f64,f64), (f64,f32,f64) and (f32,f64,f32).
template <typename TA, typename TX, typename T>
void spmm_mixed(hipStream_t s, const Csr<TA> &A, int nc, const TX *x, int ldx, T *y, int ldy, T alpha, const T *rs,
                T gamma, const T *w, int ldw, T beta, const T *z, int ldz);


// spmm_mixed with two optional per-row scalings: the input is read as
// xs[col] * x[col][:] (the operator A diag(xs)) and z as zs[row] * z[row][:]; either
// may be null (= 1); and an optional fused Chebyshev step.
template <typename TA, typename TX, typename T>
void spmm_ex(hipStream_t s, const Csr<TA> &A, int nc, const TX *x, int ldx, T *y, int ldy, T alpha, const T *rs,
             T gamma, const T *w, int ldw, T beta, const T *z, int ldz, const T *xs, const T *zs,
             const ChebFuse<T> &cf = ChebFuse<T>());

template <typename T>
inline void spmm(hipStream_t s, const Csr<T> &A, int nc, const T *x, int ldx, T *y, int ldy, T alpha, const T *rs,
                 T gamma, const T *w, int ldw, T beta, const T *z, int ldz) {
  spmm_mixed<T, T, T>(s, A, nc, x, ldx, y, ldy, alpha, rs, gamma, w, ldw, beta, z, ldz);
}

// Complex operator of the reference's MatShell A (src/em_la.cc:65-87) on nrhs
// interleaved pairs: for each pair (xr, xi)
//   yr = C xr - M xi,  yi = -M xr - C xi ;  y = alpha * (A x) + beta * z.
// K holds (C, M) per nonzero.  x may be fp32 (preconditioned directions) or fp64.
template <typename TX>
void op_spmm(hipStream_t s, int nrows, int nnz, const int *rp, const int *ci, const double2 *K, int nrhs, const TX *x,
             int ldx, double *y, int ldy, double alpha, double beta, const double *z, int ldz);

// Element assembly of one colour of tets (atomic-free read-modify-write scatter).
void assemble_colour(hipStream_t s, int ntets, const int *tet_ids, const int *tets, const int *tet2edge,
                     const double *pts, const int *attr, const double *sig6, const uint8_t *edge_bdr, double omega,
                     const int *rp, const int *ci, double2 *K);

// B = C + M in precision T (the SPD operator of the inner solve, src/em_la.cc:122-128).
template <typename T>
void make_b(hipStream_t s, int nnz, const double2 *K, T *B);

// Rows split into independent segments (the models of a multi-model batch, stacked
// block-diagonally).  Per-column scalars are then per (segment, column) at
// seg * nc + c ("logical columns", nl = nseg * nc), and reductions run over chunks
// of rows that never straddle a segment.  The default is one segment with no tables.
struct Seg {
  int nseg = 1;
  int nchunks = 0;                 // reduction chunks (0: an even split of all rows)
  const int *rseg = nullptr;       // [rows] segment of each row (null: all 0)
  const int *chunk_row = nullptr;  // [nchunks + 1] first row of each chunk
  const int *seg_chunk = nullptr;  // [nseg + 1] first chunk of each segment
};

// Per-column dot products sum_r x[r][c] * y[r][c], written (in double) to
// out[seg * nc + c] on the device.  `work` must hold dot_work_size(rows, nc, sg)
// doubles.  Results are bitwise reproducible (fixed summation order).
size_t dot_work_size(int rows, int nc, const Seg &sg = Seg());
template <typename TX, typename TY = TX>
void coldot(hipStream_t s, int rows, int nc, const TX *x, int ldx, const TY *y, int ldy, double *out, double *work,
            const Seg &sg = Seg());

// Device copy of the mesh-only matrices of the auxiliary-space Galerkin products: the
// restrictions R_k = (G0^T, Pi_x^T, Pi_y^T, Pi_z^T) share one pattern (rrp, rci), the
// interpolations P_k = (G0, Pi_x, Pi_y, Pi_z) another (prp, pci), and the four
// products the pattern Apat (arp, aci) on the nv0 interior vertices.
struct GalerkinDev {
  int nv0 = 0;
  const int *rrp = nullptr, *rci = nullptr, *prp = nullptr, *pci = nullptr, *arp = nullptr, *aci = nullptr;
  const double *rv[4] = {nullptr, nullptr, nullptr, nullptr}, *pv[4] = {nullptr, nullptr, nullptr, nullptr};
};
// vals[k] (Apat layout) = P_k^T B P_k for B on the edge pattern (brp, bci, bv): one
// thread per vertex row, each value summed in the order of the host's galerkin_blocks
// with separately rounded products and sums (bitwise equal to it).  missing (device
// int, zeroed by the call) counts couplings of B outside Apat.
void galerkin_dev(hipStream_t s, const GalerkinDev &G, const int *brp, const int *bci, const double *bv,
                  double *const vals[4], int *missing);
// l1[row] = T(1 / sum_j |v_j|) over the row in order (0 for a zero row), as the host's
// l1_inverse followed by a conversion to T.
template <typename T>
void l1_inverse_dev(hipStream_t s, int rows, const int *rp, const double *v, T *l1);

// The auxiliary-space transfers through the structure of the Nedelec interpolation
// Pi = [G0 | Pi_x | Pi_y | Pi_z] (nblk = 4; [Pi_x | Pi_y | Pi_z] when nblk = 3): edge
// e = (a, b) couples the interior vertices among its endpoints with -1 (a) / +1 (b) in
// G0 and with t_d / 2 (t = x_b - x_a) in Pi_d, so the 8-entry rows of Pi need only
// the endpoints and the half tangent.
//   ev[e]    interior indices (a', b') of the endpoints, -1 on the boundary
//   th[e][4] (t_x, t_y, t_z) / 2 and a zero pad
//   vrp/vcode  the edges of each interior vertex in ascending order (vrp [nv0 + 1]),
//              bit 31 set where the vertex is the edge's first endpoint
template <typename T>
struct AuxMap {
  int ne = 0, nv0 = 0, nblk = 0;
  const int2 *ev = nullptr;
  const int *vrp = nullptr, *vcode = nullptr;
  const T *th = nullptr;
};
// vb = Pi^T r: vb [nblk nv0][nc] (block k at rows k nv0), r fp64 [ne][nc]; each vertex
// sums its edges in ascending order, in T.
template <typename T>
void aux_restrict(hipStream_t s, const AuxMap<T> &M, int nc, const double *r, T *vb);
// y += Pi vx: vx [nblk nv0][nc], y [ne][nc].
template <typename T>
void aux_prolong(hipStream_t s, const AuxMap<T> &M, int nc, const T *vx, T *y);

// out1 = x . y1 and out2 = x . y2 per column in one pass over x (each bitwise equal to
// the corresponding coldot); work as for coldot, twice.
template <typename TX>
void coldot2(hipStream_t s, int rows, int nc, const TX *x, int ldx, const double *y1, const double *y2, int ldy,
             double *out1, double *out2, double *work, const Seg &sg = Seg());

// Multi-dot for Gram-Schmidt: out[i*nl + seg*nc + c] = sum_r V_i[r][c] w[r][c] for
// i < m, where V_i = V + i * vstride.  work must hold dot_work_size(rows, nc, sg) * 8
// doubles.
void mdot(hipStream_t s, int rows, int nc, const double *V, size_t vstride, int m, const double *w, double *out,
          double *work, const Seg &sg = Seg());

// w[r][c] -= sign * sum_{i<m} h[i*hld + seg(r)*nc + c] * V_i[r][c]
template <typename TB>
void mupdate(hipStream_t s, int rows, int nc, const TB *V, size_t vstride, int m, const double *h, int hld,
             double *w, double sign, const Seg &sg = Seg());

// y[r][c] = a[k] * x[r][c] + b[k] * y[r][c], k = seg(r)*nc + c (a, b: device
// doubles; b may be null = 0)
template <typename TX, typename TY>
void axpby_cols(hipStream_t s, int rows, int nc, const double *a, const TX *x, int ldx, const double *b, TY *y,
                int ldy, const Seg &sg = Seg());

// y[r][c] = alpha * x[r][c] + beta * y[r][c] with host scalars (beta = 0 ignores y).
template <typename TX, typename TY>
void axpby(hipStream_t s, int rows, int nc, double alpha, const TX *x, int ldx, double beta, TY *y, int ldy);

// Two columns of y [rows][ldy]: zero every row, then y[idx[i]][0..1] = (v[i].x, v[i].y)
// for i < n (distinct idx).
void put_rows(hipStream_t s, int rows, int n, const int *idx, const double2 *v, double *y, int ldy);

// Split {re, im} blocks of n vectors (vector j at offset j * rows) to interleaved
// [rows][2 n] pairs, and back.
void split_to_interleaved(hipStream_t s, int rows, int n, const double *re, const double *im, double *y);
void interleaved_to_split(hipStream_t s, int rows, int n, const double *x, double *re, double *im);

// The residual step of the flexible CG on nc columns with device scalars, per
// logical column k = seg(row) * nc + column:
//   alpha[k] = active[k] && pq[k] != 0 ? rz[k] / pq[k] : 0 (stored for the delayed
//   solution update), r -= alpha[k] q, rr[k] = r . r (fused, fixed-order reduction),
// then the convergence test on the device: active[k] = 0 once rr[k] <= tol2 * bb[k].
// rr and the updated flags are also written to h_rr and h_active, pinned host arrays
// the caller reads after synchronising s.
// With cf non-null the same pass also writes the fp32 auxiliary-space cycle's first
// inputs for the new r: cf[0] = (float) r and cf[1] = (float) (c0 * (double) (dinv[row]
// * (float) r)), the cycle's copy of its right-hand side and first Chebyshev iterate.
struct CycleFill {
  float *bt = nullptr, *x1 = nullptr;
  const float *dinv = nullptr;
  double c0 = 0.0;
};
void fcg_residual(hipStream_t s, int rows, int nc, const double *rz, const double *pq, uint8_t *active,
                  double *alpha, const double *q, double *r, double *rr, const double *bb, double tol2, double *h_rr,
                  uint8_t *h_active, double *work, const Seg &sg = Seg(), const CycleFill &cf = CycleFill());
// x += alpha[k] p ; p = z + beta p with beta = active[k] && pq[k] != 0 ? -zq[k] / pq[k]
// : 0 formed per element (k = seg(row) * nc + column)
template <typename TZ>
void fcg_px(hipStream_t s, int rows, int nc, const double *alpha, const double *zq, const double *pq,
            const uint8_t *active, const TZ *z, double *p, double *x, const Seg &sg = Seg());

// x = inv_theta * dinv[row] .* T(b)  (the first Chebyshev iterate from zero), and
// bt = T(b) when bt is not null (the right-hand side in the cycle's precision)
template <typename T>
void cheb_first(hipStream_t s, int rows, int nc, const T *dinv, const double *b, T *bt, T *x, double inv_theta);

// x = Dinv .* b (one l1-Jacobi sweep from a zero guess)
template <typename T>
void diag_scale(hipStream_t s, int rows, int nc, const T *dinv, const T *b, T *x);

// Dense coarse solve in fp64: x[i][c] = sum_j AinvT[j][i] b[j][c], AinvT being the
// transposed inverse (row j holds column j of the inverse); n <= kDenseMax.  range
// (null: all of j) gives for every 64-row block i0 = 64 k the j-interval
// [range[2k], range[2k + 1]) outside which AinvT[j][i] is zero for all of its rows.
constexpr int kDenseMax = 4096;
template <typename T>
void dense_solve(hipStream_t s, int n, int nc, const double *Ainv, const int *range, const T *b, T *x);

// beta[k] of fcg_px stored (for products that take it per column).
void fcg_beta(hipStream_t s, int nc, const double *zq, const double *pq, const uint8_t *active, double *beta);

// Estimator kernel: squared cell + face indicators per tet for ns solutions whose
// interleaved (re, im) columns are at sol + 2*slot[s] with row stride ld.
void estimate(hipStream_t s, int nt, const int *tets, const int *tet2edge, const int *neighbor, const int *non,
              const double *pts, const int *attr, const double *sig6, double omega, int ns, const int *slots,
              const double *sol, int ld, double *errors);

}  // namespace emfem
