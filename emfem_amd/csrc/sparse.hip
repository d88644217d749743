// Sparse and block-vector kernels for gfx950 (64-wide wavefronts).
//
// SpMM strategy: rows of the edge matrices hold ~16 nonzeros.  For few columns a
// group of LPR lanes shares a row (lanes split the nonzeros, each lane keeps all
// columns in registers, then a butterfly reduction); for many columns a whole
// wavefront owns a row and lanes own columns, so the column index and the matrix
// value of a nonzero are wavefront-uniform (scalar loads) and the gathered block
// row x[col][:] is one coalesced 64-lane access.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.h"
#include "device.h"
#include "kernels.h"

namespace emfem {

namespace {

constexpr int kBlock = 256;

inline int grid_for(int64_t threads) { return (int)std::max<int64_t>(1, (threads + kBlock - 1) / kBlock); }

template <typename T>
__device__ __forceinline__ T shfl_xor(T v, int mask, int width) {
  return __shfl_xor(v, mask, width);
}

// ------------------------------------------------------------------------------
// real CSR x block
// ------------------------------------------------------------------------------

// SpMM epilogue of one output entry in two steps: epi_load issues every operand load
// (w, z, the row scale and the fused Chebyshev d and x) before any arithmetic, so they
// are in flight together; epi_store writes y = v or, with a fused Chebyshev step
// (ChebFuse in kernels.h), y = x + c1 (x - d) + c2 v.
template <typename T>
struct EpiIn {
  T w, z, s, d, x;
};

template <typename T>
__device__ __forceinline__ EpiIn<T> epi_load(const T *w, size_t iw, bool gw, const T *z, size_t iz, bool gz,
                                             const T *rs, int row, const ChebFuse<T> &cf, size_t ic) {
  EpiIn<T> e;
  e.w = gw ? w[iw] : T(0);
  e.z = gz ? z[iz] : T(0);
  if (gz && cf.zcol) e.z *= (T)cf.zcol[ic - (size_t)row * cf.ld];
  e.s = rs ? rs[row] : T(1);
  e.d = cf.on() && cf.d ? cf.d[ic] : T(0);
  e.x = cf.on() ? cf.x[ic] : T(0);
  return e;
}

template <typename T>
__device__ __forceinline__ void epi_store(T *y, size_t iy, T v, const EpiIn<T> &e, const ChebFuse<T> &cf, size_t ic) {
  (void)ic;
  y[iy] = cf.on() ? e.x + T(cf.c1) * (e.x - e.d) + T(cf.c2) * v : v;
}

// Sums the NC per-lane partials over the LPR lanes of a row.  For power-of-two NC
// the sum is reduce-scattered (each exchange halves the values a lane keeps, so
// NC/2 + NC/4 + ... shuffles instead of NC per step); the lane then owns columns
// [col0, col0 + KEND).  NC = 1 uses a full butterfly.  (The SpMM pads 3 and 6
// columns to 4 and 8 zero-initialised partials to take the reduce-scatter.)
template <typename T, int NC, int LPR>
__device__ __forceinline__ int row_reduce(T (&acc)[NC], int lane) {
  constexpr bool kScatter = (NC & (NC - 1)) == 0 && NC > 1;
  int col0 = 0;
  if constexpr (kScatter) {
    int K = NC;
#pragma unroll
    for (int off = LPR / 2; off > 0; off >>= 1) {
      if (K > 1) {
        const bool hi = (lane & off) != 0;
        const int h = K / 2;
#pragma unroll
        for (int i = 0; i < NC / 2; ++i) {
          if (i < h) {
            const T send = hi ? acc[i] : acc[h + i];
            const T keep = hi ? acc[h + i] : acc[i];
            acc[i] = keep + shfl_xor(send, off, LPR);
          }
        }
        col0 += hi ? h : 0;
        K = h;
      } else {
        acc[0] += shfl_xor(acc[0], off, LPR);
      }
    }
  } else {
#pragma unroll
    for (int off = LPR / 2; off > 0; off >>= 1) {
#pragma unroll
      for (int j = 0; j < NC; ++j) acc[j] += shfl_xor(acc[j], off, LPR);
    }
  }
  return col0;
}

// Matrix values TA, input block TX, output/accumulation T.  ROLE only separates
// the edge-space smoother launches (ROLE 1) from the rest in kernel traces.
template <typename TA, typename TX, typename T, int NC, int LPR, int ROLE>
__global__ __launch_bounds__(kBlock) void k_spmm_narrow(int nrows, const int *__restrict__ rp,
                                                        const int *__restrict__ ci, const TA *__restrict__ av,
                                                        const TX *__restrict__ x, int ldx, T *y, int ldy, T alpha,
                                                        const T *__restrict__ rs, T gamma, const T *w, int ldw,
                                                        T beta, const T *z, int ldz, const T *__restrict__ xs,
                                                        const T *__restrict__ zs, ChebFuse<T> cf) {
  const int gid = xcd_block() * kBlock + threadIdx.x;
  const int row = gid / LPR;
  const int lane = gid % LPR;
  if (row >= nrows) return;
  // partials padded to a power of two (the padding stays zero)
  constexpr int NCP = NC <= 1 ? 1 : (NC <= 2 ? 2 : (NC <= 4 ? 4 : (NC <= 8 ? 8 : NC)));
  T acc[NCP];
#pragma unroll
  for (int j = 0; j < NCP; ++j) acc[j] = T(0);
  const int e = rp[row + 1];
  // kU nonzeros per lane per trip: all column/value loads, then all row gathers,
  // are in flight together (a typical 16-nonzero edge row is one trip at LPR 4)
  constexpr int kU = 4;
  for (int p0 = rp[row] + lane; p0 < e; p0 += kU * LPR) {
    int c[kU];
    T a[kU];
    // unconditional loads from clamped indices (e - 1 is in this row), masked
    // afterwards: no branch and no wait per load
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int p = p0 + u * LPR;
      const int pc = min(p, e - 1);
      c[u] = ci[pc];
      const T av_ = (T)av[pc];
      a[u] = p < e ? av_ : T(0);
    }
    if (xs) {
#pragma unroll
      for (int u = 0; u < kU; ++u) a[u] *= xs[c[u]];
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const TX *xr = x + (size_t)c[u] * ldx;
#pragma unroll
      for (int j = 0; j < NC; ++j) acc[j] += a[u] * (T)xr[j];
    }
  }
  const int col0 = row_reduce<T, NCP, LPR>(acc, lane);
  const T scale = rs ? rs[row] : T(1);
  const T zscale = zs ? beta * zs[row] : beta;
  constexpr bool kScatter = (NCP & (NCP - 1)) == 0 && NCP > 1;
  if constexpr (kScatter) {
    // lanes that own a column after the reduce-scatter
    constexpr int kOwn = NCP >= LPR ? NCP / LPR : 1;
    constexpr int kRep = NCP >= LPR ? 1 : LPR / NCP;
    if ((lane & (kRep - 1)) != 0) return;
    // the operands of all owned columns first, then the arithmetic and stores
    EpiIn<T> e_[kOwn];
#pragma unroll
    for (int i = 0; i < kOwn; ++i) {
      const int j = col0 + i;
      if (NCP == NC || j < NC)  // not a padding column
        e_[i] = epi_load<T>(w, (size_t)row * ldw + j, gamma != T(0), z, (size_t)row * ldz + j, beta != T(0),
                            nullptr, row, cf, (size_t)row * cf.ld + j);
    }
#pragma unroll
    for (int i = 0; i < kOwn; ++i) {
      const int j = col0 + i;
      if (NCP != NC && j >= NC) break;  // padding columns
      T v = alpha * acc[i];
      if (gamma != T(0)) v += gamma * e_[i].w;
      v *= scale;
      if (beta != T(0)) v += zscale * e_[i].z;
      epi_store(y, (size_t)row * ldy + j, v, e_[i], cf, (size_t)row * cf.ld + j);
    }
  } else {
#pragma unroll
    for (int j = 0; j < NC; ++j) {
      if (lane == (j % LPR)) {
        const EpiIn<T> e_ = epi_load<T>(w, (size_t)row * ldw + j, gamma != T(0), z, (size_t)row * ldz + j,
                                        beta != T(0), nullptr, row, cf, (size_t)row * cf.ld + j);
        T v = alpha * acc[j];
        if (gamma != T(0)) v += gamma * e_.w;
        v *= scale;
        if (beta != T(0)) v += zscale * e_.z;
        epi_store(y, (size_t)row * ldy + j, v, e_, cf, (size_t)row * cf.ld + j);
      }
    }
  }
}

// Narrow x-blocks: a workgroup takes one row block (at most kXnCols distinct
// columns), stages the distinct x rows (NC values each) in LDS once, then sums its
// rows with LPR lanes per row as k_spmm_narrow does, gathering from LDS.  Consecutive
// rows of an edge matrix share most columns, so the global gathers fall to the
// distinct rows of the block.
template <typename TA, typename TX, typename T, int NC, int LPR>
__global__ __launch_bounds__(kBlock) void k_spmm_xnarrow(const int *__restrict__ rp, const TA *__restrict__ av,
                                                         const int *__restrict__ xb_row,
                                                         const int *__restrict__ xb_uptr,
                                                         const int *__restrict__ xb_ucol,
                                                         const int *__restrict__ lci, const TX *__restrict__ x,
                                                         int ldx, T *y, int ldy, T alpha, const T *__restrict__ rs,
                                                         T gamma, const T *w, int ldw, T beta, const T *z, int ldz, ChebFuse<T> cf) {
  __shared__ TX xl[kXnCols * NC];
  const int b = xcd_block();
  const int u0 = xb_uptr[b], nu = xb_uptr[b + 1] - u0;
#pragma unroll 4
  for (int i = threadIdx.x; i < nu * NC; i += kBlock) {
    const int k = i / NC, j = i - k * NC;
    xl[i] = x[(size_t)xb_ucol[u0 + k] * ldx + j];
  }
  __syncthreads();
  constexpr int NCP = NC <= 1 ? 1 : (NC <= 2 ? 2 : (NC <= 4 ? 4 : (NC <= 8 ? 8 : NC)));
  const int r0 = xb_row[b], r1 = xb_row[b + 1];
  const int lane = threadIdx.x % LPR;
  // the LPR lanes of a row group share the row, so a group enters and leaves the
  // loop together (the reduction's shuffles stay within the group)
  for (int row = r0 + threadIdx.x / LPR; row < r1; row += kBlock / LPR) {
    T acc[NCP];
#pragma unroll
    for (int j = 0; j < NCP; ++j) acc[j] = T(0);
    const int e = rp[row + 1];
    constexpr int kU = 4;
    for (int p0 = rp[row] + lane; p0 < e; p0 += kU * LPR) {
      int c[kU];
      T a[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int p = p0 + u * LPR;
        const int pc = min(p, e - 1);  // unconditional loads, masked
        c[u] = lci[pc];
        const T av_ = (T)av[pc];
        a[u] = p < e ? av_ : T(0);
      }
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const TX *xr = xl + c[u] * NC;
#pragma unroll
        for (int j = 0; j < NC; ++j) acc[j] += a[u] * (T)xr[j];
      }
    }
    const int col0 = row_reduce<T, NCP, LPR>(acc, lane);
    const T scale = rs ? rs[row] : T(1);
    constexpr bool kScatter = (NCP & (NCP - 1)) == 0 && NCP > 1;
    if constexpr (kScatter) {
      constexpr int kOwn = NCP >= LPR ? NCP / LPR : 1;
      constexpr int kRep = NCP >= LPR ? 1 : LPR / NCP;
      if ((lane & (kRep - 1)) == 0) {
        EpiIn<T> e_[kOwn];
#pragma unroll
        for (int i = 0; i < kOwn; ++i) {
          const int j = col0 + i;
          if (NCP == NC || j < NC)
            e_[i] = epi_load<T>(w, (size_t)row * ldw + j, gamma != T(0), z, (size_t)row * ldz + j, beta != T(0),
                                nullptr, row, cf, (size_t)row * cf.ld + j);
        }
#pragma unroll
        for (int i = 0; i < kOwn; ++i) {
          const int j = col0 + i;
          if (NCP != NC && j >= NC) break;
          T v = alpha * acc[i];
          if (gamma != T(0)) v += gamma * e_[i].w;
          v *= scale;
          if (beta != T(0)) v += beta * e_[i].z;
          epi_store(y, (size_t)row * ldy + j, v, e_[i], cf, (size_t)row * cf.ld + j);
        }
      }
    } else {
#pragma unroll
      for (int j = 0; j < NC; ++j) {
        if (lane == (j % LPR)) {
          const EpiIn<T> e_ = epi_load<T>(w, (size_t)row * ldw + j, gamma != T(0), z, (size_t)row * ldz + j,
                                          beta != T(0), nullptr, row, cf, (size_t)row * cf.ld + j);
          T v = alpha * acc[j];
          if (gamma != T(0)) v += gamma * e_.w;
          v *= scale;
          if (beta != T(0)) v += beta * e_.z;
          epi_store(y, (size_t)row * ldy + j, v, e_, cf, (size_t)row * cf.ld + j);
        }
      }
    }
  }
}

// Many columns (nc >= 16, batched right-hand sides): LW lanes per row, lane l owning
// columns l, l + LW, ..., l + (CPL-1) LW (each load instruction reads LW consecutive
// columns of a row: coalesced), 64 / LW rows per wavefront; kU nonzeros per trip with
// all column/value loads, then all row gathers, in flight together.
template <typename TA, typename TX, typename T, int LW, int CPL>
__global__ __launch_bounds__(kBlock) void k_spmm_wide(int nrows, int rblocks, const int *__restrict__ rp,
                                                      const int *__restrict__ ci, const TA *__restrict__ av,
                                                      const TX *__restrict__ x, int ldx, T *y, int ldy, int nc,
                                                      T alpha, const T *__restrict__ rs, T gamma, const T *w,
                                                      int ldw, T beta, const T *z, int ldz, const T *__restrict__ xs,
                                                      const T *__restrict__ zs, ChebFuse<T> cf) {
  // slab-major block order: with the XCD-aware remap each XCD sweeps all rows of one
  // slab (or a contiguous share of one), so its L2 holds only that slab's window of x
  const int bid = xcd_block();
  const int slab = bid / rblocks;
  const int gid = (bid - slab * rblocks) * kBlock + threadIdx.x;
  const int row = gid / LW;
  const int c0 = slab * LW * CPL + gid % LW;  // first column of this lane
  if (row >= nrows) return;
  T acc[CPL];
#pragma unroll
  for (int j = 0; j < CPL; ++j) acc[j] = T(0);
  const int b = rp[row], e = rp[row + 1];
  constexpr int kU = (sizeof(T) == 8 || sizeof(TX) == 8) ? 4 : 8;  // registers of the gathers
  for (int p0 = b; p0 < e; p0 += kU) {
    int cu[kU];
    T au[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const bool ok = p0 + u < e;
      const int pc = min(p0 + u, e - 1);
      cu[u] = ci[pc];
      const T av_ = (T)av[pc];
      au[u] = ok ? av_ : T(0);
      if (xs && ok) au[u] *= xs[cu[u]];
    }
    TX xv[kU][CPL];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const TX *xr = x + (size_t)cu[u] * ldx + c0;
#pragma unroll
      for (int j = 0; j < CPL; ++j) xv[u][j] = c0 + j * LW < nc ? xr[j * LW] : TX(0);
    }
#pragma unroll
    for (int u = 0; u < kU; ++u)
#pragma unroll
      for (int j = 0; j < CPL; ++j) acc[j] += au[u] * (T)xv[u][j];
  }
  const T scale = rs ? rs[row] : T(1);
  const T zscale = zs ? beta * zs[row] : beta;
  EpiIn<T> e_[CPL];
#pragma unroll
  for (int j = 0; j < CPL; ++j) {
    const int c = c0 + j * LW;
    if (c < nc)
      e_[j] = epi_load<T>(w, (size_t)row * ldw + c, gamma != T(0), z, (size_t)row * ldz + c, beta != T(0), nullptr,
                          row, cf, (size_t)row * cf.ld + c);
  }
#pragma unroll
  for (int j = 0; j < CPL; ++j) {
    const int c = c0 + j * LW;
    if (c < nc) {
      T v = alpha * acc[j];
      if (gamma != T(0)) v += gamma * e_[j].w;
      v *= scale;
      if (beta != T(0)) v += zscale * e_[j].z;
      epi_store(y, (size_t)row * ldy + c, v, e_[j], cf, (size_t)row * cf.ld + c);
    }
  }
}

// Wide SpMM with 16-byte loads: LW lanes per row, each lane owns VEC = 16 / sizeof(TX)
// consecutive columns of a slab of LW * VEC columns and gathers them with one vector
// load per nonzero (nc a multiple of VEC, x rows 16-byte aligned).  Slab-major block
// order as in k_spmm_wide.
template <typename TX>
struct Vec16;
template <>
struct Vec16<float> {
  using type = float4;
  static constexpr int n = 4;
};
template <>
struct Vec16<double> {
  using type = double2;
  static constexpr int n = 2;
};

// VEC consecutive values of T moved as one aligned vector (two for 32 bytes)
template <typename T, int VEC>
struct alignas(sizeof(T) * VEC > 16 ? 16 : sizeof(T) * VEC) VecT {
  T v[VEC];
};

template <typename TA, typename TX, typename T, int LW>
__global__ __launch_bounds__(kBlock) void k_spmm_wvec(int nrows, int rblocks, const int *__restrict__ rp,
                                                      const int *__restrict__ ci, const TA *__restrict__ av,
                                                      const TX *__restrict__ x, int ldx, T *y, int ldy, int nc,
                                                      T alpha, const T *__restrict__ rs, T gamma, const T *w,
                                                      int ldw, T beta, const T *z, int ldz,
                                                      const T *__restrict__ xs, const T *__restrict__ zs, int vepi, ChebFuse<T> cf) {
  using V = typename Vec16<TX>::type;
  constexpr int VEC = Vec16<TX>::n;
  const int bid = xcd_block();
  const int slab = bid / rblocks;
  const int gid = (bid - slab * rblocks) * kBlock + threadIdx.x;
  const int row = gid / LW;
  const int c0 = slab * LW * VEC + (gid % LW) * VEC;
  if (row >= nrows) return;
  T acc[VEC];
#pragma unroll
  for (int j = 0; j < VEC; ++j) acc[j] = T(0);
  const int b = rp[row], e = rp[row + 1];
  {
    // The LW lanes of the row load up to LW of its column indices and values at once
    // (one per lane) and pass them round by shuffles; eight 16-byte gathers are then
    // in flight per lane instead of a load -> gather dependency every four nonzeros.
    // Lanes past the last column slab (c0 >= nc) take part in the loads and shuffles
    // and gather a valid column block they discard.
    const int lane = gid % LW;
    const int cx = c0 < nc ? c0 : 0;
    constexpr int kG = 8;
    for (int q0 = b; q0 < e; q0 += LW) {
      const int n = min(LW, e - q0);
      const int pl = q0 + min(lane, n - 1);
      const int cl = ci[pl];
      T al = (T)av[pl];
      if (xs) al *= xs[cl];
      for (int u0 = 0; u0 < n; u0 += kG) {
        int cu[kG];
        T au[kG];
#pragma unroll
        for (int u = 0; u < kG; ++u) {
          const int src = min(u0 + u, n - 1);
          cu[u] = __shfl(cl, src, LW);
          const T a_ = __shfl(al, src, LW);
          au[u] = u0 + u < n ? a_ : T(0);
        }
        V xv[kG];
#pragma unroll
        for (int u = 0; u < kG; ++u) xv[u] = *reinterpret_cast<const V *>(x + (size_t)cu[u] * ldx + cx);
#pragma unroll
        for (int u = 0; u < kG; ++u) {
          const TX *xe = reinterpret_cast<const TX *>(&xv[u]);
#pragma unroll
          for (int j = 0; j < VEC; ++j) acc[j] += au[u] * (T)xe[j];
        }
      }
    }
    if (c0 >= nc) return;
  }
  const T scale = rs ? rs[row] : T(1);
  const T zscale = zs ? beta * zs[row] : beta;
  if (vepi) {
    // the lane's VEC outputs (and the fused Chebyshev operands) as aligned vectors
    using TV = VecT<T, VEC>;
    // all operand vectors first, then the arithmetic
    const size_t ic = (size_t)row * cf.ld + c0;
    TV wv{}, zv{}, dv{}, xv{};
    if (gamma != T(0)) wv = *reinterpret_cast<const TV *>(w + (size_t)row * ldw + c0);
    if (beta != T(0)) {
      zv = *reinterpret_cast<const TV *>(z + (size_t)row * ldz + c0);
      if (cf.zcol) {
#pragma unroll
        for (int j = 0; j < VEC; ++j) zv.v[j] *= (T)cf.zcol[c0 + j];
      }
    }
    if (cf.on()) {
      if (cf.d) dv = *reinterpret_cast<const TV *>(cf.d + ic);
      xv = *reinterpret_cast<const TV *>(cf.x + ic);
    }
    TV v;
#pragma unroll
    for (int j = 0; j < VEC; ++j) v.v[j] = alpha * acc[j];
    if (gamma != T(0)) {
#pragma unroll
      for (int j = 0; j < VEC; ++j) v.v[j] += gamma * wv.v[j];
    }
#pragma unroll
    for (int j = 0; j < VEC; ++j) v.v[j] *= scale;
    if (beta != T(0)) {
#pragma unroll
      for (int j = 0; j < VEC; ++j) v.v[j] += zscale * zv.v[j];
    }
    if (cf.on()) {
#pragma unroll
      for (int j = 0; j < VEC; ++j) v.v[j] = xv.v[j] + T(cf.c1) * (xv.v[j] - dv.v[j]) + T(cf.c2) * v.v[j];
    }
    *reinterpret_cast<TV *>(y + (size_t)row * ldy + c0) = v;
    return;
  }
  EpiIn<T> e_[VEC];
#pragma unroll
  for (int j = 0; j < VEC; ++j) {
    const int c = c0 + j;
    e_[j] = epi_load<T>(w, (size_t)row * ldw + c, gamma != T(0), z, (size_t)row * ldz + c, beta != T(0), nullptr,
                        row, cf, (size_t)row * cf.ld + c);
  }
#pragma unroll
  for (int j = 0; j < VEC; ++j) {
    const int c = c0 + j;
    T v = alpha * acc[j];
    if (gamma != T(0)) v += gamma * e_[j].w;
    v *= scale;
    if (beta != T(0)) v += zscale * e_[j].z;
    epi_store(y, (size_t)row * ldy + c, v, e_[j], cf, (size_t)row * cf.ld + c);
  }
}

// LDS x-blocks for many columns: a workgroup takes one row block of one 64-column slab,
// stages the block's distinct x rows for the slab in LDS once (consecutive rows of an
// edge matrix share most of their columns), and each wavefront then sums its rows
// from LDS with lanes on columns.  Rows are wavefront-uniform, so a nonzero's value
// and LDS slot are scalar loads; each row is summed in nonzero order, as in the other
// SpMM kernels.
template <typename TA, typename TX, typename T>
__global__ __launch_bounds__(kBlock) void k_spmm_xblk(int nblk, const int *__restrict__ rp,
                                                      const TA *__restrict__ av, const int *__restrict__ xb_row,
                                                      const int *__restrict__ xb_uptr,
                                                      const int *__restrict__ xb_ucol,
                                                      const int *__restrict__ lci, const TX *__restrict__ x,
                                                      int ldx, T *y, int ldy, int nc, T alpha,
                                                      const T *__restrict__ rs, T gamma, const T *w, int ldw,
                                                      T beta, const T *z, int ldz, int vec_ok, ChebFuse<T> cf) {
  extern __shared__ __align__(16) unsigned char xblk_lds[];
  TX *xl = reinterpret_cast<TX *>(xblk_lds);
  using V = typename Vec16<TX>::type;
  constexpr int VEC = Vec16<TX>::n;
  constexpr int W = 64;
  constexpr int kPer = W / VEC;  // threads per staged x row
  const int bid = xcd_block();
  const int slab = bid / nblk;
  const int b = bid - slab * nblk;
  const int cb = slab * W;
  const int u0 = xb_uptr[b], nu = xb_uptr[b + 1] - u0;
  if (vec_ok) {
    // 16-byte loads, all of a thread's loads issued before the LDS stores
    constexpr int kRowsPerPass = kBlock / kPer;
    constexpr int kPass = kXBlockCols / kRowsPerPass;
    const int k0 = threadIdx.x / kPer, cc = (threadIdx.x % kPer) * VEC;
    V v[kPass];
#pragma unroll
    for (int q = 0; q < kPass; ++q) {
      const int k = k0 + q * kRowsPerPass;
      const int c = cb + cc;
      if (k < nu && c < nc) v[q] = *reinterpret_cast<const V *>(x + (size_t)xb_ucol[u0 + k] * ldx + c);
      else v[q] = V{};
    }
#pragma unroll
    for (int q = 0; q < kPass; ++q) {
      const int k = k0 + q * kRowsPerPass;
      if (k < nu) *reinterpret_cast<V *>(xl + k * W + cc) = v[q];
    }
  } else {
    for (int i = threadIdx.x; i < nu * W; i += kBlock) {
      const int k = i / W, c = cb + (i % W);
      xl[i] = c < nc ? x[(size_t)xb_ucol[u0 + k] * ldx + c] : TX(0);
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int r1 = xb_row[b + 1];
  for (int row = xb_row[b] + wave; row < r1; row += kBlock / 64) {
    T acc = T(0);
    const int pb = rp[row], pe = rp[row + 1];
#pragma unroll 4
    for (int p = pb; p < pe; ++p) acc += (T)av[p] * (T)xl[lci[p] * W + lane];
    const int c = cb + lane;
    if (c < nc) {
      const EpiIn<T> e_ = epi_load<T>(w, (size_t)row * ldw + c, gamma != T(0), z, (size_t)row * ldz + c,
                                      beta != T(0), rs, row, cf, (size_t)row * cf.ld + c);
      T v = alpha * acc;
      if (gamma != T(0)) v += gamma * e_.w;
      v *= e_.s;
      if (beta != T(0)) v += beta * e_.z;
      epi_store(y, (size_t)row * ldy + c, v, e_, cf, (size_t)row * cf.ld + c);
    }
  }
}

// LDS-staged row blocks: a workgroup takes the rows [blk_row[b], blk_row[b+1]), whose
// nonzeros (at most TILE) are streamed with fully coalesced column/value loads, and
// their products a * x[col][:] are staged in LDS; each (row, column) is then summed
// from LDS by one thread, in nonzero order.
// Persistent row blocks: workgroup g takes the consecutive tiles [g per, (g+1) per)
// of the row-block table and runs them as a three-stage software pipeline: while
// tile t is summed from LDS (and its epilogue loads are in flight), the x gathers of
// tile t+1 and the column/value/row-pointer loads of tile t+2 are already issued, so
// a tile costs about one memory latency instead of three.  Products and summation
// order are those of k_spmm_stream.
template <typename TA, typename TX, typename T, int NC, int TILE, int BS = kBlock>
__global__ __launch_bounds__(BS) void k_spmm_pstream(const int *__restrict__ rp, const int *__restrict__ ci,
                                                         const TA *__restrict__ av, const TX *__restrict__ x,
                                                         int ldx, T *y, int ldy, T alpha, const T *__restrict__ rs,
                                                         T gamma, const T *w, int ldw, T beta, const T *z, int ldz,
                                                         const int *__restrict__ blk_row, int nblk, int per,
                                                         ChebFuse<T> cf) {
  __shared__ T prod[TILE * NC];
  __shared__ int srp[TILE + 1];
  static_assert(TILE % BS == 0, "a tile is a whole number of passes of the block");
  constexpr int kTrips = TILE / BS;
  constexpr int kRp = (TILE + 1 + BS - 1) / BS;
  struct Tile {
    int r0, r1, p0, nnz;
    int c[kTrips];
    T a[kTrips];
    int rpv[kRp];
  };
  const int t0 = xcd_block() * per;
  const int t1 = min(nblk, t0 + per);
  if (t0 >= t1) return;
  // all loads are unconditional from clamped (valid) indices and masked afterwards, so
  // the compiler keeps them in flight instead of branching and waiting per load
  auto load_idx = [&](int t, Tile &I) {
    I.r0 = blk_row[t];
    I.r1 = blk_row[t + 1];
    I.p0 = rp[I.r0];
    I.nnz = rp[I.r1] - I.p0;
#pragma unroll
    for (int k = 0; k < kRp; ++k) {
      const int i = threadIdx.x + k * BS;
      I.rpv[k] = rp[I.r0 + min(i, I.r1 - I.r0)] - I.p0;
    }
    const int last = max(I.nnz - 1, 0);
#pragma unroll
    for (int u = 0; u < kTrips; ++u) {
      const int idx = threadIdx.x + u * BS;
      const int pc = I.p0 + min(idx, last);
      const int cv = __builtin_nontemporal_load(ci + pc);
      const T aval = (T)__builtin_nontemporal_load(av + pc);
      I.c[u] = cv;
      I.a[u] = idx < I.nnz ? aval : T(0);
    }
  };
  auto gather = [&](const Tile &I, TX (&xv)[kTrips][NC]) {
#pragma unroll
    for (int u = 0; u < kTrips; ++u) {
      const TX *xr = x + (size_t)I.c[u] * ldx;
#pragma unroll
      for (int j = 0; j < NC; ++j) xv[u][j] = xr[j];
    }
  };
  auto stage = [&](const Tile &I, const TX (&xv)[kTrips][NC]) {
#pragma unroll
    for (int k = 0; k < kRp; ++k) {
      const int i = threadIdx.x + k * BS;
      if (i <= I.r1 - I.r0) srp[i] = I.rpv[k];
    }
#pragma unroll
    for (int u = 0; u < kTrips; ++u) {
      const int idx = threadIdx.x + u * BS;
      if (idx < I.nnz) {
#pragma unroll
        for (int j = 0; j < NC; ++j) prod[idx * NC + j] = I.a[u] * (T)xv[u][j];
      }
    }
  };
  Tile cur, nxt, nn;
  TX xv[kTrips][NC];
  load_idx(t0, cur);
  gather(cur, xv);
  stage(cur, xv);
  if (t0 + 1 < t1) load_idx(t0 + 1, nxt);
  __syncthreads();
  for (int t = t0; t < t1; ++t) {
    const bool has1 = t + 1 < t1, has2 = t + 2 < t1;
    if (has1) gather(nxt, xv);
    if (has2) load_idx(t + 2, nn);
    const int nr = cur.r1 - cur.r0;
    for (int q = threadIdx.x; q < nr * NC; q += BS) {
      const int lr = q / NC, j = q % NC, row = cur.r0 + lr;
      const EpiIn<T> e_ = epi_load<T>(w, (size_t)row * ldw + j, gamma != T(0), z, (size_t)row * ldz + j,
                                      beta != T(0), rs, row, cf, (size_t)row * cf.ld + j);
      T acc = T(0);
      for (int k = srp[lr]; k < srp[lr + 1]; ++k) acc += prod[k * NC + j];
      T v = alpha * acc;
      if (gamma != T(0)) v += gamma * e_.w;
      v *= e_.s;
      if (beta != T(0)) v += beta * e_.z;
      epi_store(y, (size_t)row * ldy + j, v, e_, cf, (size_t)row * cf.ld + j);
    }
    __syncthreads();
    if (has1) {
      stage(nxt, xv);
      cur = nxt;
      nxt = nn;
    }
    __syncthreads();
  }
}

template <typename TA, typename TX, typename T, int NC, int TILE, int BS = kBlock>
__global__ __launch_bounds__(BS) void k_spmm_stream(const int *__restrict__ rp, const int *__restrict__ ci,
                                                        const TA *__restrict__ av, const TX *__restrict__ x, int ldx,
                                                        T *y, int ldy, T alpha, const T *__restrict__ rs, T gamma,
                                                        const T *w, int ldw, T beta, const T *z, int ldz,
                                                        const int *__restrict__ blk_row, ChebFuse<T> cf) {
  __shared__ T prod[TILE * NC];
  __shared__ int srp[TILE + 1];
  const int b = xcd_block();
  const int r0 = blk_row[b], r1 = blk_row[b + 1];
  const int p0 = rp[r0];
  const int nnz = rp[r1] - p0;
  for (int i = threadIdx.x; i <= r1 - r0; i += BS) srp[i] = rp[r0 + i] - p0;
  static_assert(TILE % BS == 0, "a tile is a whole number of passes of the block");
  constexpr int kTrips = TILE / BS;
  // unconditional loads from clamped indices (masked at the LDS store), so all of a
  // thread's loads are in flight together
  int c[kTrips];
  T a[kTrips];
  const int last = max(nnz - 1, 0);
#pragma unroll
  for (int u = 0; u < kTrips; ++u) {
    const int pc = p0 + min((int)threadIdx.x + u * BS, last);
    c[u] = __builtin_nontemporal_load(ci + pc);
    a[u] = (T)__builtin_nontemporal_load(av + pc);
  }
  TX xv[kTrips][NC];
#pragma unroll
  for (int u = 0; u < kTrips; ++u) {
    const TX *xr = x + (size_t)c[u] * ldx;
#pragma unroll
    for (int j = 0; j < NC; ++j) xv[u][j] = xr[j];
  }
#pragma unroll
  for (int u = 0; u < kTrips; ++u) {
    const int idx = threadIdx.x + u * BS;
    if (idx < nnz) {
#pragma unroll
      for (int j = 0; j < NC; ++j) prod[idx * NC + j] = a[u] * (T)xv[u][j];
    }
  }
  __syncthreads();
  const int nr = r1 - r0;
  for (int q = threadIdx.x; q < nr * NC; q += BS) {
    const int lr = q / NC, j = q % NC, row = r0 + lr;
    const EpiIn<T> e_ = epi_load<T>(w, (size_t)row * ldw + j, gamma != T(0), z, (size_t)row * ldz + j,
                                    beta != T(0), rs, row, cf, (size_t)row * cf.ld + j);
    T acc = T(0);
    for (int k = srp[lr]; k < srp[lr + 1]; ++k) acc += prod[k * NC + j];
    T v = alpha * acc;
    if (gamma != T(0)) v += gamma * e_.w;
    v *= e_.s;
    if (beta != T(0)) v += beta * e_.z;
    epi_store(y, (size_t)row * ldy + j, v, e_, cf, (size_t)row * cf.ld + j);
  }
}

// Sum of the W-wide chunks prod[k * NC + j0 .. + W) over k in [k0, k1), added in
// increasing k: the LDS reads are issued eight rows at a time so that a row's sum
// waits for about one LDS latency per eight nonzeros instead of one per nonzero.
template <typename T, int NC, int W>
__device__ __forceinline__ void lds_row_sum(const T *prod, int k0, int k1, int j0, T (&acc)[W]) {
  constexpr int kU = 8;
  int k = k0;
  for (; k + kU <= k1; k += kU) {
    T v[kU][W];
#pragma unroll
    for (int u = 0; u < kU; ++u)
#pragma unroll
      for (int i = 0; i < W; ++i) v[u][i] = prod[(k + u) * NC + j0 + i];
#pragma unroll
    for (int u = 0; u < kU; ++u)
#pragma unroll
      for (int i = 0; i < W; ++i) acc[i] += v[u][i];
  }
  for (; k < k1; ++k) {
#pragma unroll
    for (int i = 0; i < W; ++i) acc[i] += prod[k * NC + j0 + i];
  }
}

// Vector chunk of W consecutive values, loaded and stored as one W-wide access
// (dwordx2 / dwordx3 / dwordx4); W = 3 needs only the element alignment.
template <typename T, int W>
struct alignas(sizeof(T) * (W == 3 ? 1 : W)) ChunkT {
  T v[W];
};

// Widest chunk that divides NC: 4, 3 or 2 values (fp32), 2 or 3 values (fp64).
template <typename T, int NC>
constexpr int row_chunk() {
  return sizeof(T) == 8 ? (NC % 2 == 0 ? 2 : (NC % 3 == 0 ? 3 : 1))
                        : (NC % 4 == 0 ? 4 : (NC % 3 == 0 ? 3 : (NC % 2 == 0 ? 2 : 1)));
}

// k_spmm_stream on compact blocks (every leading dimension = NC): each gathered
// x row is read as NC / W vector chunks, and in the epilogue a thread owns one W-wide
// chunk of a row, loading its operands and storing its result as vectors.  The
// products and the summation order are those of k_spmm_stream.
// Every load of the gather phase is issued before its first wait and every LDS store
// is unconditional (entries past the tile's nonzeros hold a zero product and are
// never summed): the compiler otherwise sinks a load into the conditional block of
// its only use, behind a wait for every load issued before it.
// FORM selects the epilogue at compile time: kFormAny (runtime operands, loaded after
// the barrier), kFormPlain (y = alpha A x), kFormCheb (the fused Chebyshev step
// y = x + c1 (x - d) + c2 rs .* (alpha A x + gamma w) with cf.x, cf.d, rs and w
// given), kFormCheb0 (the same with d = 0, cf.d null), kFormAxpy (y = alpha A x +
// gamma w, the residual form): the operands of the first pass are loaded before the
// barrier, so their latency overlaps the gathers'.  BS threads per workgroup, TILE a
// multiple of BS.  C16: the columns are read in their 16-bit form (ci16, cbase; see
// compress_columns) instead of from ci.
constexpr int kFormAny = 0, kFormPlain = 1, kFormCheb = 2, kFormCheb0 = 3, kFormAxpy = 4;
template <typename TA, typename TX, typename T, int NC, int TILE, int FORM, int BS = kBlock, bool C16 = false>
__global__ __launch_bounds__(BS) void k_spmm_stream_v(const int *__restrict__ rp, const int *__restrict__ ci,
                                                          const uint16_t *__restrict__ ci16,
                                                          const int *__restrict__ cbase,
                                                          const TA *__restrict__ av, const TX *__restrict__ x,
                                                          T *y, T alpha, const T *__restrict__ rs, T gamma,
                                                          const T *w, T beta, const T *z,
                                                          const int *__restrict__ blk_row, ChebFuse<T> cf) {
  constexpr int WX = row_chunk<TX, NC>(), W = row_chunk<T, NC>();
  constexpr int CX = NC / WX, CY = NC / W;
  using VX = ChunkT<TX, WX>;
  using VY = ChunkT<T, W>;
  __shared__ T prod[TILE * NC];
  __shared__ int srp[TILE + 1];
  static_assert(BS <= TILE + 1, "row pointers are staged one per thread");
  static_assert(TILE % BS == 0, "a tile is a whole number of passes of the block");
  const int b = xcd_block();
  const int r0 = blk_row[b], r1 = blk_row[b + 1];
  const int *blk_ptr = blk_row + gridDim.x + 1;  // one workgroup per block
  const int p0 = blk_ptr[b];
  const int nnz = blk_ptr[b + 1] - p0;
  const int nr = r1 - r0;
  constexpr int kTrips = TILE / BS;
  int c[kTrips];
  T a[kTrips];
  const int last = max(nnz - 1, 0);
#pragma unroll
  for (int u = 0; u < kTrips; ++u) {
    const int idx = (int)threadIdx.x + u * BS;
    const int pc = p0 + min(idx, last);
    // the matrix is streamed once per launch: non-temporal loads, so that it does not
    // evict the gathered x rows from the caches
    const T aval = (T)__builtin_nontemporal_load(av + pc);
    if constexpr (C16)
      c[u] = cbase[pc / kCol16Group] + (int)__builtin_nontemporal_load(ci16 + pc);
    else
      c[u] = __builtin_nontemporal_load(ci + pc);
    a[u] = idx < nnz ? aval : T(0);
  }
  const int rpv = rp[r0 + min((int)threadIdx.x, nr)];
  VX xv[kTrips][CX];
#pragma unroll
  for (int u = 0; u < kTrips; ++u) {
    const VX *xr = reinterpret_cast<const VX *>(x + (size_t)c[u] * NC);
#pragma unroll
    for (int k = 0; k < CX; ++k) xv[u][k] = xr[k];
  }
  // the Chebyshev operands of this thread's first (row, chunk), from a clamped row
  const int nq = nr * CY;
  const int q0 = min((int)threadIdx.x, nq - 1);
  const int lr0 = q0 / CY, j00 = (q0 % CY) * W;
  constexpr bool kCheb = FORM == kFormCheb || FORM == kFormCheb0;
  VY w0, xk0, xd0;
  T sc0 = T(1);
  if constexpr (FORM == kFormAxpy) w0 = *reinterpret_cast<const VY *>(w + (size_t)(r0 + lr0) * NC + j00);
  if constexpr (kCheb) {
    const size_t o = (size_t)(r0 + lr0) * NC + j00;
    w0 = *reinterpret_cast<const VY *>(w + o);
    xk0 = *reinterpret_cast<const VY *>(cf.x + o);
    if constexpr (FORM == kFormCheb) xd0 = *reinterpret_cast<const VY *>(cf.d + o);
    sc0 = rs[r0 + lr0];
  }
  // entries past nr (clamped to the last row pointer) are never read
  srp[threadIdx.x] = rpv - p0;
  for (int i = threadIdx.x + BS; i <= nr; i += BS) srp[i] = rp[r0 + i] - p0;  // tiles of short rows
#pragma unroll
  for (int u = 0; u < kTrips; ++u) {
    const int idx = threadIdx.x + u * BS;
#pragma unroll
    for (int k = 0; k < CX; ++k)
#pragma unroll
      for (int i = 0; i < WX; ++i) prod[idx * NC + k * WX + i] = a[u] * (T)xv[u][k].v[i];
  }
  __syncthreads();
  if constexpr (FORM != kFormAny) {
    auto finish = [&](int lr, int j0, const VY &wv, const VY &xk, const VY &xd, T sc) {
      T acc[W];
#pragma unroll
      for (int i = 0; i < W; ++i) acc[i] = T(0);
      lds_row_sum<T, NC, W>(prod, srp[lr], srp[lr + 1], j0, acc);
      VY out;
#pragma unroll
      for (int i = 0; i < W; ++i) {
        T v = alpha * acc[i];
        if constexpr (FORM == kFormAxpy) v += gamma * wv.v[i];
        if constexpr (kCheb) {
          v += gamma * wv.v[i];
          v *= sc;
          const T xo = xk.v[i], dv = FORM == kFormCheb ? xd.v[i] : T(0);
          v = xo + T(cf.c1) * (xo - dv) + T(cf.c2) * v;
        }
        out.v[i] = v;
      }
      *reinterpret_cast<VY *>(y + (size_t)(r0 + lr) * NC + j0) = out;
    };
    if ((int)threadIdx.x < nq) finish(lr0, j00, w0, xk0, xd0, sc0);
    for (int q = threadIdx.x + BS; q < nq; q += BS) {
      const int lr = q / CY, j0 = (q % CY) * W;
      VY wv, xk, xd;
      T sc = T(1);
      if constexpr (FORM == kFormAxpy) wv = *reinterpret_cast<const VY *>(w + (size_t)(r0 + lr) * NC + j0);
      if constexpr (kCheb) {
        const size_t o = (size_t)(r0 + lr) * NC + j0;
        wv = *reinterpret_cast<const VY *>(w + o);
        xk = *reinterpret_cast<const VY *>(cf.x + o);
        if constexpr (FORM == kFormCheb) xd = *reinterpret_cast<const VY *>(cf.d + o);
        sc = rs[r0 + lr];
      }
      finish(lr, j0, wv, xk, xd, sc);
    }
    return;
  }
  for (int q = threadIdx.x; q < nq; q += BS) {
    const int lr = q / CY, j0 = (q % CY) * W, row = r0 + lr;
    const size_t o = (size_t)row * NC + j0;
    VY wv, zv, xk, xd;
    if (gamma != T(0)) wv = *reinterpret_cast<const VY *>(w + o);
    if (beta != T(0)) zv = *reinterpret_cast<const VY *>(z + o);
    if (cf.on()) {
      xk = *reinterpret_cast<const VY *>(cf.x + o);
      if (cf.d) xd = *reinterpret_cast<const VY *>(cf.d + o);
    }
    const T sc = rs ? rs[row] : T(1);
    T acc[W];
#pragma unroll
    for (int i = 0; i < W; ++i) acc[i] = T(0);
    lds_row_sum<T, NC, W>(prod, srp[lr], srp[lr + 1], j0, acc);
    VY out;
#pragma unroll
    for (int i = 0; i < W; ++i) {
      T v = alpha * acc[i];
      if (gamma != T(0)) v += gamma * wv.v[i];
      v *= sc;
      if (beta != T(0)) v += beta * zv.v[i];
      if (cf.on()) {
        const T xo = xk.v[i], dv = cf.d ? xd.v[i] : T(0);
        v = xo + T(cf.c1) * (xo - dv) + T(cf.c2) * v;
      }
      out.v[i] = v;
    }
    *reinterpret_cast<VY *>(y + o) = out;
  }
}

template <typename TA, typename TX, typename T, int NC>
void launch_narrow(hipStream_t s, const Csr<TA> &A, const TX *x, int ldx, T *y, int ldy, T alpha, const T *rs,
                   T gamma, const T *w, int ldw, T beta, const T *z, int ldz, const T *xs, const T *zs,
                   const ChebFuse<T> &cf) {
  const double avg = A.nrows ? double(A.nnz) / A.nrows : 0.0;
  // compact blocks with vector-aligned bases take the vector-chunk row-block kernels
  constexpr int WX = row_chunk<TX, NC>(), WY = row_chunk<T, NC>();
  auto aligned = [](const void *p, size_t a) { return p == nullptr || (reinterpret_cast<uintptr_t>(p) % a) == 0; };
  const size_t ax = sizeof(TX) * (WX == 3 ? 1 : WX), ay = sizeof(T) * (WY == 3 ? 1 : WY);
  const bool vec = WX > 1 && WY > 1 && ldx == NC && ldy == NC && (gamma == T(0) || ldw == NC) &&
                   (beta == T(0) || ldz == NC) && (!cf.on() || cf.ld == NC) && cf.zcol == nullptr &&
                   aligned(x, ax) && aligned(y, ay) && aligned(gamma == T(0) ? nullptr : w, ay) &&
                   aligned(beta == T(0) ? nullptr : z, ay) && aligned(cf.on() ? cf.x : nullptr, ay) &&
                   aligned(cf.on() ? cf.d : nullptr, ay);
  if (A.lpr == -3) {
    // persistent pipelined row blocks: about 8 workgroups per CU
    EMF_REQUIRE(A.blk_row != nullptr && A.nblk > 0, kErrArg, "row-block SpMM without its block table");
    EMF_REQUIRE(xs == nullptr && zs == nullptr, kErrArg, "row-block SpMM has no scaled operands");
    static int n_cu = 0;
    if (n_cu == 0) {
      int dev = 0;
      EMF_HIP(hipGetDevice(&dev));
      EMF_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
      n_cu = std::max(1, n_cu);
    }
    const int groups = std::min(A.nblk, 8 * n_cu);
    const int per = (A.nblk + groups - 1) / groups;
    const int used = (A.nblk + per - 1) / per;
    hipLaunchKernelGGL((k_spmm_pstream<TA, TX, T, NC, stream_tile<T, NC>(), stream_block<T, NC>()>), dim3(used),
                       dim3(stream_block<T, NC>()), 0, s, A.rp,
                       A.ci, A.v, x, ldx, y, ldy, alpha, rs, gamma, w, ldw, beta, z, ldz, A.blk_row, A.nblk, per, cf);
    return;
  }
  if (A.lpr <= -2) {
    // narrow x-blocks with -lpr lanes per row
    EMF_REQUIRE(A.xn_row != nullptr && A.xn_nblk > 0, kErrArg, "x-block SpMM without its block tables");
    EMF_REQUIRE(xs == nullptr && zs == nullptr, kErrArg, "x-block SpMM has no scaled operands");
#define EMF_XN(L)                                                                                                 \
  hipLaunchKernelGGL((k_spmm_xnarrow<TA, TX, T, NC, L>), dim3(A.xn_nblk), dim3(kBlock), 0, s, A.rp, A.v, A.xn_row, \
                     A.xn_uptr, A.xn_ucol, A.xn_lci, x, ldx, y, ldy, alpha, rs, gamma, w, ldw, beta, z, ldz, cf)
    switch (-A.lpr) {
      case 2: EMF_XN(2); break;
      case 4: EMF_XN(4); break;
      case 8: EMF_XN(8); break;
      default: EMF_REQUIRE(false, kErrArg, "x-block SpMM takes 2, 4 or 8 lanes per row");
    }
#undef EMF_XN
    return;
  }
  if (A.lpr < 0) {
    // LDS-staged row blocks (tables built by row_blocks for this NC and T)
    EMF_REQUIRE(A.blk_row != nullptr && A.nblk > 0, kErrArg, "row-block SpMM without its block table");
    EMF_REQUIRE(xs == nullptr && zs == nullptr, kErrArg, "row-block SpMM has no scaled operands");
    if (vec) {
      // the epilogue forms compiled without runtime operands: y = alpha A x, and the
      // fused Chebyshev step with its row scaling and right-hand side
      const bool plain = gamma == T(0) && beta == T(0) && rs == nullptr && !cf.on();
      const bool cheb = gamma != T(0) && beta == T(0) && rs != nullptr && cf.on();
      const bool axpy = gamma != T(0) && beta == T(0) && rs == nullptr && !cf.on();
      // 16-bit columns at the column counts the auxiliary-space cycle and its FCG run
      // on (3 and 6); the kFormAny fallback keeps 32-bit columns
      constexpr bool kC16 = NC == 3 || NC == 6;
#define EMF_SV(F, C)                                                                                                \
  hipLaunchKernelGGL((k_spmm_stream_v<TA, TX, T, NC, stream_tile<T, NC>(), F, stream_block<T, NC>(), C>),          \
                     dim3(A.nblk), dim3(stream_block<T, NC>()), 0, s, A.rp, A.ci, A.ci16, A.ci16_base, A.v, x, y,   \
                     alpha, rs, gamma, w, beta, z, A.blk_row, cf)
#define EMF_SV2(F)          \
  if (kC16 && A.ci16)       \
    EMF_SV(F, kC16);        \
  else                      \
    EMF_SV(F, false)
      if (cheb && cf.d) {
        EMF_SV2(kFormCheb);
      } else if (cheb) {
        EMF_SV2(kFormCheb0);
      } else if (plain) {
        EMF_SV2(kFormPlain);
      } else if (axpy) {
        EMF_SV2(kFormAxpy);
      } else {
        EMF_SV(kFormAny, false);
      }
#undef EMF_SV2
#undef EMF_SV
    } else
      hipLaunchKernelGGL((k_spmm_stream<TA, TX, T, NC, stream_tile<T, NC>(), stream_block<T, NC>()>), dim3(A.nblk),
                         dim3(stream_block<T, NC>()), 0, s,
                         A.rp, A.ci, A.v, x, ldx, y, ldy, alpha, rs, gamma, w, ldw, beta, z, ldz, A.blk_row, cf);
    return;
  }
  // one trip of 4 nonzeros per lane covers a row of average length
  const int lpr = A.lpr > 0 ? A.lpr : (avg <= 8 ? 2 : (avg <= 16 ? 4 : (avg <= 32 ? 8 : (avg <= 64 ? 16 : 32))));
  const int64_t threads = (int64_t)A.nrows * lpr;
  const int g = grid_for(threads);
#define EMF_NARROW(L)                                                                                        \
  if (A.role == 1)                                                                                           \
    hipLaunchKernelGGL((k_spmm_narrow<TA, TX, T, NC, L, 1>), dim3(g), dim3(kBlock), 0, s, A.nrows, A.rp, A.ci, \
                       A.v, x, ldx, y, ldy, alpha, rs, gamma, w, ldw, beta, z, ldz, xs, zs, cf);             \
  else                                                                                                       \
    hipLaunchKernelGGL((k_spmm_narrow<TA, TX, T, NC, L, 0>), dim3(g), dim3(kBlock), 0, s, A.nrows, A.rp, A.ci, \
                       A.v, x, ldx, y, ldy, alpha, rs, gamma, w, ldw, beta, z, ldz, xs, zs, cf)
  switch (lpr) {
    case 2: EMF_NARROW(2); break;
    case 4: EMF_NARROW(4); break;
    case 8: EMF_NARROW(8); break;
    case 16: EMF_NARROW(16); break;
    default: EMF_NARROW(32); break;
  }
#undef EMF_NARROW
}

}  // namespace

int stream_tile_for(int nc, bool fp64) {
  switch (nc) {
#define EMF_TILE(K) \
  case K: return fp64 ? stream_tile<double, K>() : stream_tile<float, K>();
    EMF_TILE(1) EMF_TILE(2) EMF_TILE(3) EMF_TILE(4) EMF_TILE(6) EMF_TILE(8)
#undef EMF_TILE
    default: return 0;
  }
}

std::vector<int> row_blocks(const std::vector<int> &rp, int tile) {
  const int n = (int)rp.size() - 1;
  std::vector<int> blk{0};
  if (tile <= 0) return {};
  for (int r = 0; r < n;) {
    if (rp[r + 1] - rp[r] > tile) return {};
    int e = r;
    while (e < n && rp[e + 1] - rp[r] <= tile && e - r < tile) ++e;
    blk.push_back(e);
    r = e;
  }
  // first nonzero of every block, so a workgroup's bounds need no row-pointer load
  const size_t nb = blk.size();
  for (size_t b = 0; b < nb; ++b) blk.push_back(rp[blk[b]]);
  return blk;
}

namespace {

// One thread per group of kCol16Group nonzeros: the group's smallest column, and a
// 1 in *flag when its columns span more than 16 bits.
__global__ void k_col16_base(int nnz, const int *__restrict__ ci, int *base, int *flag) {
  const int g = blockIdx.x * kBlock + threadIdx.x;
  const int p0 = g * kCol16Group;
  if (p0 >= nnz) return;
  const int p1 = min(p0 + kCol16Group, nnz);
  int lo = ci[p0], hi = lo;
  for (int p = p0 + 1; p < p1; ++p) {
    lo = min(lo, ci[p]);
    hi = max(hi, ci[p]);
  }
  base[g] = lo;
  if (hi - lo > 65535) flag[0] = 1;
}

template <typename T>
__global__ void k_prolong_single(int64_t n, int nc, const int *__restrict__ pci, const T *__restrict__ pv,
                                 const T *__restrict__ xc, const T *__restrict__ zs, const T *__restrict__ z, T *y) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int row = (int)(i / nc), c = (int)(i % nc);
  y[i] = pv[row] * xc[(size_t)pci[row] * nc + c] + zs[row] * z[i];
}

template <typename T>
__global__ void k_scale_columns(int nnz, const int *__restrict__ ci, const T *__restrict__ v, const T *__restrict__ d,
                                T *out) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p < nnz) out[p] = v[p] * d[ci[p]];
}

__global__ void k_col16_off(int nnz, const int *__restrict__ ci, const int *__restrict__ base, uint16_t *off) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= nnz) return;
  off[p] = (uint16_t)(ci[p] - base[p / kCol16Group]);
}

}  // namespace

template <typename T>
void prolong_single(hipStream_t s, int rows, int nc, const int *pci, const T *pv, const T *xc, const T *zs,
                    const T *z, T *y) {
  const int64_t n = (int64_t)rows * nc;
  if (n == 0) return;
  hipLaunchKernelGGL((k_prolong_single<T>), dim3(grid_for(n)), dim3(kBlock), 0, s, n, nc, pci, pv, xc, zs, z, y);
  EMF_HIP(hipGetLastError());
}
template void prolong_single<float>(hipStream_t, int, int, const int *, const float *, const float *, const float *,
                                    const float *, float *);
template void prolong_single<double>(hipStream_t, int, int, const int *, const double *, const double *,
                                     const double *, const double *, double *);

template <typename T>
void scale_columns(hipStream_t s, int nnz, const int *ci, const T *v, const T *d, T *out) {
  if (nnz <= 0) return;
  hipLaunchKernelGGL((k_scale_columns<T>), dim3(grid_for(nnz)), dim3(kBlock), 0, s, nnz, ci, v, d, out);
  EMF_HIP(hipGetLastError());
}
template void scale_columns<float>(hipStream_t, int, const int *, const float *, const float *, float *);
template void scale_columns<double>(hipStream_t, int, const int *, const double *, const double *, double *);

bool compress_columns(hipStream_t s, int nnz, const int *ci, int *base, uint16_t *off, int *flag) {
  if (nnz <= 0) return false;
  const int ng = (nnz + kCol16Group - 1) / kCol16Group;
  EMF_HIP(hipMemsetAsync(flag, 0, sizeof(int), s));
  hipLaunchKernelGGL(k_col16_base, dim3((ng + kBlock - 1) / kBlock), dim3(kBlock), 0, s, nnz, ci, base, flag);
  EMF_HIP(hipGetLastError());
  int wide = 0;
  EMF_HIP(hipMemcpyAsync(&wide, flag, sizeof(int), hipMemcpyDeviceToHost, s));
  EMF_HIP(hipStreamSynchronize(s));
  if (wide) return false;
  hipLaunchKernelGGL(k_col16_off, dim3(grid_for(nnz)), dim3(kBlock), 0, s, nnz, ci, base, off);
  EMF_HIP(hipGetLastError());
  EMF_HIP(hipStreamSynchronize(s));
  return true;
}

XBlocks x_blocks(const std::vector<int> &rp, const std::vector<int> &ci, int ncols, int umax, int rmax) {
  XBlocks X;
  X.umax = umax;
  const int n = (int)rp.size() - 1;
  if (umax <= 0 || umax > 65536) return X;
  std::vector<int> mark(ncols, -1), slot(ncols, 0);
  std::vector<int> cols;
  X.row.push_back(0);
  X.uptr.push_back(0);
  X.lci.resize(ci.size());
  int blk = 0;
  for (int r = 0; r < n;) {
    cols.clear();
    int e = r;
    while (e < n && e - r < rmax) {
      int fresh = 0;
      for (int p = rp[e]; p < rp[e + 1]; ++p)
        if (mark[ci[p]] != blk) {
          mark[ci[p]] = blk;
          slot[ci[p]] = -1;  // provisional
          ++fresh;
        }
      if ((int)cols.size() + fresh > umax) {
        // roll back this row's provisional marks
        for (int p = rp[e]; p < rp[e + 1]; ++p)
          if (slot[ci[p]] == -1) mark[ci[p]] = blk - 1 < 0 ? -2 : blk - 1;
        break;
      }
      for (int p = rp[e]; p < rp[e + 1]; ++p)
        if (slot[ci[p]] == -1) {
          slot[ci[p]] = 0;
          cols.push_back(ci[p]);
        }
      ++e;
    }
    if (e == r) return XBlocks();  // one row has more than umax distinct columns
    std::sort(cols.begin(), cols.end());
    for (int k = 0; k < (int)cols.size(); ++k) slot[cols[k]] = k;
    for (int p = rp[r]; p < rp[e]; ++p) X.lci[p] = slot[ci[p]];
    X.ucol.insert(X.ucol.end(), cols.begin(), cols.end());
    X.row.push_back(e);
    X.uptr.push_back((int)X.ucol.size());
    ++blk;
    r = e;
  }
  return X;
}

template <typename TA, typename TX, typename T>
void spmm_ex(hipStream_t s, const Csr<TA> &A, int nc, const TX *x, int ldx, T *y, int ldy, T alpha, const T *rs,
             T gamma, const T *w, int ldw, T beta, const T *z, int ldz, const T *xs, const T *zs,
             const ChebFuse<T> &cf) {
  if (A.nrows == 0 || nc == 0) return;
  EMF_REQUIRE((A.lpr != -1 && A.lpr != -3) || (nc <= 8 && nc != 5 && nc != 7), kErrArg,
              "row-block SpMM supports 1, 2, 3, 4, 6 or 8 columns");
  EMF_REQUIRE(A.lpr >= -1 || A.lpr == -3 || nc >= 16 || A.xn_nblk > 0, kErrArg,
              "narrow x-block SpMM without its tables");
  if (nc >= 16 && A.lpr == -2) {
    EMF_REQUIRE(A.xb_row != nullptr && A.xb_nblk > 0 && A.xb_umax <= kXBlockCols, kErrArg,
                "x-block SpMM without its block tables");
    EMF_REQUIRE(xs == nullptr && zs == nullptr, kErrArg, "x-block SpMM has no scaled operands");
    const int nslab = (nc + 63) / 64;
    const size_t lds = (size_t)kXBlockCols * 64 * sizeof(TX);
    const int vec_ok = ((uintptr_t)x % 16 == 0) && ((size_t)ldx * sizeof(TX)) % 16 == 0 && nc % 4 == 0;
    hipLaunchKernelGGL((k_spmm_xblk<TA, TX, T>), dim3(A.xb_nblk * nslab), dim3(kBlock), lds, s, A.xb_nblk, A.rp, A.v,
                       A.xb_row, A.xb_uptr, A.xb_ucol, A.xb_lci, x, ldx, y, ldy, nc, alpha, rs, gamma, w, ldw, beta,
                       z, ldz, vec_ok, cf);
    EMF_HIP(hipGetLastError());
    return;
  }
  const int vec = 16 / (int)sizeof(TX);
  const bool vec_ok = nc % vec == 0 && (uintptr_t)x % 16 == 0 && ((size_t)ldx * sizeof(TX)) % 16 == 0;
  // untuned: 16-byte gathers with 16 lanes per row where the operands allow
  const int code = A.lpr != 0 || nc < 16 ? A.lpr : (vec_ok ? 16 * vec + 1 : 0);
  if (nc >= 16 && code >= 16 && (code & 1) && vec_ok) {
    // 16-byte gathers: slab of W = code - 1 columns, 16 / sizeof(TX) per lane (other
    // operands take the strided kernel with the same slab)
    const int W = code - 1;
    const int lw = W / vec;
    // vector epilogue when every row segment it moves is aligned to its vector size
    const size_t va = std::min<size_t>(16, sizeof(T) * vec);
    auto al = [&](const void *p, int ld) { return (uintptr_t)p % va == 0 && ((size_t)ld * sizeof(T)) % va == 0; };
    const int vepi = al(y, ldy) && (gamma == T(0) || al(w, ldw)) && (beta == T(0) || al(z, ldz)) &&
                     (!cf.on() || ((cf.d == nullptr || al(cf.d, cf.ld)) && al(cf.x, cf.ld)));
    const int rb = grid_for((int64_t)A.nrows * lw);
    const int nslab = (nc + W - 1) / W;
#define EMF_WVEC(LW)                                                                                             \
  hipLaunchKernelGGL((k_spmm_wvec<TA, TX, T, LW>), dim3(rb * nslab), dim3(kBlock), 0, s, A.nrows, rb, A.rp, A.ci, \
                     A.v, x, ldx, y, ldy, nc, alpha, rs, gamma, w, ldw, beta, z, ldz, xs, zs, vepi, cf)
    switch (lw) {
      case 4: EMF_WVEC(4); break;
      case 8: EMF_WVEC(8); break;
      case 16: EMF_WVEC(16); break;
      case 32: EMF_WVEC(32); break;
      case 64: EMF_WVEC(64); break;
      default: EMF_REQUIRE(false, kErrArg, "vector wide SpMM slab width out of range");
    }
#undef EMF_WVEC
    EMF_HIP(hipGetLastError());
    return;
  }
  if (nc >= 16) {
    // column slabs of W = LW * CPL columns (A.lpr = W when tuned), CPL = 4 columns per
    // lane; 16 lanes per row up to 64 columns, 32 up to 128, 64 beyond
    const int wdef = nc <= 64 ? 64 : (nc <= 128 ? 128 : 256);
    const int W = code >= 16 ? (code & ~1) : wdef;
    const int rb = grid_for((int64_t)A.nrows * (W / 4));
    const int nslab = (nc + W - 1) / W;
#define EMF_WIDE(LW)                                                                                            \
  hipLaunchKernelGGL((k_spmm_wide<TA, TX, T, LW, 4>), dim3(rb * nslab), dim3(kBlock), 0, s, A.nrows, rb, A.rp, \
                     A.ci, A.v, x, ldx, y, ldy, nc, alpha, rs, gamma, w, ldw, beta, z, ldz, xs, zs, cf)
    switch (W) {
      case 16: EMF_WIDE(4); break;
      case 32: EMF_WIDE(8); break;
      case 64: EMF_WIDE(16); break;
      case 128: EMF_WIDE(32); break;
      case 256: EMF_WIDE(64); break;
      default: EMF_REQUIRE(false, kErrArg, "wide SpMM slab width must be 16, 32, 64, 128 or 256");
    }
#undef EMF_WIDE
    EMF_HIP(hipGetLastError());
    return;
  }
  int c0 = 0;
  while (c0 < nc) {
    const int left = nc - c0;
    const int n = left >= 8 ? 8 : (left >= 6 ? 6 : (left >= 4 ? 4 : (left >= 3 ? 3 : (left >= 2 ? 2 : 1))));
    const T *wc = w ? w + c0 : nullptr;
    const T *zc = z ? z + c0 : nullptr;
    ChebFuse<T> cfc = cf;
    if (cf.on()) {
      if (cfc.d) cfc.d += c0;
      cfc.x += c0;
    }
    if (cfc.zcol) {
      cfc.zcol += c0;
    }
#define EMF_CASE(K)                                                                                             \
  case K:                                                                                                       \
    launch_narrow<TA, TX, T, K>(s, A, x + c0, ldx, y + c0, ldy, alpha, rs, gamma, wc, ldw, beta, zc, ldz, xs, zs, \
                                cfc);                                                                           \
    break;
    switch (n) {
      EMF_CASE(8)
      EMF_CASE(6)
      EMF_CASE(4)
      EMF_CASE(3)
      EMF_CASE(2)
      default:
        launch_narrow<TA, TX, T, 1>(s, A, x + c0, ldx, y + c0, ldy, alpha, rs, gamma, wc, ldw, beta, zc, ldz, xs, zs,
                                    cfc);
        break;
    }
#undef EMF_CASE
    c0 += n;
  }
  EMF_HIP(hipGetLastError());
}

template <typename TA, typename TX, typename T>
void spmm_mixed(hipStream_t s, const Csr<TA> &A, int nc, const TX *x, int ldx, T *y, int ldy, T alpha, const T *rs,
                T gamma, const T *w, int ldw, T beta, const T *z, int ldz) {
  spmm_ex<TA, TX, T>(s, A, nc, x, ldx, y, ldy, alpha, rs, gamma, w, ldw, beta, z, ldz, nullptr, nullptr,
                     ChebFuse<T>());
}

template <typename TA, typename TX, typename T>
int tune_lpr(hipStream_t s, const Csr<TA> &A, int nc, const TX *x, T *y, const T *rs, const ChebFuse<T> &cf) {
  if (A.nrows == 0) return 0;
  hipEvent_t e0, e1;
  EMF_HIP(hipEventCreate(&e0));
  EMF_HIP(hipEventCreate(&e1));
  int best = 0;
  float best_ms = 1e30f;
  std::vector<int> cand{2, 4, 8, 16};
  if (nc >= 16) {
    // wide: column slab widths; slabs narrower than nc trade matrix re-reads for L2
    // reuse of x
    cand.clear();
    for (int W = 16; W <= 256; W *= 2) {
      cand.push_back(W);
      if (W >= nc) break;
    }
    // 16-byte gathers (slab W, code W + 1) where nc and the row stride allow
    const int vec = 16 / (int)sizeof(TX);
    if (nc % vec == 0 && (uintptr_t)x % 16 == 0 && ((size_t)nc * sizeof(TX)) % 16 == 0)
      for (int W = 16; W <= 256; W *= 2) {
        if (W / vec >= 4 && W / vec <= 64) cand.push_back(W + 1);
        if (W >= nc) break;
      }
    if (A.xb_row && A.xb_nblk > 0) cand.push_back(-2);
  } else {
    // LDS-staged row blocks (the persistent pipelined variant, -3, only on request: it
    // can win an isolated timing yet move more bytes and lose inside the solve)
    if (A.blk_row && A.nblk > 0) cand.push_back(-1);
    if (A.xn_row && A.xn_nblk > 0)
      for (int l : {-2, -4, -8}) cand.push_back(l);  // x-blocks, 2 / 4 / 8 lanes per row
  }
  for (int lpr : cand) {
    Csr<TA> B = A;
    B.lpr = lpr;
    // the form the caller runs: y = rs .* (-A x) + y with the optional fused step when
    // rs is given, else y = A x
    auto run = [&]() {
      if (rs)
        spmm_ex<TA, TX, T>(s, B, nc, x, nc, y, nc, T(-1), rs, T(0), nullptr, 0, T(1), y, nc, nullptr, nullptr, cf);
      else
        spmm_mixed<TA, TX, T>(s, B, nc, x, nc, y, nc, T(1), nullptr, T(0), nullptr, 0, T(0), nullptr, 0);
    };
    run();
    EMF_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < 5; ++r) run();
    EMF_HIP(hipEventRecord(e1, s));
    EMF_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    EMF_HIP(hipEventElapsedTime(&ms, e0, e1));
    // the block variants must win by 5% (timing noise would otherwise flip a near
    // tie from run to run)
    if (lpr < 0) ms *= 1.05f;
    if (ms < best_ms) {
      best_ms = ms;
      best = lpr;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return best;
}

#define EMF_SPMM_INST(TA, TX, T)                                                                                    \
  template void spmm_mixed<TA, TX, T>(hipStream_t, const Csr<TA> &, int, const TX *, int, T *, int, T, const T *, T, \
                                      const T *, int, T, const T *, int);                                           \
  template void spmm_ex<TA, TX, T>(hipStream_t, const Csr<TA> &, int, const TX *, int, T *, int, T, const T *, T,    \
                                   const T *, int, T, const T *, int, const T *, const T *,                         \
                                   const ChebFuse<T> &);                                                            \
  template int tune_lpr<TA, TX, T>(hipStream_t, const Csr<TA> &, int, const TX *, T *, const T *,                   \
                                   const ChebFuse<T> &);
EMF_SPMM_INST(float, float, float)
EMF_SPMM_INST(double, double, double)
EMF_SPMM_INST(double, float, double)
EMF_SPMM_INST(float, double, float)
#undef EMF_SPMM_INST

// ------------------------------------------------------------------------------
// complex operator (C, M) on interleaved pairs
// ------------------------------------------------------------------------------

namespace {

template <typename TX, int NP, int LPR>
__global__ __launch_bounds__(kBlock) void k_op_narrow(int nrows, const int *__restrict__ rp, const int *__restrict__ ci,
                                                      const double2 *__restrict__ K, const TX *__restrict__ x, int ldx,
                                                      double *y, int ldy, double alpha, double beta, const double *z,
                                                      int ldz) {
  const int gid = xcd_block() * kBlock + threadIdx.x;
  const int row = gid / LPR;
  const int lane = gid % LPR;
  if (row >= nrows) return;
  double ar[NP], ai[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) ar[j] = ai[j] = 0.0;
  const int e = rp[row + 1];
  for (int p = rp[row] + lane; p < e; p += LPR) {
    const double2 cm = K[p];
    const TX *xr = x + (size_t)ci[p] * ldx;
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      const double re = (double)xr[2 * j], im = (double)xr[2 * j + 1];
      ar[j] += cm.x * re - cm.y * im;
      ai[j] += -cm.y * re - cm.x * im;
    }
  }
#pragma unroll
  for (int off = LPR / 2; off > 0; off >>= 1) {
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      ar[j] += shfl_xor(ar[j], off, LPR);
      ai[j] += shfl_xor(ai[j], off, LPR);
    }
  }
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    if (lane == (j % LPR)) {
      double vr = alpha * ar[j], vi = alpha * ai[j];
      if (beta != 0.0) {
        vr += beta * z[(size_t)row * ldz + 2 * j];
        vi += beta * z[(size_t)row * ldz + 2 * j + 1];
      }
      y[(size_t)row * ldy + 2 * j] = vr;
      y[(size_t)row * ldy + 2 * j + 1] = vi;
    }
  }
}

template <typename TX, int PPL>
__global__ __launch_bounds__(kBlock) void k_op_wide(int nrows, const int *__restrict__ rp, const int *__restrict__ ci,
                                                    const double2 *__restrict__ K, const TX *__restrict__ x, int ldx,
                                                    double *y, int ldy, int np, double alpha, double beta,
                                                    const double *z, int ldz) {
  int row = xcd_block() * (kBlock / 64) + threadIdx.x / 64;
  row = __builtin_amdgcn_readfirstlane(row);
  const int lane = threadIdx.x % 64;
  if (row >= nrows) return;
  double ar[PPL], ai[PPL];
#pragma unroll
  for (int j = 0; j < PPL; ++j) ar[j] = ai[j] = 0.0;
  const int b = rp[row], e = rp[row + 1];
  constexpr int kU = 4;  // four nonzeros per trip, all gathers in flight together
  for (int p0 = b; p0 < e; p0 += kU) {
    double2 cm[kU];
    int cu[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const bool ok = p0 + u < e;
      const int pc = min(p0 + u, e - 1);  // unconditional loads, masked
      cu[u] = ci[pc];
      const double2 k_ = K[pc];
      cm[u] = ok ? k_ : make_double2(0.0, 0.0);
    }
    double xr_[kU][PPL], xi_[kU][PPL];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const TX *xr = x + (size_t)cu[u] * ldx;
#pragma unroll
      for (int j = 0; j < PPL; ++j) {
        const int q = lane + 64 * j;
        xr_[u][j] = q < np ? (double)xr[2 * q] : 0.0;
        xi_[u][j] = q < np ? (double)xr[2 * q + 1] : 0.0;
      }
    }
#pragma unroll
    for (int u = 0; u < kU; ++u)
#pragma unroll
      for (int j = 0; j < PPL; ++j) {
        ar[j] += cm[u].x * xr_[u][j] - cm[u].y * xi_[u][j];
        ai[j] += -cm[u].y * xr_[u][j] - cm[u].x * xi_[u][j];
      }
  }
#pragma unroll
  for (int j = 0; j < PPL; ++j) {
    const int q = lane + 64 * j;
    if (q < np) {
      double vr = alpha * ar[j], vi = alpha * ai[j];
      if (beta != 0.0) {
        vr += beta * z[(size_t)row * ldz + 2 * q];
        vi += beta * z[(size_t)row * ldz + 2 * q + 1];
      }
      y[(size_t)row * ldy + 2 * q] = vr;
      y[(size_t)row * ldy + 2 * q + 1] = vi;
    }
  }
}

template <typename TX, int NP>
void launch_op_narrow(hipStream_t s, int nrows, int nnz, const int *rp, const int *ci, const double2 *K, const TX *x,
                      int ldx, double *y, int ldy, double alpha, double beta, const double *z, int ldz) {
  const double avg = nrows ? double(nnz) / nrows : 0.0;
  const int lpr = avg <= 6 ? 4 : (avg <= 20 ? 8 : 16);
  const int g = grid_for((int64_t)nrows * lpr);
  if (lpr == 4)
    hipLaunchKernelGGL((k_op_narrow<TX, NP, 4>), dim3(g), dim3(kBlock), 0, s, nrows, rp, ci, K, x, ldx, y, ldy, alpha,
                       beta, z, ldz);
  else if (lpr == 8)
    hipLaunchKernelGGL((k_op_narrow<TX, NP, 8>), dim3(g), dim3(kBlock), 0, s, nrows, rp, ci, K, x, ldx, y, ldy, alpha,
                       beta, z, ldz);
  else
    hipLaunchKernelGGL((k_op_narrow<TX, NP, 16>), dim3(g), dim3(kBlock), 0, s, nrows, rp, ci, K, x, ldx, y, ldy,
                       alpha, beta, z, ldz);
}

}  // namespace

template <typename TX>
void op_spmm(hipStream_t s, int nrows, int nnz, const int *rp, const int *ci, const double2 *K, int nrhs, const TX *x,
             int ldx, double *y, int ldy, double alpha, double beta, const double *z, int ldz) {
  if (nrows == 0 || nrhs == 0) return;
  if (nrhs >= 8) {
    const int g = (nrows + kBlock / 64 - 1) / (kBlock / 64);
    for (int q0 = 0; q0 < nrhs; q0 += 128) {
      const int np = std::min(128, nrhs - q0);
      const double *zq = z ? z + 2 * q0 : nullptr;
      if (np <= 64)
        hipLaunchKernelGGL((k_op_wide<TX, 1>), dim3(g), dim3(kBlock), 0, s, nrows, rp, ci, K, x + 2 * q0, ldx,
                           y + 2 * q0, ldy, np, alpha, beta, zq, ldz);
      else
        hipLaunchKernelGGL((k_op_wide<TX, 2>), dim3(g), dim3(kBlock), 0, s, nrows, rp, ci, K, x + 2 * q0, ldx,
                           y + 2 * q0, ldy, np, alpha, beta, zq, ldz);
    }
    EMF_HIP(hipGetLastError());
    return;
  }
  int q0 = 0;
  while (q0 < nrhs) {
    const int left = nrhs - q0;
    const int n = left >= 4 ? 4 : left;
    const double *zq = z ? z + 2 * q0 : nullptr;
    switch (n) {
      case 4: launch_op_narrow<TX, 4>(s, nrows, nnz, rp, ci, K, x + 2 * q0, ldx, y + 2 * q0, ldy, alpha, beta, zq, ldz); break;
      case 3: launch_op_narrow<TX, 3>(s, nrows, nnz, rp, ci, K, x + 2 * q0, ldx, y + 2 * q0, ldy, alpha, beta, zq, ldz); break;
      case 2: launch_op_narrow<TX, 2>(s, nrows, nnz, rp, ci, K, x + 2 * q0, ldx, y + 2 * q0, ldy, alpha, beta, zq, ldz); break;
      default: launch_op_narrow<TX, 1>(s, nrows, nnz, rp, ci, K, x + 2 * q0, ldx, y + 2 * q0, ldy, alpha, beta, zq, ldz); break;
    }
    q0 += n;
  }
  EMF_HIP(hipGetLastError());
}

template void op_spmm<float>(hipStream_t, int, int, const int *, const int *, const double2 *, int, const float *,
                             int, double *, int, double, double, const double *, int);
template void op_spmm<double>(hipStream_t, int, int, const int *, const int *, const double2 *, int, const double *,
                              int, double *, int, double, double, const double *, int);

// ------------------------------------------------------------------------------
// B = C + M
// ------------------------------------------------------------------------------

namespace {
template <typename T>
__global__ void k_make_b(int nnz, const double2 *__restrict__ K, T *__restrict__ B) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < nnz) B[i] = (T)(K[i].x + K[i].y);
}
}  // namespace

template <typename T>
void make_b(hipStream_t s, int nnz, const double2 *K, T *B) {
  hipLaunchKernelGGL((k_make_b<T>), dim3(grid_for(nnz)), dim3(kBlock), 0, s, nnz, K, B);
  EMF_HIP(hipGetLastError());
}
template void make_b<float>(hipStream_t, int, const double2 *, float *);
template void make_b<double>(hipStream_t, int, const double2 *, double *);

namespace {
__global__ void k_put_rows(int rows, int n, const int *__restrict__ idx, const double2 *__restrict__ v,
                           double *y, int ldy, int phase) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (phase == 0) {
    // zero the two columns of every row
    if (i < rows) {
      y[(size_t)i * ldy] = 0.0;
      y[(size_t)i * ldy + 1] = 0.0;
    }
  } else if (i < n) {
    y[(size_t)idx[i] * ldy] = v[i].x;
    y[(size_t)idx[i] * ldy + 1] = v[i].y;
  }
}
}  // namespace

// Split {re, im} blocks (n vectors of `rows` entries each, vector j at j * rows) <->
// interleaved [rows][2 n]; one thread per (row, vector), rows fastest so the split side
// is read and written coalesced.
__global__ void k_split_interleave(int rows, int n, const double *__restrict__ re, const double *__restrict__ im,
                                   double *__restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= (int64_t)rows * n) return;
  const int r = (int)(i % rows), j = (int)(i / rows);
  const double a = re[i], b = im[i];
  y[(size_t)r * 2 * n + 2 * j] = a;
  y[(size_t)r * 2 * n + 2 * j + 1] = b;
}

__global__ void k_interleave_split(int rows, int n, const double *__restrict__ x, double *__restrict__ re,
                                   double *__restrict__ im) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= (int64_t)rows * n) return;
  const int r = (int)(i % rows), j = (int)(i / rows);
  re[i] = x[(size_t)r * 2 * n + 2 * j];
  im[i] = x[(size_t)r * 2 * n + 2 * j + 1];
}

void split_to_interleaved(hipStream_t s, int rows, int n, const double *re, const double *im, double *y) {
  const int64_t m = (int64_t)rows * n;
  hipLaunchKernelGGL(k_split_interleave, dim3(grid_for(m)), dim3(kBlock), 0, s, rows, n, re, im, y);
  EMF_HIP(hipGetLastError());
}

void interleaved_to_split(hipStream_t s, int rows, int n, const double *x, double *re, double *im) {
  const int64_t m = (int64_t)rows * n;
  hipLaunchKernelGGL(k_interleave_split, dim3(grid_for(m)), dim3(kBlock), 0, s, rows, n, x, re, im);
  EMF_HIP(hipGetLastError());
}

void put_rows(hipStream_t s, int rows, int n, const int *idx, const double2 *v, double *y, int ldy) {
  hipLaunchKernelGGL(k_put_rows, dim3(grid_for(rows)), dim3(kBlock), 0, s, rows, n, idx, v, y, ldy, 0);
  if (n > 0) hipLaunchKernelGGL(k_put_rows, dim3(grid_for(n)), dim3(kBlock), 0, s, rows, n, idx, v, y, ldy, 1);
  EMF_HIP(hipGetLastError());
}

// ------------------------------------------------------------------------------
// reductions
// ------------------------------------------------------------------------------

namespace {

constexpr int kDotBlocks = 512;

// Rows [r0, r1) of reduction block b: the segment chunks, or an even split of all rows.
__device__ __forceinline__ void chunk_rows(const Seg &sg, int rows, int b, int nb, int &r0, int &r1) {
  if (sg.chunk_row) {
    r0 = sg.chunk_row[b];
    r1 = sg.chunk_row[b + 1];
  } else {
    r0 = (int)((int64_t)rows * b / nb);
    r1 = (int)((int64_t)rows * (b + 1) / nb);
  }
}

// Threads of a block are assigned to columns c = tid % nc and stride over the
// block's rows; the per-column block sums go to work[block][c].
template <typename TX, typename TY>
__global__ __launch_bounds__(kBlock) void k_coldot(int rows, int nc, const TX *__restrict__ x, int ldx,
                                                   const TY *__restrict__ y, int ldy, double *work, Seg sg) {
  __shared__ double sh[kBlock];
  const int rpp = kBlock / nc;
  const int per_block = rpp * nc;
  const int t = threadIdx.x;
  int r0, r1;
  chunk_rows(sg, rows, blockIdx.x, gridDim.x, r0, r1);
  double acc = 0.0;
  if (t < per_block) {
    const int c = t % nc;
    // four rows per trip with every load issued first (clamped, masked), summed in
    // row order
    for (int r = r0 + t / nc; r < r1; r += 4 * rpp) {
      double xv[4], yv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int ru = min(r + u * rpp, r1 - 1);
        xv[u] = (double)x[(size_t)ru * ldx + c];
        yv[u] = (double)y[(size_t)ru * ldy + c];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (r + u * rpp < r1) acc += xv[u] * yv[u];
    }
  }
  sh[t] = acc;
  __syncthreads();
  if (t < nc) {
    double sum = 0.0;
    for (int k = t; k < per_block; k += nc) sum += sh[k];
    work[(size_t)blockIdx.x * nc + t] = sum;
  }
}

// Two dot products against one shared operand: x . y1 and x . y2 per column, in the
// order of k_coldot (so each equals the corresponding coldot bitwise); the partials
// go to work[0][block][c] and work[1][block][c].
template <typename TX>
__global__ __launch_bounds__(kBlock) void k_coldot2(int rows, int nc, const TX *__restrict__ x, int ldx,
                                                    const double *__restrict__ y1, const double *__restrict__ y2,
                                                    int ldy, double *work, Seg sg) {
  __shared__ double sh[2][kBlock];
  const int rpp = kBlock / nc;
  const int per_block = rpp * nc;
  const int t = threadIdx.x;
  int r0, r1;
  chunk_rows(sg, rows, blockIdx.x, gridDim.x, r0, r1);
  double acc1 = 0.0, acc2 = 0.0;
  if (t < per_block) {
    const int c = t % nc;
    for (int r = r0 + t / nc; r < r1; r += 4 * rpp) {
      double xv[4], av[4], bv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int ru = min(r + u * rpp, r1 - 1);
        xv[u] = (double)x[(size_t)ru * ldx + c];
        av[u] = y1[(size_t)ru * ldy + c];
        bv[u] = y2[(size_t)ru * ldy + c];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (r + u * rpp < r1) {
          acc1 += xv[u] * av[u];
          acc2 += xv[u] * bv[u];
        }
    }
  }
  sh[0][t] = acc1;
  sh[1][t] = acc2;
  __syncthreads();
  if (t < nc) {
    double s1 = 0.0, s2 = 0.0;
    for (int k = t; k < per_block; k += nc) {
      s1 += sh[0][k];
      s2 += sh[1][k];
    }
    work[(size_t)blockIdx.x * nc + t] = s1;
    work[((size_t)gridDim.x + blockIdx.x) * nc + t] = s2;
  }
}

// Fixed-order block sum of sh[0..kBlock) (tree over LDS); result in sh[0].
__device__ __forceinline__ void block_tree_sum(double *sh) {
  const int t = threadIdx.x;
#pragma unroll
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    __syncthreads();
    if (t < w) sh[t] += sh[t + w];
  }
  __syncthreads();
}

// One block per output (basis vector i, segment s, column c): the partials of the
// segment's reduction blocks are summed in a fixed order (strided per thread, then a
// tree), so results do not depend on scheduling.  work is [i][block][c]; out is
// [i][s * nc + c].
__global__ __launch_bounds__(kBlock) void k_finish(int nblocks, int nc, const double *__restrict__ work, double *out,
                                                   Seg sg) {
  __shared__ double sh[kBlock];
  const int nl = sg.nseg * nc;
  const int idx = blockIdx.x;
  const int i = idx / nl, rem = idx % nl, sgi = rem / nc, c = rem % nc;
  const int b0 = sg.seg_chunk ? sg.seg_chunk[sgi] : 0, b1 = sg.seg_chunk ? sg.seg_chunk[sgi + 1] : nblocks;
  double s = 0.0;
  for (int b = b0 + threadIdx.x; b < b1; b += kBlock) s += work[((size_t)i * nblocks + b) * nc + c];
  sh[threadIdx.x] = s;
  block_tree_sum(sh);
  if (threadIdx.x == 0) out[idx] = sh[0];
}

int reduce_blocks(int rows, int nc, const Seg &sg) {
  if (sg.chunk_row) return sg.nchunks;
  const int rpp = kBlock / nc;
  return std::max(1, std::min(kDotBlocks, (rows + rpp - 1) / rpp));
}

}  // namespace

size_t dot_work_size(int rows, int nc, const Seg &sg) {
  (void)rows;
  return (size_t)std::max(kDotBlocks, sg.nchunks) * std::max(nc, 1) + 64;
}

template <typename TX, typename TY>
void coldot(hipStream_t s, int rows, int nc, const TX *x, int ldx, const TY *y, int ldy, double *out, double *work,
            const Seg &sg) {
  if (nc > kBlock) {
    EMF_REQUIRE(sg.nseg == 1, kErrArg, "segmented reductions support at most 256 columns");
    for (int c0 = 0; c0 < nc; c0 += kBlock)
      coldot<TX, TY>(s, rows, std::min(kBlock, nc - c0), x + c0, ldx, y + c0, ldy, out + c0, work, sg);
    return;
  }
  const int nb = reduce_blocks(rows, nc, sg);
  hipLaunchKernelGGL((k_coldot<TX, TY>), dim3(nb), dim3(kBlock), 0, s, rows, nc, x, ldx, y, ldy, work, sg);
  hipLaunchKernelGGL(k_finish, dim3(sg.nseg * nc), dim3(kBlock), 0, s, nb, nc, work, out, sg);
  EMF_HIP(hipGetLastError());
}
namespace {

// One lane per (interior vertex, column): the vertex's edges four at a time (codes,
// then the residual entries and half tangents, all loads before the sums).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_aux_restrict(int nv0, int nblk, int nc, const int *__restrict__ vrp,
                                                         const int *__restrict__ vcode, const T *__restrict__ th,
                                                         const double *__restrict__ r, T *vb) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= (int64_t)nv0 * nc) return;
  const int v = (int)(i / nc), c = (int)(i % nc);
  T g = T(0), px = T(0), py = T(0), pz = T(0);
  const int p0 = vrp[v], p1 = vrp[v + 1];
  // sixteen edges per trip (an interior vertex of a Kuhn mesh has 14): all codes,
  // then all gathers in flight together
  constexpr int kU = 16;
  for (int p = p0; p < p1; p += kU) {
    int code[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) code[u] = vcode[min(p + u, p1 - 1)];
    double rv[kU];
    VecT<T, 4> t[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int e = code[u] & 0x7fffffff;
      rv[u] = r[(size_t)e * nc + c];
      t[u] = *reinterpret_cast<const VecT<T, 4> *>(th + (size_t)e * 4);
    }
#pragma unroll
    for (int u = 0; u < kU; ++u)
      if (p + u < p1) {
        const T x = (T)rv[u];
        g += code[u] < 0 ? -x : x;
        px += t[u].v[0] * x;
        py += t[u].v[1] * x;
        pz += t[u].v[2] * x;
      }
  }
  const size_t o = (size_t)v * nc + c, blk = (size_t)nv0 * nc;
  int k = 0;
  if (nblk == 4) vb[o] = g, k = 1;
  vb[k * blk + o] = px;
  vb[(k + 1) * blk + o] = py;
  vb[(k + 2) * blk + o] = pz;
}

// One lane per (edge, column): the blocks' entries of both interior endpoints, all
// loads first.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_aux_prolong(int ne, int nv0, int nblk, int nc, const int2 *__restrict__ ev,
                                                        const T *__restrict__ th, const T *__restrict__ vx, T *y) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= (int64_t)ne * nc) return;
  const int e = (int)(i / nc), c = (int)(i % nc);
  const int2 ab = ev[e];
  const VecT<T, 4> t = *reinterpret_cast<const VecT<T, 4> *>(th + (size_t)e * 4);
  const size_t blk = (size_t)nv0 * nc;
  const int k0 = nblk == 4 ? 1 : 0;
  T v[2][4];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int w = q == 0 ? ab.x : ab.y;
    const size_t o = (size_t)max(w, 0) * nc + c;
    v[q][0] = nblk == 4 ? vx[o] : T(0);
    v[q][1] = vx[k0 * blk + o];
    v[q][2] = vx[(k0 + 1) * blk + o];
    v[q][3] = vx[(k0 + 2) * blk + o];
  }
  const size_t iy = (size_t)e * nc + c;
  const T y0 = y[iy];
  T acc = T(0);
  if (ab.x >= 0) acc += -v[0][0] + t.v[0] * v[0][1] + t.v[1] * v[0][2] + t.v[2] * v[0][3];
  if (ab.y >= 0) acc += v[1][0] + t.v[0] * v[1][1] + t.v[1] * v[1][2] + t.v[2] * v[1][3];
  y[iy] = y0 + acc;
}

// k_aux_prolong with one lane per edge for all NC columns (NC a multiple of 3): the
// endpoints and half tangent once per edge, the vertex rows and y as 3-wide vector
// chunks; per (edge, column) the operations and their order are those of
// k_aux_prolong.
template <typename T, int NC>
__global__ __launch_bounds__(kBlock) void k_aux_prolong_v(int ne, int nv0, int nblk, const int2 *__restrict__ ev,
                                                          const T *__restrict__ th, const T *__restrict__ vx, T *y) {
  static_assert(NC % 3 == 0, "3-wide chunks");
  constexpr int CH = NC / 3;
  using VC = ChunkT<T, 3>;
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= ne) return;
  const int2 ab = ev[e];
  const VecT<T, 4> t = *reinterpret_cast<const VecT<T, 4> *>(th + (size_t)e * 4);
  const size_t blk = (size_t)nv0 * NC;
  const int k0 = nblk == 4 ? 1 : 0;
  VC v[2][4][CH];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int w = q == 0 ? ab.x : ab.y;
    const size_t o = (size_t)max(w, 0) * NC;
#pragma unroll
    for (int k = 0; k < CH; ++k) {
      if (nblk == 4)
        v[q][0][k] = *reinterpret_cast<const VC *>(vx + o + 3 * k);
      else
        v[q][0][k] = VC{{T(0), T(0), T(0)}};
      v[q][1][k] = *reinterpret_cast<const VC *>(vx + k0 * blk + o + 3 * k);
      v[q][2][k] = *reinterpret_cast<const VC *>(vx + (k0 + 1) * blk + o + 3 * k);
      v[q][3][k] = *reinterpret_cast<const VC *>(vx + (k0 + 2) * blk + o + 3 * k);
    }
  }
  VC yv[CH];
#pragma unroll
  for (int k = 0; k < CH; ++k) yv[k] = *reinterpret_cast<const VC *>(y + (size_t)e * NC + 3 * k);
#pragma unroll
  for (int k = 0; k < CH; ++k) {
    VC out;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      T acc = T(0);
      if (ab.x >= 0)
        acc += -v[0][0][k].v[i] + t.v[0] * v[0][1][k].v[i] + t.v[1] * v[0][2][k].v[i] + t.v[2] * v[0][3][k].v[i];
      if (ab.y >= 0)
        acc += v[1][0][k].v[i] + t.v[0] * v[1][1][k].v[i] + t.v[1] * v[1][2][k].v[i] + t.v[2] * v[1][3][k].v[i];
      out.v[i] = yv[k].v[i] + acc;
    }
    *reinterpret_cast<VC *>(y + (size_t)e * NC + 3 * k) = out;
  }
}

}  // namespace

namespace {

// One thread per (vertex row, block k): the four blocks of a row are independent sums.
__global__ __launch_bounds__(kBlock) void k_galerkin(GalerkinDev G, const int *__restrict__ brp,
                                                     const int *__restrict__ bci, const double *__restrict__ bv,
                                                     double *v0, double *v1, double *v2, double *v3, int *missing) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= (int64_t)G.nv0 * 4) return;
  const int u = (int)(i >> 2), k = (int)(i & 3);
  double *vals = k == 0 ? v0 : (k == 1 ? v1 : (k == 2 ? v2 : v3));
  const double *rv = G.rv[k], *pv = G.pv[k];
  const int a0 = G.arp[u], a1 = G.arp[u + 1];
  int miss = 0;
  // rows of up to kSlots entries (an interior vertex of a Kuhn mesh has 15) keep their
  // columns and sums in registers: each product is added to its slot by a select over
  // the unrolled slots, so every sum takes the same additions in the same order as
  // the loop below (and the host pass)
  constexpr int kSlots = 24;
  if (a1 - a0 <= kSlots) {
    const int len = a1 - a0;
    int cols[kSlots];
    double acc[kSlots];
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      cols[j] = j < len ? G.aci[a0 + j] : -1;
      acc[j] = 0.0;
    }
    for (int p = G.rrp[u]; p < G.rrp[u + 1]; ++p) {
      const int ed = G.rci[p];
      const double r = rv[p];
      for (int q = brp[ed]; q < brp[ed + 1]; ++q) {
        const int f = bci[q];
        const double ra = r * bv[q];
        for (int s2 = G.prp[f]; s2 < G.prp[f + 1]; ++s2) {
          const int col = G.pci[s2];
          const double prod = ra * pv[s2];
          bool hit = false;
#pragma unroll
          for (int j = 0; j < kSlots; ++j) {
            const bool m = cols[j] == col;
            acc[j] = m ? acc[j] + prod : acc[j];
            hit |= m;
          }
          miss += hit ? 0 : 1;
        }
      }
    }
#pragma unroll
    for (int j = 0; j < kSlots; ++j)
      if (j < len) vals[a0 + j] = acc[j];
    if (miss && k == 0) atomicAdd(missing, miss);
    return;
  }
  for (int t = a0; t < a1; ++t) vals[t] = 0.0;
  for (int p = G.rrp[u]; p < G.rrp[u + 1]; ++p) {
    const int ed = G.rci[p];
    const double r = rv[p];
    for (int q = brp[ed]; q < brp[ed + 1]; ++q) {
      const int f = bci[q];
      const double ra = r * bv[q];
      for (int s2 = G.prp[f]; s2 < G.prp[f + 1]; ++s2) {
        const int col = G.pci[s2];
        int t = a0;
        while (t < a1 && G.aci[t] != col) ++t;
        if (t == a1) {
          ++miss;
          continue;
        }
        const double prod = ra * pv[s2];
        vals[t] = vals[t] + prod;
      }
    }
  }
  if (miss && k == 0) atomicAdd(missing, miss);
}

template <typename T>
__global__ void k_l1_inverse(int rows, const int *__restrict__ rp, const double *__restrict__ v, T *l1) {
#pragma clang fp contract(off)
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= rows) return;
  double s = 0.0;
  for (int p = rp[r]; p < rp[r + 1]; ++p) s += fabs(v[p]);
  l1[r] = (T)(s > 0.0 ? 1.0 / s : 0.0);
}

}  // namespace

void galerkin_dev(hipStream_t s, const GalerkinDev &G, const int *brp, const int *bci, const double *bv,
                  double *const vals[4], int *missing) {
  EMF_HIP(hipMemsetAsync(missing, 0, sizeof(int), s));
  if (G.nv0 == 0) return;
  hipLaunchKernelGGL(k_galerkin, dim3(((int64_t)G.nv0 * 4 + kBlock - 1) / kBlock), dim3(kBlock), 0, s, G, brp, bci, bv,
                     vals[0], vals[1], vals[2], vals[3], missing);
  EMF_HIP(hipGetLastError());
}
template <typename T>
void l1_inverse_dev(hipStream_t s, int rows, const int *rp, const double *v, T *l1) {
  if (rows == 0) return;
  hipLaunchKernelGGL((k_l1_inverse<T>), dim3((rows + kBlock - 1) / kBlock), dim3(kBlock), 0, s, rows, rp, v, l1);
  EMF_HIP(hipGetLastError());
}
template void l1_inverse_dev<float>(hipStream_t, int, const int *, const double *, float *);
template void l1_inverse_dev<double>(hipStream_t, int, const int *, const double *, double *);

template <typename T>
void aux_restrict(hipStream_t s, const AuxMap<T> &M, int nc, const double *r, T *vb) {
  EMF_REQUIRE(M.nblk == 3 || M.nblk == 4, kErrArg, "auxiliary transfer: 3 or 4 blocks");
  const int64_t n = (int64_t)M.nv0 * nc;
  if (n == 0) return;
  // a lane per (vertex, column): a lane per vertex for all columns was measured slower
  // (62.3 against 32.7 us at 1.55M edges and 3 columns: a quarter of the lanes, each
  // with a longer chain of dependent sums)
  hipLaunchKernelGGL((k_aux_restrict<T>), dim3(grid_for(n)), dim3(kBlock), 0, s, M.nv0, M.nblk, nc, M.vrp, M.vcode,
                     M.th, r, vb);
  EMF_HIP(hipGetLastError());
}
template <typename T>
void aux_prolong(hipStream_t s, const AuxMap<T> &M, int nc, const T *vx, T *y) {
  EMF_REQUIRE(M.nblk == 3 || M.nblk == 4, kErrArg, "auxiliary transfer: 3 or 4 blocks");
  const int64_t n = (int64_t)M.ne * nc;
  if (n == 0) return;
  // a lane per edge for 3 and 6 columns, else a lane per (edge, column)
  if (nc == 3)
    hipLaunchKernelGGL((k_aux_prolong_v<T, 3>), dim3(grid_for(M.ne)), dim3(kBlock), 0, s, M.ne, M.nv0, M.nblk, M.ev,
                       M.th, vx, y);
  else if (nc == 6)
    hipLaunchKernelGGL((k_aux_prolong_v<T, 6>), dim3(grid_for(M.ne)), dim3(kBlock), 0, s, M.ne, M.nv0, M.nblk, M.ev,
                       M.th, vx, y);
  else
    hipLaunchKernelGGL((k_aux_prolong<T>), dim3(grid_for(n)), dim3(kBlock), 0, s, M.ne, M.nv0, M.nblk, nc, M.ev, M.th,
                       vx, y);
  EMF_HIP(hipGetLastError());
}
template void aux_restrict<float>(hipStream_t, const AuxMap<float> &, int, const double *, float *);
template void aux_restrict<double>(hipStream_t, const AuxMap<double> &, int, const double *, double *);
template void aux_prolong<float>(hipStream_t, const AuxMap<float> &, int, const float *, float *);
template void aux_prolong<double>(hipStream_t, const AuxMap<double> &, int, const double *, double *);

template <typename TX>
void coldot2(hipStream_t s, int rows, int nc, const TX *x, int ldx, const double *y1, const double *y2, int ldy,
             double *out1, double *out2, double *work, const Seg &sg) {
  if (nc > kBlock) {
    EMF_REQUIRE(sg.nseg == 1, kErrArg, "segmented reductions support at most 256 columns");
    for (int c0 = 0; c0 < nc; c0 += kBlock)
      coldot2<TX>(s, rows, std::min(kBlock, nc - c0), x + c0, ldx, y1 + c0, y2 + c0, ldy, out1 + c0, out2 + c0, work,
                  sg);
    return;
  }
  const int nb = reduce_blocks(rows, nc, sg);
  const int nl = sg.nseg * nc;
  hipLaunchKernelGGL((k_coldot2<TX>), dim3(nb), dim3(kBlock), 0, s, rows, nc, x, ldx, y1, y2, ldy, work, sg);
  // k_finish writes [i][nl]: the two results land in out1 when out2 == out1 + nl,
  // otherwise each half is finished on its own
  if (out2 == out1 + nl) {
    hipLaunchKernelGGL(k_finish, dim3(2 * nl), dim3(kBlock), 0, s, nb, nc, work, out1, sg);
  } else {
    hipLaunchKernelGGL(k_finish, dim3(nl), dim3(kBlock), 0, s, nb, nc, work, out1, sg);
    hipLaunchKernelGGL(k_finish, dim3(nl), dim3(kBlock), 0, s, nb, nc, work + (size_t)nb * nc, out2, sg);
  }
  EMF_HIP(hipGetLastError());
}
template void coldot2<float>(hipStream_t, int, int, const float *, int, const double *, const double *, int, double *,
                             double *, double *, const Seg &);
template void coldot2<double>(hipStream_t, int, int, const double *, int, const double *, const double *, int,
                              double *, double *, double *, const Seg &);
template void coldot<float, float>(hipStream_t, int, int, const float *, int, const float *, int, double *, double *,
                                   const Seg &);
template void coldot<double, double>(hipStream_t, int, int, const double *, int, const double *, int, double *,
                                     double *, const Seg &);
template void coldot<double, float>(hipStream_t, int, int, const double *, int, const float *, int, double *,
                                    double *, const Seg &);

namespace {

// Up to 8 basis vectors per pass; w is read once per pass.
template <int M>
__global__ __launch_bounds__(kBlock) void k_mdot(int rows, int nc, const double *__restrict__ V, size_t vstride,
                                                 const double *__restrict__ w, double *work, Seg sg) {
  __shared__ double sh[M][kBlock];
  const int rpp = kBlock / nc;
  const int per_block = rpp * nc;
  const int t = threadIdx.x;
  int r0, r1;
  chunk_rows(sg, rows, blockIdx.x, gridDim.x, r0, r1);
  double acc[M];
#pragma unroll
  for (int i = 0; i < M; ++i) acc[i] = 0.0;
  if (t < per_block) {
    const int c = t % nc;
    for (int r = r0 + t / nc; r < r1; r += rpp) {
      const size_t o = (size_t)r * nc + c;
      const double wv = w[o];
#pragma unroll
      for (int i = 0; i < M; ++i) acc[i] += V[i * vstride + o] * wv;
    }
  }
#pragma unroll
  for (int i = 0; i < M; ++i) sh[i][t] = acc[i];
  __syncthreads();
  if (t < nc) {
#pragma unroll
    for (int i = 0; i < M; ++i) {
      double sum = 0.0;
      for (int k = t; k < per_block; k += nc) sum += sh[i][k];
      work[((size_t)i * gridDim.x + blockIdx.x) * nc + t] = sum;
    }
  }
}

}  // namespace

void mdot(hipStream_t s, int rows, int nc, const double *V, size_t vstride, int m, const double *w, double *out,
          double *work, const Seg &sg) {
  EMF_REQUIRE(nc <= kBlock, kErrArg, "mdot supports at most 256 columns per call");
  const int nb = reduce_blocks(rows, nc, sg);
  const int nl = sg.nseg * nc;
  for (int i0 = 0; i0 < m; i0 += 8) {
    const int mm = std::min(8, m - i0);
    const double *Vi = V + (size_t)i0 * vstride;
    switch (mm) {
#define EMF_MDOT(K)                                                                                    \
  case K:                                                                                              \
    hipLaunchKernelGGL((k_mdot<K>), dim3(nb), dim3(kBlock), 0, s, rows, nc, Vi, vstride, w, work, sg); \
    break;
      EMF_MDOT(1) EMF_MDOT(2) EMF_MDOT(3) EMF_MDOT(4) EMF_MDOT(5) EMF_MDOT(6) EMF_MDOT(7) EMF_MDOT(8)
#undef EMF_MDOT
    }
    hipLaunchKernelGGL(k_finish, dim3(nl * mm), dim3(kBlock), 0, s, nb, nc, work, out + (size_t)i0 * nl, sg);
  }
  EMF_HIP(hipGetLastError());
}

// ------------------------------------------------------------------------------
// block updates
// ------------------------------------------------------------------------------

namespace {

// Per-column scalars of row r, column c sit at [segment(r) * nc + c].
__device__ __forceinline__ int scalar_index(const int *rseg, int r, int nc, int c) {
  return (rseg ? rseg[r] * nc : 0) + c;
}

template <typename TB>
__global__ void k_mupdate(int64_t n, int nc, const TB *__restrict__ V, size_t vstride, int m,
                          const double *__restrict__ h, int hld, double *w, double sign, const int *rseg) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int c = scalar_index(rseg, (int)(i / nc), nc, (int)(i % nc));
  double acc = 0.0;
  for (int k = 0; k < m; ++k) acc += h[(size_t)k * hld + c] * (double)V[k * vstride + i];
  w[i] -= sign * acc;
}

template <typename TX, typename TY>
__global__ void k_axpby_cols(int rows, int nc, const double *__restrict__ a, const TX *__restrict__ x, int ldx,
                             const double *__restrict__ b, TY *y, int ldy, const int *rseg) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= (int64_t)rows * nc) return;
  const int r = (int)(i / nc), c = (int)(i % nc);
  const int k = scalar_index(rseg, r, nc, c);
  double v = a[k] * (double)x[(size_t)r * ldx + c];
  if (b) v += b[k] * (double)y[(size_t)r * ldy + c];
  y[(size_t)r * ldy + c] = (TY)v;
}

template <typename TX, typename TY>
__global__ void k_axpby(int rows, int nc, double alpha, const TX *__restrict__ x, int ldx, double beta, TY *y,
                        int ldy) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= (int64_t)rows * nc) return;
  const int r = (int)(i / nc), c = (int)(i % nc);
  double v = alpha * (double)x[(size_t)r * ldx + c];
  if (beta != 0.0) v += beta * (double)y[(size_t)r * ldy + c];
  y[(size_t)r * ldy + c] = (TY)v;
}

// The residual step of the flexible CG: alpha from rz, pq and the active flag (every
// block of a segment forms the same value and stores it), r -= alpha q, and the block
// partials of r . r in work[block][c].  A chunk of rows lies in one segment.  With
// cf.bt set, also the fp32 cycle's copy of r and its first Chebyshev iterate (the
// values k_cheb_first forms from r).
__global__ __launch_bounds__(kBlock) void k_fcg_res(int rows, int nc, const double *__restrict__ rz,
                                                    const double *__restrict__ pq,
                                                    const uint8_t *__restrict__ active, double *alpha,
                                                    const double *__restrict__ q, double *r, double *work, Seg sg,
                                                    CycleFill cf) {
  __shared__ double sh[kBlock];
  const int rpp = kBlock / nc;
  const int per_block = rpp * nc;
  const int t = threadIdx.x;
  int r0, r1;
  chunk_rows(sg, rows, blockIdx.x, gridDim.x, r0, r1);
  double acc = 0.0;
  if (t < per_block && r0 < r1) {
    const int c = t % nc;
    const int k = scalar_index(sg.rseg, r0, nc, c);
    const double a = (active[k] && pq[k] != 0.0) ? rz[k] / pq[k] : 0.0;
    if (t < nc) alpha[k] = a;
    // four rows per trip, every load issued first (clamped rows, masked updates)
    for (int row = r0 + t / nc; row < r1; row += 4 * rpp) {
      double rv[4], qv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const size_t o = (size_t)min(row + u * rpp, r1 - 1) * nc + c;
        rv[u] = r[o];
        qv[u] = q[o];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (row + u * rpp < r1) {
          const size_t o = (size_t)(row + u * rpp) * nc + c;
          const double nr = rv[u] - a * qv[u];
          r[o] = nr;
          acc += nr * nr;
          if (cf.bt) {
            const float bv = (float)nr;
            cf.bt[o] = bv;
            cf.x1[o] = (float)(cf.c0 * (double)(cf.dinv[row + u * rpp] * bv));
          }
        }
    }
  }
  sh[t] = acc;
  __syncthreads();
  if (t < nc) {
    double sum = 0.0;
    for (int k = t; k < per_block; k += nc) sum += sh[k];
    work[(size_t)blockIdx.x * nc + t] = sum;
  }
}

// k_finish for r . r followed by the convergence test of its logical column: the
// flag drops once rr <= tol2 * bb; rr and the flag also go to the pinned host arrays.
__global__ __launch_bounds__(kBlock) void k_finish_conv(int nblocks, int nc, const double *__restrict__ work,
                                                        double *out, Seg sg, const double *__restrict__ bb,
                                                        double tol2, uint8_t *active, double *h_out,
                                                        uint8_t *h_active) {
  __shared__ double sh[kBlock];
  const int idx = blockIdx.x;
  const int sgi = idx / nc, c = idx % nc;
  const int b0 = sg.seg_chunk ? sg.seg_chunk[sgi] : 0, b1 = sg.seg_chunk ? sg.seg_chunk[sgi + 1] : nblocks;
  double s = 0.0;
  for (int b = b0 + threadIdx.x; b < b1; b += kBlock) s += work[(size_t)b * nc + c];
  sh[threadIdx.x] = s;
  block_tree_sum(sh);
  if (threadIdx.x == 0) {
    const double v = sh[0];
    uint8_t a = active[idx];
    if (a && v <= tol2 * bb[idx]) a = 0;
    out[idx] = v;
    active[idx] = a;
    h_out[idx] = v;
    h_active[idx] = a;
  }
}

// The delayed solution update of the flexible CG and the next search direction in
// one pass: x += alpha p ; p = z + beta p (per (segment, column) scalars; beta formed
// here from zq, pq and the active flag).
template <typename TZ>
__global__ void k_fcg_px(int rows, int nc, const double *__restrict__ alpha, const double *__restrict__ zq,
                         const double *__restrict__ pq, const uint8_t *__restrict__ active,
                         const TZ *__restrict__ z, double *p, double *x, const int *rseg) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= (int64_t)rows * nc) return;
  const int r = (int)(i / nc), c = (int)(i % nc);
  const int k = scalar_index(rseg, r, nc, c);
  const double pv = p[i], xv = x[i], zv = (double)z[i];
  const double beta = (active[k] && pq[k] != 0.0) ? -zq[k] / pq[k] : 0.0;
  x[i] = xv + alpha[k] * pv;
  p[i] = zv + beta * pv;
}

template <typename T>
__global__ void k_cheb_first(int64_t n, int nc, const T *__restrict__ dinv, const double *__restrict__ b, T *bt, T *x,
                             double inv_theta) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const T bv = (T)b[i];
  if (bt) bt[i] = bv;
  x[i] = (T)(inv_theta * (double)(dinv[i / nc] * bv));
}

template <typename T>
__global__ void k_diag_scale(int64_t n, int nc, const T *__restrict__ dinv, const T *__restrict__ b, T *x) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  x[i] = dinv[i / nc] * b[i];
}

// One block per (64 rows, group of CG columns c0 + [0, CG)): lane l of wavefront w
// sums row i0 + l over the w-th quarter of [jl, jh) for all CG columns, each load of
// the transposed inverse AinvT[j][i] (consecutive lanes, consecutive i: coalesced)
// serving the CG columns, whose b[j][c] are the same for all lanes; the four quarters
// are added in a fixed order.  [jl, jh) is the block's column range (all of j
// without a range table).  Columns past nc read the last column and are not stored.
template <typename T, int CG>
__global__ __launch_bounds__(kBlock) void k_dense_solve(int n, int nc, const double *__restrict__ AinvT,
                                                        const int *__restrict__ range, const T *__restrict__ b,
                                                        T *x) {
  static_assert(kBlock == 256, "four wavefronts per block");
  __shared__ double part[4][CG][64];
  const int c0 = blockIdx.y * CG;
  const int jl = range ? range[2 * blockIdx.x] : 0, jh = range ? range[2 * blockIdx.x + 1] : n;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + lane;
  const int q = (jh - jl + 3) / 4, j0 = jl + w * q, j1 = min(jh, j0 + q);
  double s[CG];
#pragma unroll
  for (int k = 0; k < CG; ++k) s[k] = 0.0;
  if (i < n) {
    const double *a = AinvT + i;
    const T *bc = b + c0;
    const int last = nc - 1 - c0;
#pragma unroll 4
    for (int j = j0; j < j1; ++j) {
      const double av = a[(size_t)j * n];
#pragma unroll
      for (int k = 0; k < CG; ++k) s[k] += av * (double)bc[(size_t)j * nc + min(k, last)];
    }
  }
#pragma unroll
  for (int k = 0; k < CG; ++k) part[w][k][lane] = s[k];
  __syncthreads();
  if (w == 0 && i < n) {
#pragma unroll
    for (int k = 0; k < CG; ++k)
      if (c0 + k < nc)
        x[(size_t)i * nc + c0 + k] =
            (T)(((part[0][k][lane] + part[1][k][lane]) + part[2][k][lane]) + part[3][k][lane]);
  }
}

__global__ void k_fcg_beta(int nc, const double *zq, const double *pq, const uint8_t *active, double *beta) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= nc) return;
  beta[c] = (active[c] && pq[c] != 0.0) ? -zq[c] / pq[c] : 0.0;
}

}  // namespace

template <typename TB>
void mupdate(hipStream_t s, int rows, int nc, const TB *V, size_t vstride, int m, const double *h, int hld, double *w,
             double sign, const Seg &sg) {
  if (m <= 0) return;
  const int64_t n = (int64_t)rows * nc;
  hipLaunchKernelGGL((k_mupdate<TB>), dim3(grid_for(n)), dim3(kBlock), 0, s, n, nc, V, vstride, m, h, hld, w, sign,
                     sg.rseg);
  EMF_HIP(hipGetLastError());
}
template void mupdate<float>(hipStream_t, int, int, const float *, size_t, int, const double *, int, double *, double,
                             const Seg &);
template void mupdate<double>(hipStream_t, int, int, const double *, size_t, int, const double *, int, double *,
                              double, const Seg &);

template <typename TX, typename TY>
void axpby_cols(hipStream_t s, int rows, int nc, const double *a, const TX *x, int ldx, const double *b, TY *y,
                int ldy, const Seg &sg) {
  const int64_t n = (int64_t)rows * nc;
  hipLaunchKernelGGL((k_axpby_cols<TX, TY>), dim3(grid_for(n)), dim3(kBlock), 0, s, rows, nc, a, x, ldx, b, y, ldy,
                     sg.rseg);
  EMF_HIP(hipGetLastError());
}
template void axpby_cols<float, float>(hipStream_t, int, int, const double *, const float *, int, const double *,
                                       float *, int, const Seg &);
template void axpby_cols<double, double>(hipStream_t, int, int, const double *, const double *, int, const double *,
                                         double *, int, const Seg &);
template void axpby_cols<float, double>(hipStream_t, int, int, const double *, const float *, int, const double *,
                                        double *, int, const Seg &);

template <typename TX, typename TY>
void axpby(hipStream_t s, int rows, int nc, double alpha, const TX *x, int ldx, double beta, TY *y, int ldy) {
  const int64_t n = (int64_t)rows * nc;
  if (n == 0) return;
  hipLaunchKernelGGL((k_axpby<TX, TY>), dim3(grid_for(n)), dim3(kBlock), 0, s, rows, nc, alpha, x, ldx, beta, y, ldy);
  EMF_HIP(hipGetLastError());
}
template void axpby<float, float>(hipStream_t, int, int, double, const float *, int, double, float *, int);
template void axpby<double, double>(hipStream_t, int, int, double, const double *, int, double, double *, int);
template void axpby<double, float>(hipStream_t, int, int, double, const double *, int, double, float *, int);
template void axpby<float, double>(hipStream_t, int, int, double, const float *, int, double, double *, int);

void fcg_residual(hipStream_t s, int rows, int nc, const double *rz, const double *pq, uint8_t *active,
                  double *alpha, const double *q, double *r, double *rr, const double *bb, double tol2, double *h_rr,
                  uint8_t *h_active, double *work, const Seg &sg, const CycleFill &cf) {
  EMF_REQUIRE(nc <= kBlock, kErrArg, "inner solve supports at most 256 columns");
  EMF_REQUIRE((cf.bt == nullptr) == (cf.x1 == nullptr) && (cf.bt == nullptr || cf.dinv != nullptr), kErrArg,
              "cycle fill needs its copy, first iterate and scaling");
  const int nb = reduce_blocks(rows, nc, sg);
  hipLaunchKernelGGL(k_fcg_res, dim3(nb), dim3(kBlock), 0, s, rows, nc, rz, pq, active, alpha, q, r, work, sg, cf);
  hipLaunchKernelGGL(k_finish_conv, dim3(sg.nseg * nc), dim3(kBlock), 0, s, nb, nc, work, rr, sg, bb, tol2, active,
                     h_rr, h_active);
  EMF_HIP(hipGetLastError());
}

template <typename TZ>
void fcg_px(hipStream_t s, int rows, int nc, const double *alpha, const double *zq, const double *pq,
            const uint8_t *active, const TZ *z, double *p, double *x, const Seg &sg) {
  const int64_t n = (int64_t)rows * nc;
  if (n == 0) return;
  hipLaunchKernelGGL((k_fcg_px<TZ>), dim3(grid_for(n)), dim3(kBlock), 0, s, rows, nc, alpha, zq, pq, active, z, p, x,
                     sg.rseg);
  EMF_HIP(hipGetLastError());
}
template void fcg_px<float>(hipStream_t, int, int, const double *, const double *, const double *, const uint8_t *,
                            const float *, double *, double *, const Seg &);
template void fcg_px<double>(hipStream_t, int, int, const double *, const double *, const double *, const uint8_t *,
                             const double *, double *, double *, const Seg &);

template <typename T>
void cheb_first(hipStream_t s, int rows, int nc, const T *dinv, const double *b, T *bt, T *x, double inv_theta) {
  const int64_t n = (int64_t)rows * nc;
  hipLaunchKernelGGL((k_cheb_first<T>), dim3(grid_for(n)), dim3(kBlock), 0, s, n, nc, dinv, b, bt, x, inv_theta);
  EMF_HIP(hipGetLastError());
}
template <typename T>
void diag_scale(hipStream_t s, int rows, int nc, const T *dinv, const T *b, T *x) {
  const int64_t n = (int64_t)rows * nc;
  hipLaunchKernelGGL((k_diag_scale<T>), dim3(grid_for(n)), dim3(kBlock), 0, s, n, nc, dinv, b, x);
  EMF_HIP(hipGetLastError());
}
template <typename T>
void dense_solve(hipStream_t s, int n, int nc, const double *Ainv, const int *range, const T *b, T *x) {
  EMF_REQUIRE(n <= kDenseMax, kErrArg, "dense coarse solve larger than kDenseMax");
  EMF_REQUIRE(nc <= 65535, kErrArg, "dense coarse solve: at most 65535 columns");
  if (n <= 0 || nc <= 0) return;
  const dim3 rows((n + 63) / 64);
#define EMF_DS(CG)                                                                                              \
  hipLaunchKernelGGL((k_dense_solve<T, CG>), dim3(rows.x, (nc + CG - 1) / CG), dim3(kBlock), 0, s, n, nc, Ainv, \
                     range, b, x)
  if (nc == 1)
    EMF_DS(1);
  else if (nc == 2)
    EMF_DS(2);
  else if (nc <= 4)
    EMF_DS(4);
  else
    EMF_DS(8);
#undef EMF_DS
  EMF_HIP(hipGetLastError());
}

#define EMF_INST(T)                                                                                        \
  template void cheb_first<T>(hipStream_t, int, int, const T *, const double *, T *, T *, double);         \
  template void diag_scale<T>(hipStream_t, int, int, const T *, const T *, T *);                           \
  template void dense_solve<T>(hipStream_t, int, int, const double *, const int *, const T *, T *);
EMF_INST(float)
EMF_INST(double)
#undef EMF_INST

void fcg_beta(hipStream_t s, int nc, const double *zq, const double *pq, const uint8_t *active, double *beta) {
  hipLaunchKernelGGL(k_fcg_beta, dim3(grid_for(nc)), dim3(kBlock), 0, s, nc, zq, pq, active, beta);
  EMF_HIP(hipGetLastError());
}

}  // namespace emfem
