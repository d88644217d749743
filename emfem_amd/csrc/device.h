// Device-side helpers shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace emfem {

// MI355X dispatches consecutive workgroups round-robin over its 8 XCDs, each with
// its own L2.  Row-parallel kernels whose rows gather neighbouring rows (SpMM x
// gathers, per-tet vertex and edge data) renumber their workgroups so that every
// XCD works on one contiguous range of rows: the gathered rows then stay within one
// L2 instead of being fetched by all eight.  A bijection of [0, gridDim.x).
__device__ __forceinline__ int xcd_block() {
  constexpr int kXcd = 8;
  const int n = gridDim.x, b = blockIdx.x;
  const int k = b % kXcd, q = n / kXcd, r = n % kXcd;
  return k * q + (k < r ? k : r) + b / kXcd;
}

}  // namespace emfem
