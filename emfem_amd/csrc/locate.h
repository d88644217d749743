// Device point location (see locate.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace emfem {

// Tets the breadth-first search may queue per point before it gives up (-2).
constexpr int kLocateQueue = 4096;

// d_result[i] = tet containing point i (d_q [nq][3]), -1 if no tet contains it, -2 if
// the search queued kLocateQueue tets without an answer.  d_nearest holds nq ints and
// d_queue nq * kLocateQueue ints of scratch.
void locate_points(hipStream_t s, int nq, const double *d_q, int nv, const double *d_pts, const int *d_tets,
                   const int *d_neighbor, const int *d_v2t, int *d_nearest, int *d_queue, int *d_result);

}  // namespace emfem
