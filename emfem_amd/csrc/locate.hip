// Point location on the device: Mesh::find_closest_vertex + find_cell_around_point
// (src/em_mesh.cc:518-566) for a batch of query points.
//
// Nearest vertex: one workgroup per query scans all vertices (squared Euclidean
// distance, the kd-tree's L2 metric, src/em_mesh.h:60-61) and reduces (distance,
// index) pairs in a fixed tree, ties to the lowest index.  Search: one lane per
// query runs the reference's breadth-first search from the last tet listing that
// vertex, over face neighbours in order; the first tet whose four orientation tests
// are >= 0 (TetAccessor::contains_point, src/em_mesh.h:208-230) wins.  The queue
// holds each tet once (a tet is enqueued only if not yet queued), which visits tets
// in the same order as the reference's queue with repeats skipped at pop.
#include <hip/hip_runtime.h>

#include "common.h"
#include "device.h"
#include "locate.h"

namespace emfem {
namespace {

constexpr int kBlock = 256;

__global__ __launch_bounds__(kBlock) void k_nearest_vertex(int nv, const double *__restrict__ pts,
                                                           const double *__restrict__ q, int *nearest) {
  __shared__ double sd[kBlock];
  __shared__ int si[kBlock];
  const double px = q[blockIdx.x * 3], py = q[blockIdx.x * 3 + 1], pz = q[blockIdx.x * 3 + 2];
  double bd = 1e300;
  int bi = 0x7fffffff;
  for (int v = threadIdx.x; v < nv; v += kBlock) {
    const double dx = pts[(size_t)v * 3] - px, dy = pts[(size_t)v * 3 + 1] - py, dz = pts[(size_t)v * 3 + 2] - pz;
    const double d = dx * dx + dy * dy + dz * dz;
    if (d < bd) {  // strided scan in increasing v: keeps the first minimum
      bd = d;
      bi = v;
    }
  }
  sd[threadIdx.x] = bd;
  si[threadIdx.x] = bi;
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    __syncthreads();
    if (threadIdx.x < w) {
      const double od = sd[threadIdx.x + w];
      const int oi = si[threadIdx.x + w];
      if (od < sd[threadIdx.x] || (od == sd[threadIdx.x] && oi < si[threadIdx.x])) {
        sd[threadIdx.x] = od;
        si[threadIdx.x] = oi;
      }
    }
  }
  if (threadIdx.x == 0) nearest[blockIdx.x] = si[0] == 0x7fffffff ? 0 : si[0];
}

// orient3d(a, b, c, d) = det[a - d; b - d; c - d], evaluated without contraction so
// that it rounds exactly as the host test does (and is exact for whole-metre input).
__device__ double orient3d(const double *a, const double *b, const double *c, const double *d) {
#pragma clang fp contract(off)
  const double adx = a[0] - d[0], bdx = b[0] - d[0], cdx = c[0] - d[0];
  const double ady = a[1] - d[1], bdy = b[1] - d[1], cdy = c[1] - d[1];
  const double adz = a[2] - d[2], bdz = b[2] - d[2], cdz = c[2] - d[2];
  return adx * (bdy * cdz - bdz * cdy) + bdx * (cdy * adz - cdz * ady) + cdx * (ady * bdz - adz * bdy);
}

__device__ bool contains(const double *__restrict__ pts, const int *__restrict__ tets, int t, const double *p) {
  const double *v0 = pts + (size_t)tets[t * 4 + 0] * 3, *v1 = pts + (size_t)tets[t * 4 + 1] * 3;
  const double *v2 = pts + (size_t)tets[t * 4 + 2] * 3, *v3 = pts + (size_t)tets[t * 4 + 3] * 3;
  return orient3d(v0, v1, v3, p) >= 0.0 && orient3d(v1, v2, v3, p) >= 0.0 && orient3d(v1, v0, v2, p) >= 0.0 &&
         orient3d(v0, v3, v2, p) >= 0.0;
}

// result[i] = containing tet, -1 when the search exhausted the mesh, -2 when the
// queue reached `cap` tets before either.
__global__ void k_locate_bfs(int nq, const double *__restrict__ q, const int *__restrict__ nearest,
                             const double *__restrict__ pts, const int *__restrict__ tets,
                             const int *__restrict__ neighbor, const int *__restrict__ v2t, int cap, int *queue,
                             int *result) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  const double p[3] = {q[i * 3], q[i * 3 + 1], q[i * 3 + 2]};
  int *Q = queue + (size_t)i * cap;
  int head = 0, tail = 0;
  Q[tail++] = v2t[nearest[i]];
  int found = -1;
  bool overflow = false;
  while (head < tail) {
    const int t = Q[head++];
    if (contains(pts, tets, t, p)) {
      found = t;
      break;
    }
    for (int f = 0; f < 4; ++f) {
      const int n = neighbor[t * 4 + f];
      if (n < 0) continue;
      bool seen = false;
      for (int k = 0; k < tail && !seen; ++k) seen = Q[k] == n;
      if (seen) continue;
      if (tail == cap) {
        overflow = true;
        break;
      }
      Q[tail++] = n;
    }
    if (overflow) break;
  }
  result[i] = found >= 0 ? found : (overflow ? -2 : -1);
}

}  // namespace

void locate_points(hipStream_t s, int nq, const double *d_q, int nv, const double *d_pts, const int *d_tets,
                   const int *d_neighbor, const int *d_v2t, int *d_nearest, int *d_queue, int *d_result) {
  if (nq == 0) return;
  hipLaunchKernelGGL(k_nearest_vertex, dim3(nq), dim3(kBlock), 0, s, nv, d_pts, d_q, d_nearest);
  hipLaunchKernelGGL(k_locate_bfs, dim3((nq + 63) / 64), dim3(64), 0, s, nq, d_q, d_nearest, d_pts, d_tets,
                     d_neighbor, d_v2t, kLocateQueue, d_queue, d_result);
  EMF_HIP(hipGetLastError());
}

}  // namespace emfem
