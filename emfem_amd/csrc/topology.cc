// Host-side mesh topology.  See topology.h for the reference mapping.
#include "topology.h"

#include <algorithm>
#include <cmath>
#include <deque>
#include <numeric>
#include <string>
#include <thread>

#include "common.h"
#include "fe.h"

namespace emfem {

namespace {

// Parallel sort: sort chunks on threads, then merge pairwise.
template <typename T, typename Cmp>
void psort(std::vector<T> &v, Cmp cmp) {
  int nth = host_threads();
  if (v.size() < (1u << 16) || nth == 1) {
    std::sort(v.begin(), v.end(), cmp);
    return;
  }
  int parts = 1;
  while (parts * 2 <= nth) parts *= 2;
  std::vector<size_t> bnd(parts + 1);
  for (int i = 0; i <= parts; ++i) bnd[i] = v.size() * i / parts;
  {
    std::vector<std::thread> th;
    for (int i = 0; i < parts; ++i)
      th.emplace_back([&, i]() { std::sort(v.begin() + bnd[i], v.begin() + bnd[i + 1], cmp); });
    for (auto &t : th) t.join();
  }
  for (int w = 1; w < parts; w *= 2) {
    std::vector<std::thread> th;
    for (int i = 0; i + w < parts; i += 2 * w) {
      size_t b = bnd[i], m = bnd[i + w], e = bnd[std::min(i + 2 * w, parts)];
      th.emplace_back([&, b, m, e]() { std::inplace_merge(v.begin() + b, v.begin() + m, v.begin() + e, cmp); });
    }
    for (auto &t : th) t.join();
  }
}

inline uint64_t edge_key(int a, int b) {
  if (a > b) std::swap(a, b);
  return (uint64_t(uint32_t(a)) << 32) | uint32_t(b);
}

struct FaceKey {
  uint64_t hi;
  uint32_t lo;
  int id;  // t*4 + f
};

void build_neighbors(Topology &T) {
  const int nt = T.nt;
  std::vector<FaceKey> faces((size_t)nt * 4);
  parallel_ranges(nt, [&](int64_t b, int64_t e, int) {
    for (int64_t t = b; t < e; ++t)
      for (int f = 0; f < 4; ++f) {
        int v[3];
        for (int i = 0; i < 3; ++i) v[i] = T.tets[t * 4 + (f + i + 1) % 4];
        std::sort(v, v + 3);
        faces[t * 4 + f] = FaceKey{(uint64_t(uint32_t(v[0])) << 32) | uint32_t(v[1]), uint32_t(v[2]),
                                   int(t * 4 + f)};
      }
  });
  psort(faces, [](const FaceKey &x, const FaceKey &y) {
    if (x.hi != y.hi) return x.hi < y.hi;
    if (x.lo != y.lo) return x.lo < y.lo;
    return x.id < y.id;
  });
  T.neighbor.assign((size_t)nt * 4, -1);
  T.non.assign((size_t)nt * 4, -1);
  for (size_t i = 0; i + 1 < faces.size(); ++i) {
    const FaceKey &p = faces[i], &q = faces[i + 1];
    if (p.hi == q.hi && p.lo == q.lo) {
      T.neighbor[p.id] = q.id / 4;
      T.neighbor[q.id] = p.id / 4;
      T.non[p.id] = q.id % 4;
      T.non[q.id] = p.id % 4;
      ++i;
    }
  }
}

void build_boundary(Topology &T) {
  T.edge_bdr.assign(T.ne, 0);
  T.vertex_bdr.assign(T.nv, 0);
  for (int t = 0; t < T.nt; ++t)
    for (int f = 0; f < 4; ++f) {
      if (T.neighbor[t * 4 + f] >= 0) continue;
      for (int i = 0; i < 3; ++i) T.vertex_bdr[T.tets[t * 4 + (f + i + 1) % 4]] = 1;
      for (int i = 0; i < 6; ++i) {
        int p, q;
        local_edge(i, p, q);
        if (p != f && q != f) T.edge_bdr[T.tet2edge[t * 6 + i]] = 1;
      }
    }
  T.bdr_cells.clear();
  for (int t = 0; t < T.nt; ++t) {
    bool any = false;
    for (int i = 0; i < 6; ++i) any |= T.edge_bdr[T.tet2edge[t * 6 + i]] != 0;
    if (any) T.bdr_cells.push_back(t);
  }
  T.v2t.assign(T.nv, -1);
  for (int t = 0; t < T.nt; ++t)
    for (int v = 0; v < 4; ++v) T.v2t[T.tets[t * 4 + v]] = t;
}

// Greedy edge-conflict colouring in tet order with 128-bit colour masks per edge.
void build_colouring(Topology &T) {
  std::vector<uint64_t> m0(T.ne, 0), m1(T.ne, 0);
  std::vector<int> colour(T.nt);
  int ncol = 0;
  for (int t = 0; t < T.nt; ++t) {
    uint64_t a = 0, b = 0;
    for (int i = 0; i < 6; ++i) {
      a |= m0[T.tet2edge[t * 6 + i]];
      b |= m1[T.tet2edge[t * 6 + i]];
    }
    int c;
    if (~a) c = __builtin_ctzll(~a);
    else if (~b) c = 64 + __builtin_ctzll(~b);
    else throw Error(kErrUser, "edge-conflict colouring needs more than 128 colours");
    colour[t] = c;
    ncol = std::max(ncol, c + 1);
    for (int i = 0; i < 6; ++i) {
      if (c < 64) m0[T.tet2edge[t * 6 + i]] |= uint64_t(1) << c;
      else m1[T.tet2edge[t * 6 + i]] |= uint64_t(1) << (c - 64);
    }
  }
  T.color_ptr.assign(ncol + 1, 0);
  for (int t = 0; t < T.nt; ++t) T.color_ptr[colour[t] + 1]++;
  for (int c = 0; c < ncol; ++c) T.color_ptr[c + 1] += T.color_ptr[c];
  std::vector<int> fill(T.color_ptr.begin(), T.color_ptr.end() - 1);
  T.color_order.resize(T.nt);
  for (int t = 0; t < T.nt; ++t) T.color_order[fill[colour[t]]++] = t;
}

void finish_topology(Topology &T) {
  build_boundary(T);
  build_colouring(T);
}

}  // namespace

void build_topology(Topology &T) {
  const int nt = T.nt;
  for (int t = 0; t < nt; ++t)
    for (int v = 0; v < 4; ++v) {
      int id = T.tets[t * 4 + v];
      EMF_REQUIRE(id >= 0 && id < T.nv, kErrUser, "tet " + std::to_string(t) + " references vertex out of range");
    }
  std::vector<uint64_t> keys((size_t)nt * 6);
  parallel_ranges(nt, [&](int64_t b, int64_t e, int) {
    for (int64_t t = b; t < e; ++t)
      for (int i = 0; i < 6; ++i) {
        int p, q;
        local_edge(i, p, q);
        keys[t * 6 + i] = edge_key(T.tets[t * 4 + p], T.tets[t * 4 + q]);
      }
  });
  std::vector<uint64_t> uniq(keys);
  psort(uniq, std::less<uint64_t>());
  uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
  T.ne = (int)uniq.size();
  T.edgelist.resize((size_t)T.ne * 2);
  for (int e = 0; e < T.ne; ++e) {
    T.edgelist[e * 2] = int(uniq[e] >> 32);
    T.edgelist[e * 2 + 1] = int(uniq[e] & 0xffffffffu);
  }
  T.tet2edge.resize((size_t)nt * 6);
  parallel_ranges((int64_t)nt * 6, [&](int64_t b, int64_t e, int) {
    for (int64_t i = b; i < e; ++i)
      T.tet2edge[i] = int(std::lower_bound(uniq.begin(), uniq.end(), keys[i]) - uniq.begin());
  });
  build_neighbors(T);
  finish_topology(T);
}

void set_user_edges(Topology &T, int ne, const int *edgelist, const int *tet2edge) {
  EMF_REQUIRE(ne > 0, kErrArg, "n_edges must be positive");
  T.ne = ne;
  T.edgelist.resize((size_t)ne * 2);
  for (int e = 0; e < ne; ++e) {
    int a = edgelist[e * 2], b = edgelist[e * 2 + 1];
    EMF_REQUIRE(a >= 0 && a < T.nv && b >= 0 && b < T.nv && a != b, kErrUser,
                "edge " + std::to_string(e) + " has invalid endpoints");
    T.edgelist[e * 2] = std::min(a, b);
    T.edgelist[e * 2 + 1] = std::max(a, b);
  }
  std::vector<int> out((size_t)T.nt * 6);
  for (int t = 0; t < T.nt; ++t) {
    for (int i = 0; i < 6; ++i) {
      int p, q;
      local_edge(i, p, q);
      uint64_t want = edge_key(T.tets[t * 4 + p], T.tets[t * 4 + q]);
      int found = -1;
      for (int j = 0; j < 6; ++j) {
        int g = tet2edge[t * 6 + j];
        EMF_REQUIRE(g >= 0 && g < ne, kErrUser, "tet2edge entry out of range");
        if (edge_key(T.edgelist[g * 2], T.edgelist[g * 2 + 1]) == want) found = g;
      }
      EMF_REQUIRE(found >= 0, kErrUser, "tet " + std::to_string(t) + ": tet2edge does not match its vertices");
      out[t * 6 + i] = found;
    }
  }
  T.tet2edge.swap(out);
  finish_topology(T);
}

void build_pattern(const Topology &T, std::vector<int> &rowptr, std::vector<int> &col) {
  const int ne = T.ne, nt = T.nt;
  // edge -> tets (ascending tet order)
  std::vector<int> eptr(ne + 1, 0), etet((size_t)nt * 6);
  for (int i = 0; i < nt * 6; ++i) eptr[T.tet2edge[i] + 1]++;
  for (int e = 0; e < ne; ++e) eptr[e + 1] += eptr[e];
  {
    std::vector<int> fill(eptr.begin(), eptr.end() - 1);
    for (int t = 0; t < nt; ++t)
      for (int i = 0; i < 6; ++i) etet[fill[T.tet2edge[t * 6 + i]]++] = t;
  }
  std::vector<int> cnt(ne + 1, 0);
  auto row_cols = [&](int e, std::vector<int> &buf) {
    buf.clear();
    for (int k = eptr[e]; k < eptr[e + 1]; ++k)
      for (int j = 0; j < 6; ++j) buf.push_back(T.tet2edge[etet[k] * 6 + j]);
    std::sort(buf.begin(), buf.end());
    buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
  };
  parallel_ranges(ne, [&](int64_t b, int64_t e, int) {
    std::vector<int> buf;
    for (int64_t r = b; r < e; ++r) {
      row_cols((int)r, buf);
      cnt[r + 1] = (int)buf.size();
    }
  });
  for (int e = 0; e < ne; ++e) cnt[e + 1] += cnt[e];
  EMF_REQUIRE(cnt[ne] >= 0, kErrUser, "pattern exceeds int32 nonzeros");
  rowptr.assign(cnt.begin(), cnt.end());
  col.resize(cnt[ne]);
  parallel_ranges(ne, [&](int64_t b, int64_t e, int) {
    std::vector<int> buf;
    for (int64_t r = b; r < e; ++r) {
      row_cols((int)r, buf);
      std::copy(buf.begin(), buf.end(), col.begin() + rowptr[r]);
    }
  });
}

namespace {

// orient3d(a,b,c,d) = det[a-d; b-d; c-d] in double; exact for whole-metre coordinates.
double orient3d(const double *a, const double *b, const double *c, const double *d) {
  double adx = a[0] - d[0], bdx = b[0] - d[0], cdx = c[0] - d[0];
  double ady = a[1] - d[1], bdy = b[1] - d[1], cdy = c[1] - d[1];
  double adz = a[2] - d[2], bdz = b[2] - d[2], cdz = c[2] - d[2];
  return adx * (bdy * cdz - bdz * cdy) + bdx * (cdy * adz - cdz * ady) + cdx * (ady * bdz - adz * bdy);
}

// TetAccessor::contains_point (src/em_mesh.h:208-230).
bool contains(const Topology &T, int t, const double *p) {
  const double *v0 = &T.pts[T.tets[t * 4 + 0] * 3], *v1 = &T.pts[T.tets[t * 4 + 1] * 3];
  const double *v2 = &T.pts[T.tets[t * 4 + 2] * 3], *v3 = &T.pts[T.tets[t * 4 + 3] * 3];
  return orient3d(v0, v1, v3, p) >= 0.0 && orient3d(v1, v2, v3, p) >= 0.0 && orient3d(v1, v0, v2, p) >= 0.0 &&
         orient3d(v0, v3, v2, p) >= 0.0;
}

}  // namespace

int nearest_vertex(const Topology &T, const double *p) {
  int best = 0;
  double bd = 1e300;
  for (int v = 0; v < T.nv; ++v) {
    // squared distance: the kd-tree's L2 metric (src/em_mesh.h:60-61)
    double dx = T.pts[v * 3] - p[0], dy = T.pts[v * 3 + 1] - p[1], dz = T.pts[v * 3 + 2] - p[2];
    double d = dx * dx + dy * dy + dz * dz;
    if (d < bd) {
      bd = d;
      best = v;
    }
  }
  return best;
}

int find_cell(const Topology &T, const double *p) {
  if (T.nt == 0) return -1;
  std::vector<uint8_t> touched(T.nt, 0);
  std::deque<int> q;
  q.push_back(T.v2t[nearest_vertex(T, p)]);
  while (!q.empty()) {
    int t = q.front();
    q.pop_front();
    if (touched[t]) continue;
    touched[t] = 1;
    if (contains(T, t, p)) return t;
    for (int i = 0; i < 4; ++i) {
      int n = T.neighbor[t * 4 + i];
      if (n >= 0 && !touched[n]) q.push_back(n);
    }
  }
  return -1;
}

}  // namespace emfem
