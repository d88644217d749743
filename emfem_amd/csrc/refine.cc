// Mesh refinement without TetGen.  See refine.h.
#include "refine.h"

#include <algorithm>
#include <array>
#include <unordered_map>

#include "common.h"

namespace emfem {

namespace {

inline uint64_t ekey(int a, int b) {
  if (a > b) std::swap(a, b);
  return (uint64_t(uint32_t(a)) << 32) | uint32_t(b);
}

double det4(const RawMesh &m, const int *v) {
  const double *p0 = &m.pts[(size_t)v[0] * 3], *p1 = &m.pts[(size_t)v[1] * 3];
  const double *p2 = &m.pts[(size_t)v[2] * 3], *p3 = &m.pts[(size_t)v[3] * 3];
  double a[3], b[3], c[3];
  for (int k = 0; k < 3; ++k) {
    a[k] = p1[k] - p0[k];
    b[k] = p2[k] - p0[k];
    c[k] = p3[k] - p0[k];
  }
  return a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) + a[2] * (b[0] * c[1] - b[1] * c[0]);
}

const int kPairs[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};

int add_midpoint(RawMesh &m, std::unordered_map<uint64_t, int> &mid, int a, int b) {
  const uint64_t k = ekey(a, b);
  auto it = mid.find(k);
  if (it != mid.end()) return it->second;
  const int id = m.nv();
  for (int j = 0; j < 3; ++j) m.pts.push_back(0.5 * (m.pts[(size_t)a * 3 + j] + m.pts[(size_t)b * 3 + j]));
  mid.emplace(k, id);
  return id;
}

}  // namespace

void refine_bisect(RawMesh &m, const std::vector<uint8_t> &flag) {
  EMF_REQUIRE((int)flag.size() == m.nt(), kErrArg, "refinement flags do not match the tet count");
  // target edges: the longest edge of each flagged tet (ties: smallest vertex pair)
  std::vector<uint64_t> targets;
  for (int t = 0; t < m.nt(); ++t) {
    if (!flag[t]) continue;
    const int *v = &m.tets[(size_t)t * 4];
    double best = -1.0;
    uint64_t bk = 0;
    for (auto &p : kPairs) {
      const int a = v[p[0]], b = v[p[1]];
      double l = 0.0;
      for (int j = 0; j < 3; ++j) {
        const double d = m.pts[(size_t)a * 3 + j] - m.pts[(size_t)b * 3 + j];
        l += d * d;
      }
      const uint64_t k = ekey(a, b);
      if (l > best || (l == best && k < bk)) {
        best = l;
        bk = k;
      }
    }
    targets.push_back(bk);
  }
  std::sort(targets.begin(), targets.end());
  targets.erase(std::unique(targets.begin(), targets.end()), targets.end());
  if (targets.empty()) return;
  // edge -> tets
  std::unordered_map<uint64_t, std::vector<int>> around;
  around.reserve((size_t)m.nt() * 2);
  auto add_tet = [&](int t) {
    const int *v = &m.tets[(size_t)t * 4];
    for (auto &p : kPairs) around[ekey(v[p[0]], v[p[1]])].push_back(t);
  };
  auto drop_tet = [&](int t) {
    const int *v = &m.tets[(size_t)t * 4];
    for (auto &p : kPairs) {
      auto &lst = around[ekey(v[p[0]], v[p[1]])];
      lst.erase(std::remove(lst.begin(), lst.end(), t), lst.end());
    }
  };
  for (int t = 0; t < m.nt(); ++t) add_tet(t);
  std::unordered_map<uint64_t, int> mid;
  for (uint64_t e : targets) {
    auto it = around.find(e);
    if (it == around.end() || it->second.empty()) continue;
    const int a = int(e >> 32), b = int(e & 0xffffffffu);
    const int mv = add_midpoint(m, mid, a, b);
    std::vector<int> tl = it->second;
    std::sort(tl.begin(), tl.end());
    for (int t : tl) {
      drop_tet(t);
      std::array<int, 4> v = {m.tets[(size_t)t * 4], m.tets[(size_t)t * 4 + 1], m.tets[(size_t)t * 4 + 2],
                              m.tets[(size_t)t * 4 + 3]};
      std::array<int, 4> ca = v, cb = v;
      for (int k = 0; k < 4; ++k) {
        if (v[k] == b) ca[k] = mv;  // child keeping a
        if (v[k] == a) cb[k] = mv;  // child keeping b
      }
      const int nt = m.nt();
      for (int k = 0; k < 4; ++k) m.tets[(size_t)t * 4 + k] = ca[k];
      for (int k = 0; k < 4; ++k) m.tets.push_back(cb[k]);
      m.attr.push_back(m.attr[t]);
      add_tet(t);
      add_tet(nt);
    }
    around.erase(e);
  }
}

void refine_uniform(RawMesh &m) {
  const int nt = m.nt();
  std::unordered_map<uint64_t, int> mid;
  mid.reserve((size_t)nt * 2);
  std::vector<int> tets;
  std::vector<int> attr;
  tets.reserve((size_t)nt * 32);
  attr.reserve((size_t)nt * 8);
  for (int t = 0; t < nt; ++t) {
    const int v0 = m.tets[(size_t)t * 4], v1 = m.tets[(size_t)t * 4 + 1], v2 = m.tets[(size_t)t * 4 + 2],
              v3 = m.tets[(size_t)t * 4 + 3];
    const int m01 = add_midpoint(m, mid, v0, v1), m02 = add_midpoint(m, mid, v0, v2);
    const int m03 = add_midpoint(m, mid, v0, v3), m12 = add_midpoint(m, mid, v1, v2);
    const int m13 = add_midpoint(m, mid, v1, v3), m23 = add_midpoint(m, mid, v2, v3);
    // four corner tets and the octahedron split along its m02-m13 diagonal
    const int child[8][4] = {{v0, m01, m02, m03},   {m01, v1, m12, m13},  {m02, m12, v2, m23},
                             {m03, m13, m23, v3},   {m02, m13, m01, m12}, {m02, m13, m12, m23},
                             {m02, m13, m23, m03},  {m02, m13, m03, m01}};
    for (auto &c : child) {
      int v[4] = {c[0], c[1], c[2], c[3]};
      if (det4(m, v) < 0) std::swap(v[2], v[3]);
      for (int k = 0; k < 4; ++k) tets.push_back(v[k]);
      attr.push_back(m.attr[t]);
    }
  }
  m.tets.swap(tets);
  m.attr.swap(attr);
}

}  // namespace emfem
