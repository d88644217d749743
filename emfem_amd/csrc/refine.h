// Mesh refinement without TetGen.
//
// Reference: Mesh::refine_tetgen (src/em_mesh.cc:101-139) gives every marked tet a
// volume bound of V/2 and lets TetGen remesh with quality constraints;
// Mesh::refine_uniform (src/em_mesh.cc:141-182) splits every tet into 8 at its edge
// midpoints.  TetGen is not available, so marked tets are refined here by
// conforming longest-edge bisection: the longest edge of each marked tet is split at
// its midpoint, together with every tet sharing that edge (no hanging nodes), which
// halves the marked tet's volume as the reference's bound asks.  Attributes are
// inherited by the children.
#pragma once

#include <cstdint>
#include <vector>

namespace emfem {

struct RawMesh {
  std::vector<double> pts;  // [nv*3]
  std::vector<int> tets;    // [nt*4], positive orientation
  std::vector<int> attr;    // [nt]
  int nv() const { return (int)pts.size() / 3; }
  int nt() const { return (int)attr.size(); }
};

// Bisects the longest edge of every flagged tet (and all tets around it).
void refine_bisect(RawMesh &m, const std::vector<uint8_t> &flag);

// 1 -> 8 split of every tet at its edge midpoints.
void refine_uniform(RawMesh &m);

}  // namespace emfem
