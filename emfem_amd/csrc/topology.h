// Host-side mesh topology: the arrays the device kernels consume.
//
// Reference: Mesh / TetAccessor (src/em_mesh.h:24-283) and the topology builders
// build_edge_info, build_neighbor_info, build_boundary_info
// (src/em_mesh.cc:448-502), the sparsity pattern (src/em_fwd.cc:191-246) and point
// location (src/em_mesh.cc:504-566).
#pragma once

#include <cstdint>
#include <vector>

namespace emfem {

struct Topology {
  int nv = 0, nt = 0, ne = 0;
  std::vector<double> pts;         // [nv*3]
  std::vector<int> tets;           // [nt*4]
  std::vector<int> attr;           // [nt]
  std::vector<int> edgelist;       // [ne*2], lower vertex id first
  std::vector<int> tet2edge;       // [nt*6], canonical local edge order (fe.h local_edge)
  std::vector<int> neighbor, non;  // [nt*4], face f opposite vertex f, -1 on the boundary
  std::vector<uint8_t> edge_bdr;   // [ne]
  std::vector<uint8_t> vertex_bdr; // [nv]
  std::vector<int> v2t;            // [nv] last tet listing the vertex (src/em_mesh.cc:510-515)
  std::vector<int> bdr_cells;      // ascending tets with a boundary edge
  // Atomic-free assembly schedule: tets grouped by colour so that no two tets of a
  // colour share an edge; colour c is color_order[color_ptr[c] .. color_ptr[c+1]).
  std::vector<int> color_order, color_ptr;
};

// Derives edges (lexicographic numbering), neighbours, boundary markers, the vertex
// -> tet map, bdr cells and the colouring from pts/tets.
void build_topology(Topology &T);

// Replaces the edge numbering by a caller-supplied (TetGen) one; tet2edge may list a
// tet's edges in any order and is re-ordered into the canonical local order.
void set_user_edges(Topology &T, int ne, const int *edgelist, const int *tet2edge);

// CSR pattern: row = edge, columns = sorted union of the edges of its tets.
void build_pattern(const Topology &T, std::vector<int> &rowptr, std::vector<int> &col);

// Tet containing p by the reference's rule (nearest vertex, then breadth-first
// search over face neighbours from the last tet that lists it); -1 if none.
int find_cell(const Topology &T, const double *p);

// Nearest vertex (lowest id on ties).
int nearest_vertex(const Topology &T, const double *p);

}  // namespace emfem
