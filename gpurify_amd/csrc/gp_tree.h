// Host-side panel geometry and the uniform-depth octree the FMM walks.
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/gpurify.h"

namespace gp {

// Solve-time panel set (integrals/panel.rs:78-119): centroids, areas and
// conductor ids of the caller's panels, in the caller's order.  The vertices
// and conductor ids are the caller's arrays, borrowed for the call.
struct PanelGeom {
  int n = 0;
  int ncond = 0;
  const double *verts = nullptr;   // n * 12, metres
  std::vector<double> cx, cy, cz, area;
  const uint32_t *cond = nullptr;  // n
  // Optional +1 / -1 per panel (caller order): the right-hand side of
  // conductor j is sign on its panels and its charge 1/2 sum sign sigma a
  // (a panel set extended by its mirror images, the mirrors at -1).
  const double *sign = nullptr;
};

// Validates the boundary input and computes centroids/areas with the
// reference formulas (panel.rs:29-67), on the set-up threads when parallel
// is set.  Throws gp::Error on bad input (the first bad panel's message).
void panel_geometry(const gp_panels &in, PanelGeom &g, bool parallel = true);
// The checks of panel_geometry for panel i (conductor id, finite vertices).
void check_panel(const gp_panels &in, int i);
// Area-weighted centroid c and area of one panel's 4 vertices (q: 12
// doubles), panel.rs:29-47 and 58-67; false when the area is zero.
bool panel_props(const double *q, double c[3], double &area);

// Octree over the panel centroids with all leaves at depth L, empty boxes not
// stored (fmm/tree.rs:41-106).  Boxes of each level are ordered by Morton code,
// panels are re-ordered so every leaf owns a contiguous range.
struct Tree {
  int L = 0;
  double center[3] = {0, 0, 0};
  double half = 1.0;             // root half-width, metres (tree.rs:53-62)
  std::vector<uint32_t> perm;    // sorted position -> caller's panel index
  std::vector<int> lvl_off;      // boxes of level l: [lvl_off[l], lvl_off[l+1])
  std::vector<int> parent;       // per box, -1 for the root
  std::vector<int> octant;       // per box, octant within its parent (ox + 2 oy + 4 oz)
  std::vector<int32_t> cell;     // per box, 3 ints
  std::vector<int> child_off, child_idx;  // CSR of children per box
  std::vector<int> leaf_lo;      // per leaf (+1): sorted panel range
  // Near field (bbfmm.rs:189-206): per leaf, the leaf itself then its
  // existing neighbours in (dx, dy, dz) loop order (tree.rs:166-184).
  std::vector<int> near_off, near_leaf;
  // Interaction lists (tree.rs:188-209) per box of level >= 2:
  // (source box, flattened offset index (o0+3)*49 + (o1+3)*7 + (o2+3)).
  std::vector<int> il_off, il_src, il_oidx;
  int n_boxes() const { return lvl_off.empty() ? 0 : lvl_off.back(); }
  int n_leaves() const { return lvl_off[L + 1] - lvl_off[L]; }
  int leaf_box(int leaf) const { return lvl_off[L] + leaf; }
  double level_half(int l) const { return half / (double)(1u << l); }
};

// Root cube of the centroids (tree.rs:43-62): centre and half-width, metres.
void root_cube(const PanelGeom &g, double center[3], double &half);
void root_cube(int n, const double *cx, const double *cy, const double *cz, double center[3],
               double &half);

// Builds the tree at depth max_level over the centroids of g.
void build_tree(const PanelGeom &g, int max_level, Tree &t);

// Smallest depth >= 2 whose panel-weighted mean leaf population is <= leaf_target
// (the reference fixes depth 4; its uniform cube degenerates on flat layouts).
int choose_depth(const PanelGeom &g, int leaf_target);

}  // namespace gp
