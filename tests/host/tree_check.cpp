// Host check of the octree builder (gpurify_amd/csrc/gp_tree.cpp), built
// with AddressSanitizer and UndefinedBehaviorSanitizer by
// `make -C gpurify_amd/csrc sanitize` and run by tests/test_host_sanitizers.py.
//
// Seeded panel sets (uniform clouds, flat IC-like slabs, tight clusters with
// outliers, coincident centroids, a single panel) go through panel_geometry,
// choose_depth and build_tree; every structure is compared with a brute-force
// construction from the definitions of fmm/tree.rs:41-226:
//   * each centroid lies in its leaf's cell (cell = floor((x - origin) /
//     leaf size), clamped), leaves partition the panels, and panels keep
//     their input order within a leaf (stable sort);
//   * parent cell = child cell / 2, octant = the child's cell parity bits,
//     children lists are exactly the boxes whose parent is the box;
//   * near list of a leaf = the leaf first, then every existing leaf with
//     |dcell|_inf == 1;
//   * interaction list of a box of level >= 2 = every existing box of its
//     level with |dcell|_inf >= 2 whose parent is within |dcell|_inf <= 1 of
//     its parent, with the offset index (o0+3)*49 + (o1+3)*7 + (o2+3);
//   * choose_depth returns the smallest depth >= 2 whose panel-weighted mean
//     leaf population is at most the target.
// Exit status 0 when every check holds; the first failure is printed.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <set>
#include <string>
#include <tuple>
#include <vector>

#include "../../gpurify_amd/csrc/gp_common.h"
#include "../../gpurify_amd/csrc/gp_tree.h"

namespace {

int failures = 0;

void fail(const std::string &what) {
  if (failures++ < 10) std::fprintf(stderr, "FAIL: %s\n", what.c_str());
}

std::vector<double> square(double x, double y, double z, double h, int orient) {
  // a small square panel around (x, y, z) in one of the three planes
  std::vector<double> q(12);
  const double d[4][2] = {{-h, -h}, {h, -h}, {h, h}, {-h, h}};
  for (int k = 0; k < 4; ++k) {
    double p[3] = {x, y, z};
    const int a = orient == 0 ? 0 : orient == 1 ? 1 : 0, b = orient == 0 ? 1 : 2;
    p[a] += d[k][0];
    p[b] += d[k][1];
    for (int c = 0; c < 3; ++c) q[3 * k + c] = p[c];
  }
  return q;
}

struct Case {
  std::string name;
  std::vector<double> verts;
  std::vector<uint32_t> cond;
};

Case make_case(int kind, int n, unsigned seed) {
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  Case c;
  c.name = "kind " + std::to_string(kind) + " n " + std::to_string(n) + " seed " + std::to_string(seed);
  for (int i = 0; i < n; ++i) {
    double x, y, z;
    if (kind == 0) {  // uniform cloud
      x = U(rng), y = U(rng), z = U(rng);
    } else if (kind == 1) {  // flat slab (IC layout): wide in x, y, thin in z
      x = 200e-6 * U(rng), y = 200e-6 * U(rng), z = 1e-6 * (0.6 + 0.8 * U(rng));
    } else if (kind == 2) {  // clusters plus a few far outliers
      const int cl = (int)(U(rng) * 4);
      x = 0.1 * cl + 1e-3 * U(rng), y = 1e-3 * U(rng), z = 1e-3 * U(rng);
      if (i % 97 == 0) x += 10.0;
    } else {  // coincident centroids in small groups
      const int g = i / 5;
      x = 0.01 * (g % 7), y = 0.01 * ((g / 7) % 7), z = 0.01 * (g / 49);
    }
    const auto q = square(x, y, z, kind == 1 ? 5e-8 : 1e-4, i % 3);
    c.verts.insert(c.verts.end(), q.begin(), q.end());
    c.cond.push_back((uint32_t)(i % 3));
  }
  return c;
}

void check_tree(const Case &cs, const gp::PanelGeom &g, int L) {
  gp::Tree t;
  gp::build_tree(g, L, t);
  const std::string tag = cs.name + " L " + std::to_string(L) + ": ";
  const int n = g.n;
  // perm is a permutation
  {
    std::vector<int> seen(n, 0);
    for (uint32_t p : t.perm)
      if (p >= (uint32_t)n || seen[p]++) return fail(tag + "perm is not a permutation");
  }
  // leaf cells of the centroids (definition)
  const uint32_t ncell = 1u << L;
  const double ls = 2.0 * t.half / ncell;
  const double org[3] = {t.center[0] - t.half, t.center[1] - t.half, t.center[2] - t.half};
  auto cell_of = [&](int i, int64_t c[3]) {
    const double a[3] = {g.cx[i], g.cy[i], g.cz[i]};
    for (int d = 0; d < 3; ++d) {
      int64_t v = (int64_t)std::floor((a[d] - org[d]) / ls);
      c[d] = std::min<int64_t>(std::max<int64_t>(v, 0), (int64_t)ncell - 1);
    }
  };
  const int nl = t.n_leaves();
  if ((int)t.leaf_lo.size() != nl + 1 || t.leaf_lo[0] != 0 || t.leaf_lo[nl] != n)
    return fail(tag + "leaf ranges do not cover the panels");
  for (int li = 0; li < nl; ++li) {
    if (t.leaf_lo[li + 1] <= t.leaf_lo[li]) return fail(tag + "empty leaf");
    const int b = t.leaf_box(li);
    for (int s = t.leaf_lo[li]; s < t.leaf_lo[li + 1]; ++s) {
      int64_t c[3];
      cell_of((int)t.perm[s], c);
      for (int d = 0; d < 3; ++d)
        if (c[d] != t.cell[3 * b + d]) return fail(tag + "panel outside its leaf cell");
      if (s > t.leaf_lo[li] && t.perm[s] < t.perm[s - 1]) return fail(tag + "leaf order not stable");
    }
  }
  // levels: cells unique and sorted per level, parents and octants
  const int nb = t.n_boxes();
  std::vector<int> lvl(nb);
  std::vector<std::map<std::tuple<int, int, int>, int>> at(L + 1);
  for (int l = 0; l <= L; ++l)
    for (int b = t.lvl_off[l]; b < t.lvl_off[l + 1]; ++b) {
      lvl[b] = l;
      const auto key = std::make_tuple(t.cell[3 * b], t.cell[3 * b + 1], t.cell[3 * b + 2]);
      if (at[l].count(key)) return fail(tag + "duplicate box");
      at[l][key] = b;
    }
  if (t.lvl_off[1] != 1) return fail(tag + "not one root");
  for (int b = 0; b < nb; ++b) {
    const int l = lvl[b];
    if (l == 0) {
      if (t.parent[b] != -1) return fail(tag + "root has a parent");
      continue;
    }
    const int p = t.parent[b];
    if (p < 0 || lvl[p] != l - 1) return fail(tag + "parent on the wrong level");
    int oct = 0;
    for (int d = 0; d < 3; ++d) {
      if (t.cell[3 * p + d] != t.cell[3 * b + d] / 2) return fail(tag + "parent cell != child cell / 2");
      oct |= (t.cell[3 * b + d] & 1) << d;
    }
    if (t.octant[b] != oct) return fail(tag + "octant");
  }
  for (int b = 0; b < nb; ++b) {
    std::set<int> kids(t.child_idx.begin() + t.child_off[b], t.child_idx.begin() + t.child_off[b + 1]);
    std::set<int> want;
    for (int c = 0; c < nb; ++c)
      if (t.parent[c] == b) want.insert(c);
    if (kids != want) return fail(tag + "children list");
  }
  // every non-leaf box has children; leaves are exactly level L
  for (int b = 0; b < t.lvl_off[L]; ++b)
    if (t.child_off[b + 1] == t.child_off[b]) return fail(tag + "inner box without children");
  // near lists (brute force over the leaf level)
  for (int li = 0; li < nl; ++li) {
    const int b = t.leaf_box(li);
    std::vector<int> got(t.near_leaf.begin() + t.near_off[li], t.near_leaf.begin() + t.near_off[li + 1]);
    if (got.empty() || got[0] != li) return fail(tag + "near list does not start with the leaf");
    std::set<int> gs(got.begin() + 1, got.end()), want;
    if (gs.size() + 1 != got.size()) return fail(tag + "duplicate near leaf");
    for (int lj = 0; lj < nl; ++lj) {
      const int c = t.leaf_box(lj);
      int m = 0;
      for (int d = 0; d < 3; ++d) m = std::max(m, std::abs(t.cell[3 * c + d] - t.cell[3 * b + d]));
      if (m == 1) want.insert(lj);
    }
    if (gs != want) return fail(tag + "near list");
  }
  // interaction lists (brute force per level)
  for (int b = 0; b < nb; ++b) {
    const int l = lvl[b];
    std::set<std::pair<int, int>> got, want;
    for (int q = t.il_off[b]; q < t.il_off[b + 1]; ++q)
      if (!got.insert({t.il_src[q], t.il_oidx[q]}).second) return fail(tag + "duplicate interaction");
    if (l >= 2) {
      const int p = t.parent[b];
      for (int c = t.lvl_off[l]; c < t.lvl_off[l + 1]; ++c) {
        int m = 0, mp = 0, o[3];
        for (int d = 0; d < 3; ++d) {
          o[d] = t.cell[3 * c + d] - t.cell[3 * b + d];
          m = std::max(m, std::abs(o[d]));
          mp = std::max(mp, std::abs(t.cell[3 * t.parent[c] + d] - t.cell[3 * p + d]));
        }
        if (m >= 2 && mp <= 1) want.insert({c, (o[0] + 3) * 49 + (o[1] + 3) * 7 + (o[2] + 3)});
      }
    }
    if (got != want) return fail(tag + "interaction list of box " + std::to_string(b));
  }
}

void check_depth(const Case &cs, const gp::PanelGeom &g, int target) {
  const int L = gp::choose_depth(g, target);
  auto mean_pop = [&](int l) {
    gp::Tree t;
    gp::build_tree(g, l, t);
    double sq = 0.0;
    for (int li = 0; li < t.n_leaves(); ++li) {
      const double m = t.leaf_lo[li + 1] - t.leaf_lo[li];
      sq += m * m;
    }
    return sq / g.n;
  };
  const std::string tag = cs.name + " target " + std::to_string(target) + ": ";
  if (L < 2 || L > 20) return fail(tag + "depth out of range");
  if (L < 20 && mean_pop(L) > target) return fail(tag + "depth above the target population");
  if (L > 2 && mean_pop(L - 1) <= target) return fail(tag + "a shallower depth meets the target");
}

}  // namespace

int main() {
  int cases = 0;
  const int sizes[] = {1, 2, 7, 64, 513, 3000};
  for (int kind = 0; kind < 4; ++kind)
    for (int n : sizes)
      for (unsigned seed : {1u, 2u}) {
        const Case cs = make_case(kind, n, 1000u * kind + 17u * n + seed);
        gp_panels in{(uint32_t)n, 3u, cs.verts.data(), cs.cond.data()};
        gp::PanelGeom g;
        gp::panel_geometry(in, g, seed == 1);
        for (int L : {1, 2, 3, 5}) check_tree(cs, g, L);
        if (n <= 513) check_depth(cs, g, 8);
        check_depth(cs, g, 64);
        ++cases;
      }
  // a bad panel is reported, not read past
  {
    Case cs = make_case(0, 10, 5);
    cs.cond[7] = 3;
    gp_panels in{10u, 3u, cs.verts.data(), cs.cond.data()};
    gp::PanelGeom g;
    try {
      gp::panel_geometry(in, g);
      fail("an out-of-range conductor id was accepted");
    } catch (const gp::Error &e) {
      if (std::string(e.what()).find("panel 7") == std::string::npos) fail("wrong panel named");
    }
  }
  std::printf("tree_check: %d panel sets, %d failures\n", cases, failures);
  return failures ? 1 : 0;
}
