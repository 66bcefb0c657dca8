// Panel geometry and octree construction (host).
//
// The root cube and the leaf-cell formula follow fmm/tree.rs:41-95 exactly,
// in metres, so that with the same depth the leaf of every panel equals the
// reference's.  The data structure is built for the device: Morton-ordered
// boxes per level, panels sorted by leaf (stable, so a leaf keeps the
// reference's per-leaf point order), CSR lists.

#include "gp_tree.h"

#include <atomic>

#include "gp_parallel.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <string>

#include "gp_common.h"

namespace gp {

namespace {

struct V3 {
  double x, y, z;
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double norm(V3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

// Spread the low 21 bits of v to every third bit.
inline uint64_t spread3(uint64_t v) {
  v &= 0x1fffff;
  v = (v | v << 32) & 0x1f00000000ffffULL;
  v = (v | v << 16) & 0x1f0000ff0000ffULL;
  v = (v | v << 8) & 0x100f00f00f00f00fULL;
  v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
  v = (v | v << 2) & 0x1249249249249249ULL;
  return v;
}
inline uint64_t morton(uint32_t x, uint32_t y, uint32_t z) {
  return spread3(x) | (spread3(y) << 1) | (spread3(z) << 2);
}
inline uint32_t compact3(uint64_t v) {
  v &= 0x1249249249249249ULL;
  v = (v ^ (v >> 2)) & 0x10c30c30c30c30c3ULL;
  v = (v ^ (v >> 4)) & 0x100f00f00f00f00fULL;
  v = (v ^ (v >> 8)) & 0x1f0000ff0000ffULL;
  v = (v ^ (v >> 16)) & 0x1f00000000ffffULL;
  v = (v ^ (v >> 32)) & 0x1fffff;
  return (uint32_t)v;
}

}  // namespace

void root_cube(int n, const double *cx, const double *cy, const double *cz, double center[3],
               double &half) {
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = 0; i < n; ++i) {
    double a[3] = {cx[i], cy[i], cz[i]};
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::min(lo[d], a[d]);
      hi[d] = std::max(hi[d], a[d]);
    }
  }
  half = 0.0;
  for (int d = 0; d < 3; ++d) {
    center[d] = 0.5 * (lo[d] + hi[d]);
    half = std::max(half, 0.5 * (hi[d] - lo[d]));
  }
  half *= 1.0 + 1e-6;
  if (half == 0.0) half = 1.0;
}

void root_cube(const PanelGeom &g, double center[3], double &half) {
  root_cube(g.n, g.cx.data(), g.cy.data(), g.cz.data(), center, half);
}

namespace {

// Leaf cell of a point at depth L (tree.rs:81-92).
inline void leaf_cell(const double a[3], const double origin[3], double leaf_size, uint32_t ncell,
                      uint32_t cell[3]) {
  for (int d = 0; d < 3; ++d) {
    int64_t idx = (int64_t)std::floor((a[d] - origin[d]) / leaf_size);
    idx = std::min<int64_t>(std::max<int64_t>(idx, 0), (int64_t)ncell - 1);
    cell[d] = (uint32_t)idx;
  }
}

}  // namespace

bool panel_props(const double *q, double c[3], double &area) {
  V3 v[4];
  for (int k = 0; k < 4; ++k) v[k] = {q[3 * k], q[3 * k + 1], q[3 * k + 2]};
  // panel.rs:29-47 (fan triangulation, area-weighted centroid) and 58-67.
  double asum = 0.0, cs[3] = {0, 0, 0};
  for (int k = 1; k < 3; ++k) {
    double a = norm(cross(sub(v[k], v[0]), sub(v[k + 1], v[0]))) * 0.5;
    double tc[3] = {(v[0].x + v[k].x + v[k + 1].x) * (1.0 / 3.0),
                    (v[0].y + v[k].y + v[k + 1].y) * (1.0 / 3.0),
                    (v[0].z + v[k].z + v[k + 1].z) * (1.0 / 3.0)};
    cs[0] = cs[0] + tc[0] * a;
    cs[1] = cs[1] + tc[1] * a;
    cs[2] = cs[2] + tc[2] * a;
    asum += a;
  }
  if (asum == 0.0) return false;
  for (int d = 0; d < 3; ++d) c[d] = cs[d] * (1.0 / asum);
  area = asum;
  return true;
}

void check_panel(const gp_panels &in, int i) {
  if (in.conductor[i] >= in.n_conductors)
    throw Error(GP_ERR_INVALID_INPUT, "panel " + std::to_string(i) + " conductor id " +
                                          std::to_string(in.conductor[i]) + " out of range");
  const double *q = in.vertices + (size_t)i * 12;
  for (int k = 0; k < 12; ++k)
    if (!std::isfinite(q[k]))
      throw Error(GP_ERR_INVALID_INPUT, "panel " + std::to_string(i) + " has a non-finite vertex");
}

void panel_geometry(const gp_panels &in, PanelGeom &g, bool parallel) {
  if (in.n_panels == 0) throw Error(GP_ERR_EMPTY, "no panels in geometry");
  if (!in.vertices || !in.conductor) throw Error(GP_ERR_INVALID_INPUT, "null panel arrays");
  if (in.n_conductors == 0) throw Error(GP_ERR_INVALID_INPUT, "n_conductors must be > 0");
  g.n = (int)in.n_panels;
  g.ncond = (int)in.n_conductors;
  g.verts = in.vertices;
  g.cond = in.conductor;
  g.cx.resize(g.n);
  g.cy.resize(g.n);
  g.cz.resize(g.n);
  g.area.resize(g.n);
  // chunks of panels on the set-up threads; a bad panel only raises a flag,
  // and a sequential pass then reports the first one
  constexpr int CH = 4096;
  std::atomic<bool> bad(false);
  auto chunk = [&](int q) {
    const int i1 = std::min(g.n, (q + 1) * CH);
    for (int i = q * CH; i < i1; ++i) {
      const double *v = in.vertices + (size_t)i * 12;
      bool ok = in.conductor[i] < in.n_conductors;
      for (int k = 0; k < 12; ++k) ok &= std::isfinite(v[k]);
      double c[3] = {0, 0, 0};
      ok = ok && panel_props(v, c, g.area[i]);
      if (!ok) {
        bad = true;
        return;
      }
      g.cx[i] = c[0];
      g.cy[i] = c[1];
      g.cz[i] = c[2];
    }
  };
  const int nch = (g.n + CH - 1) / CH;
  if (parallel)
    parallel_for(nch, chunk, 1);
  else
    for (int q = 0; q < nch; ++q) chunk(q);
  if (bad)
    for (int i = 0; i < g.n; ++i) {
      check_panel(in, i);
      double c[3], a;
      if (!panel_props(in.vertices + (size_t)i * 12, c, a))
        throw Error(GP_ERR_INVALID_INPUT, "panel " + std::to_string(i) + " has zero area");
    }
}

int choose_depth(const PanelGeom &g, int leaf_target) {
  if (leaf_target <= 0) leaf_target = 64;
  double center[3], half;
  root_cube(g, center, half);
  const int LMAX = 20;
  uint32_t ncell = 1u << LMAX;
  double leaf_size = 2.0 * half / ncell;
  double origin[3] = {center[0] - half, center[1] - half, center[2] - half};
  std::vector<uint64_t> keys(g.n);
  for (int i = 0; i < g.n; ++i) {
    double a[3] = {g.cx[i], g.cy[i], g.cz[i]};
    uint32_t c[3];
    leaf_cell(a, origin, leaf_size, ncell, c);
    keys[i] = morton(c[0], c[1], c[2]);
  }
  std::sort(keys.begin(), keys.end());
  // Panel-weighted mean leaf population sum(n_leaf^2) / N: the population
  // seen by a typical panel, which sets the near-field row length.  (The
  // plain mean is dominated by the many sparse leaves of the substrate plate
  // on IC layouts.)
  for (int L = 2; L <= LMAX; ++L) {
    int shift = 3 * (LMAX - L);
    double sq = 0.0;
    size_t i = 0;
    while (i < keys.size()) {
      size_t j = i;
      while (j < keys.size() && (keys[j] >> shift) == (keys[i] >> shift)) ++j;
      sq += (double)(j - i) * (double)(j - i);
      i = j;
    }
    if (sq / (double)g.n <= (double)leaf_target) return L;
  }
  return LMAX;
}

void build_tree(const PanelGeom &g, int L, Tree &t) {
  if (L < 1 || L > 20) throw Error(GP_ERR_INVALID_INPUT, "octree depth must be in [1, 20]");
  t.L = L;
  root_cube(g, t.center, t.half);
  const uint32_t ncell = 1u << L;
  const double leaf_size = 2.0 * t.half / ncell;
  const double origin[3] = {t.center[0] - t.half, t.center[1] - t.half, t.center[2] - t.half};

  // Leaf key per panel, stable sort by key.
  std::vector<uint64_t> key(g.n);
  for (int i = 0; i < g.n; ++i) {
    double a[3] = {g.cx[i], g.cy[i], g.cz[i]};
    uint32_t c[3];
    leaf_cell(a, origin, leaf_size, ncell, c);
    key[i] = morton(c[0], c[1], c[2]);
  }
  t.perm.resize(g.n);
  std::iota(t.perm.begin(), t.perm.end(), 0u);
  std::stable_sort(t.perm.begin(), t.perm.end(),
                   [&](uint32_t a, uint32_t b) { return key[a] < key[b]; });

  // Unique keys per level, finest first.
  std::vector<std::vector<uint64_t>> lk(L + 1);
  t.leaf_lo.clear();
  for (int s = 0; s < g.n; ++s) {
    uint64_t k = key[t.perm[s]];
    if (lk[L].empty() || lk[L].back() != k) {
      lk[L].push_back(k);
      t.leaf_lo.push_back(s);
    }
  }
  t.leaf_lo.push_back(g.n);
  for (int l = L - 1; l >= 0; --l) {
    for (uint64_t k : lk[l + 1]) {
      uint64_t p = k >> 3;
      if (lk[l].empty() || lk[l].back() != p) lk[l].push_back(p);
    }
  }
  t.lvl_off.assign(L + 2, 0);
  for (int l = 0; l <= L; ++l) t.lvl_off[l + 1] = t.lvl_off[l] + (int)lk[l].size();
  int nb = t.lvl_off[L + 1];
  t.parent.assign(nb, -1);
  t.octant.assign(nb, 0);
  t.cell.assign((size_t)nb * 3, 0);
  for (int l = 0; l <= L; ++l)
    for (size_t i = 0; i < lk[l].size(); ++i) {
      int b = t.lvl_off[l] + (int)i;
      uint64_t k = lk[l][i];
      t.cell[3 * b] = (int32_t)compact3(k);
      t.cell[3 * b + 1] = (int32_t)compact3(k >> 1);
      t.cell[3 * b + 2] = (int32_t)compact3(k >> 2);
      if (l > 0) {
        uint64_t pk = k >> 3;
        auto it = std::lower_bound(lk[l - 1].begin(), lk[l - 1].end(), pk);
        t.parent[b] = t.lvl_off[l - 1] + (int)(it - lk[l - 1].begin());
        t.octant[b] = (int)(k & 7);  // bit0 = x, bit1 = y, bit2 = z
      }
    }
  // children CSR (children of a box are contiguous in the next level)
  t.child_off.assign(nb + 1, 0);
  for (int b = 0; b < nb; ++b)
    if (t.parent[b] >= 0) t.child_off[t.parent[b] + 1]++;
  for (int b = 0; b < nb; ++b) t.child_off[b + 1] += t.child_off[b];
  t.child_idx.assign(t.child_off[nb], 0);
  {
    std::vector<int> fill(t.child_off.begin(), t.child_off.end() - 1);
    for (int b = 0; b < nb; ++b)
      if (t.parent[b] >= 0) t.child_idx[fill[t.parent[b]]++] = b;
  }

  auto find_box = [&](int l, int64_t cx, int64_t cy, int64_t cz) -> int {
    int64_t nc = (int64_t)1 << l;
    if (cx < 0 || cy < 0 || cz < 0 || cx >= nc || cy >= nc || cz >= nc) return -1;
    uint64_t k = morton((uint32_t)cx, (uint32_t)cy, (uint32_t)cz);
    auto it = std::lower_bound(lk[l].begin(), lk[l].end(), k);
    if (it == lk[l].end() || *it != k) return -1;
    return t.lvl_off[l] + (int)(it - lk[l].begin());
  };

  // Near lists at the leaf level.
  int nl = t.n_leaves();
  t.near_off.assign(nl + 1, 0);
  t.near_leaf.clear();
  for (int li = 0; li < nl; ++li) {
    int b = t.lvl_off[L] + li;
    t.near_leaf.push_back(li);
    for (int dx = -1; dx <= 1; ++dx)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dz = -1; dz <= 1; ++dz) {
          if (!dx && !dy && !dz) continue;
          int nbx = find_box(L, (int64_t)t.cell[3 * b] + dx, (int64_t)t.cell[3 * b + 1] + dy,
                             (int64_t)t.cell[3 * b + 2] + dz);
          if (nbx >= 0) t.near_leaf.push_back(nbx - t.lvl_off[L]);
        }
    t.near_off[li + 1] = (int)t.near_leaf.size();
  }

  // Interaction lists: boxes c of the same level with |c - b|_inf >= 2 whose
  // parent is adjacent to (or equal to) b's parent (tree.rs:188-209).
  t.il_off.assign(nb + 1, 0);
  t.il_src.clear();
  t.il_oidx.clear();
  for (int b = 0; b < nb; ++b) {
    int l = 0;
    while (b >= t.lvl_off[l + 1]) ++l;
    if (l >= 2) {
      int p = t.parent[b];
      for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dz = -1; dz <= 1; ++dz) {
            int pn = find_box(l - 1, (int64_t)t.cell[3 * p] + dx, (int64_t)t.cell[3 * p + 1] + dy,
                              (int64_t)t.cell[3 * p + 2] + dz);
            if (pn < 0) continue;
            for (int q = t.child_off[pn]; q < t.child_off[pn + 1]; ++q) {
              int c = t.child_idx[q];
              int o0 = t.cell[3 * c] - t.cell[3 * b], o1 = t.cell[3 * c + 1] - t.cell[3 * b + 1],
                  o2 = t.cell[3 * c + 2] - t.cell[3 * b + 2];
              if (std::max(std::abs(o0), std::max(std::abs(o1), std::abs(o2))) <= 1) continue;
              t.il_src.push_back(c);
              t.il_oidx.push_back((o0 + 3) * 49 + (o1 + 3) * 7 + (o2 + 3));
            }
          }
    }
    t.il_off[b + 1] = (int)t.il_src.size();
  }
}

}  // namespace gp
