// Layout -> panel bridge: the capacitance half of extract_quasistatic
// (crates/pex/src/bridge.rs:307-443 with the mesher of 445-652).
//
// Produces exactly the reference's panels (same order, same vertex
// arithmetic) at refine = 1.  The union-grid coverage test is a 2-D
// difference array over the compressed grid (O(rects + cells)) instead of the
// reference's per-cell scan over all rectangles (O(cells x rects),
// bridge.rs:476-486); both mark a cell filled iff some rectangle contains it.

#include <algorithm>
#include <cmath>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "gp_bridge.h"
#include "gp_common.h"

namespace gp {

namespace {

constexpr double kMetresPerNm = 1.0e-9;           // bridge.rs:26
constexpr size_t kMaxGridCells = 262144;          // bridge.rs:28
constexpr size_t kMaxPanelsRef = 20000;           // bridge.rs:29

struct Quad {
  double v[12];
};

// bridge.rs:559-563
int subdivisions(double length, double transverse) {
  double v = std::ceil(length / (4.0 * std::max(transverse, 2.2250738585072014e-308)));
  return (int)std::min(std::max(v, 1.0), 32.0);
}

void push_quad(std::vector<Quad> &out, const double a[3], const double b[3], const double c[3],
               const double d[3], bool reverse) {
  Quad q;
  const double *src[4] = {a, b, c, d};
  for (int k = 0; k < 4; ++k) {
    const double *p = reverse ? src[3 - k] : src[k];
    q.v[3 * k] = p[0];
    q.v[3 * k + 1] = p[1];
    q.v[3 * k + 2] = p[2];
  }
  out.push_back(q);
}

// Top/bottom face, bridge.rs:566-596; r splits each reference panel r x r.
void xy_face(std::vector<Quad> &out, double x0, double x1, double y0, double y1, double z,
             bool reverse, int r) {
  const int nx = subdivisions(x1 - x0, y1 - y0) * r, ny = subdivisions(y1 - y0, x1 - x0) * r;
  for (int ix = 0; ix < nx; ++ix)
    for (int iy = 0; iy < ny; ++iy) {
      double xa = x0 + (x1 - x0) * ix / nx, xb = x0 + (x1 - x0) * (ix + 1) / nx;
      double ya = y0 + (y1 - y0) * iy / ny, yb = y0 + (y1 - y0) * (iy + 1) / ny;
      double a[3] = {xa, ya, z}, b[3] = {xb, ya, z}, c[3] = {xb, yb, z}, d[3] = {xa, yb, z};
      push_quad(out, a, b, c, d, reverse);
    }
}

// x = const sidewall, bridge.rs:599-624 (r strips along z when refined).
void yz_face(std::vector<Quad> &out, double x, double y0, double y1, double z0, double z1,
             bool reverse, int r) {
  const int ny = subdivisions(y1 - y0, z1 - z0) * r;
  for (int iy = 0; iy < ny; ++iy) {
    double ya = y0 + (y1 - y0) * iy / ny, yb = y0 + (y1 - y0) * (iy + 1) / ny;
    for (int iz = 0; iz < r; ++iz) {
      double za = r == 1 ? z0 : z0 + (z1 - z0) * iz / r;
      double zb = r == 1 ? z1 : z0 + (z1 - z0) * (iz + 1) / r;
      double a[3] = {x, ya, za}, b[3] = {x, yb, za}, c[3] = {x, yb, zb}, d[3] = {x, ya, zb};
      push_quad(out, a, b, c, d, reverse);
    }
  }
}

// y = const sidewall, bridge.rs:627-652.
void xz_face(std::vector<Quad> &out, double x0, double x1, double y, double z0, double z1,
             bool reverse, int r) {
  const int nx = subdivisions(x1 - x0, z1 - z0) * r;
  for (int ix = 0; ix < nx; ++ix) {
    double xa = x0 + (x1 - x0) * ix / nx, xb = x0 + (x1 - x0) * (ix + 1) / nx;
    for (int iz = 0; iz < r; ++iz) {
      double za = r == 1 ? z0 : z0 + (z1 - z0) * iz / r;
      double zb = r == 1 ? z1 : z0 + (z1 - z0) * (iz + 1) / r;
      double a[3] = {xa, y, za}, b[3] = {xb, y, za}, c[3] = {xb, y, zb}, d[3] = {xa, y, zb};
      push_quad(out, a, b, c, d, reverse);
    }
  }
}

// cap/mesh.rs:12-43
void plate(std::vector<Quad> &out, double cx, double cy, double z0, double lx, double ly, int nx,
           int ny) {
  double x0 = cx - lx / 2.0, y0 = cy - ly / 2.0, dx = lx / nx, dy = ly / ny;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      double xa = x0 + i * dx, xb = xa + dx, ya = y0 + j * dy, yb = ya + dy;
      double a[3] = {xa, ya, z0}, b[3] = {xb, ya, z0}, c[3] = {xb, yb, z0}, d[3] = {xa, yb, z0};
      push_quad(out, a, b, c, d, false);
    }
}

// bridge.rs:445-557 for one (net, layer) group.
void group_panels(const gp_volume_group &g, double z0, double z1, double xy, int r,
                  std::vector<Quad> &out) {
  std::vector<int32_t> xs, ys;
  xs.reserve(2 * g.n_rects);
  ys.reserve(2 * g.n_rects);
  for (uint32_t k = 0; k < g.n_rects; ++k) {
    xs.push_back(g.rects[4 * k]);
    xs.push_back(g.rects[4 * k + 2]);
    ys.push_back(g.rects[4 * k + 1]);
    ys.push_back(g.rects[4 * k + 3]);
  }
  std::sort(xs.begin(), xs.end());
  xs.erase(std::unique(xs.begin(), xs.end()), xs.end());
  std::sort(ys.begin(), ys.end());
  ys.erase(std::unique(ys.begin(), ys.end()), ys.end());
  const size_t nx = xs.empty() ? 0 : xs.size() - 1, ny = ys.empty() ? 0 : ys.size() - 1;
  const size_t cells = nx * ny;
  if (cells > kMaxGridCells)
    throw Error(GP_ERR_GRID_CAPACITY, "conductor union grid needs " + std::to_string(cells) +
                                          " cells; limit is " + std::to_string(kMaxGridCells));
  // Coverage by a 2-D difference array over the compressed grid.
  std::vector<int32_t> diff((nx + 1) * (ny + 1), 0);
  for (uint32_t k = 0; k < g.n_rects; ++k) {
    const int32_t *rc = g.rects + 4 * k;
    size_t ix0 = std::lower_bound(xs.begin(), xs.end(), rc[0]) - xs.begin();
    size_t ix1 = std::lower_bound(xs.begin(), xs.end(), rc[2]) - xs.begin();
    size_t iy0 = std::lower_bound(ys.begin(), ys.end(), rc[1]) - ys.begin();
    size_t iy1 = std::lower_bound(ys.begin(), ys.end(), rc[3]) - ys.begin();
    if (ix1 <= ix0 || iy1 <= iy0) continue;
    diff[ix0 * (ny + 1) + iy0] += 1;
    diff[ix1 * (ny + 1) + iy0] -= 1;
    diff[ix0 * (ny + 1) + iy1] -= 1;
    diff[ix1 * (ny + 1) + iy1] += 1;
  }
  std::vector<char> filled(cells, 0), consumed(cells, 0);
  {
    std::vector<int32_t> col(ny + 1, 0);
    for (size_t ix = 0; ix < nx; ++ix) {
      int32_t run = 0;
      for (size_t iy = 0; iy < ny; ++iy) {
        col[iy] += diff[ix * (ny + 1) + iy];
        run += col[iy];
        filled[ix * ny + iy] = run > 0;
      }
    }
  }
  // Greedy maximal rectangles for the top and bottom faces (bridge.rs:488-524).
  for (size_t ix = 0; ix < nx; ++ix)
    for (size_t iy = 0; iy < ny; ++iy) {
      size_t cell = ix * ny + iy;
      if (!filled[cell] || consumed[cell]) continue;
      size_t x_end = ix + 1;
      while (x_end < nx && filled[x_end * ny + iy] && !consumed[x_end * ny + iy]) ++x_end;
      size_t y_end = iy + 1;
      for (bool grow = true; grow && y_end < ny;) {
        for (size_t x = ix; x < x_end; ++x) {
          size_t c = x * ny + y_end;
          if (!filled[c] || consumed[c]) {
            grow = false;
            break;
          }
        }
        if (grow) ++y_end;
      }
      for (size_t x = ix; x < x_end; ++x)
        for (size_t y = iy; y < y_end; ++y) consumed[x * ny + y] = 1;
      double x0 = (double)xs[ix] * xy, x1 = (double)xs[x_end] * xy;
      double y0 = (double)ys[iy] * xy, y1 = (double)ys[y_end] * xy;
      xy_face(out, x0, x1, y0, y1, z1, false, r);
      xy_face(out, x0, x1, y0, y1, z0, true, r);
    }
  // Sidewalls on exposed atomic cell edges (bridge.rs:526-548).
  for (size_t ix = 0; ix < nx; ++ix)
    for (size_t iy = 0; iy < ny; ++iy) {
      if (!filled[ix * ny + iy]) continue;
      double x0 = (double)xs[ix] * xy, x1 = (double)xs[ix + 1] * xy;
      double y0 = (double)ys[iy] * xy, y1 = (double)ys[iy + 1] * xy;
      if (ix == 0 || !filled[(ix - 1) * ny + iy]) yz_face(out, x0, y0, y1, z0, z1, true, r);
      if (ix + 1 == nx || !filled[(ix + 1) * ny + iy]) yz_face(out, x1, y0, y1, z0, z1, false, r);
      if (iy == 0 || !filled[ix * ny + iy - 1]) xz_face(out, x0, x1, y0, z0, z1, false, r);
      if (iy + 1 == ny || !filled[ix * ny + iy + 1]) xz_face(out, x0, x1, y1, z0, z1, true, r);
    }
}

// validate_stack (bridge.rs:149-169) applied to the group's resolved layer.
void validate_group(const gp_volume_group &g, uint32_t index) {
  if (g.n_rects > 0 && !g.rects)
    throw Error(GP_ERR_INVALID_INPUT, "group " + std::to_string(index) + " has null rects");
  struct {
    const char *field;
    double v;
    bool positive;
  } checks[3] = {{"thickness_nm", g.thickness_nm, true},
                 {"height_nm", g.height_nm, false},
                 {"dielectric_k", g.dielectric_k, true}};
  for (auto &c : checks)
    if (!std::isfinite(c.v) || (c.positive && c.v <= 0.0) || (!c.positive && c.v < 0.0))
      throw Error(GP_ERR_INVALID_INPUT, "PEX layer of group " + std::to_string(index) +
                                            " has invalid " + c.field + " value " +
                                            std::to_string(c.v));
  for (uint32_t k = 0; k < g.n_rects; ++k)
    if (g.rects[4 * k] > g.rects[4 * k + 2] || g.rects[4 * k + 1] > g.rects[4 * k + 3])
      throw Error(GP_ERR_INVALID_INPUT, "group " + std::to_string(index) + " rect " +
                                            std::to_string(k) + " has x1 < x0 or y1 < y0");
}

}  // namespace

void mesh_groups(const gp_volume_group *groups, uint32_t n_groups, double xy_scale, int refine,
                 size_t max_panels, Meshed &m) {
  if (n_groups == 0 || !groups) throw Error(GP_ERR_EMPTY, "no volume groups");
  if (!std::isfinite(xy_scale) || xy_scale <= 0.0)
    throw Error(GP_ERR_INVALID_INPUT, "deck database unit must be finite and positive, got " +
                                          std::to_string(xy_scale / kMetresPerNm) + " nm");
  if (refine < 1) refine = 1;
  if (max_panels == 0) max_panels = kMaxPanelsRef;
  for (uint32_t i = 0; i < n_groups; ++i) validate_group(groups[i], i);
  // Conductor order: ascending net id, substrate last (bridge.rs:316-321, 361).
  std::set<uint32_t> nets;
  for (uint32_t i = 0; i < n_groups; ++i) nets.insert(groups[i].net);
  std::map<uint32_t, uint32_t> index;
  m.nets.assign(nets.begin(), nets.end());
  for (size_t i = 0; i < m.nets.size(); ++i) index[m.nets[i]] = (uint32_t)i;
  std::vector<Quad> quads;
  m.cond.clear();
  m.eps.clear();
  double b[4] = {INFINITY, INFINITY, -INFINITY, -INFINITY};
  double min_z = INFINITY, max_thick = 0.0;
  for (uint32_t i = 0; i < n_groups; ++i) {
    const gp_volume_group &g = groups[i];
    const double z0 = g.height_nm * kMetresPerNm;                     // bridge.rs:135
    const double z1 = (g.height_nm + g.thickness_nm) * kMetresPerNm;  // bridge.rs:136
    size_t before = quads.size();
    group_panels(g, z0, z1, xy_scale, refine, quads);
    m.cond.insert(m.cond.end(), quads.size() - before, index[g.net]);
    m.eps.insert(m.eps.end(), quads.size() - before, g.dielectric_k);
    if (quads.size() > max_panels)
      throw Error(GP_ERR_PANEL_CAPACITY, "extruded layout needs " + std::to_string(quads.size()) +
                                             " panels; limit is " + std::to_string(max_panels));
    for (uint32_t k = 0; k < g.n_rects; ++k) {
      b[0] = std::min(b[0], (double)g.rects[4 * k] * xy_scale);
      b[1] = std::min(b[1], (double)g.rects[4 * k + 1] * xy_scale);
      b[2] = std::max(b[2], (double)g.rects[4 * k + 2] * xy_scale);
      b[3] = std::max(b[3], (double)g.rects[4 * k + 3] * xy_scale);
    }
    min_z = std::min(min_z, z0);
    max_thick = std::max(max_thick, z1 - z0);
  }
  // Substrate plate (bridge.rs:346-379).
  const double gap = min_z > 0.0 ? min_z : std::max(max_thick, 100.0 * kMetresPerNm);
  const double gz = min_z > 0.0 ? 0.0 : min_z - gap;
  const double sx = b[2] - b[0], sy = b[3] - b[1];
  const double pad = 0.5 * std::max(std::max(sx, sy), 8.0 * gap);
  const double lx = sx + 2.0 * pad, ly = sy + 2.0 * pad;
  double geps = groups[0].dielectric_k, bestz = groups[0].height_nm * kMetresPerNm;
  for (uint32_t i = 1; i < n_groups; ++i) {
    double z0 = groups[i].height_nm * kMetresPerNm;
    if (z0 < bestz) {
      bestz = z0;
      geps = groups[i].dielectric_k;
    }
  }
  const int gnx = (int)std::min(std::max(std::ceil(lx / (4.0 * gap)), 4.0), 32.0) * refine;
  const int gny = (int)std::min(std::max(std::ceil(ly / (4.0 * gap)), 4.0), 32.0) * refine;
  size_t before = quads.size();
  plate(quads, 0.5 * (b[0] + b[2]), 0.5 * (b[1] + b[3]), gz, lx, ly, gnx, gny);
  m.cond.insert(m.cond.end(), quads.size() - before, (uint32_t)m.nets.size());
  m.eps.insert(m.eps.end(), quads.size() - before, geps);
  m.ground_z = gz;
  m.n_plate = quads.size() - before;
  if (quads.size() > max_panels)
    throw Error(GP_ERR_PANEL_CAPACITY, "extruded layout needs " + std::to_string(quads.size()) +
                                           " panels; limit is " + std::to_string(max_panels));
  m.verts.resize(quads.size() * 12);
  for (size_t i = 0; i < quads.size(); ++i)
    std::copy(quads[i].v, quads[i].v + 12, m.verts.begin() + 12 * i);
}

}  // namespace gp
