// DC resistance half of extract_quasistatic (crates/pex/src/bridge.rs:73-147,
// 180-299).
//
// The reference turns every polygon of a layer with a sheet resistance into
// one FastHenry segment between two fresh nodes (bridge.rs:245-273), with its
// own port, and solves the whole netlist at f = 0 with the nodal direct
// method (henry/solver.rs:336-584): branch impedances Zb = R + j*0*L, LU of
// Zb, Yn = A Zb^-1 A^T over the non-ground nodes, LU of Yn, one unit current
// per port.  Each segment is its own connected component, grounded at its
// port's second node (solver.rs:366-371), so at f = 0 Zb and Yn are diagonal
// and the chain reduces, operation for operation, to
//     g = 1 / R  (num-complex division: R / (R * R)),
//     Z = 1 / g  (g / (g * g)),
// with R = length / (sigma * (w * h)) from build_filaments (solver.rs:
// 162-198, one filament: filament_partition gives (0, w) and (0, h)).  The
// off-diagonal entries j*0*L are signed zeros, which change no real part
// (partial inductances are finite for distinct filaments).  This file
// evaluates that chain per polygon in O(1) instead of the reference's dense
// O(segments^3) complex LU; the oracle restates the full nodal solve and the
// tests compare the two bit for bit.
#include "gp_resistance.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <map>
#include <numeric>
#include <string>

#include "../../include/gpurify.h"
#include "gp_common.h"

namespace gp {
namespace {

constexpr double kMetresPerNm = 1e-9;

__int128 orient(int64_t ax, int64_t ay, int64_t bx, int64_t by, int64_t cx, int64_t cy) {
  return (__int128)(bx - ax) * (__int128)(cy - ay) - (__int128)(by - ay) * (__int128)(cx - ax);
}

// poly_self_intersects (crates/core/src/geometry/ops.rs:216-298): a proper
// crossing of two non-adjacent, non-degenerate edges, found by a sweep along
// the wider bounding-box axis.
bool self_intersects(const int32_t *xy, int n, int64_t bw, int64_t bh) {
  if (n < 4) return false;
  struct E {
    int64_t x0, y0, x1, y1;
  };
  std::vector<E> e(n);
  for (int i = 0; i < n; ++i) {
    const int j = (i + 1) % n;
    e[i] = {xy[2 * i], xy[2 * i + 1], xy[2 * j], xy[2 * j + 1]};
  }
  const bool sx = bw >= bh;
  auto lo = [&](const E &a) { return sx ? std::min(a.x0, a.x1) : std::min(a.y0, a.y1); };
  auto hi = [&](const E &a) { return sx ? std::max(a.x0, a.x1) : std::max(a.y0, a.y1); };
  auto zero = [](const E &a) { return a.x0 == a.x1 && a.y0 == a.y1; };
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return lo(e[a]) < lo(e[b]); });
  for (int w = 0; w < n; ++w) {
    const int i = order[w];
    const E &a = e[i];
    if (zero(a)) continue;
    const int64_t ahi = hi(a);
    for (int q = w + 1; q < n; ++q) {
      const int j = order[q];
      const E &b = e[j];
      if (lo(b) > ahi) break;
      if (j == (i + 1) % n || i == (j + 1) % n) continue;
      if (zero(b)) continue;
      const __int128 d1 = orient(b.x0, b.y0, b.x1, b.y1, a.x0, a.y0);
      const __int128 d2 = orient(b.x0, b.y0, b.x1, b.y1, a.x1, a.y1);
      const __int128 d3 = orient(a.x0, a.y0, a.x1, a.y1, b.x0, b.y0);
      const __int128 d4 = orient(a.x0, a.y0, a.x1, a.y1, b.x1, b.y1);
      if (((d1 > 0) != (d2 > 0)) && ((d3 > 0) != (d4 > 0)) && d1 != 0 && d2 != 0 && d3 != 0 &&
          d4 != 0)
        return true;
    }
  }
  return false;
}

}  // namespace

// rectilinear_metrics (crates/pex/src/analytical/mod.rs:577-628): the
// rectangle with the polygon's area and perimeter.
void rectilinear_metrics(const int32_t *xy, int n, uint32_t poly, double &length_nm,
                         double &width_nm) {
  auto fail = [&](const std::string &m) {
    throw Error(GP_ERR_INVALID_INPUT, "polygon " + std::to_string(poly) + ": " + m);
  };
  if (n < 4)
    fail("polygon has " + std::to_string(n) + " vertices; at least four are required");
  int32_t xmin = xy[0], xmax = xy[0], ymin = xy[1], ymax = xy[1];
  for (int i = 1; i < n; ++i) {
    xmin = std::min(xmin, xy[2 * i]);
    xmax = std::max(xmax, xy[2 * i]);
    ymin = std::min(ymin, xy[2 * i + 1]);
    ymax = std::max(ymax, xy[2 * i + 1]);
  }
  if (self_intersects(xy, n, (int64_t)xmax - xmin, (int64_t)ymax - ymin))
    fail("self-intersecting polygon is unsupported");
  __int128 area2 = 0;
  double perimeter = 0.0;
  for (int i = 0; i < n; ++i) {
    const int j = (i + 1) % n;
    const int64_t x0 = xy[2 * i], y0 = xy[2 * i + 1], x1 = xy[2 * j], y1 = xy[2 * j + 1];
    const int64_t dx = x1 - x0, dy = y1 - y0;
    if (dx == 0 && dy == 0) fail("zero-length edge at vertex " + std::to_string(i));
    if (dx != 0 && dy != 0) fail("non-Manhattan edge at vertex " + std::to_string(i));
    perimeter += (double)(std::llabs(dx) + std::llabs(dy));
    area2 += (__int128)x0 * y1 - (__int128)x1 * y0;
  }
  const double area = (double)(area2 < 0 ? -area2 : area2) / 2.0;
  if (!std::isfinite(area) || area <= 0.0 || perimeter <= 0.0)
    fail("polygon has zero or non-finite area/perimeter");
  const double s = perimeter / 2.0;
  const double disc = s * s - 4.0 * area;
  const double tol = s * s * 1e-12;
  if (disc < -tol) fail("area/perimeter cannot form a rectilinear equivalent rectangle");
  const double root = std::sqrt(std::max(disc, 0.0));
  length_nm = (s + root) / 2.0;
  width_nm = (s - root) / 2.0;
  if (width_nm <= 0.0 || !std::isfinite(length_nm))
    fail("polygon has no positive equivalent conductor width");
}

double segment_dc_ohm(const int32_t *xy, int n, uint32_t poly, double dbu_nm,
                      double sheet_res_ohm_sq, double thickness_nm, double height_nm) {
  const double xy_scale = dbu_nm * kMetresPerNm;
  double len_nm, wid_nm;
  rectilinear_metrics(xy, n, poly, len_nm, wid_nm);
  const double length = len_nm * xy_scale;
  const double width = wid_nm * xy_scale;
  const double thickness = thickness_nm * kMetresPerNm;
  const double sigma = 1.0 / (sheet_res_ohm_sq * thickness);
  if (!std::isfinite(length) || !std::isfinite(width) || !std::isfinite(sigma) || length <= 0.0 ||
      width <= 0.0 || sigma <= 0.0)
    throw Error(GP_ERR_INVALID_INPUT, "polygon " + std::to_string(poly) +
                                          ": invalid process value sheet-resistance volume");
  // Node positions of bridge.rs:229-250 and the filament of solver.rs:162-198.
  int32_t xmin = xy[0], xmax = xy[0], ymin = xy[1], ymax = xy[1];
  for (int i = 1; i < n; ++i) {
    xmin = std::min(xmin, xy[2 * i]);
    xmax = std::max(xmax, xy[2 * i]);
    ymin = std::min(ymin, xy[2 * i + 1]);
    ymax = std::max(ymax, xy[2 * i + 1]);
  }
  const double cx = 0.5 * ((double)xmin + (double)xmax) * xy_scale;
  const double cy = 0.5 * ((double)ymin + (double)ymax) * xy_scale;
  const double cz = (height_nm + 0.5 * thickness_nm) * kMetresPerNm;
  const bool along_x = (int64_t)xmax - xmin >= (int64_t)ymax - ymin;
  double p1[3], p2[3];
  if (along_x) {
    p1[0] = cx - 0.5 * length, p1[1] = cy, p1[2] = cz;
    p2[0] = cx + 0.5 * length, p2[1] = cy, p2[2] = cz;
  } else {
    p1[0] = cx, p1[1] = cy - 0.5 * length, p1[2] = cz;
    p2[0] = cx, p2[1] = cy + 0.5 * length, p2[2] = cz;
  }
  const double d[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
  const double seg_len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);  // Vec3::norm
  const double r = seg_len / (sigma * (width * thickness));
  // x = 1/Zb and v = 1/Yn through num-complex division of (1 + 0i) by a
  // real-valued complex: re = (1 * c + 0 * d) / (c * c + d * d), d = +-0.
  const double g = r / (r * r);
  return g / (g * g);
}

void solve_resistance(const gp_res_polygon *polys, uint32_t n_polys, double dbu_nm,
                      std::map<uint32_t, double> &r_of_net) {
  if (!std::isfinite(dbu_nm) || dbu_nm <= 0.0)
    throw Error(GP_ERR_INVALID_INPUT, "invalid database unit " + std::to_string(dbu_nm) + " nm");
  if (n_polys && !polys) throw Error(GP_ERR_INVALID_INPUT, "polygons are NULL");
  // extract_quasistatic's polygon loop (bridge.rs:97-129): via resistance
  // first, map entries for volume conductors.
  for (uint32_t p = 0; p < n_polys; ++p) {
    const gp_res_polygon &q = polys[p];
    if (!q.has_pex) continue;
    if (q.via_res_ohm != 0.0) r_of_net[q.net] += q.via_res_ohm;
    if (q.volume_conductor) r_of_net.emplace(q.net, 0.0);
  }
  // solve_resistance (bridge.rs:180-299): one port per sheet-resistance
  // polygon, summed into its net in polygon order.
  for (uint32_t p = 0; p < n_polys; ++p) {
    const gp_res_polygon &q = polys[p];
    if (!q.has_pex || q.sheet_res_ohm_sq == 0.0) continue;
    if (!q.xy && q.n_vertices) throw Error(GP_ERR_INVALID_INPUT, "polygon vertices are NULL");
    const double ohm = segment_dc_ohm(q.xy, (int)q.n_vertices, p, dbu_nm, q.sheet_res_ohm_sq,
                                      q.thickness_nm, q.height_nm);
    if (!std::isfinite(ohm))
      throw Error(GP_ERR_NON_FINITE, "net " + std::to_string(q.net) + ": non-finite resistance");
    r_of_net[q.net] += ohm;
  }
}

}  // namespace gp
