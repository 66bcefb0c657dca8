// gp_oracle.cpp - CPU FP64 restatement of the reference capacitance path.
//
// TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg; never by the
// product library.  Each routine names the reference file:line it restates
// (paths relative to the reference repository root, crate crates/pex unless
// stated).  OpenMP stands in for rayon where the reference parallelises.
//
// Parity status per function (how each is pinned) is listed in the tests that
// exercise it; see tests/test_oracle_*.py.

#include "gp_oracle.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <set>
#include <tuple>
#include <unordered_map>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------- Vec3
// quasistatic/geometry.rs:11-110
struct V3 {
  double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator/(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 normalized(V3 a) {
  double n = norm(a);
  return n == 0.0 ? V3{0, 0, 0} : a / n;
}

const double EPS0 = 8.8541878128e-12;                        // constants.rs:13
const double ONE_OVER_4PI_EPS0 = 1.0 / (4.0 * M_PI * EPS0);  // constants.rs:20

// A panel as the reference's Vec<Vec3> (3 or 4 vertices).  The boundary
// format stores 4 vertices; a triangle repeats its last vertex, which adds a
// zero-length edge (skipped by the Wilton loop, panel.rs:163) and a
// zero-area fan triangle, so every quantity is unchanged.
struct Pan {
  V3 v[4];
  int nv;
};

Pan load_panel(const double *verts4, long i) {
  Pan p;
  const double *q = verts4 + 12 * i;
  for (int k = 0; k < 4; ++k) p.v[k] = {q[3 * k], q[3 * k + 1], q[3 * k + 2]};
  p.nv = 4;
  return p;
}

// panel.rs:256-269 (Newell normal)
V3 polygon_normal(const V3 *v, int nv) {
  double nx = 0, ny = 0, nz = 0;
  for (int i = 0; i < nv; ++i) {
    V3 c = v[i], d = v[(i + 1) % nv];
    nx += (c.y - d.y) * (c.z + d.z);
    ny += (c.z - d.z) * (c.x + d.x);
    nz += (c.x - d.x) * (c.y + d.y);
  }
  return normalized(V3{nx, ny, nz});
}

// panel.rs:29-47
V3 centroid(const V3 *v, int nv) {
  V3 v0 = v[0];
  double area_sum = 0.0;
  V3 c{0, 0, 0};
  for (int i = 1; i < nv - 1; ++i) {
    V3 v1 = v[i], v2 = v[i + 1];
    double a = norm(cross(v1 - v0, v2 - v0)) * 0.5;
    V3 tri_c = (v0 + v1 + v2) * (1.0 / 3.0);
    c = c + tri_c * a;
    area_sum += a;
  }
  return area_sum == 0.0 ? v0 : c * (1.0 / area_sum);
}

// panel.rs:58-67
double area(const V3 *v, int nv) {
  V3 v0 = v[0];
  double a = 0.0;
  for (int i = 1; i < nv - 1; ++i) a += norm(cross(v[i] - v0, v[i + 1] - v0)) * 0.5;
  return a;
}

// panel.rs:50-55
V3 panel_normal(const V3 *v) { return normalized(cross(v[1] - v[0], v[2] - v[0])); }

// panel.rs:276-285
inline double asinh_over(double s, double r0_sq) {
  double r0 = std::sqrt(r0_sq);
  return r0 == 0.0 ? 0.0 : std::asinh(s / r0);
}

// panel.rs:147-208
double potential_integral(const V3 *vertices, int nv, V3 p) {
  V3 n = polygon_normal(vertices, nv);
  V3 v0 = vertices[0];
  double d = dot(p - v0, n);
  V3 p0 = p - n * d;
  double absd = std::fabs(d);
  double total = 0.0;
  for (int i = 0; i < nv; ++i) {
    V3 a = vertices[i], b = vertices[(i + 1) % nv];
    V3 edge = b - a;
    double edge_len = norm(edge);
    if (edge_len == 0.0) continue;
    V3 lhat = edge / edge_len;
    V3 uhat = cross(lhat, n);
    double p0_perp = dot(a - p0, uhat);
    double s_minus = dot(a - p0, lhat);
    double s_plus = dot(b - p0, lhat);
    double r0_sq = p0_perp * p0_perp + d * d;
    double r_plus = std::sqrt(s_plus * s_plus + r0_sq);
    double r_minus = std::sqrt(s_minus * s_minus + r0_sq);
    double num = r_plus + s_plus;
    double den = r_minus + s_minus;
    double log_term;
    if (num > 0.0 && den > 0.0)
      log_term = p0_perp * std::log(num / den);
    else if (std::fabs(den) > 0.0)
      log_term = p0_perp * (asinh_over(s_plus, r0_sq) - asinh_over(s_minus, r0_sq));
    else
      log_term = 0.0;
    double atan_term = 0.0;
    if (absd > 0.0 && p0_perp != 0.0) {
      double t_plus = (p0_perp * s_plus) / (r0_sq + absd * r_plus);
      double t_minus = (p0_perp * s_minus) / (r0_sq + absd * r_minus);
      atan_term = absd * (std::atan(t_plus) - std::atan(t_minus));
    }
    total += log_term - atan_term;
  }
  return total;
}

// panel.rs:213-221
const double TRI7[7][4] = {
    {0.3333333333333333, 0.3333333333333333, 0.3333333333333333, 0.225},
    {0.059715871789770, 0.470142064105115, 0.470142064105115, 0.132394152788506},
    {0.470142064105115, 0.059715871789770, 0.470142064105115, 0.132394152788506},
    {0.470142064105115, 0.470142064105115, 0.059715871789770, 0.132394152788506},
    {0.797426985353087, 0.101286507323456, 0.101286507323456, 0.125939180544827},
    {0.101286507323456, 0.797426985353087, 0.101286507323456, 0.125939180544827},
    {0.101286507323456, 0.101286507323456, 0.797426985353087, 0.125939180544827},
};

// panel.rs:232-253
double field_integral(const V3 *v, int nv, V3 p, V3 n_obs) {
  V3 v0 = v[0];
  double acc = 0.0;
  for (int i = 1; i < nv - 1; ++i) {
    V3 v1 = v[i], v2 = v[i + 1];
    double ar = norm(cross(v1 - v0, v2 - v0)) * 0.5;
    if (ar == 0.0) continue;
    for (int q = 0; q < 7; ++q) {
      V3 rp = v0 * TRI7[q][0] + v1 * TRI7[q][1] + v2 * TRI7[q][2];
      V3 dd = rp - p;
      double r2 = dot(dd, dd);
      if (r2 > 0.0) {
        double r = std::sqrt(r2);
        acc += TRI7[q][3] * ar * dot(dd, n_obs) / (r2 * r);
      }
    }
  }
  return acc;
}

// PanelSet (panel.rs:78-119): SoA lanes.
struct PanelSet {
  std::vector<Pan> pans;
  std::vector<V3> c;
  std::vector<double> area;
  std::vector<uint32_t> cond;
  long n() const { return (long)pans.size(); }
};

PanelSet build_set(int n, const double *verts4, const uint32_t *cond) {
  PanelSet s;
  s.pans.resize(n);
  s.c.resize(n);
  s.area.resize(n);
  s.cond.resize(n, 0);
  for (int i = 0; i < n; ++i) {
    s.pans[i] = load_panel(verts4, i);
    s.c[i] = centroid(s.pans[i].v, 4);
    s.area[i] = area(s.pans[i].v, 4);
    if (cond) s.cond[i] = cond[i];
  }
  return s;
}

// ---------------------------------------------------------------- dense LU
// linalg.rs:172-252 (Doolittle, partial pivot, row-major; OpenMP rows in
// place of rayon when the trailing block exceeds 64^2).
struct LU {
  long n;
  std::vector<double> a;
  std::vector<long> piv;
  bool factor(const double *A, long nn) {
    n = nn;
    a.assign(A, A + n * n);
    piv.resize(n);
    for (long i = 0; i < n; ++i) piv[i] = i;
    for (long k = 0; k < n; ++k) {
      long p = k;
      double mx = std::fabs(a[k * n + k]);
      for (long i = k + 1; i < n; ++i) {
        double m = std::fabs(a[i * n + k]);
        if (m > mx) {
          mx = m;
          p = i;
        }
      }
      if (mx == 0.0) return false;
      if (p != k) {
        for (long j = 0; j < n; ++j) std::swap(a[k * n + j], a[p * n + j]);
        std::swap(piv[k], piv[p]);
      }
      const double pivot = a[k * n + k];
      const double *prow = &a[k * n];
      long rows_left = n - k - 1;
      bool par = rows_left * (n - k) > 64 * 64;
#pragma omp parallel for schedule(static) if (par)
      for (long i = k + 1; i < n; ++i) {
        double *row = &a[i * n];
        double f = row[k] / pivot;
        row[k] = f;
        for (long j = k + 1; j < n; ++j) row[j] = row[j] - f * prow[j];
      }
    }
    return true;
  }
  void solve(const double *b, double *x) const {
    std::vector<double> t(n);
    for (long i = 0; i < n; ++i) t[i] = b[piv[i]];
    for (long i = 0; i < n; ++i) {
      double s = t[i];
      for (long j = 0; j < i; ++j) s = s - a[i * n + j] * t[j];
      t[i] = s;
    }
    for (long i = n - 1; i >= 0; --i) {
      double s = t[i];
      for (long j = i + 1; j < n; ++j) s = s - a[i * n + j] * t[j];
      t[i] = s / a[i * n + i];
    }
    std::copy(t.begin(), t.end(), x);
  }
};

// ---------------------------------------------------------------- GMRES
// simd.rs:20-70: 16 partial sums (4 x f64x4 FMA accumulators).
double sdot(const double *a, const double *b, long n) {
  double acc[16] = {0};
  long n16 = n & ~15L, j = 0;
  for (; j < n16; j += 16)
    for (int l = 0; l < 16; ++l) acc[l] = std::fma(a[j + l], b[j + l], acc[l]);
  long n4 = n & ~3L;
  for (; j < n4; j += 4)
    for (int l = 0; l < 4; ++l) acc[l] = std::fma(a[j + l], b[j + l], acc[l]);
  double v[4];
  for (int l = 0; l < 4; ++l) v[l] = (acc[l] + acc[4 + l]) + (acc[8 + l] + acc[12 + l]);
  double s = (v[0] + v[1]) + (v[2] + v[3]);
  for (; j < n; ++j) s += a[j] * b[j];
  return s;
}
double snrm2(const double *a, long n) { return std::sqrt(sdot(a, a, n)); }
void saxpy(double alpha, const double *x, double *y, long n) {
  for (long j = 0; j < n; ++j) y[j] = std::fma(alpha, x[j], y[j]);
}

struct Op {
  virtual long dim() const = 0;
  virtual void apply(const double *x, double *y) const = 0;
  virtual ~Op() {}
};

struct GmresResult {
  std::vector<double> x;
  int iterations;
  double residual;
  bool converged;
};

// krylov.rs:195-207
void givens(double a, double b, double &c, double &s) {
  if (b == 0.0) {
    c = 1.0;
    s = 0.0;
  } else if (std::fabs(b) > std::fabs(a)) {
    double t = a / b;
    s = 1.0 / std::sqrt(1.0 + t * t);
    c = t * s;
  } else {
    double t = b / a;
    c = 1.0 / std::sqrt(1.0 + t * t);
    s = t * c;
  }
}

// krylov.rs:50-191 (restarted GMRES(m), right preconditioning, MGS)
GmresResult gmres(const Op &a, const double *b, const Op *precond, double tol, int restart,
                  int max_restarts) {
  long n = a.dim();
  double bnorm = std::max(snrm2(b, n), 2.2250738585072014e-308);
  std::vector<double> x(n, 0.0), z(n), ax(n);
  int total_iters = 0;
  for (int cycle = 0; cycle < max_restarts; ++cycle) {
    a.apply(x.data(), ax.data());
    std::vector<double> r(n);
    for (long i = 0; i < n; ++i) r[i] = b[i] - ax[i];
    double beta = snrm2(r.data(), n);
    if (beta / bnorm <= tol) return {x, total_iters, beta / bnorm, true};
    int m = restart;
    std::vector<std::vector<double>> v;
    v.reserve(m + 1);
    for (auto &vi : r) vi /= beta;
    v.push_back(r);
    std::vector<std::vector<double>> h(m + 1, std::vector<double>(m, 0.0));
    std::vector<double> cs(m, 0.0), sn(m, 0.0), g(m + 1, 0.0);
    g[0] = beta;
    int k_used = 0;
    for (int k = 0; k < m; ++k) {
      const double *zk;
      if (precond) {
        precond->apply(v[k].data(), z.data());
        zk = z.data();
      } else {
        zk = v[k].data();
      }
      std::vector<double> w(n, 0.0);
      a.apply(zk, w.data());
      for (int i = 0; i <= k; ++i) {
        h[i][k] = sdot(w.data(), v[i].data(), n);
        saxpy(-h[i][k], v[i].data(), w.data(), n);
      }
      h[k + 1][k] = snrm2(w.data(), n);
      if (h[k + 1][k] > 1e-14)
        for (auto &wj : w) wj /= h[k + 1][k];
      v.push_back(w);
      for (int i = 0; i < k; ++i) {
        double temp = cs[i] * h[i][k] + sn[i] * h[i + 1][k];
        h[i + 1][k] = -sn[i] * h[i][k] + cs[i] * h[i + 1][k];
        h[i][k] = temp;
      }
      double c, s;
      givens(h[k][k], h[k + 1][k], c, s);
      cs[k] = c;
      sn[k] = s;
      h[k][k] = c * h[k][k] + s * h[k + 1][k];
      h[k + 1][k] = 0.0;
      double temp = c * g[k] + s * g[k + 1];
      g[k + 1] = -s * g[k] + c * g[k + 1];
      g[k] = temp;
      total_iters += 1;
      k_used = k + 1;
      if (std::fabs(g[k + 1]) / bnorm <= tol) break;
    }
    std::vector<double> y(k_used, 0.0);
    for (int i = k_used - 1; i >= 0; --i) {
      double sum = g[i];
      for (int j = i + 1; j < k_used; ++j) sum -= h[i][j] * y[j];
      y[i] = sum / h[i][i];
    }
    std::vector<double> dx(n, 0.0);
    for (int i = 0; i < k_used; ++i)
      for (long j = 0; j < n; ++j) dx[j] += y[i] * v[i][j];
    const double *dz;
    if (precond) {
      precond->apply(dx.data(), z.data());
      dz = z.data();
    } else {
      dz = dx.data();
    }
    for (long j = 0; j < n; ++j) x[j] += dz[j];
    double est = std::fabs(g[k_used]) / bnorm;
    if (est <= tol * 10.0) {
      a.apply(x.data(), ax.data());
      std::vector<double> r2(n);
      for (long i = 0; i < n; ++i) r2[i] = b[i] - ax[i];
      double resid = snrm2(r2.data(), n) / bnorm;
      if (resid <= tol) return {x, total_iters, resid, true};
    }
  }
  a.apply(x.data(), ax.data());
  std::vector<double> r(n);
  for (long i = 0; i < n; ++i) r[i] = b[i] - ax[i];
  return {x, total_iters, snrm2(r.data(), n) / bnorm, false};
}

// ---------------------------------------------------------------- dense P
// cap/solver.rs:81-97
void assemble_p(const PanelSet &s, double *P) {
  long n = s.n();
#pragma omp parallel for schedule(dynamic, 8)
  for (long k = 0; k < n; ++k) {
    V3 rk = s.c[k];
    for (long l = 0; l < n; ++l)
      P[k * n + l] = ONE_OVER_4PI_EPS0 * potential_integral(s.pans[l].v, 4, rk);
  }
}

struct DenseOp : Op {
  const double *P;
  long n;
  long dim() const override { return n; }
  // operator.rs:24-35 (row dot)
  void apply(const double *x, double *y) const override {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) y[i] = sdot(P + i * n, x, n);
  }
};

// ---------------------------------------------------------------- Chebyshev
// fmm/chebyshev.rs:11-17
std::vector<double> cheb_nodes(int n) {
  std::vector<double> v(n);
  for (int m = 0; m < n; ++m) v[m] = std::cos((2.0 * m + 1.0) * M_PI / (2.0 * n));
  return v;
}

// chebyshev.rs:40-74
void interp_weights(double x, int n, const double *nodes, double *w) {
  double tx[64];
  tx[0] = 1.0;
  if (n > 1) {
    tx[1] = x;
    for (int k = 2; k < n; ++k) tx[k] = 2.0 * x * tx[k - 1] - tx[k - 2];
  }
  double inv_n = 1.0 / n;
  for (int m = 0; m < n; ++m) {
    double xm = nodes[m];
    double sum = 0.5, tkm1 = 1.0, tk = xm;
    for (int k = 1; k < n; ++k) {
      double tkm;
      if (k == 1) {
        tkm = xm;
      } else {
        double t = 2.0 * xm * tk - tkm1;
        tkm1 = tk;
        tk = t;
        tkm = t;
      }
      sum += tkm * tx[k];
    }
    w[m] = 2.0 * inv_n * sum;
  }
}

// d/dx of the 1-D interpolation weights (no reference counterpart: the
// gradient of the far-field interpolant, for the FMM form of the dielectric
// interface rows): w_m'(x) = (2/n) sum_{k=1}^{n-1} T_k(x_m) T_k'(x), with
// T_k' = k U_{k-1} (Chebyshev polynomials of the second kind).
void interp_dweights(double x, int n, const double *nodes, double *dw) {
  double u[64], dt[64];  // U_k(x), T_k'(x)
  u[0] = 1.0;
  if (n > 1) u[1] = 2.0 * x;
  for (int k = 2; k < n; ++k) u[k] = 2.0 * x * u[k - 1] - u[k - 2];
  dt[0] = 0.0;
  for (int k = 1; k < n; ++k) dt[k] = k * u[k - 1];
  for (int m = 0; m < n; ++m) {
    const double xm = nodes[m];
    double s = 0.0, tkm1 = 1.0, tk = xm;  // T_0, T_1 at the node
    for (int k = 1; k < n; ++k) {
      if (k >= 2) {
        const double t = 2.0 * xm * tk - tkm1;
        tkm1 = tk;
        tk = t;
      }
      s += tk * dt[k];
    }
    dw[m] = 2.0 / n * s;
  }
}

// chebyshev.rs:77-94
void interp_weights_3d(const double p[3], int n, const double *nodes, double *w) {
  double w1[64], w2[64], w3[64];
  interp_weights(p[0], n, nodes, w1);
  interp_weights(p[1], n, nodes, w2);
  interp_weights(p[2], n, nodes, w3);
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < n; ++b) {
      double wab = w1[a] * w2[b];
      int base = (a * n + b) * n;
      for (int c = 0; c < n; ++c) w[base + c] = wab * w3[c];
    }
}

// ---------------------------------------------------------------- octree
// fmm/tree.rs:12-37
struct Node {
  int level;
  uint32_t cell[3];
  double center[3];
  double half;
  long parent;
  std::vector<long> children, points, neighbors, interaction;
};

struct CellKey {
  uint32_t a, b, c;
  bool operator==(const CellKey &o) const { return a == o.a && b == o.b && c == o.c; }
};
struct CellHash {
  size_t operator()(const CellKey &k) const {
    return ((size_t)k.a * 73856093u) ^ ((size_t)k.b * 19349663u) ^ ((size_t)k.c * 83492791u);
  }
};

struct Octree {
  std::vector<Node> nodes;
  int max_level;
  std::vector<std::vector<long>> levels;
  std::vector<std::unordered_map<CellKey, long, CellHash>> maps;

  // tree.rs:109-151
  long ensure_box(int level, const uint32_t cell[3], const double rc[3], double rh) {
    auto it = maps[level].find({cell[0], cell[1], cell[2]});
    if (it != maps[level].end()) return it->second;
    long parent = -1;
    if (level > 0) {
      uint32_t pc[3] = {cell[0] / 2, cell[1] / 2, cell[2] / 2};
      parent = ensure_box(level - 1, pc, rc, rh);
    }
    double half = rh / (double)(1u << level);
    double origin[3] = {rc[0] - rh, rc[1] - rh, rc[2] - rh};
    Node nd;
    nd.level = level;
    for (int d = 0; d < 3; ++d) {
      nd.cell[d] = cell[d];
      nd.center[d] = origin[d] + (cell[d] + 0.5) * 2.0 * half;
    }
    nd.half = half;
    nd.parent = parent;
    long id = (long)nodes.size();
    nodes.push_back(nd);
    maps[level][{cell[0], cell[1], cell[2]}] = id;
    if (parent >= 0) nodes[parent].children.push_back(id);
    return id;
  }

  // tree.rs:41-106
  void build(const std::vector<V3> &pts, int L) {
    max_level = L;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (auto &p : pts) {
      double a[3] = {p.x, p.y, p.z};
      for (int d = 0; d < 3; ++d) {
        lo[d] = std::min(lo[d], a[d]);
        hi[d] = std::max(hi[d], a[d]);
      }
    }
    double half = 0.0, center[3];
    for (int d = 0; d < 3; ++d) {
      center[d] = 0.5 * (lo[d] + hi[d]);
      half = std::max(half, 0.5 * (hi[d] - lo[d]));
    }
    half *= 1.0 + 1e-6;
    if (half == 0.0) half = 1.0;
    nodes.clear();
    maps.assign(L + 1, {});
    Node root;
    root.level = 0;
    root.cell[0] = root.cell[1] = root.cell[2] = 0;
    for (int d = 0; d < 3; ++d) root.center[d] = center[d];
    root.half = half;
    root.parent = -1;
    nodes.push_back(root);
    maps[0][{0, 0, 0}] = 0;
    uint32_t ncell = 1u << L;
    double leaf_size = 2.0 * half / ncell;
    double origin[3] = {center[0] - half, center[1] - half, center[2] - half};
    for (size_t pi = 0; pi < pts.size(); ++pi) {
      double a[3] = {pts[pi].x, pts[pi].y, pts[pi].z};
      uint32_t cell[3];
      for (int d = 0; d < 3; ++d) {
        int64_t idx = (int64_t)std::floor((a[d] - origin[d]) / leaf_size);
        idx = std::min<int64_t>(std::max<int64_t>(idx, 0), (int64_t)ncell - 1);
        cell[d] = (uint32_t)idx;
      }
      long leaf = ensure_box(L, cell, center, half);
      nodes[leaf].points.push_back((long)pi);
    }
    levels.assign(L + 1, {});
    for (size_t id = 0; id < nodes.size(); ++id) levels[nodes[id].level].push_back((long)id);
    build_lists();
  }

  // tree.rs:154-210
  void build_lists() {
    for (size_t id = 0; id < nodes.size(); ++id) {
      int level = nodes[id].level;
      int64_t nc = (int64_t)1 << level;
      std::vector<long> nb;
      for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dz = -1; dz <= 1; ++dz) {
            if (!dx && !dy && !dz) continue;
            int64_t c[3] = {(int64_t)nodes[id].cell[0] + dx, (int64_t)nodes[id].cell[1] + dy,
                            (int64_t)nodes[id].cell[2] + dz};
            if (c[0] < 0 || c[0] >= nc || c[1] < 0 || c[1] >= nc || c[2] < 0 || c[2] >= nc) continue;
            auto it = maps[level].find({(uint32_t)c[0], (uint32_t)c[1], (uint32_t)c[2]});
            if (it != maps[level].end()) nb.push_back(it->second);
          }
      nodes[id].neighbors = nb;
    }
    for (size_t id = 0; id < nodes.size(); ++id) {
      if (nodes[id].level < 2) continue;
      long parent = nodes[id].parent;
      std::set<long> own(nodes[id].neighbors.begin(), nodes[id].neighbors.end());
      std::vector<long> fam = nodes[parent].neighbors;
      fam.push_back(parent);
      std::vector<long> inter;
      for (long pn : fam)
        for (long child : nodes[pn].children)
          if (child != (long)id && !own.count(child)) inter.push_back(child);
      nodes[id].interaction = inter;
    }
  }
  const std::vector<long> &leaves() const { return levels[max_level]; }
};

// ---------------------------------------------------------------- low rank
// linalg.rs:264-400 (column-pivoted MGS QR truncation)
struct LowRank {
  std::vector<double> u, vt;
  int m = 0, n = 0, k = 0;
  void compress(const double *a, int mm, int nn, double eps) {
    m = mm;
    n = nn;
    std::vector<double> q(a, a + (size_t)m * n), r((size_t)n * n, 0.0);
    std::vector<int> perm(n);
    for (int j = 0; j < n; ++j) perm[j] = j;
    std::vector<double> col(n);
    for (int j = 0; j < n; ++j) {
      double s = 0.0;
      for (int i = 0; i < m; ++i) s += a[(size_t)i * n + j] * a[(size_t)i * n + j];
      col[j] = s;
    }
    int max_k = std::min(m, n);
    k = 0;
    double mx = 0.0;
    for (double c : col) mx = std::max(mx, c);
    double thr = eps * mx;
    for (int step = 0; step < max_k; ++step) {
      int best = step;
      for (int j = step + 1; j < n; ++j)
        if (col[j] > col[best]) best = j;
      if (col[best] < thr) break;
      if (best != step) {
        std::swap(perm[step], perm[best]);
        std::swap(col[step], col[best]);
        for (int i = 0; i < m; ++i) std::swap(q[(size_t)i * n + step], q[(size_t)i * n + best]);
        for (int row = 0; row < step; ++row) std::swap(r[(size_t)row * n + step], r[(size_t)row * n + best]);
      }
      double nr = 0.0;
      for (int i = 0; i < m; ++i) nr += q[(size_t)i * n + step] * q[(size_t)i * n + step];
      nr = std::sqrt(nr);
      if (nr < 1e-15) break;
      r[(size_t)step * n + step] = nr;
      double inv = 1.0 / nr;
      for (int i = 0; i < m; ++i) q[(size_t)i * n + step] *= inv;
      for (int j = step + 1; j < n; ++j) {
        double d = 0.0;
        for (int i = 0; i < m; ++i) d += q[(size_t)i * n + step] * q[(size_t)i * n + j];
        r[(size_t)step * n + j] = d;
        for (int i = 0; i < m; ++i) q[(size_t)i * n + j] -= d * q[(size_t)i * n + step];
        col[j] -= d * d;
        if (col[j] < 0.0) col[j] = 0.0;
      }
      k = step + 1;
    }
    if (k == 0) return;
    u.assign((size_t)m * k, 0.0);
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < k; ++j) u[(size_t)i * k + j] = q[(size_t)i * n + j];
    vt.assign((size_t)k * n, 0.0);
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < n; ++j) vt[(size_t)i * n + perm[j]] = r[(size_t)i * n + j];
  }
  // linalg.rs:386-400
  void apply_add(const double *x, double *y, double scale, double *scratch) const {
    if (k == 0) return;
    for (int i = 0; i < k; ++i) {
      double s = 0.0;
      const double *row = &vt[(size_t)i * n];
      for (int j = 0; j < n; ++j) s += row[j] * x[j];
      scratch[i] = s;
    }
    for (int i = 0; i < m; ++i) {
      const double *row = &u[(size_t)i * k];
      double v = 0.0;
      for (int j = 0; j < k; ++j) v += row[j] * scratch[j];
      y[i] += scale * v;
    }
  }
};

// ---------------------------------------------------------------- bbFMM
const int M2L_SPAN = 7;  // bbfmm.rs:20-25
inline int m2l_index(int o0, int o1, int o2) {
  return ((o0 + 3) * M2L_SPAN + o1 + 3) * M2L_SPAN + o2 + 3;
}

struct Fmm {
  Octree tree;
  int n = 0, n3 = 0;
  std::vector<double> nodes1d;
  std::vector<std::array<double, 3>> ref3d;
  std::vector<V3> points;
  std::vector<std::vector<double>> pw;  // point weights n^3 per point (bbfmm.rs:33-34)
  std::vector<std::vector<double>> m2m_1d;
  std::vector<LowRank> m2l_lr;      // mode 0
  std::vector<std::vector<double>> m2l_dense;  // mode 1
  std::vector<char> m2l_have;
  int mode = 0;

  // bbfmm.rs:46-128
  void init(const std::vector<V3> &pts, int order, int max_level, int m2l_mode) {
    mode = m2l_mode;
    tree.build(pts, max_level);
    n = order;
    n3 = n * n * n;
    nodes1d = cheb_nodes(n);
    ref3d.resize(n3);
    for (int a = 0; a < n; ++a)
      for (int b = 0; b < n; ++b)
        for (int c = 0; c < n; ++c) ref3d[(a * n + b) * n + c] = {nodes1d[a], nodes1d[b], nodes1d[c]};
    points = pts;
    pw.assign(pts.size(), {});
    for (long leaf : tree.leaves()) {
      const Node &nd = tree.nodes[leaf];
      for (long pi : nd.points) {
        double loc[3] = {(pts[pi].x - nd.center[0]) / nd.half, (pts[pi].y - nd.center[1]) / nd.half,
                         (pts[pi].z - nd.center[2]) / nd.half};
        pw[pi].resize(n3);
        interp_weights_3d(loc, n, nodes1d.data(), pw[pi].data());
      }
    }
    // bbfmm.rs:89-112
    m2m_1d.assign(24, std::vector<double>(n * n, 0.0));
    for (int oz = 0; oz < 2; ++oz)
      for (int oy = 0; oy < 2; ++oy)
        for (int ox = 0; ox < 2; ++ox) {
          int o = ox + 2 * oy + 4 * oz;
          double cc[3] = {(2.0 * ox - 1.0) * 0.5, (2.0 * oy - 1.0) * 0.5, (2.0 * oz - 1.0) * 0.5};
          for (int axis = 0; axis < 3; ++axis) {
            auto &mat = m2m_1d[o * 3 + axis];
            for (int k = 0; k < n; ++k) {
              double w[64];
              interp_weights(cc[axis] + 0.5 * nodes1d[k], n, nodes1d.data(), w);
              for (int m = 0; m < n; ++m) mat[m * n + k] = w[m];
            }
          }
        }
    precompute_m2l();
  }

  // bbfmm.rs:132-174
  void precompute_m2l() {
    std::set<std::tuple<int, int, int>> offs;
    for (auto &nd : tree.nodes)
      for (long c : nd.interaction) {
        const Node &cn = tree.nodes[c];
        offs.insert({(int)cn.cell[0] - (int)nd.cell[0], (int)cn.cell[1] - (int)nd.cell[1],
                     (int)cn.cell[2] - (int)nd.cell[2]});
      }
    double eps = std::max(1e-3 * std::pow(0.1, n), 1e-14);
    std::vector<std::tuple<int, int, int>> ol(offs.begin(), offs.end());
    int S = M2L_SPAN * M2L_SPAN * M2L_SPAN;
    m2l_have.assign(S, 0);
    if (mode == 0) m2l_lr.assign(S, LowRank());
    else m2l_dense.assign(S, {});
#pragma omp parallel for schedule(dynamic, 1)
    for (long t = 0; t < (long)ol.size(); ++t) {
      int o0, o1, o2;
      std::tie(o0, o1, o2) = ol[t];
      std::vector<double> g((size_t)n3 * n3);
      double sh[3] = {2.0 * o0, 2.0 * o1, 2.0 * o2};
      for (int m = 0; m < n3; ++m)
        for (int l = 0; l < n3; ++l) {
          double dx = ref3d[m][0] - ref3d[l][0] - sh[0];
          double dy = ref3d[m][1] - ref3d[l][1] - sh[1];
          double dz = ref3d[m][2] - ref3d[l][2] - sh[2];
          double r = std::sqrt(dx * dx + dy * dy + dz * dz);
          g[(size_t)m * n3 + l] = r > 0.0 ? 1.0 / r : 0.0;
        }
      int idx = m2l_index(o0, o1, o2);
      if (mode == 0) m2l_lr[idx].compress(g.data(), n3, n3, eps);
      else m2l_dense[idx] = g;
      m2l_have[idx] = 1;
    }
  }

  int octant(long child, long parent) const {
    const Node &c = tree.nodes[child], &p = tree.nodes[parent];
    return (int)(c.cell[0] - 2 * p.cell[0]) + 2 * (int)(c.cell[1] - 2 * p.cell[1]) +
           4 * (int)(c.cell[2] - 2 * p.cell[2]);
  }

  // bbfmm.rs:189-206
  std::vector<std::pair<std::vector<long>, std::vector<long>>> near_blocks() const {
    std::vector<std::pair<std::vector<long>, std::vector<long>>> out;
    for (long leaf : tree.leaves()) {
      const auto &t = tree.nodes[leaf].points;
      if (t.empty()) continue;
      std::vector<long> s = t;
      for (long c : tree.nodes[leaf].neighbors)
        s.insert(s.end(), tree.nodes[c].points.begin(), tree.nodes[c].points.end());
      out.push_back({t, s});
    }
    return out;
  }

  // bbfmm.rs:209-448
  // gnrm/gout (optional): for the points with a nonzero gnrm[pi], gout[pi]
  // = gnrm[pi] . grad of the far-field interpolant at the point (the local
  // expansion differentiated; no reference counterpart).
  std::vector<double> run(const double *q, bool with_p2p, const V3 *gnrm = nullptr,
                          double *gout = nullptr) const {
    size_t nb = tree.nodes.size();
    std::vector<std::vector<double>> w(nb, std::vector<double>(n3, 0.0)),
        local(nb, std::vector<double>(n3, 0.0));
    // P2M (215-228)
    for (long leaf : tree.leaves()) {
      auto &wl = w[leaf];
      for (long pi : tree.nodes[leaf].points) {
        double qi = q[pi];
        if (qi == 0.0) continue;
        const auto &p = pw[pi];
        for (int m = 0; m < n3; ++m) wl[m] += p[m] * qi;
      }
    }
    std::vector<double> t1(n3), t2(n3);
    // M2M (230-294)
    for (int level = tree.max_level; level >= 1; --level) {
      for (long id : tree.levels[level]) {
        long parent = tree.nodes[id].parent;
        if (parent < 0) continue;
        int o = octant(id, parent);
        const double *mx = m2m_1d[o * 3].data(), *my = m2m_1d[o * 3 + 1].data(),
                     *mz = m2m_1d[o * 3 + 2].data();
        const auto &wc = w[id];
        auto &wp = w[parent];
        for (int a = 0; a < n; ++a)
          for (int b = 0; b < n; ++b)
            for (int cp = 0; cp < n; ++cp) {
              double s = 0.0;
              for (int c = 0; c < n; ++c) s += mz[cp * n + c] * wc[a * n * n + b * n + c];
              t1[a * n * n + b * n + cp] = s;
            }
        for (int a = 0; a < n; ++a)
          for (int bp = 0; bp < n; ++bp)
            for (int cp = 0; cp < n; ++cp) {
              double s = 0.0;
              for (int b = 0; b < n; ++b) s += my[bp * n + b] * t1[a * n * n + b * n + cp];
              t2[a * n * n + bp * n + cp] = s;
            }
        for (int ap = 0; ap < n; ++ap)
          for (int bp = 0; bp < n; ++bp)
            for (int cp = 0; cp < n; ++cp) {
              double s = 0.0;
              for (int a = 0; a < n; ++a) s += mx[ap * n + a] * t2[a * n * n + bp * n + cp];
              wp[ap * n * n + bp * n + cp] += s;
            }
      }
    }
    // M2L (296-335)
    for (int level = 2; level <= tree.max_level; ++level) {
      const auto &ids = tree.levels[level];
      if (ids.empty()) continue;
      double inv_half = 1.0 / tree.nodes[ids[0]].half;
#pragma omp parallel for schedule(dynamic, 4) if (ids.size() >= 512)
      for (long t = 0; t < (long)ids.size(); ++t) {
        long id = ids[t];
        const Node &nd = tree.nodes[id];
        std::vector<double> acc(n3, 0.0), scratch(n3, 0.0);
        for (long c : nd.interaction) {
          const Node &cn = tree.nodes[c];
          int idx = m2l_index((int)cn.cell[0] - (int)nd.cell[0], (int)cn.cell[1] - (int)nd.cell[1],
                              (int)cn.cell[2] - (int)nd.cell[2]);
          if (!m2l_have[idx]) continue;
          const auto &wc = w[c];
          if (mode == 0) {
            m2l_lr[idx].apply_add(wc.data(), acc.data(), inv_half, scratch.data());
          } else {
            const auto &g = m2l_dense[idx];
            for (int m = 0; m < n3; ++m) {
              double s = 0.0;
              for (int l = 0; l < n3; ++l) s += g[(size_t)m * n3 + l] * wc[l];
              acc[m] += inv_half * s;
            }
          }
        }
        auto &lc = const_cast<std::vector<double> &>(local[id]);
        for (int m = 0; m < n3; ++m) lc[m] += acc[m];
      }
    }
    // L2L (337-398)
    for (int level = 2; level < tree.max_level; ++level) {
      for (long id : tree.levels[level]) {
        for (long child : tree.nodes[id].children) {
          int o = octant(child, id);
          const double *mx = m2m_1d[o * 3].data(), *my = m2m_1d[o * 3 + 1].data(),
                       *mz = m2m_1d[o * 3 + 2].data();
          const auto &lp = local[id];
          auto &lch = local[child];
          for (int ap = 0; ap < n; ++ap)
            for (int b = 0; b < n; ++b)
              for (int c = 0; c < n; ++c) {
                double s = 0.0;
                for (int a = 0; a < n; ++a) s += mx[a * n + ap] * lp[a * n * n + b * n + c];
                t1[ap * n * n + b * n + c] = s;
              }
          for (int a = 0; a < n; ++a)
            for (int bp = 0; bp < n; ++bp)
              for (int c = 0; c < n; ++c) {
                double s = 0.0;
                for (int b = 0; b < n; ++b) s += my[b * n + bp] * t1[a * n * n + b * n + c];
                t2[a * n * n + bp * n + c] = s;
              }
          for (int a = 0; a < n; ++a)
            for (int b = 0; b < n; ++b)
              for (int cp = 0; cp < n; ++cp) {
                double s = 0.0;
                for (int c = 0; c < n; ++c) s += mz[c * n + cp] * t2[a * n * n + b * n + c];
                lch[a * n * n + b * n + cp] += s;
              }
        }
      }
    }
    // L2P + P2P (400-447)
    std::vector<double> y(points.size(), 0.0);
    const auto &lv = tree.leaves();
#pragma omp parallel for schedule(dynamic, 4) if (lv.size() >= 512)
    for (long t = 0; t < (long)lv.size(); ++t) {
      long leaf = lv[t];
      const auto &ll = local[leaf];
      const Node &nd = tree.nodes[leaf];
      for (long pi : nd.points) {
        const auto &p = pw[pi];
        double acc = 0.0;
        for (int m = 0; m < n3; ++m) acc += p[m] * ll[m];
        if (gnrm && gout && (gnrm[pi].x != 0.0 || gnrm[pi].y != 0.0 || gnrm[pi].z != 0.0)) {
          double loc[3] = {(points[pi].x - nd.center[0]) / nd.half,
                           (points[pi].y - nd.center[1]) / nd.half,
                           (points[pi].z - nd.center[2]) / nd.half};
          double w[3][64], dw[3][64];
          for (int d = 0; d < 3; ++d) {
            interp_weights(loc[d], n, nodes1d.data(), w[d]);
            interp_dweights(loc[d], n, nodes1d.data(), dw[d]);
          }
          double g[3] = {0.0, 0.0, 0.0};
          for (int a = 0; a < n; ++a)
            for (int b = 0; b < n; ++b)
              for (int c = 0; c < n; ++c) {
                const double l = ll[(a * n + b) * n + c];
                g[0] += dw[0][a] * w[1][b] * w[2][c] * l;
                g[1] += w[0][a] * dw[1][b] * w[2][c] * l;
                g[2] += w[0][a] * w[1][b] * dw[2][c] * l;
              }
          const V3 nn = gnrm[pi];
          gout[pi] = (nn.x * g[0] + nn.y * g[1] + nn.z * g[2]) / nd.half;
        }
        if (with_p2p) {
          V3 xi = points[pi];
          auto add_box = [&](long c) {
            for (long sj : tree.nodes[c].points) {
              if (sj == pi) continue;
              double r = norm(xi - points[sj]);
              if (r > 0.0) acc += q[sj] / r;
            }
          };
          for (long c : nd.neighbors) add_box(c);
          add_box(leaf);
        }
        y[pi] += acc;
      }
    }
    return y;
  }
};

// ---------------------------------------------------------------- FMM cap op
// cap/fmm_solver.rs:23-130
struct FmmCapOp : Op {
  Fmm fmm;
  std::vector<double> areas;
  std::vector<long> off;
  std::vector<uint32_t> idx;
  std::vector<double> val;
  long nn;
  mutable std::vector<double> qbuf;

  void init(const PanelSet &s, int order, int max_level, int m2l_mode, bool values = true) {
    nn = s.n();
    areas = s.area;
    fmm.init(s.c, order, max_level, m2l_mode);
    auto blocks = fmm.near_blocks();
    std::vector<std::vector<uint32_t>> rows_i(nn);
    std::vector<std::vector<double>> rows_v(nn);
#pragma omp parallel for schedule(dynamic, 1)
    for (long b = 0; b < (long)blocks.size(); ++b) {
      const auto &tg = blocks[b].first;
      const auto &sr = blocks[b].second;
      for (long i : tg) {
        V3 ri = s.c[i];
        rows_i[i].resize(sr.size());
        if (values) rows_v[i].resize(sr.size());
        for (size_t k = 0; k < sr.size(); ++k) {
          long j = sr[k];
          rows_i[i][k] = (uint32_t)j;
          if (values) rows_v[i][k] = ONE_OVER_4PI_EPS0 * potential_integral(s.pans[j].v, 4, ri);
        }
      }
    }
    off.assign(nn + 1, 0);
    for (long i = 0; i < nn; ++i) off[i + 1] = off[i] + (long)rows_i[i].size();
    idx.resize(off[nn]);
    if (values) val.resize(off[nn]);
    for (long i = 0; i < nn; ++i) {
      std::copy(rows_i[i].begin(), rows_i[i].end(), idx.begin() + off[i]);
      if (values) std::copy(rows_v[i].begin(), rows_v[i].end(), val.begin() + off[i]);
    }
    qbuf.assign(nn, 0.0);
  }
  long dim() const override { return nn; }
  void apply(const double *sigma, double *y) const override {
    for (long j = 0; j < nn; ++j) qbuf[j] = areas[j] * sigma[j];
    std::vector<double> far = fmm.run(qbuf.data(), false);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < nn; ++i) {
      double acc = ONE_OVER_4PI_EPS0 * far[i];
      for (long k = off[i]; k < off[i + 1]; ++k) acc += val[k] * sigma[idx[k]];
      y[i] = acc;
    }
  }
  // fmm_solver.rs:95-101
  bool diagonal(long i, double &d) const {
    for (long k = off[i]; k < off[i + 1]; ++k)
      if ((long)idx[k] == i) {
        d = val[k];
        return true;
      }
    return false;
  }
};

// fmm_solver.rs:133-144
struct DiagJacobi : Op {
  std::vector<double> inv;
  long dim() const override { return (long)inv.size(); }
  void apply(const double *x, double *y) const override {
    for (size_t i = 0; i < inv.size(); ++i) y[i] = x[i] * inv[i];
  }
};

// cap/solver.rs:55-71
std::vector<double> rhs_unit(const PanelSet &s, uint32_t j) {
  std::vector<double> b(s.n(), 0.0);
  for (long k = 0; k < s.n(); ++k)
    if (s.cond[k] == j) b[k] = 1.0;
  return b;
}
void accumulate_c(const PanelSet &s, const double *sigma, int j, int ncond, double *C,
                  const double *eps_panel) {
  for (long k = 0; k < s.n(); ++k) {
    double w = eps_panel ? eps_panel[k] : 1.0;
    if (eps_panel)
      C[(size_t)s.cond[k] * ncond + j] += w * sigma[k] * s.area[k];
    else
      C[(size_t)s.cond[k] * ncond + j] += sigma[k] * s.area[k];
  }
}

// cap/solver.rs:100-134 (+ dielectric.rs:76-151 restricted to conductor
// roles when eps_panel != NULL, the bridge's inhomogeneous case).
int cap_dense(const PanelSet &s, int ncond, int method, double *C, double *P_out,
              const double *eps_panel) {
  long n = s.n();
  if (n == 0) return -6;
  std::vector<double> P((size_t)n * n);
  assemble_p(s, P.data());
  if (P_out) std::copy(P.begin(), P.end(), P_out);
  std::fill(C, C + (size_t)ncond * ncond, 0.0);
  if (method == 0) {
    LU lu;
    if (!lu.factor(P.data(), n)) return -1;
    std::vector<double> sig(n);
    for (int j = 0; j < ncond; ++j) {
      auto b = rhs_unit(s, j);
      lu.solve(b.data(), sig.data());
      accumulate_c(s, sig.data(), j, ncond, C, eps_panel);
    }
  } else {
    DenseOp op;
    op.P = P.data();
    op.n = n;
    for (int j = 0; j < ncond; ++j) {
      auto b = rhs_unit(s, j);
      auto r = gmres(op, b.data(), nullptr, 1e-8, 100, 50);
      accumulate_c(s, r.x.data(), j, ncond, C, eps_panel);
    }
  }
  return 0;
}

// cap/fmm_solver.rs:147-172
int cap_fmm(const PanelSet &s, int ncond, int order, int max_level, double tol, int m2l_mode,
            double *C, int *iters, double *resid) {
  long n = s.n();
  if (n == 0) return -6;
  FmmCapOp op;
  op.init(s, order, max_level, m2l_mode);
  DiagJacobi pre;
  pre.inv.assign(n, 1.0);
  for (long i = 0; i < n; ++i) {
    double d = 1.0;
    op.diagonal(i, d);
    pre.inv[i] = std::fabs(d) > 1e-30 ? 1.0 / d : 1.0;
  }
  std::fill(C, C + (size_t)ncond * ncond, 0.0);
  for (int j = 0; j < ncond; ++j) {
    auto b = rhs_unit(s, j);
    auto r = gmres(op, b.data(), &pre, tol, 200, 50);
    if (iters) iters[j] = r.iterations;
    if (resid) resid[j] = r.residual;
    accumulate_c(s, r.x.data(), j, ncond, C, nullptr);
  }
  return 0;
}

// ---------------------------------------------------------------- ground plane
// (No reference counterpart: the reference models the substrate as a meshed
// finite plate, bridge.rs:346-379, cap/mesh.rs:12-43.)  A perfectly
// conducting infinite plane z = zg by the method of images: the Green's
// function 1/|x - y| - 1/|x - y*| with y* = (y.x, y.y, 2 zg - y.z), so a panel
// l contributes I_l(c_k) - I_{l*}(c_k), l* its mirror image (the integral is
// independent of the vertex orientation the mirror reverses).
Pan mirror_panel(const Pan &p, double zg) {
  Pan m = p;
  for (int v = 0; v < 4; ++v) m.v[v].z = 2.0 * zg - p.v[v].z;
  return m;
}

bool above_plane(const PanelSet &s, double zg) {
  for (const Pan &p : s.pans)
    for (int v = 0; v < 4; ++v)
      if (!(p.v[v].z > zg)) return false;
  return true;
}

void assemble_p_image(const PanelSet &s, double zg, double *P) {
  long n = s.n();
  std::vector<Pan> mir(n);
  for (long l = 0; l < n; ++l) mir[l] = mirror_panel(s.pans[l], zg);
#pragma omp parallel for schedule(dynamic, 8)
  for (long k = 0; k < n; ++k) {
    V3 rk = s.c[k];
    for (long l = 0; l < n; ++l)
      P[k * n + l] = ONE_OVER_4PI_EPS0 * (potential_integral(s.pans[l].v, 4, rk) -
                                          potential_integral(mir[l].v, 4, rk));
  }
}

// Dense LU of the image system, one solve per conductor (the conductors'
// Maxwell matrix with the plane as the grounded reference).
int cap_dense_image(const PanelSet &s, int ncond, double zg, double *C) {
  long n = s.n();
  if (n == 0) return -6;
  if (!above_plane(s, zg)) return -7;
  std::vector<double> P((size_t)n * n);
  assemble_p_image(s, zg, P.data());
  std::fill(C, C + (size_t)ncond * ncond, 0.0);
  LU lu;
  if (!lu.factor(P.data(), n)) return -1;
  std::vector<double> sig(n);
  for (int j = 0; j < ncond; ++j) {
    auto b = rhs_unit(s, j);
    lu.solve(b.data(), sig.data());
    accumulate_c(s, sig.data(), j, ncond, C, nullptr);
  }
  return 0;
}

// The FMM form: the mirror panels join the panel set as explicit sources and
// targets (2N panels, mirrors after the originals, same conductor ids); the
// right-hand side of conductor j is +1 on its panels and -1 on their mirrors,
// whose solution is antisymmetric (sigma* = -sigma), so the first N rows are
// the image system.  C_ij = 1/2 sum_{k in i} s_k sigma_k a_k over both halves
// (s = +1 / -1).  Same Jacobi preconditioner and GMRES as cap_fmm.
int cap_fmm_image(const PanelSet &s, int ncond, double zg, int order, int max_level, double tol,
                  int m2l_mode, double *C, int *iters, double *resid) {
  long n = s.n();
  if (n == 0) return -6;
  if (!above_plane(s, zg)) return -7;
  PanelSet d;
  d.pans = s.pans;
  d.cond = s.cond;
  for (long l = 0; l < n; ++l) d.pans.push_back(mirror_panel(s.pans[l], zg));
  d.cond.insert(d.cond.end(), s.cond.begin(), s.cond.end());
  for (const Pan &p : d.pans) {
    d.c.push_back(centroid(p.v, 4));
    d.area.push_back(area(p.v, 4));
  }
  FmmCapOp op;
  op.init(d, order, max_level, m2l_mode);
  DiagJacobi pre;
  pre.inv.assign(2 * n, 1.0);
  for (long i = 0; i < 2 * n; ++i) {
    double dg = 1.0;
    op.diagonal(i, dg);
    pre.inv[i] = std::fabs(dg) > 1e-30 ? 1.0 / dg : 1.0;
  }
  std::fill(C, C + (size_t)ncond * ncond, 0.0);
  for (int j = 0; j < ncond; ++j) {
    auto b = rhs_unit(d, j);
    for (long k = n; k < 2 * n; ++k) b[k] = -b[k];
    auto r = gmres(op, b.data(), &pre, tol, 200, 50);
    if (iters) iters[j] = r.iterations;
    if (resid) resid[j] = r.residual;
    for (long k = 0; k < 2 * n; ++k)
      C[(size_t)d.cond[k] * ncond + j] += 0.5 * (k < n ? 1.0 : -1.0) * r.x[k] * d.area[k];
  }
  return 0;
}

// ---------------------------------------------------------------- meshers
// cap/mesh.rs:12-43
void plate(double cx, double cy, double z0, double lx, double ly, int nx, int ny,
           std::vector<Pan> &out) {
  double x0 = cx - lx / 2.0, y0 = cy - ly / 2.0, dx = lx / nx, dy = ly / ny;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      double xa = x0 + i * dx, xb = xa + dx, ya = y0 + j * dy, yb = ya + dy;
      Pan p;
      p.nv = 4;
      p.v[0] = {xa, ya, z0};
      p.v[1] = {xb, ya, z0};
      p.v[2] = {xb, yb, z0};
      p.v[3] = {xa, yb, z0};
      out.push_back(p);
    }
}

// cap/mesh.rs:50-116
void icosphere(V3 center, double r, int subdivisions, std::vector<Pan> &out) {
  double t = (1.0 + std::sqrt(5.0)) / 2.0;
  std::vector<V3> verts = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                           {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
  for (auto &v : verts) v = normalized(v);
  std::vector<std::array<int, 3>> faces = {
      {0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
      {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
      {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
  for (int s = 0; s < subdivisions; ++s) {
    std::vector<std::array<int, 3>> nf;
    std::map<std::pair<int, int>, int> cache;
    auto mid = [&](int a, int b) {
      auto key = a < b ? std::make_pair(a, b) : std::make_pair(b, a);
      auto it = cache.find(key);
      if (it != cache.end()) return it->second;
      V3 m = normalized((verts[a] + verts[b]) * 0.5);
      int idx = (int)verts.size();
      verts.push_back(m);
      cache[key] = idx;
      return idx;
    };
    for (auto &f : faces) {
      int a = mid(f[0], f[1]);
      int b = mid(f[1], f[2]);
      int c = mid(f[2], f[0]);
      nf.push_back({f[0], a, c});
      nf.push_back({f[1], b, a});
      nf.push_back({f[2], c, b});
      nf.push_back({a, b, c});
    }
    faces = nf;
  }
  for (auto &f : faces) {
    Pan p;
    p.nv = 3;
    p.v[0] = center + verts[f[0]] * r;
    p.v[1] = center + verts[f[1]] * r;
    p.v[2] = center + verts[f[2]] * r;
    p.v[3] = p.v[2];
    out.push_back(p);
  }
}

// bridge.rs:559-563
int subdivisions(double length, double transverse) {
  double v = std::ceil(length / (4.0 * std::max(transverse, 2.2250738585072014e-308)));
  v = std::min(std::max(v, 1.0), 32.0);
  return (int)v;
}

Pan quad(V3 a, V3 b, V3 c, V3 d, bool reverse) {
  Pan p;
  p.nv = 4;
  if (reverse) {
    p.v[0] = d;
    p.v[1] = c;
    p.v[2] = b;
    p.v[3] = a;
  } else {
    p.v[0] = a;
    p.v[1] = b;
    p.v[2] = c;
    p.v[3] = d;
  }
  return p;
}

// bridge.rs:566-596; refinement r multiplies both counts.
void xy_face(std::vector<Pan> &out, double x0, double x1, double y0, double y1, double z,
             bool reverse, int r) {
  int nx = subdivisions(x1 - x0, y1 - y0) * r, ny = subdivisions(y1 - y0, x1 - x0) * r;
  for (int ix = 0; ix < nx; ++ix)
    for (int iy = 0; iy < ny; ++iy) {
      double xa = x0 + (x1 - x0) * ix / nx, xb = x0 + (x1 - x0) * (ix + 1) / nx;
      double ya = y0 + (y1 - y0) * iy / ny, yb = y0 + (y1 - y0) * (iy + 1) / ny;
      out.push_back(quad({xa, ya, z}, {xb, ya, z}, {xb, yb, z}, {xa, yb, z}, reverse));
    }
}
// bridge.rs:599-624; refinement adds r strips along z (r = 1: one strip).
void yz_face(std::vector<Pan> &out, double x, double y0, double y1, double z0, double z1,
             bool reverse, int r) {
  int ny = subdivisions(y1 - y0, z1 - z0) * r;
  for (int iy = 0; iy < ny; ++iy) {
    double ya = y0 + (y1 - y0) * iy / ny, yb = y0 + (y1 - y0) * (iy + 1) / ny;
    for (int iz = 0; iz < r; ++iz) {
      double za = r == 1 ? z0 : z0 + (z1 - z0) * iz / r;
      double zb = r == 1 ? z1 : z0 + (z1 - z0) * (iz + 1) / r;
      out.push_back(quad({x, ya, za}, {x, yb, za}, {x, yb, zb}, {x, ya, zb}, reverse));
    }
  }
}
// bridge.rs:627-652
void xz_face(std::vector<Pan> &out, double x0, double x1, double y, double z0, double z1,
             bool reverse, int r) {
  int nx = subdivisions(x1 - x0, z1 - z0) * r;
  for (int ix = 0; ix < nx; ++ix) {
    double xa = x0 + (x1 - x0) * ix / nx, xb = x0 + (x1 - x0) * (ix + 1) / nx;
    for (int iz = 0; iz < r; ++iz) {
      double za = r == 1 ? z0 : z0 + (z1 - z0) * iz / r;
      double zb = r == 1 ? z1 : z0 + (z1 - z0) * (iz + 1) / r;
      out.push_back(quad({xa, y, za}, {xb, y, za}, {xb, y, zb}, {xa, y, zb}, reverse));
    }
  }
}

struct Group {
  uint32_t net;
  double z0, z1, k;
  std::vector<std::array<int32_t, 4>> rects;
};

const double METRES_PER_NM = 1.0e-9;
const size_t MAX_GRID_CELLS = 262144;  // bridge.rs:28
const size_t MAX_PANELS = 20000;       // bridge.rs:29

// bridge.rs:445-557
int append_group_panels(const Group &g, uint32_t conductor, double xy, int r, size_t max_panels,
                        std::vector<Pan> &pans, std::vector<uint32_t> &cond,
                        std::vector<double> &peps) {
  std::vector<int32_t> xs, ys;
  for (auto &rc : g.rects) {
    xs.push_back(rc[0]);
    xs.push_back(rc[2]);
    ys.push_back(rc[1]);
    ys.push_back(rc[3]);
  }
  std::sort(xs.begin(), xs.end());
  xs.erase(std::unique(xs.begin(), xs.end()), xs.end());
  std::sort(ys.begin(), ys.end());
  ys.erase(std::unique(ys.begin(), ys.end()), ys.end());
  size_t nx = xs.empty() ? 0 : xs.size() - 1, ny = ys.empty() ? 0 : ys.size() - 1;
  size_t cells = nx * ny;
  if (cells > MAX_GRID_CELLS) return -2;
  std::vector<char> filled(cells, 0), consumed(cells, 0);
  for (size_t ix = 0; ix < nx; ++ix)
    for (size_t iy = 0; iy < ny; ++iy) {
      bool any = false;
      for (auto &rc : g.rects)
        if (rc[0] <= xs[ix] && rc[2] >= xs[ix + 1] && rc[1] <= ys[iy] && rc[3] >= ys[iy + 1]) {
          any = true;
          break;
        }
      filled[ix * ny + iy] = any;
    }
  size_t before = pans.size();
  for (size_t ix = 0; ix < nx; ++ix)
    for (size_t iy = 0; iy < ny; ++iy) {
      size_t cell = ix * ny + iy;
      if (!filled[cell] || consumed[cell]) continue;
      size_t x_end = ix + 1;
      while (x_end < nx && filled[x_end * ny + iy] && !consumed[x_end * ny + iy]) ++x_end;
      size_t y_end = iy + 1;
      bool stop = false;
      while (y_end < ny && !stop) {
        for (size_t x = ix; x < x_end; ++x) {
          size_t nx2 = x * ny + y_end;
          if (!filled[nx2] || consumed[nx2]) {
            stop = true;
            break;
          }
        }
        if (!stop) ++y_end;
      }
      for (size_t x = ix; x < x_end; ++x)
        for (size_t y = iy; y < y_end; ++y) consumed[x * ny + y] = 1;
      double x0 = (double)xs[ix] * xy, x1 = (double)xs[x_end] * xy;
      double y0 = (double)ys[iy] * xy, y1 = (double)ys[y_end] * xy;
      xy_face(pans, x0, x1, y0, y1, g.z1, false, r);
      xy_face(pans, x0, x1, y0, y1, g.z0, true, r);
    }
  for (size_t ix = 0; ix < nx; ++ix)
    for (size_t iy = 0; iy < ny; ++iy) {
      if (!filled[ix * ny + iy]) continue;
      double x0 = (double)xs[ix] * xy, x1 = (double)xs[ix + 1] * xy;
      double y0 = (double)ys[iy] * xy, y1 = (double)ys[iy + 1] * xy;
      if (ix == 0 || !filled[(ix - 1) * ny + iy]) yz_face(pans, x0, y0, y1, g.z0, g.z1, true, r);
      if (ix + 1 == nx || !filled[(ix + 1) * ny + iy]) yz_face(pans, x1, y0, y1, g.z0, g.z1, false, r);
      if (iy == 0 || !filled[ix * ny + iy - 1]) xz_face(pans, x0, x1, y0, g.z0, g.z1, false, r);
      if (iy + 1 == ny || !filled[ix * ny + iy + 1]) xz_face(pans, x0, x1, y1, g.z0, g.z1, true, r);
    }
  for (size_t i = before; i < pans.size(); ++i) {
    cond.push_back(conductor);
    peps.push_back(g.k);
  }
  if (pans.size() > max_panels) return -3;
  return 0;
}

struct Meshed {
  std::vector<Pan> pans;
  std::vector<uint32_t> cond;
  std::vector<double> peps;
  std::vector<uint32_t> nets;
};

// bridge.rs:307-386 (panelization half of solve_capacitance)
int mesh_groups(const std::vector<Group> &groups, double xy, int r, size_t max_panels, Meshed &m) {
  if (groups.empty()) return -6;
  std::set<uint32_t> nets;
  for (auto &g : groups) nets.insert(g.net);
  std::map<uint32_t, uint32_t> net_index;
  for (uint32_t nt : nets) {
    net_index[nt] = (uint32_t)m.nets.size();
    m.nets.push_back(nt);
  }
  double b[4] = {INFINITY, INFINITY, -INFINITY, -INFINITY};
  double min_z = INFINITY, max_thick = 0.0;
  for (auto &g : groups) {
    int rc = append_group_panels(g, net_index[g.net], xy, r, max_panels, m.pans, m.cond, m.peps);
    if (rc) return rc;
    for (auto &rt : g.rects) {
      b[0] = std::min(b[0], (double)rt[0] * xy);
      b[1] = std::min(b[1], (double)rt[1] * xy);
      b[2] = std::max(b[2], (double)rt[2] * xy);
      b[3] = std::max(b[3], (double)rt[3] * xy);
    }
    min_z = std::min(min_z, g.z0);
    max_thick = std::max(max_thick, g.z1 - g.z0);
  }
  double gap = min_z > 0.0 ? min_z : std::max(max_thick, 100.0 * METRES_PER_NM);
  double gz = min_z > 0.0 ? 0.0 : min_z - gap;
  double sx = b[2] - b[0], sy = b[3] - b[1];
  double pad = 0.5 * std::max(std::max(sx, sy), 8.0 * gap);
  double lx = sx + 2.0 * pad, ly = sy + 2.0 * pad;
  uint32_t gid = (uint32_t)nets.size();
  // ground_eps: dielectric of the group with the smallest z0 (first on ties, bridge.rs:362-365)
  double geps = groups[0].k, bestz = groups[0].z0;
  for (auto &g : groups)
    if (g.z0 < bestz) {
      bestz = g.z0;
      geps = g.k;
    }
  int gnx = (int)std::min(std::max(std::ceil(lx / (4.0 * gap)), 4.0), 32.0) * r;
  int gny = (int)std::min(std::max(std::ceil(ly / (4.0 * gap)), 4.0), 32.0) * r;
  size_t before = m.pans.size();
  plate(0.5 * (b[0] + b[2]), 0.5 * (b[1] + b[3]), gz, lx, ly, gnx, gny, m.pans);
  for (size_t i = before; i < m.pans.size(); ++i) {
    m.cond.push_back(gid);
    m.peps.push_back(geps);
  }
  if (m.pans.size() > max_panels) return -3;
  return 0;
}

std::vector<Group> load_groups(int ng, const uint32_t *net, const double *h, const double *t,
                               const double *k, const int64_t *off, const int32_t *rects) {
  std::vector<Group> gs(ng);
  for (int i = 0; i < ng; ++i) {
    gs[i].net = net[i];
    gs[i].z0 = h[i] * METRES_PER_NM;                 // bridge.rs:135
    gs[i].z1 = (h[i] + t[i]) * METRES_PER_NM;        // bridge.rs:136
    gs[i].k = k[i];
    for (int64_t q = off[i]; q < off[i + 1]; ++q)
      gs[i].rects.push_back({rects[4 * q], rects[4 * q + 1], rects[4 * q + 2], rects[4 * q + 3]});
  }
  return gs;
}

void store_panels(const std::vector<Pan> &pans, double *verts4) {
  for (size_t i = 0; i < pans.size(); ++i)
    for (int k = 0; k < 4; ++k) {
      verts4[12 * i + 3 * k] = pans[i].v[k].x;
      verts4[12 * i + 3 * k + 1] = pans[i].v[k].y;
      verts4[12 * i + 3 * k + 2] = pans[i].v[k].z;
    }
}

PanelSet set_from(const std::vector<Pan> &pans, const std::vector<uint32_t> &cond) {
  PanelSet s;
  s.pans = pans;
  s.cond = cond;
  s.c.resize(pans.size());
  s.area.resize(pans.size());
  for (size_t i = 0; i < pans.size(); ++i) {
    s.c[i] = centroid(pans[i].v, 4);
    s.area[i] = area(pans[i].v, 4);
  }
  return s;
}

// The dielectric formulation of cap/dielectric.rs:76-151 on the FMM
// operator (no reference counterpart: the reference solves it dense only).
// Near field (the leaf and its neighbours, the near blocks of the conductor
// operator): conductor rows k_e * potential_integral, interface rows
// -k_e * field_integral (l != k) and the jump (eps_out + eps_in) /
// (2 eps0 (eps_out - eps_in)) on the diagonal, as the dense rows.  Far
// field: the potential of the panels' centroid charges area * sigma for
// conductor rows (FmmCapOp), and -k_e n_k . grad of that potential, the
// differentiated local expansion, for interface rows.
struct FmmDielOp : Op {
  Fmm fmm;
  std::vector<double> areas;
  std::vector<V3> gnrm;  // interface rows: oriented normal; conductor rows: 0
  std::vector<long> off;
  std::vector<uint32_t> idx;
  std::vector<double> val;
  long nn = 0;
  mutable std::vector<double> qbuf, gbuf;

  void init(const PanelSet &s, const int *role, const V3 *nrm, const double *diag, int order,
            int max_level, int m2l_mode) {
    nn = s.n();
    areas = s.area;
    gnrm.assign(nn, V3{0.0, 0.0, 0.0});
    for (long i = 0; i < nn; ++i)
      if (role[i] == 1) gnrm[i] = nrm[i];
    fmm.init(s.c, order, max_level, m2l_mode);
    auto blocks = fmm.near_blocks();
    std::vector<std::vector<uint32_t>> rows_i(nn);
    std::vector<std::vector<double>> rows_v(nn);
#pragma omp parallel for schedule(dynamic, 1)
    for (long b = 0; b < (long)blocks.size(); ++b) {
      const auto &tg = blocks[b].first;
      const auto &sr = blocks[b].second;
      for (long i : tg) {
        rows_i[i].resize(sr.size());
        rows_v[i].resize(sr.size());
        for (size_t k = 0; k < sr.size(); ++k) {
          long j = sr[k];
          rows_i[i][k] = (uint32_t)j;
          if (role[i] == 0)
            rows_v[i][k] = ONE_OVER_4PI_EPS0 * potential_integral(s.pans[j].v, 4, s.c[i]);
          else
            rows_v[i][k] = j == i ? diag[i]
                                  : -ONE_OVER_4PI_EPS0 * field_integral(s.pans[j].v, 4, s.c[i], nrm[i]);
        }
      }
    }
    off.assign(nn + 1, 0);
    for (long i = 0; i < nn; ++i) off[i + 1] = off[i] + (long)rows_i[i].size();
    idx.resize(off[nn]);
    val.resize(off[nn]);
    for (long i = 0; i < nn; ++i) {
      std::copy(rows_i[i].begin(), rows_i[i].end(), idx.begin() + off[i]);
      std::copy(rows_v[i].begin(), rows_v[i].end(), val.begin() + off[i]);
    }
    qbuf.assign(nn, 0.0);
    gbuf.assign(nn, 0.0);
  }
  long dim() const override { return nn; }
  void apply(const double *sigma, double *y) const override {
    for (long j = 0; j < nn; ++j) qbuf[j] = areas[j] * sigma[j];
    std::vector<double> far = fmm.run(qbuf.data(), false, gnrm.data(), gbuf.data());
#pragma omp parallel for schedule(static)
    for (long i = 0; i < nn; ++i) {
      const bool iface = gnrm[i].x != 0.0 || gnrm[i].y != 0.0 || gnrm[i].z != 0.0;
      double acc = iface ? -ONE_OVER_4PI_EPS0 * gbuf[i] : ONE_OVER_4PI_EPS0 * far[i];
      for (long k = off[i]; k < off[i + 1]; ++k) acc += val[k] * sigma[idx[k]];
      y[i] = acc;
    }
  }
  double diagonal(long i) const {
    for (long k = off[i]; k < off[i + 1]; ++k)
      if ((long)idx[k] == i) return val[k];
    return 1.0;
  }
};

}  // namespace

// ================================================================ C ABI
// cap/dielectric.rs:47-151: mixed conductor / dielectric-interface rows.
// role[k] = 0 conductor (cond[k], eps[k] = eps_surrounding) or 1 dielectric
// (eps_out[k], eps_in[k], reference point ref[3k..3k+2]).  C: ncond^2 F.
int solve_dielectric(int n, const double *verts4, const int *role, const uint32_t *cond,
                     const double *eps, const double *eps_out, const double *eps_in,
                     const double *ref, int ncond, double *C) {
  if (n == 0) return -6;
  PanelSet s = build_set(n, verts4, cond);
  const double eps0 = EPS0, k_e = ONE_OVER_4PI_EPS0;
  // Panel::normal (panel.rs:50-55), flipped toward the outer side for
  // dielectric panels (oriented_normal, dielectric.rs:39-46)
  std::vector<V3> nrm(n);
  for (int k = 0; k < n; ++k) {
    const V3 *v = s.pans[k].v;
    V3 nk = normalized(cross(v[1] - v[0], v[2] - v[0]));
    if (role[k] == 1) {
      V3 r{ref[3 * k], ref[3 * k + 1], ref[3 * k + 2]};
      if (dot(nk, s.c[k] - r) < 0.0) nk = nk * -1.0;
    }
    nrm[k] = nk;
  }
  std::vector<double> A((size_t)n * n, 0.0);
#pragma omp parallel for schedule(dynamic, 16)
  for (int k = 0; k < n; ++k) {
    double *row = &A[(size_t)k * n];
    if (role[k] == 0) {
      for (int l = 0; l < n; ++l) row[l] = k_e * potential_integral(s.pans[l].v, 4, s.c[k]);
    } else {
      for (int l = 0; l < n; ++l)
        if (l != k) row[l] = -k_e * field_integral(s.pans[l].v, 4, s.c[k], nrm[k]);
      row[k] = (eps_out[k] + eps_in[k]) / (2.0 * eps0 * (eps_out[k] - eps_in[k]));
    }
  }
  LU lu;
  if (!lu.factor(A.data(), n)) return -1;
  std::fill(C, C + (size_t)ncond * ncond, 0.0);
  std::vector<double> b(n), sig(n);
  for (int j = 0; j < ncond; ++j) {
    for (int k = 0; k < n; ++k) b[k] = (role[k] == 0 && (int)cond[k] == j) ? 1.0 : 0.0;
    lu.solve(b.data(), sig.data());
    for (int k = 0; k < n; ++k)
      if (role[k] == 0) C[(size_t)cond[k] * ncond + j] += eps[k] * sig[k] * s.area[k];
  }
  return 0;
}

// solve_dielectric's problem on FmmDielOp: Jacobi-preconditioned GMRES
// (restart 200, 50 restarts) per conductor, charges weighted as the dense
// solve.  Returns 0, or -6 for no panels.
int solve_dielectric_fmm(int n, const double *verts4, const int *role, const uint32_t *cond,
                         const double *eps, const double *eps_out, const double *eps_in,
                         const double *ref, int ncond, int order, int max_level, double tol,
                         int m2l_mode, double *C) {
  if (n == 0) return -6;
  PanelSet s = build_set(n, verts4, cond);
  std::vector<V3> nrm(n);
  std::vector<double> diag(n, 0.0);
  for (int k = 0; k < n; ++k) {
    const V3 *v = s.pans[k].v;
    V3 nk = normalized(cross(v[1] - v[0], v[2] - v[0]));
    if (role[k] == 1) {
      V3 r{ref[3 * k], ref[3 * k + 1], ref[3 * k + 2]};
      if (dot(nk, s.c[k] - r) < 0.0) nk = nk * -1.0;
      diag[k] = (eps_out[k] + eps_in[k]) / (2.0 * EPS0 * (eps_out[k] - eps_in[k]));
    }
    nrm[k] = nk;
  }
  FmmDielOp op;
  op.init(s, role, nrm.data(), diag.data(), order, max_level, m2l_mode);
  DiagJacobi pre;
  pre.inv.assign(n, 1.0);
  for (long i = 0; i < n; ++i) {
    const double d = op.diagonal(i);
    pre.inv[i] = std::fabs(d) > 1e-30 ? 1.0 / d : 1.0;
  }
  std::fill(C, C + (size_t)ncond * ncond, 0.0);
  std::vector<double> b(n);
  for (int j = 0; j < ncond; ++j) {
    for (int k = 0; k < n; ++k) b[k] = (role[k] == 0 && (int)cond[k] == j) ? 1.0 : 0.0;
    auto r = gmres(op, b.data(), &pre, tol, 200, 50);
    for (int k = 0; k < n; ++k)
      if (role[k] == 0) C[(size_t)cond[k] * ncond + j] += eps[k] * r.x[k] * s.area[k];
  }
  return 0;
}

extern "C" {

double or_potential_integral(const double *verts, int nv, const double *p) {
  V3 v[8];
  for (int i = 0; i < nv && i < 8; ++i) v[i] = {verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]};
  return potential_integral(v, nv, {p[0], p[1], p[2]});
}

// panel.rs:213-221: the 7-point table as rows (l0, l1, l2, weight).
void or_tri7(double *tab) {
  for (int q = 0; q < 7; ++q)
    for (int c = 0; c < 4; ++c) tab[q * 4 + c] = TRI7[q][c];
}

// linalg.rs:277-381 (LowRankApprox::compress).  u: m*min(m,n), vt: min(m,n)*n
// (the first k columns / rows written); returns k.
int or_lowrank_compress(const double *a, int m, int n, double eps, double *u, double *vt) {
  LowRank lr;
  lr.compress(a, m, n, eps);
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < lr.k; ++j) u[(size_t)i * lr.k + j] = lr.u[(size_t)i * lr.k + j];
  for (size_t e = 0; e < (size_t)lr.k * n; ++e) vt[e] = lr.vt[e];
  return lr.k;
}

// bbfmm.rs:150 (eps) and 156-166 (G of one offset on the reference box).
double or_m2l_eps(int order) { return std::max(1e-3 * std::pow(0.1, order), 1e-14); }
void or_m2l_matrix(int order, int o0, int o1, int o2, double *G) {
  const int n = order, n3 = n * n * n;
  std::vector<double> nd = cheb_nodes(n);
  for (int m = 0; m < n3; ++m)
    for (int l = 0; l < n3; ++l) {
      double dx = nd[m / (n * n)] - nd[l / (n * n)] - 2.0 * o0;
      double dy = nd[(m / n) % n] - nd[(l / n) % n] - 2.0 * o1;
      double dz = nd[m % n] - nd[l % n] - 2.0 * o2;
      double r = std::sqrt(dx * dx + dy * dy + dz * dz);
      G[(size_t)m * n3 + l] = r > 0.0 ? 1.0 / r : 0.0;
    }
}

double or_field_integral(const double *verts, int nv, const double *p, const double *n_obs) {
  V3 v[8];
  for (int i = 0; i < nv && i < 8; ++i) v[i] = {verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]};
  return field_integral(v, nv, {p[0], p[1], p[2]}, {n_obs[0], n_obs[1], n_obs[2]});
}

void or_panel_props(int n, const double *verts4, double *cx, double *cy, double *cz, double *ar,
                    double *nrm) {
  for (int i = 0; i < n; ++i) {
    Pan p = load_panel(verts4, i);
    V3 c = centroid(p.v, 4);
    cx[i] = c.x;
    cy[i] = c.y;
    cz[i] = c.z;
    ar[i] = area(p.v, 4);
    if (nrm) {
      V3 nn = panel_normal(p.v);
      nrm[3 * i] = nn.x;
      nrm[3 * i + 1] = nn.y;
      nrm[3 * i + 2] = nn.z;
    }
  }
}

void or_assemble_p(int n, const double *verts4, double *P) {
  PanelSet s = build_set(n, verts4, nullptr);
  assemble_p(s, P);
}

int or_lu_solve(int n, const double *A, int nrhs, double *B) {
  LU lu;
  if (!lu.factor(A, n)) return -1;
  std::vector<double> x(n);
  for (int c = 0; c < nrhs; ++c) {
    lu.solve(B + (size_t)c * n, x.data());
    std::copy(x.begin(), x.end(), B + (size_t)c * n);
  }
  return 0;
}

int or_cap_solve_dense(int n, const double *verts4, const uint32_t *cond, int ncond, int method,
                       double *C, double *P_out) {
  PanelSet s = build_set(n, verts4, cond);
  return cap_dense(s, ncond, method, C, P_out, nullptr);
}

int or_cap_solve_fmm(int n, const double *verts4, const uint32_t *cond, int ncond, int order,
                     int max_level, double tol, int m2l_mode, double *C, int *iters_out,
                     double *resid_out) {
  PanelSet s = build_set(n, verts4, cond);
  return cap_fmm(s, ncond, order, max_level, tol, m2l_mode, C, iters_out, resid_out);
}

int or_cap_solve_dense_image(int n, const double *verts4, const uint32_t *cond, int ncond,
                             double zg, double *C) {
  PanelSet s = build_set(n, verts4, cond);
  return cap_dense_image(s, ncond, zg, C);
}

int or_cap_solve_fmm_image(int n, const double *verts4, const uint32_t *cond, int ncond, double zg,
                           int order, int max_level, double tol, int m2l_mode, double *C,
                           int *iters_out, double *resid_out) {
  PanelSet s = build_set(n, verts4, cond);
  return cap_fmm_image(s, ncond, zg, order, max_level, tol, m2l_mode, C, iters_out, resid_out);
}

int or_cap_solve_fmm_cols(int n, const double *verts4, const uint32_t *cond, int ncond,
                          int order, int max_level, double tol, int m2l_mode, int j0, int nj,
                          double *C, int *iters_out, double *seconds_out) {
  // cap/fmm_solver.rs:147-172 for the conductor columns [j0, j0 + nj) only:
  // operator setup once, then one GMRES per column.
  auto t0 = std::chrono::steady_clock::now();
  PanelSet s = build_set(n, verts4, cond);
  FmmCapOp op;
  op.init(s, order, max_level, m2l_mode);
  DiagJacobi pre;
  pre.inv.assign(n, 1.0);
  for (long i = 0; i < n; ++i) {
    double d = 1.0;
    op.diagonal(i, d);
    pre.inv[i] = std::fabs(d) > 1e-30 ? 1.0 / d : 1.0;
  }
  auto t1 = std::chrono::steady_clock::now();
  std::fill(C, C + (size_t)ncond * ncond, 0.0);
  for (int j = j0; j < j0 + nj && j < ncond; ++j) {
    auto b = rhs_unit(s, j);
    auto r = gmres(op, b.data(), &pre, tol, 200, 50);
    if (iters_out) iters_out[j - j0] = r.iterations;
    accumulate_c(s, r.x.data(), j, ncond, C, nullptr);
  }
  auto t2 = std::chrono::steady_clock::now();
  if (seconds_out) {
    seconds_out[0] = std::chrono::duration<double>(t1 - t0).count();
    seconds_out[1] = std::chrono::duration<double>(t2 - t1).count();
  }
  return 0;
}

int or_fmm_cap_apply(int n, const double *verts4, int order, int max_level, int m2l_mode,
                     const double *sigma, double *y) {
  PanelSet s = build_set(n, verts4, nullptr);
  FmmCapOp op;
  op.init(s, order, max_level, m2l_mode);
  op.apply(sigma, y);
  return 0;
}

int or_fmm_matvec(int npts, const double *pts, const double *q, int order, int max_level,
                  int with_p2p, int m2l_mode, double *y) {
  std::vector<V3> p(npts);
  for (int i = 0; i < npts; ++i) p[i] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
  Fmm f;
  f.init(p, order, max_level, m2l_mode);
  auto r = f.run(q, with_p2p != 0);
  std::copy(r.begin(), r.end(), y);
  return 0;
}

void or_dense_p2p(int npts, const double *pts, const double *q, double *y) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < npts; ++i) {
    V3 xi{pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
    double acc = 0.0;
    for (int j = 0; j < npts; ++j)
      if (i != j) acc += q[j] / norm(xi - V3{pts[3 * j], pts[3 * j + 1], pts[3 * j + 2]});
    y[i] = acc;
  }
}

void or_lcg_points(int n, uint64_t seed, double *pts) {
  uint64_t s = seed;
  for (int i = 0; i < 3 * n; ++i) {
    s = s * 6364136223846793005ULL + 1442695040888963407ULL;
    pts[i] = (double)(s >> 33) / (double)(1ULL << 31) - 1.0;
  }
}

void or_cheb_nodes(int n, double *nodes) {
  auto v = cheb_nodes(n);
  std::copy(v.begin(), v.end(), nodes);
}

void or_interp_weights(double x, int n, double *w) {
  auto nd = cheb_nodes(n);
  interp_weights(x, n, nd.data(), w);
}

int or_octree_stats(int npts, const double *pts, int max_level, double *out) {
  std::vector<V3> p(npts);
  for (int i = 0; i < npts; ++i) p[i] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
  Octree t;
  t.build(p, max_level);
  double tot = 0, pairs = 0, near = 0, mind = INFINITY;
  for (long l : t.leaves()) {
    tot += (double)t.nodes[l].points.size();
    near += (double)t.nodes[l].neighbors.size();
  }
  for (auto &nd : t.nodes) {
    pairs += (double)nd.interaction.size();
    for (long c : nd.interaction) {
      const Node &cn = t.nodes[c];
      double dx = nd.center[0] - cn.center[0], dy = nd.center[1] - cn.center[1],
             dz = nd.center[2] - cn.center[2];
      mind = std::min(mind, std::sqrt(dx * dx + dy * dy + dz * dz) / (2.0 * nd.half));
    }
  }
  out[0] = (double)t.nodes.size();
  out[1] = (double)t.leaves().size();
  out[2] = tot;
  out[3] = pairs;
  out[4] = mind;
  out[5] = near;
  return 0;
}

long or_fmm_near_pattern(int n, const double *verts4, int order, int max_level, int64_t *off,
                         uint32_t *idx) {
  PanelSet s = build_set(n, verts4, nullptr);
  FmmCapOp op;
  op.init(s, order, max_level, 1, false);
  if (off)
    for (long i = 0; i <= op.nn; ++i) off[i] = op.off[i];
  if (idx) std::copy(op.idx.begin(), op.idx.end(), idx);
  return op.off[op.nn];
}

int or_mesh_plate(double cx, double cy, double z0, double lx, double ly, int nx, int ny,
                  double *verts4) {
  std::vector<Pan> p;
  plate(cx, cy, z0, lx, ly, nx, ny, p);
  if (verts4) store_panels(p, verts4);
  return (int)p.size();
}

int or_mesh_icosphere(double cx, double cy, double cz, double r, int subdiv, double *verts4) {
  std::vector<Pan> p;
  icosphere({cx, cy, cz}, r, subdiv, p);
  if (verts4) store_panels(p, verts4);
  return (int)p.size();
}

void or_process_stack(int nl, const double *ti, const double *hi, const double *ki, double *t,
                      double *h, double *k) {
  const double HT = 200.0, HS = 100.0;  // process_stack.rs:32-33
  double next_z = 0.0;
  for (int i = 0; i < nl; ++i) {
    double th = ti[i] > 0.0 ? ti[i] : HT;
    double he = (hi[i] > 0.0 || ti[i] > 0.0) ? hi[i] : next_z;
    next_z = he + th + HS;
    t[i] = th;
    h[i] = he;
    k[i] = ki[i] > 0.0 ? ki[i] : 3.9;
  }
}

int or_mesh_groups(int ng, const uint32_t *net, const double *h, const double *t, const double *k,
                   const int64_t *off, const int32_t *rects, double xy, int refine, int max_panels,
                   int *n_panels_out, double *verts4, uint32_t *cond, double *eps) {
  auto gs = load_groups(ng, net, h, t, k, off, rects);
  Meshed m;
  size_t cap = max_panels > 0 ? (size_t)max_panels : MAX_PANELS;
  int rc = mesh_groups(gs, xy, refine < 1 ? 1 : refine, cap, m);
  if (n_panels_out) *n_panels_out = (int)m.pans.size();
  if (rc) return rc;
  if (verts4) store_panels(m.pans, verts4);
  if (cond) std::copy(m.cond.begin(), m.cond.end(), cond);
  if (eps) std::copy(m.peps.begin(), m.peps.end(), eps);
  return 0;
}

int or_solve_capacitance(int ng, const uint32_t *net, const double *h, const double *t,
                         const double *k, const int64_t *off, const int32_t *rects, double xy,
                         int refine, int max_panels, int mode, int order, int max_level, double tol,
                         int *n_nets_out, uint32_t *nets_out, double *cap_af_out, double *C_out,
                         int *n_panels_out) {
  auto gs = load_groups(ng, net, h, t, k, off, rects);
  Meshed m;
  size_t cap = max_panels > 0 ? (size_t)max_panels : MAX_PANELS;
  int rc = mesh_groups(gs, xy, refine < 1 ? 1 : refine, cap, m);
  if (n_panels_out) *n_panels_out = (int)m.pans.size();
  if (rc) return rc;
  int nn = (int)m.nets.size();
  int ncond = nn + 1;
  if (n_nets_out) *n_nets_out = nn;
  PanelSet s = set_from(m.pans, m.cond);
  std::vector<double> C((size_t)ncond * ncond, 0.0);
  double first = m.peps[0];
  bool homog = true;
  for (double e : m.peps)
    if (std::fabs(e - first) > std::max(std::fabs(first), 1.0) * 1.0e-12) homog = false;
  const size_t DENSE_PANEL_LIMIT = 500;  // bridge.rs:30
  if (homog) {
    bool dense = mode == 1 || (mode == 0 && s.n() < (long)DENSE_PANEL_LIMIT);
    int r2 = dense ? cap_dense(s, ncond, 0, C.data(), nullptr, nullptr)
                   : (mode == 0 ? cap_fmm(s, ncond, 4, 4, 1.0e-8, 0, C.data(), nullptr, nullptr)
                                : cap_fmm(s, ncond, order, max_level, tol, 0, C.data(), nullptr,
                                          nullptr));
    if (r2) return r2;
    for (auto &v : C) v *= first;  // bridge.rs:405-407
  } else {
    if (s.n() >= (long)DENSE_PANEL_LIMIT) return -5;  // bridge.rs:410-415
    int r2 = cap_dense(s, ncond, 0, C.data(), nullptr, m.peps.data());
    if (r2) return r2;
  }
  if (C_out) std::copy(C.begin(), C.end(), C_out);
  for (int i = 0; i < nn; ++i) {
    double af = C[(size_t)i * ncond + i] * 1.0e18;  // bridge.rs:432-441
    if (!std::isfinite(af) || af < 0.0) return -4;
    if (nets_out) nets_out[i] = m.nets[i];
    if (cap_af_out) cap_af_out[i] = af;
  }
  return 0;
}

void or_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int or_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

int or_solve_dielectric(int n, const double *verts4, const int *role, const uint32_t *cond,
                        const double *eps, const double *eps_out, const double *eps_in,
                        const double *ref, int ncond, double *C) {
  return solve_dielectric(n, verts4, role, cond, eps, eps_out, eps_in, ref, ncond, C);
}

int or_solve_dielectric_fmm(int n, const double *verts4, const int *role, const uint32_t *cond,
                            const double *eps, const double *eps_out, const double *eps_in,
                            const double *ref, int ncond, int order, int max_level, double tol,
                            int m2l_mode, double *C) {
  return solve_dielectric_fmm(n, verts4, role, cond, eps, eps_out, eps_in, ref, ncond, order,
                              max_level, tol, m2l_mode, C);
}

void or_interp_dweights(double x, int n, double *dw) {
  std::vector<double> nd = cheb_nodes(n);
  interp_dweights(x, n, nd.data(), dw);
}

}  // extern "C"
