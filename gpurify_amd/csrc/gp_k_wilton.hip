// Collocation assembly on the device: the Wilton closed-form panel integral
// in FP64, the dense potential-coefficient matrix, the near-field leaf tiles
// of the precorrected FMM operator and the Jacobi diagonal.
//
// All geometry here is in scaled units (metres divided by the root half-width
// L0, centred on the root cube), so P' = P / (k_e L0) with k_e = 1/(4 pi eps0).

#include "gp_common.h"
#include "gp_kernels.h"

namespace gp {

// Forward-mode dual number: a value and its directional derivatives along
// ND directions (geometric parameters moving the same panels), for the
// sensitivities: every operation carries d/dtheta along, the transcendental
// itself evaluated once for all directions.
template <int ND>
struct DualN {
  double v, d[ND];
  __device__ DualN(double v_ = 0.0) : v(v_) {
#pragma unroll
    for (int i = 0; i < ND; ++i) d[i] = 0.0;
  }
};
// a' = f'(a) a' for every direction
template <int ND>
__device__ inline DualN<ND> chain(double fv, double fp, const DualN<ND> &a) {
  DualN<ND> r(fv);
#pragma unroll
  for (int i = 0; i < ND; ++i) r.d[i] = fp * a.d[i];
  return r;
}
template <int ND>
__device__ inline DualN<ND> operator+(const DualN<ND> &a, const DualN<ND> &b) {
  DualN<ND> r(a.v + b.v);
#pragma unroll
  for (int i = 0; i < ND; ++i) r.d[i] = a.d[i] + b.d[i];
  return r;
}
template <int ND>
__device__ inline DualN<ND> operator-(const DualN<ND> &a, const DualN<ND> &b) {
  DualN<ND> r(a.v - b.v);
#pragma unroll
  for (int i = 0; i < ND; ++i) r.d[i] = a.d[i] - b.d[i];
  return r;
}
template <int ND>
__device__ inline DualN<ND> operator-(const DualN<ND> &a) {
  return chain(-a.v, -1.0, a);
}
template <int ND>
__device__ inline DualN<ND> operator*(const DualN<ND> &a, const DualN<ND> &b) {
  DualN<ND> r(a.v * b.v);
#pragma unroll
  for (int i = 0; i < ND; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
template <int ND>
__device__ inline DualN<ND> operator/(const DualN<ND> &a, const DualN<ND> &b) {
  const double q = a.v / b.v, ib = 1.0 / b.v;
  DualN<ND> r(q);
#pragma unroll
  for (int i = 0; i < ND; ++i) r.d[i] = (a.d[i] - q * b.d[i]) * ib;
  return r;
}
template <int ND>
__device__ inline DualN<ND> &operator+=(DualN<ND> &a, const DualN<ND> &b) { return a = a + b; }
template <int ND>
__device__ inline DualN<ND> &operator/=(DualN<ND> &a, const DualN<ND> &b) { return a = a / b; }
template <int ND>
__device__ inline DualN<ND> sqrt(const DualN<ND> &a) {
  const double r = ::sqrt(a.v);
  return chain(r, r > 0.0 ? 0.5 / r : 0.0, a);
}
template <int ND>
__device__ inline DualN<ND> log(const DualN<ND> &a) { return chain(::log(a.v), 1.0 / a.v, a); }
template <int ND>
__device__ inline DualN<ND> atan(const DualN<ND> &a) {
  return chain(::atan(a.v), 1.0 / (1.0 + a.v * a.v), a);
}
template <int ND>
__device__ inline DualN<ND> asinh(const DualN<ND> &a) {
  return chain(::asinh(a.v), 1.0 / ::sqrt(1.0 + a.v * a.v), a);
}
// |x| with derivative 0 at x = 0 (the central difference there).
template <int ND>
__device__ inline DualN<ND> fabs(const DualN<ND> &a) {
  return chain(::fabs(a.v), a.v > 0.0 ? 1.0 : (a.v < 0.0 ? -1.0 : 0.0), a);
}
template <int ND>
__device__ inline double val(const DualN<ND> &x) { return x.v; }
__device__ inline double val(double x) { return x; }
// The double forms, visible next to the Dual ones inside this namespace.
__device__ inline double sqrt(double x) { return ::sqrt(x); }
__device__ inline double log(double x) { return ::log(x); }
__device__ inline double atan(double x) { return ::atan(x); }
__device__ inline double asinh(double x) { return ::asinh(x); }
__device__ inline double fabs(double x) { return ::fabs(x); }

// Closed-form I(p) = integral over a flat quadrilateral of 1/|p - r'| da'
// (Wilton et al. 1984), following crates/pex/src/quasistatic/integrals/
// panel.rs:147-208 term by term: Newell normal (panel.rs:256-269), per edge
// the in-plane distance, the log term with its asinh fallback
// (panel.rs:271-285) and the solid-angle atan term.  v: 12 values.  T is
// double, or Dual for the directional derivative (branches follow values).
template <typename T>
__device__ T wilton_t(const T *v, T px, T py, T pz) {
  T vx[4], vy[4], vz[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    vx[k] = v[3 * k];
    vy[k] = v[3 * k + 1];
    vz[k] = v[3 * k + 2];
  }
  T nx = 0.0, ny = 0.0, nz = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    int j = (i + 1) & 3;
    nx += (vy[i] - vy[j]) * (vz[i] + vz[j]);
    ny += (vz[i] - vz[j]) * (vx[i] + vx[j]);
    nz += (vx[i] - vx[j]) * (vy[i] + vy[j]);
  }
  T nn = sqrt(nx * nx + ny * ny + nz * nz);
  if (val(nn) == 0.0) {
    nx = ny = nz = 0.0;
  } else {
    nx /= nn;
    ny /= nn;
    nz /= nn;
  }
  T d = (px - vx[0]) * nx + (py - vy[0]) * ny + (pz - vz[0]) * nz;
  T p0x = px - nx * d, p0y = py - ny * d, p0z = pz - nz * d;
  T absd = fabs(d);
  T total = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    int j = (i + 1) & 3;
    T ex = vx[j] - vx[i], ey = vy[j] - vy[i], ez = vz[j] - vz[i];
    T el = sqrt(ex * ex + ey * ey + ez * ez);
    if (val(el) == 0.0) continue;
    T lx = ex / el, ly = ey / el, lz = ez / el;
    T ux = ly * nz - lz * ny, uy = lz * nx - lx * nz, uz = lx * ny - ly * nx;
    T ax = vx[i] - p0x, ay = vy[i] - p0y, az = vz[i] - p0z;
    T bx = vx[j] - p0x, by = vy[j] - p0y, bz = vz[j] - p0z;
    T perp = ax * ux + ay * uy + az * uz;
    T sm = ax * lx + ay * ly + az * lz;
    T sp = bx * lx + by * ly + bz * lz;
    T r0sq = perp * perp + d * d;
    T rp = sqrt(sp * sp + r0sq);
    T rm = sqrt(sm * sm + r0sq);
    T num = rp + sp, den = rm + sm;
    T logt;
    if (val(num) > 0.0 && val(den) > 0.0) {
      logt = perp * log(num / den);
    } else if (::fabs(val(den)) > 0.0) {
      T r0 = sqrt(r0sq);
      logt = val(r0) == 0.0 ? T(0.0) : perp * (asinh(sp / r0) - asinh(sm / r0));
    } else {
      logt = 0.0;
    }
    T att = 0.0;
    // perp == 0 gives att == 0 (the reference skips it); its derivative in
    // perp is not 0 there, so the Dual form always evaluates the term.
    constexpr bool dual = !std::is_same<T, double>::value;
    if (val(absd) > 0.0 && (val(perp) != 0.0 || dual)) {
      T tp = (perp * sp) / (r0sq + absd * rp);
      T tm = (perp * sm) / (r0sq + absd * rm);
      att = absd * (atan(tp) - atan(tm));
    }
    total += logt - att;
  }
  return total;
}

__device__ double wilton(const double *__restrict__ v, double px, double py, double pz) {
  return wilton_t<double>(v, px, py, pz);
}

// Fan-triangulated area and area-weighted centroid of a panel (panel.rs:29-67),
// with their directional derivatives.
template <int ND>
__device__ void panel_props_dual(const DualN<ND> *v, DualN<ND> c[3], DualN<ND> &area) {
  using D = DualN<ND>;
  D asum(0.0), cx(0.0), cy(0.0), cz(0.0);
  for (int k = 1; k < 3; ++k) {
    D ax = v[3 * k] - v[0], ay = v[3 * k + 1] - v[1], az = v[3 * k + 2] - v[2];
    D bx = v[3 * k + 3] - v[0], by = v[3 * k + 4] - v[1], bz = v[3 * k + 5] - v[2];
    D qx = ay * bz - az * by, qy = az * bx - ax * bz, qz = ax * by - ay * bx;
    D a = sqrt(qx * qx + qy * qy + qz * qz) * D(0.5);
    D third(1.0 / 3.0);
    cx += (v[0] + v[3 * k] + v[3 * k + 3]) * third * a;
    cy += (v[1] + v[3 * k + 1] + v[3 * k + 4]) * third * a;
    cz += (v[2] + v[3 * k + 2] + v[3 * k + 5]) * third * a;
    asum += a;
  }
  c[0] = cx / asum;
  c[1] = cy / asum;
  c[2] = cz / asum;
  area = asum;
}

// 2 x 2 Gauss rule of a flat panel through its bilinear map from [-1, 1]^2:
// points xg and weights wg (|dx/da x dx/db| at the point), so that
// sum_g wg / |p - xg| integrates 1/|p - r'| with a relative error below
// 4e-6 once |p - centroid| exceeds 6 panel diagonals (checked against the
// closed form); used for the far entries of the sensitivities.
template <typename T>
__device__ void gauss4_t(const T *v, T xg[4][3], T wg[4]) {
  const double g = 0.57735026918962576451;  // 1 / sqrt(3)
  const double ga[4] = {-g, g, g, -g}, gb[4] = {-g, -g, g, g};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const double a = ga[q], b = gb[q];
    const double N[4] = {0.25 * (1 - a) * (1 - b), 0.25 * (1 + a) * (1 - b),
                         0.25 * (1 + a) * (1 + b), 0.25 * (1 - a) * (1 + b)};
    const double Na[4] = {-0.25 * (1 - b), 0.25 * (1 - b), 0.25 * (1 + b), -0.25 * (1 + b)};
    const double Nb[4] = {-0.25 * (1 - a), -0.25 * (1 + a), 0.25 * (1 + a), 0.25 * (1 - a)};
    T xa[3], xb[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      T x = 0.0, ta = 0.0, tb = 0.0;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        x += T(N[i]) * v[3 * i + d];
        ta += T(Na[i]) * v[3 * i + d];
        tb += T(Nb[i]) * v[3 * i + d];
      }
      xg[q][d] = x;
      xa[d] = ta;
      xb[d] = tb;
    }
    T cx = xa[1] * xb[2] - xa[2] * xb[1], cy = xa[2] * xb[0] - xa[0] * xb[2],
      cz = xa[0] * xb[1] - xa[1] * xb[0];
    wg[q] = sqrt(cx * cx + cy * cy + cz * cz);
  }
}

// Integral over the mirror image of panel v in the plane z = zm2 / 2 (its
// vertices' z reflected to zm2 - z; the integral does not depend on the
// vertex orientation the reflection reverses).
__device__ double wilton_image(const double *__restrict__ v, double zm2, double px, double py,
                               double pz) {
  double q[12];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    q[3 * k] = v[3 * k];
    q[3 * k + 1] = v[3 * k + 1];
    q[3 * k + 2] = zm2 - v[3 * k + 2];
  }
  return wilton(q, px, py, pz);
}

// P'[k][l] = I_l(c_k), cap/solver.rs:81-97 (row k = collocation at centroid
// k); with a ground plane (image != 0, the plane z = zm2 / 2 in the unit
// frame) P'[k][l] = I_l(c_k) - I_{l*}(c_k), l* the mirror image of panel l.
__global__ __launch_bounds__(256) void k_dense_assemble(int n, const double *__restrict__ verts,
                                                        const double *__restrict__ cen,
                                                        double *__restrict__ P, int image,
                                                        double zm2) {
  long total = (long)n * n;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    int k = (int)(e / n), l = (int)(e % n);
    const double *v = verts + 12L * l;
    const double px = cen[3 * k], py = cen[3 * k + 1], pz = cen[3 * k + 2];
    double pv = wilton(v, px, py, pz);
    if (image) pv -= wilton_image(v, zm2, px, py, pz);
    P[e] = pv;
  }
}

// Dielectric-interface rows of the mixed system (cap/dielectric.rs:91-101),
// in the unit frame of the conductor rows: row k = drow[i] of the dense
// matrix becomes -F(l, c_k, n_k) for l != k and diag[i] on the diagonal,
// F being the normal-field integral over panel l (integrals/panel.rs:
// 232-253, scale-free).  F: the degree-5 seven-point symmetric triangle rule
// (barycentric points and weights of the published rule) on the fan
// triangles of the quad; a triangle's repeated last vertex adds a
// zero-area fan triangle, which is skipped.  One work-group per row.
__constant__ double c_tri7[7][4] = {
    {1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0, 0.225},
    {0.059715871789770, 0.470142064105115, 0.470142064105115, 0.132394152788506},
    {0.470142064105115, 0.059715871789770, 0.470142064105115, 0.132394152788506},
    {0.470142064105115, 0.470142064105115, 0.059715871789770, 0.132394152788506},
    {0.797426985353087, 0.101286507323456, 0.101286507323456, 0.125939180544827},
    {0.101286507323456, 0.797426985353087, 0.101286507323456, 0.125939180544827},
    {0.101286507323456, 0.101286507323456, 0.797426985353087, 0.125939180544827}};

__device__ double normal_field(const double *__restrict__ q, double px, double py, double pz,
                               double nx, double ny, double nz) {
  double acc = 0.0;
  for (int t = 1; t <= 2; ++t) {
    const double *a = q, *b = q + 3 * t, *c = q + 3 * t + 3;
    const double ux = b[0] - a[0], uy = b[1] - a[1], uz = b[2] - a[2];
    const double vx = c[0] - a[0], vy = c[1] - a[1], vz = c[2] - a[2];
    const double cx = uy * vz - uz * vy, cy = uz * vx - ux * vz, cz = ux * vy - uy * vx;
    const double ar = sqrt(cx * cx + cy * cy + cz * cz) * 0.5;
    if (ar == 0.0) continue;
    for (int g = 0; g < 7; ++g) {
      const double b0 = c_tri7[g][0], b1 = c_tri7[g][1], b2 = c_tri7[g][2];
      const double dx = a[0] * b0 + b[0] * b1 + c[0] * b2 - px;
      const double dy = a[1] * b0 + b[1] * b1 + c[1] * b2 - py;
      const double dz = a[2] * b0 + b[2] * b1 + c[2] * b2 - pz;
      const double r2 = dx * dx + dy * dy + dz * dz;
      if (r2 > 0.0) acc += c_tri7[g][3] * ar * (dx * nx + dy * ny + dz * nz) / (r2 * sqrt(r2));
    }
  }
  return acc;
}

__global__ __launch_bounds__(256) void k_dielectric_rows(int n, const int *__restrict__ drow,
                                                         const double *__restrict__ nrm,
                                                         const double *__restrict__ diag,
                                                         const double *__restrict__ verts,
                                                         const double *__restrict__ cen,
                                                         double *__restrict__ P) {
  const int i = blockIdx.x, k = drow[i];
  const double px = cen[3 * k], py = cen[3 * k + 1], pz = cen[3 * k + 2];
  const double nx = nrm[3 * i], ny = nrm[3 * i + 1], nz = nrm[3 * i + 2];
  double *row = P + (size_t)k * n;
  for (int l = threadIdx.x; l < n; l += blockDim.x)
    row[l] = l == k ? diag[i] : -normal_field(verts + 12L * l, px, py, pz, nx, ny, nz);
}

// Near-field tile of target leaf T: rows = T's panels, columns = the panels of
// T and its neighbour leaves (cap/fmm_solver.rs:52-91), row-major at nf_off[T].
// Writes the FP64 values and, when v32 != nullptr, their FP32 rounding.
// The plane frame of a source panel for the near-field tiles: the Newell
// normal (panel.rs:256-269), e1 along the first non-degenerate edge,
// e2 = n x e1, the vertices' in-plane coordinates (a_k, b_k) from vertex 0,
// and whether the panel is flat (every vertex within rounding of the plane,
// non-zero normal).  Both near-field assembly kernels call it, so they agree
// on which panels each one takes.
__device__ __forceinline__ bool panel_plane(const double *__restrict__ v, double n[3],
                                            double e1[3], double e2[3], double av[4],
                                            double bv[4]) {
  double nx = 0.0, ny = 0.0, nz = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int j = (i + 1) & 3;
    nx += (v[3 * i + 1] - v[3 * j + 1]) * (v[3 * i + 2] + v[3 * j + 2]);
    ny += (v[3 * i + 2] - v[3 * j + 2]) * (v[3 * i] + v[3 * j]);
    nz += (v[3 * i] - v[3 * j]) * (v[3 * i + 1] + v[3 * j + 1]);
  }
  const double nn = sqrt(nx * nx + ny * ny + nz * nz);
  double e1x = 0.0, e1y = 0.0, e1z = 0.0, emax = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int j = (i + 1) & 3;
    const double ex = v[3 * j] - v[3 * i], ey = v[3 * j + 1] - v[3 * i + 1],
                 ez = v[3 * j + 2] - v[3 * i + 2];
    const double el = sqrt(ex * ex + ey * ey + ez * ez);
    if (el > 0.0 && emax == 0.0) e1x = ex / el, e1y = ey / el, e1z = ez / el;
    emax = fmax(emax, el);
  }
  bool flat = nn > 0.0 && emax > 0.0;
  if (flat) nx /= nn, ny /= nn, nz /= nn;
  n[0] = nx, n[1] = ny, n[2] = nz;
  e1[0] = e1x, e1[1] = e1y, e1[2] = e1z;
  e2[0] = ny * e1z - nz * e1y, e2[1] = nz * e1x - nx * e1z, e2[2] = nx * e1y - ny * e1x;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double qx = v[3 * k] - v[0], qy = v[3 * k + 1] - v[1], qz = v[3 * k + 2] - v[2];
    av[k] = qx * e1[0] + qy * e1[1] + qz * e1[2];
    bv[k] = qx * e2[0] + qy * e2[1] + qz * e2[2];
    flat = flat && fabs(qx * nx + qy * ny + qz * nz) <= 1e-14 * emax;
  }
  return flat;
}

// The arctangent and logarithm of k_near_assemble's edge loop as calls, not
// inlined: 94 VGPRs (five waves a SIMD, as many as its LDS allows); with
// both inlined the kernel takes 148 (three waves a SIMD).
__device__ __noinline__ double atan_call(double x) { return ::atan(x); }
__device__ __noinline__ double log_call(double x) { return ::log(x); }

// Near-field tiles, one work-group per target leaf: each thread owns
// columns of the leaf's tile (source panels), builds a flat source panel's
// plane frame once (panel_plane) and evaluates the closed-form integral of
// wilton() (panel.rs:147-208) for every target row of the leaf, so the
// per-panel half of wilton() (the Newell normal, edge lengths and
// directions: five square roots and fifteen divisions) is paid once per
// (leaf, column), not once per entry.  In the plane, edge i has the unit
// direction (la_i, lb_i) and the in-plane normal u = l x n = (lb_i, -la_i); a
// target at in-plane (a, b) and height d gives perp = (a_i - a) lb_i -
// (b_i - b) la_i, s- = (a_i - a) la_i + (b_i - b) lb_i and s+ likewise from
// vertex j: the projections of wilton() on the plane, up to rounding; then the
// log term with its asinh fallback and the solid-angle term.  The panel's
// in-plane data sit in this thread's LDS slot, so the edge loop runs rolled
// (fewer registers, four waves a SIMD).  Columns of panels that are not flat
// are left to k_near_assemble_skew.  A leaf narrower than the work-group
// splits its rows over groups of S threads.
__global__ __launch_bounds__(256) void k_near_assemble(
    const int *__restrict__ leaf_lo, const int64_t *__restrict__ nf_off,
    const int *__restrict__ col_off, const int *__restrict__ cols,
    const double *__restrict__ verts, const double *__restrict__ cen, double *__restrict__ v64,
    float *__restrict__ v32) {
  __shared__ double sgeo[256 * 16];
  const int T = blockIdx.x;
  const int t0 = leaf_lo[T], nt = leaf_lo[T + 1] - t0;
  const int c0 = col_off[T], S = col_off[T + 1] - c0;
  const int64_t base = nf_off[T];
  const int tid = threadIdx.x, nth = blockDim.x;  // nth == 256 (the LDS slots)
  if (S == 0) return;
  int G, g, s0, sstep;
  if (S >= nth) {
    G = 1, g = 0, s0 = tid, sstep = nth;
  } else {
    G = nth / S, g = tid / S, s0 = tid % S, sstep = S;
    if (g >= G) return;
  }
  double *st = sgeo + (size_t)tid * 16;  // a_k, b_k, la_i, lb_i
  for (int s = s0; s < S; s += sstep) {
    const double *v = verts + 12L * cols[c0 + s];
    double n[3], e1[3], e2[3], av[4], bv[4];
    if (!panel_plane(v, n, e1, e2, av, bv)) continue;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int j = (i + 1) & 3;
      const double ea = av[j] - av[i], eb = bv[j] - bv[i];
      const double el = sqrt(ea * ea + eb * eb);
      // a zero-length edge: l = 0 makes perp, s- and s+ zero and its terms vanish
      const double inv = el > 0.0 ? 1.0 / el : 0.0;
      st[i] = av[i];
      st[4 + i] = bv[i];
      st[8 + i] = ea * inv;
      st[12 + i] = eb * inv;
    }
    const double v0x = v[0], v0y = v[1], v0z = v[2];
    for (int t = g; t < nt; t += G) {
      const int it = t0 + t;
      const double qx = cen[3 * it] - v0x, qy = cen[3 * it + 1] - v0y, qz = cen[3 * it + 2] - v0z;
      const double d = qx * n[0] + qy * n[1] + qz * n[2];
      const double a = qx * e1[0] + qy * e1[1] + qz * e1[2];
      const double b = qx * e2[0] + qy * e2[1] + qz * e2[2];
      const double absd = fabs(d), dd = d * d;
      double total = 0.0;
#pragma unroll 1
      for (int i = 0; i < 4; ++i) {
        const int j = (i + 1) & 3;
        const double lai = st[8 + i], lbi = st[12 + i];
        const double ai = st[i] - a, bi = st[4 + i] - b, aj = st[j] - a, bj = st[4 + j] - b;
        const double perp = ai * lbi - bi * lai;
        const double sm = ai * lai + bi * lbi;
        const double sp = aj * lai + bj * lbi;
        const double r0sq = perp * perp + dd;
        const double rp = sqrt(sp * sp + r0sq);
        const double rm = sqrt(sm * sm + r0sq);
        const double num = rp + sp, den = rm + sm;
        double logt;
        if (num > 0.0 && den > 0.0) {
          logt = perp * log_call(num / den);
        } else if (fabs(den) > 0.0) {
          const double r0 = sqrt(r0sq);
          logt = r0 == 0.0 ? 0.0 : perp * (asinh(sp / r0) - asinh(sm / r0));
        } else {
          logt = 0.0;
        }
        double att = 0.0;
        if (absd > 0.0 && perp != 0.0) {
          // atan(perp sp / Q) - atan(perp sm / W) with Q, W > 0, as one angle:
          // atan2(perp (sp W - sm Q), Q W + perp^2 sp sm) (one arctangent
          // and one division instead of two of each; the same value up to
          // rounding, the angle's error a few ulps of pi either way)
          const double Q = r0sq + absd * rp, W = r0sq + absd * rm;
          const double ay = perp * (sp * W - sm * Q), ax = Q * W + perp * perp * sp * sm;
          double ang = atan_call(ay / ax);  // ax == +0: +-pi/2 from ay / ax = +-inf
          if (ax < 0.0) ang += copysign(3.141592653589793, ay);
          att = absd * ang;
        }
        total += logt - att;
      }
      const int64_t e = base + (int64_t)t * S + s;
      v64[e] = total;
      if (v32) v32[e] = (float)total;
    }
  }
}

// The columns k_near_assemble leaves: source panels that are not flat (a
// skew quadrilateral of a panel file), wilton() per entry.  Same thread
// mapping; a flat column costs one panel_plane() call.
__global__ __launch_bounds__(256) void k_near_assemble_skew(
    const int *__restrict__ leaf_lo, const int64_t *__restrict__ nf_off,
    const int *__restrict__ col_off, const int *__restrict__ cols,
    const double *__restrict__ verts, const double *__restrict__ cen, double *__restrict__ v64,
    float *__restrict__ v32) {
  const int T = blockIdx.x;
  const int t0 = leaf_lo[T], nt = leaf_lo[T + 1] - t0;
  const int c0 = col_off[T], S = col_off[T + 1] - c0;
  const int64_t base = nf_off[T];
  const int tid = threadIdx.x, nth = blockDim.x;
  if (S == 0) return;
  int G, g, s0, sstep;
  if (S >= nth) {
    G = 1, g = 0, s0 = tid, sstep = nth;
  } else {
    G = nth / S, g = tid / S, s0 = tid % S, sstep = S;
    if (g >= G) return;
  }
  for (int s = s0; s < S; s += sstep) {
    const double *v = verts + 12L * cols[c0 + s];
    double n[3], e1[3], e2[3], av[4], bv[4];
    if (panel_plane(v, n, e1, e2, av, bv)) continue;
    for (int t = g; t < nt; t += G) {
      const int it = t0 + t;
      const double w = wilton(v, cen[3 * it], cen[3 * it + 1], cen[3 * it + 2]);
      const int64_t e = base + (int64_t)t * S + s;
      v64[e] = w;
      if (v32) v32[e] = (float)w;
    }
  }
}

// Dielectric-interface rows of the near-field tiles (the rows of
// cap/dielectric.rs:107-121 restricted to the near leaves): entry (i, l) of
// an interface panel i becomes -F(l, c_i, n_i), the 7-point normal-field
// integral, and the diagonal the jump term diag[i] (unit frame: 2 pi
// (eps_out + eps_in) / (eps_out - eps_in)); conductor rows keep their
// potential coefficients.  Panels in device order.
__global__ __launch_bounds__(256) void k_near_dielectric_rows(
    const int *__restrict__ leaf_lo, const int64_t *__restrict__ nf_off,
    const int *__restrict__ col_off, const int *__restrict__ cols,
    const unsigned char *__restrict__ iface, const double *__restrict__ nrm,
    const double *__restrict__ diag, const double *__restrict__ verts,
    const double *__restrict__ cen, double *__restrict__ v64) {
  const int T = blockIdx.x;
  const int t0 = leaf_lo[T], nt = leaf_lo[T + 1] - t0;
  const int c0 = col_off[T], S = col_off[T + 1] - c0;
  const long cnt = (long)nt * S;
  const int64_t base = nf_off[T];
  for (long e = threadIdx.x; e < cnt; e += blockDim.x) {
    const int t = (int)(e / S), s = (int)(e % S);
    const int i = t0 + t;
    if (!iface[i]) continue;
    const int j = cols[c0 + s];
    v64[base + e] = j == i ? diag[i]
                           : -normal_field(verts + 12L * j, cen[3 * i], cen[3 * i + 1], cen[3 * i + 2],
                                           nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
  }
}

// Jacobi preconditioner of the interface rows: 1 / (their jump term).
__global__ void k_dielectric_inv_diag(int n, const unsigned char *__restrict__ iface,
                                      const double *__restrict__ diag, double *__restrict__ inv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && iface[i]) inv[i] = 1.0 / diag[i];
}

void launch_near_dielectric_rows(int n_leaves, int n, const int *leaf_lo, const int64_t *nf_off,
                                 const int *col_off, const int *cols, const unsigned char *iface,
                                 const double *nrm, const double *diag, const double *verts,
                                 const double *cen, double *v64, double *inv_diag, hipStream_t s) {
  if (n_leaves == 0) return;
  ProfScope ps("near_dielectric_rows", s, 0.0, 0.0);
  hipLaunchKernelGGL(k_near_dielectric_rows, dim3(n_leaves), dim3(256), 0, s, leaf_lo, nf_off,
                     col_off, cols, iface, nrm, diag, verts, cen, v64);
  GP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_dielectric_inv_diag, dim3((n + 255) / 256), dim3(256), 0, s, n, iface, diag,
                     inv_diag);
  GP_LAUNCH_CHECK();
}

// Self term P'_ii for the Jacobi preconditioner (cap/fmm_solver.rs:150-162).
// inv = 1/d unless |k_e L0 d| <= 1e-30, the reference's threshold in SI units;
// panel i belongs to problem p with off[p] <= i < off[p + 1] (off[0] = 0,
// off[np] = n), whose threshold scale k_e L0 is si_scale[p].
__global__ void k_self_inv_diag(int n, int np, const int *__restrict__ off,
                                const double *__restrict__ si_scale,
                                const double *__restrict__ verts,
                                const double *__restrict__ cen, double *__restrict__ inv) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int lo = 0, hi = np;  // off[lo] <= i < off[hi]
  while (hi - lo > 1) {
    const int m = (lo + hi) >> 1;
    if (off[m] <= i) lo = m;
    else hi = m;
  }
  double d = wilton(verts + 12L * i, cen[3 * i], cen[3 * i + 1], cen[3 * i + 2]);
  inv[i] = fabs(d * si_scale[lo]) > 1e-30 ? 1.0 / d : 1.0;
}

// P'_p[k][l] = I_l(c_k) for every problem p of a batch (blockIdx.y), its
// panels at verts + 12 n_off[p], cen + 3 n_off[p], matrix at P + a_off[p].
__global__ __launch_bounds__(256) void k_dense_assemble_batched(const int *__restrict__ n,
                                                                const int *__restrict__ n_off,
                                                                const long *__restrict__ a_off,
                                                                const double *__restrict__ verts,
                                                                const double *__restrict__ cen,
                                                                double *__restrict__ P,
                                                                const double *__restrict__ zm2) {
  const int p = blockIdx.y, np = n[p];
  const double *v = verts + 12L * n_off[p], *c = cen + 3L * n_off[p];
  double *Pp = P + a_off[p];
  const long total = (long)np * np;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    const int k = (int)(e / np), l = (int)(e % np);
    const double px = c[3 * k], py = c[3 * k + 1], pz = c[3 * k + 2];
    double pv = wilton(v + 12L * l, px, py, pz);
    if (zm2) pv -= wilton_image(v + 12L * l, zm2[p], px, py, pz);
    Pp[e] = pv;
  }
}


// AT = A^T (n x n, row-major), 32 x 32 tiles through LDS.
__global__ __launch_bounds__(256) void k_transpose(int n, const double *__restrict__ A,
                                                   double *__restrict__ AT) {
  __shared__ double t[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int y = ty; y < 32; y += 8)
    if (by + y < n && bx + tx < n) t[y][tx] = A[(long)(by + y) * n + bx + tx];
  __syncthreads();
  for (int y = ty; y < 32; y += 8)
    if (bx + y < n && by + tx < n) AT[(long)(bx + y) * n + by + tx] = t[tx][y];
}

void launch_transpose(int n, const double *A, double *AT, hipStream_t s) {
  dim3 grid((n + 31) / 32, (n + 31) / 32);
  hipLaunchKernelGGL(k_transpose, grid, dim3(256), 0, s, n, A, AT);
  GP_LAUNCH_CHECK();
}

void launch_dense_assemble_batched(const DenseBatch &b, hipStream_t s) {
  ProfScope ps("dense_assemble", s, 0.0, 0.0);
  hipLaunchKernelGGL(k_dense_assemble_batched, dim3(64, b.count), dim3(256), 0, s, b.n, b.n_off,
                     b.a_off, b.verts, b.cen, b.A, b.zm2);
  GP_LAUNCH_CHECK();
}

void launch_dense_assemble(int n, const double *verts, const double *cen, double *P,
                           hipStream_t s, bool image, double zm2) {
  long total = (long)n * n;
  int blocks = (int)std::min<long>((total + 255) / 256, 65536);
  ProfScope ps("dense_assemble", s, 0.0, 8.0 * total);
  hipLaunchKernelGGL(k_dense_assemble, dim3(blocks), dim3(256), 0, s, n, verts, cen, P,
                     image ? 1 : 0, zm2);
  GP_LAUNCH_CHECK();
}

void launch_dielectric_rows(int n, int nd, const int *drow, const double *nrm, const double *diag,
                            const double *verts, const double *cen, double *P, hipStream_t s) {
  if (nd == 0) return;
  ProfScope ps("dielectric_rows", s, 0.0, 8.0 * (double)nd * n);
  hipLaunchKernelGGL(k_dielectric_rows, dim3(nd), dim3(256), 0, s, n, drow, nrm, diag, verts, cen,
                     P);
  GP_LAUNCH_CHECK();
}

void launch_near_assemble(int n_leaves, const int *leaf_lo, const int64_t *nf_off,
                          const int *col_off, const int *cols, const double *verts,
                          const double *cen, double *v64, float *v32, hipStream_t s) {
  if (n_leaves == 0) return;
  ProfScope ps("near_assemble", s, 0.0, 0.0);
  hipLaunchKernelGGL(k_near_assemble, dim3(n_leaves), dim3(256), 0, s, leaf_lo, nf_off, col_off,
                     cols, verts, cen, v64, v32);
  GP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_near_assemble_skew, dim3(n_leaves), dim3(256), 0, s, leaf_lo, nf_off,
                     col_off, cols, verts, cen, v64, v32);
  GP_LAUNCH_CHECK();
}

void launch_self_inv_diag(int n, int np, const int *off, const double *si_scale,
                          const double *verts, const double *cen, double *inv, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_self_inv_diag, dim3((n + 255) / 256), dim3(256), 0, s, n, np, off,
                     si_scale, verts, cen, inv);
  GP_LAUNCH_CHECK();
}

// ---------------------------------------------------------------- sensitivities
// A group of ND geometric parameters moves the vertices of the same panels
// moved[0..nm) (device order); direction t has velocities dv + t * dstride
// (12 per moved panel: d vertex / d theta in scaled units per scaled theta).
// mv_of[k] is k's index in moved or -1.  With P' sigma_j = e_j and
// P'^T lambda_i = a_i (scaled areas of conductor i):
//   dC'_ij/dtheta = sum_(k moved, cond k = i) da_k sigma_kj
//                   - sum_(k, l) lambda_ki (dP'_kl / dtheta) sigma_lj,
// dP'_kl nonzero when panel k (its collocation point) or l moves.  Every
// entry is the forward-mode derivative of the closed-form integral.

__global__ void k_scatter_moved(int nm, const int *__restrict__ moved, int *__restrict__ mv_of,
                                int set) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nm) mv_of[moved[i]] = set ? i : -1;
}

// Gauss table of every panel: gq[k][0..11] points, [12..15] weights, [16]
// squared diagonal (the far-entry test), [17..19] unused.
__global__ void k_gauss_table(int n, const double *__restrict__ verts, double *__restrict__ gq) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double *v = verts + 12L * k;
  double xg[4][3], wg[4];
  gauss4_t<double>(v, xg, wg);
  double *o = gq + 20L * k;
  for (int q = 0; q < 4; ++q)
    for (int d = 0; d < 3; ++d) o[3 * q + d] = xg[q][d];
  for (int q = 0; q < 4; ++q) o[12 + q] = wg[q];
  double d1 = 0.0, d2 = 0.0;
  for (int d = 0; d < 3; ++d) {
    d1 += (v[6 + d] - v[d]) * (v[6 + d] - v[d]);
    d2 += (v[9 + d] - v[3 + d]) * (v[9 + d] - v[3 + d]);
  }
  o[16] = fmax(d1, d2);
}

// dc/dtheta (3 per moved panel and direction), da/dtheta and the
// derivatives of the panel's Gauss points and weights (16 per direction).
template <int ND>
__global__ void k_moved_props(int nm, const int *__restrict__ moved,
                              const double *__restrict__ verts, const double *__restrict__ dv,
                              long dstride, double *__restrict__ dcen, double *__restrict__ darea,
                              double *__restrict__ dgq) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nm) return;
  const int k = moved[i];
  DualN<ND> v[12];
  for (int e = 0; e < 12; ++e) {
    v[e] = DualN<ND>(verts[12L * k + e]);
    for (int t = 0; t < ND; ++t) v[e].d[t] = dv[t * dstride + 12L * i + e];
  }
  DualN<ND> c[3], a;
  panel_props_dual(v, c, a);
  DualN<ND> xg[4][3], wg[4];
  gauss4_t<DualN<ND>>(v, xg, wg);
  for (int t = 0; t < ND; ++t) {
    for (int d = 0; d < 3; ++d) dcen[(3L * t) * nm + 3 * i + d] = c[d].d[t];
    darea[(long)t * nm + i] = a.d[t];
    double *o = dgq + ((long)t * nm + i) * 16;
    for (int q = 0; q < 4; ++q) {
      for (int d = 0; d < 3; ++d) o[3 * q + d] = xg[q][d].d[t];
      o[12 + q] = wg[q].d[t];
    }
  }
}

// d/dtheta of sum_g w_g / |p - x_g| for the ND directions: dp (observer),
// dx / dw (source points and weights, null when the source stays).
template <int ND>
__device__ inline void far_entry(const double p[3], const double (*dp)[3], const double *gq,
                                 const double *dgq, long dgs, double out[ND]) {
#pragma unroll
  for (int t = 0; t < ND; ++t) out[t] = 0.0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const double ex = p[0] - gq[3 * q], ey = p[1] - gq[3 * q + 1], ez = p[2] - gq[3 * q + 2];
    const double ir = 1.0 / sqrt(ex * ex + ey * ey + ez * ez);
    const double w = gq[12 + q], wir3 = w * ir * ir * ir;
#pragma unroll
    for (int t = 0; t < ND; ++t) {
      double gx = dp ? dp[t][0] : 0.0, gy = dp ? dp[t][1] : 0.0, gz = dp ? dp[t][2] : 0.0;
      double dw = 0.0;
      if (dgq) {
        const double *o = dgq + t * dgs;
        gx -= o[3 * q];
        gy -= o[3 * q + 1];
        gz -= o[3 * q + 2];
        dw = o[12 + q];
      }
      out[t] += dw * ir - wir3 * (ex * gx + ey * gy + ez * gz);
    }
  }
}

// Sensitivity entries d/dtheta P'(p, panel l) along the ND directions: the
// observer p = pv moves with pd (null: it stays), the source panel l moves
// when ml >= 0 (its derivatives at index ml of the moved-panel arrays).
// Beyond sqrt(far1) of l's diagonal the centroid rule A_l / |p - c_l|,
// beyond sqrt(far2) the 2x2 Gauss rule (sens_far_entry, which gives 0
// inside sqrt(far2)); inside, the dual closed form (sens_near_entry).  The
// two live in separate kernels: the closed form in dual numbers needs far
// more registers than the rules, and only a few entries per row take it.
__device__ inline bool sens_is_near(int l, const double pv[3], const double *__restrict__ cen,
                                    const double *__restrict__ gq, double far2, double far1) {
  const double cx = cen[3L * l] - pv[0], cy = cen[3L * l + 1] - pv[1], cz = cen[3L * l + 2] - pv[2];
  const double r2 = cx * cx + cy * cy + cz * cz, d2 = gq[20L * l + 16];
  return !((far1 > 0.0 && r2 > far1 * d2) || (far2 > 0.0 && r2 > far2 * d2));
}

template <int ND>
__device__ inline void sens_far_entry(int l, int ml, int nm, const double pv[3],
                                      const double (*pd)[3], const double *__restrict__ cen,
                                      const double *__restrict__ dcen,
                                      const double *__restrict__ darea,
                                      const double *__restrict__ gq,
                                      const double *__restrict__ dgq, double far2, double far1,
                                      double out[ND]) {
  const double cx = cen[3L * l] - pv[0], cy = cen[3L * l + 1] - pv[1], cz = cen[3L * l + 2] - pv[2];
  const double r2 = cx * cx + cy * cy + cz * cz;
  const double *gl = gq + 20L * l;
  if (far1 > 0.0 && r2 > far1 * gl[16]) {
    const double A = gl[12] + gl[13] + gl[14] + gl[15], ir = 1.0 / sqrt(r2), air3 = A * ir * ir * ir;
#pragma unroll
    for (int t = 0; t < ND; ++t) {
      double gx = pd ? pd[t][0] : 0.0, gy = pd ? pd[t][1] : 0.0, gz = pd ? pd[t][2] : 0.0, dA = 0.0;
      if (ml >= 0) {
        const double *dc = dcen + (3L * t) * nm + 3L * ml;
        gx -= dc[0];
        gy -= dc[1];
        gz -= dc[2];
        dA = darea[(long)t * nm + ml];
      }
      out[t] = dA * ir + air3 * (cx * gx + cy * gy + cz * gz);
    }
  } else if (far2 > 0.0 && r2 > far2 * gl[16]) {
    far_entry<ND>(pv, pd, gl, ml >= 0 ? dgq + 16L * ml : nullptr, 16L * nm, out);
  } else {
#pragma unroll
    for (int t = 0; t < ND; ++t) out[t] = 0.0;
  }
}

template <int ND>
__device__ inline void sens_near_entry(int l, int ml, const double pv[3], const double (*pd)[3],
                                       const double *__restrict__ verts,
                                       const double *__restrict__ dv, long dstride,
                                       double out[ND]) {
  DualN<ND> v[12], P[3];
#pragma unroll
  for (int e = 0; e < 12; ++e) {
    v[e] = DualN<ND>(verts[12L * l + e]);
    if (ml >= 0)
#pragma unroll
      for (int tt = 0; tt < ND; ++tt) v[e].d[tt] = dv[tt * dstride + 12L * ml + e];
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    P[d] = DualN<ND>(pv[d]);
    if (pd)
#pragma unroll
      for (int tt = 0; tt < ND; ++tt) P[d].d[tt] = pd[tt][d];
  }
  const DualN<ND> w = wilton_t<DualN<ND>>(v, P[0], P[1], P[2]);
#pragma unroll
  for (int tt = 0; tt < ND; ++tt) out[tt] = w.d[tt];
}

// Both far-rule contractions are tiles of SR_ROWS observer rows against a
// run of source panels, in sub-chunks of SR_SC sources: the entries go to
// LDS (each thread: one source, every other row of the tile, so the
// source's data is read once per 8 entries), then thread groups of nc
// columns (256 / nc of them) contract slices of the sub-chunk against the
// sources' sigma rows into SR_ROWS x ND register accumulators, so a sigma
// row is read once per tile instead of once per row.  The groups' sums are
// added through LDS at the end, one row at a time.
constexpr int SR_ROWS = 16, SR_SC = 128;

// Rows of the moved panels: partial[((chunk * nm + mi) * ND + t) * nc + j] =
// sum over the panels l of this chunk (2048 a chunk) of dP'_kl sigma[l][j]
// along direction t, k = moved[mi] (the observer and, when l is moved, the
// source move).  Work-group (tile of SR_ROWS moved rows, chunk).
template <int ND>
__global__ __launch_bounds__(256) void k_sens_rows(int N, const int *__restrict__ moved,
                                                   const int *__restrict__ mv_of,
                                                   const double *__restrict__ verts,
                                                   const double *__restrict__ cen,
                                                   const double *__restrict__ dv, long dstride,
                                                   const double *__restrict__ dcen,
                                                   const double *__restrict__ darea,
                                                   const double *__restrict__ gq,
                                                   const double *__restrict__ dgq, double far2,
                                                   double far1, int nm,
                                                   const double *__restrict__ sigma, int lds,
                                                   int nc, double *__restrict__ partial) {
  __shared__ double sdp[ND][SR_ROWS][SR_SC];
  __shared__ double sred[ND][256];
  __shared__ double sp[SR_ROWS][3 + 3 * ND];  // observers: centroid, ND derivatives
  const int m0 = blockIdx.x * SR_ROWS, chunk = blockIdx.y, tid = threadIdx.x;
  const int mr = min(SR_ROWS, nm - m0);
  for (int e = tid; e < SR_ROWS * (3 + 3 * ND); e += 256) {
    const int r = e / (3 + 3 * ND), c = e % (3 + 3 * ND), mi = m0 + r;
    double x = 0.0;
    if (r < mr) {
      if (c < 3) x = cen[3L * moved[mi] + c];
      else x = dcen[(3L * ((c - 3) / 3)) * nm + 3L * mi + (c - 3) % 3];
    }
    sp[r][c] = x;
  }
  const int l0 = chunk * 2048, nl = min(N, l0 + 2048) - l0;
  const int ng = max(1, 256 / nc), g = tid / nc, j = tid % nc;
  double acc[SR_ROWS][ND];
#pragma unroll
  for (int r = 0; r < SR_ROWS; ++r)
#pragma unroll
    for (int t = 0; t < ND; ++t) acc[r][t] = 0.0;
  const int q = tid % SR_SC, rb = tid / SR_SC;  // entry phase: source q, rows rb + 2 i
  for (int s0 = 0; s0 < nl; s0 += SR_SC) {
    const int ns = min(SR_SC, nl - s0);
    __syncthreads();  // observers staged / previous sub-chunk contracted
    {
      const int l = l0 + s0 + q;
      const int ml = q < ns ? mv_of[l] : -1;
#pragma unroll 1
      for (int r = rb; r < SR_ROWS; r += 256 / SR_SC) {
        double f[ND];
        if (q < ns && r < mr) {
          double pd[ND][3];
#pragma unroll
          for (int t = 0; t < ND; ++t)
#pragma unroll
            for (int d = 0; d < 3; ++d) pd[t][d] = sp[r][3 + 3 * t + d];
          sens_far_entry<ND>(l, ml, nm, &sp[r][0], pd, cen, dcen, darea, gq, dgq, far2, far1, f);
        } else {
#pragma unroll
          for (int t = 0; t < ND; ++t) f[t] = 0.0;
        }
#pragma unroll
        for (int t = 0; t < ND; ++t) sdp[t][r][q] = f[t];
      }
    }
    __syncthreads();
    if (g < ng) {
      const int span = (ns + ng - 1) / ng, q1 = min(ns, (g + 1) * span);
      for (int qq = g * span; qq < q1; ++qq) {
        const double sg = sigma[(long)(l0 + s0 + qq) * lds + j];
#pragma unroll
        for (int r = 0; r < SR_ROWS; ++r)
#pragma unroll
          for (int t = 0; t < ND; ++t) acc[r][t] += sdp[t][r][qq] * sg;
      }
    }
  }
#pragma unroll
  for (int r = 0; r < SR_ROWS; ++r) {
    __syncthreads();
#pragma unroll
    for (int t = 0; t < ND; ++t) sred[t][tid] = g < ng ? acc[r][t] : 0.0;
    __syncthreads();
    if (tid < ND * nc && r < mr) {
      const int t = tid / nc, jj = tid % nc;
      double s = 0.0;
      for (int gg = 0; gg < ng; ++gg) s += sred[t][gg * nc + jj];
      partial[(((long)chunk * nm + m0 + r) * ND + t) * nc + jj] = s;
    }
  }
}

// The closed-form entries the far-rule kernels leave out (source within
// sqrt(far2) of its diagonal; every entry when far2 = 0), added to their
// outputs afterwards.  Work-group: one observer row against a run of
// sources (rows of moved panels: a chunk of 2048 panels, the partial of
// (chunk, row); rows that stay: every moved panel, R directly).  Rounds of
// 2048 sources, 8 tests per thread issued together; the hits are compacted
// in a fixed order (test slot, then thread), their dual closed forms go to
// LDS (the kernel's registers are sized for those, so one round per chunk
// keeps the latency-bound scan short), then groups of nc threads contract
// the hits against their sigma rows.  Deterministic.
// Per chunk c of 2048 sources (source s: panel s, or moved[s] with
// `moved`): the bounding box of the centroids and the largest squared
// radius within which an entry takes the closed form (min of the active far
// thresholds times the squared diagonal; +inf when neither is active), box[8 c
// + 0..6] = min xyz, max xyz, radius^2.  An observer farther than that from
// the box has no closed-form entry in the chunk.
__global__ __launch_bounds__(256) void k_sens_boxes(int n_src, const int *__restrict__ moved,
                                                    const double *__restrict__ cen,
                                                    const double *__restrict__ gq, double far2,
                                                    double far1, double *__restrict__ box) {
  __shared__ double sr[7][256];
  const int c = blockIdx.x, tid = threadIdx.x;
  double v[7] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY, 0.0};
  for (int s = c * 2048 + tid; s < min(n_src, (c + 1) * 2048); s += 256) {
    const int l = moved ? moved[s] : s;
    const double d2 = gq[20L * l + 16];
    double r2 = INFINITY;
    if (far1 > 0.0) r2 = fmin(r2, far1 * d2);
    if (far2 > 0.0) r2 = fmin(r2, far2 * d2);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      v[d] = fmin(v[d], cen[3L * l + d]);
      v[3 + d] = fmax(v[3 + d], cen[3L * l + d]);
    }
    v[6] = fmax(v[6], r2);
  }
#pragma unroll
  for (int e = 0; e < 7; ++e) sr[e][tid] = v[e];
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (tid < h) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        sr[d][tid] = fmin(sr[d][tid], sr[d][tid + h]);
        sr[3 + d][tid] = fmax(sr[3 + d][tid], sr[3 + d][tid + h]);
      }
      sr[6][tid] = fmax(sr[6][tid], sr[6][tid + h]);
    }
    __syncthreads();
  }
  if (tid < 7) box[8L * c + tid] = sr[tid][0];
}

// true when the observer p is farther than the chunk's closed-form radius
// from its box
__device__ inline bool sens_box_far(const double *__restrict__ b, const double pv[3]) {
  double d2 = 0.0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double e = fmax(fmax(b[d] - pv[d], pv[d] - b[3 + d]), 0.0);
    d2 += e * e;
  }
  return d2 > b[6];
}

constexpr int SN_T = 8;  // tests per thread per round
template <int ND>
__global__ __launch_bounds__(256) void k_sens_near(int N, int nm, int rows_moved,
                                                   const int *__restrict__ moved,
                                                   const int *__restrict__ mv_of,
                                                   const double *__restrict__ verts,
                                                   const double *__restrict__ cen,
                                                   const double *__restrict__ dv, long dstride,
                                                   const double *__restrict__ dcen,
                                                   const double *__restrict__ gq, double far2,
                                                   double far1, const double *__restrict__ sigma,
                                                   int lds, int nc, const double *__restrict__ box,
                                                   double *__restrict__ out) {
  constexpr int RS = 256 * SN_T;
  __shared__ double sdp[ND][RS];
  __shared__ int shl[RS];
  __shared__ int swc[SN_T][4];
  __shared__ double sred[ND][256];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int k, mi = -1, s_lo, s_hi, chunk = 0;
  if (rows_moved) {
    mi = blockIdx.x;
    k = moved[mi];
    chunk = blockIdx.y;
    s_lo = chunk * 2048;
    s_hi = min(N, s_lo + 2048);
  } else {
    k = blockIdx.x;
    if (mv_of[k] >= 0) return;  // the whole work-group: a moved row is k_sens_rows'
    s_lo = 0;
    s_hi = nm;
  }
  const double pv[3] = {cen[3L * k], cen[3L * k + 1], cen[3L * k + 2]};
  if (rows_moved && sens_box_far(box + 8L * chunk, pv)) return;  // no closed-form entry here
  double pd[ND][3];
#pragma unroll
  for (int t = 0; t < ND; ++t)
#pragma unroll
    for (int d = 0; d < 3; ++d) pd[t][d] = rows_moved ? dcen[(3L * t) * nm + 3L * mi + d] : 0.0;
  const int ng = max(1, 256 / nc), g = tid / nc, j = tid % nc;
  double acc[ND];
#pragma unroll
  for (int t = 0; t < ND; ++t) acc[t] = 0.0;
  for (int b0 = s_lo; b0 < s_hi; b0 += RS) {
    if (!rows_moved && sens_box_far(box + 8L * (b0 / RS), pv)) continue;  // uniform
    int cand[SN_T];
    unsigned hit = 0;
#pragma unroll
    for (int i = 0; i < SN_T; ++i) {
      const int s = b0 + i * 256 + tid;
      cand[i] = s < s_hi ? (rows_moved ? s : moved[s]) : -1;
    }
#pragma unroll
    for (int i = 0; i < SN_T; ++i)
      if (cand[i] >= 0 && sens_is_near(cand[i], pv, cen, gq, far2, far1)) hit |= 1u << i;
    __syncthreads();  // the previous round's hits are contracted
#pragma unroll
    for (int i = 0; i < SN_T; ++i) {
      const unsigned long long m = __ballot((hit >> i) & 1u);
      if (lane == 0) swc[i][wave] = __popcll(m);
    }
    __syncthreads();
    int nh = 0;
#pragma unroll
    for (int i = 0; i < SN_T; ++i) {
      const unsigned long long m = __ballot((hit >> i) & 1u);
      int base = nh;
      for (int w = 0; w < wave; ++w) base += swc[i][w];
      nh += swc[i][0] + swc[i][1] + swc[i][2] + swc[i][3];
      if ((hit >> i) & 1u) {
        const int h = base + __popcll(m & ((1ull << lane) - 1ull));
        const int l = cand[i], s = b0 + i * 256 + tid;
        const int ml = rows_moved ? mv_of[l] : s;
        double f[ND];
        sens_near_entry<ND>(l, ml, pv, rows_moved ? pd : nullptr, verts, dv, dstride, f);
#pragma unroll
        for (int t = 0; t < ND; ++t) sdp[t][h] = f[t];
        shl[h] = l;
      }
    }
    __syncthreads();
    if (g < ng) {
      const int span = (nh + ng - 1) / ng, h1 = min(nh, (g + 1) * span);
      for (int h = g * span; h < h1; ++h) {
        const double sg = sigma[(long)shl[h] * lds + j];
#pragma unroll
        for (int t = 0; t < ND; ++t) acc[t] += sdp[t][h] * sg;
      }
    }
  }
#pragma unroll
  for (int t = 0; t < ND; ++t) sred[t][tid] = g < ng ? acc[t] : 0.0;
  __syncthreads();
  if (tid < ND * nc) {
    const int t = tid / nc, jj = tid % nc;
    double sum = 0.0;
    for (int gg = 0; gg < ng; ++gg) sum += sred[t][gg * nc + jj];
    if (rows_moved)
      out[(((long)chunk * nm + mi) * ND + t) * nc + jj] += sum;
    else
      out[(long)t * N * nc + (long)k * nc + jj] += sum;
  }
}

// R_t[moved[mi]][j] = sum over chunks of the partials (fixed order); R_t at
// R + t * N * nc.
__global__ void k_sens_rows_reduce(int N, int nm, int nchunk, int nd, int nc,
                                   const int *__restrict__ moved,
                                   const double *__restrict__ partial, double *__restrict__ R) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nm * nd * nc) return;
  const int mi = e / (nd * nc), t = (e / nc) % nd, j = e % nc;
  double s = 0.0;
  for (int c = 0; c < nchunk; ++c) s += partial[(((long)c * nm + mi) * nd + t) * nc + j];
  R[(long)t * N * nc + (long)moved[mi] * nc + j] = s;
}

// Rows of the panels that stay: R_t[k][j] = sum over moved l of dP'_kl
// sigma[l][j] (fixed collocation point, moving source).  Work-group: a tile
// of SR_ROWS consecutive panels against every moved panel; rows of moved
// panels are skipped (k_sens_rows writes them).
template <int ND>
__global__ __launch_bounds__(256) void k_sens_cols(int N, int nm, const int *__restrict__ moved,
                                                   const int *__restrict__ mv_of,
                                                   const double *__restrict__ verts,
                                                   const double *__restrict__ cen,
                                                   const double *__restrict__ dv, long dstride,
                                                   const double *__restrict__ dcen,
                                                   const double *__restrict__ darea,
                                                   const double *__restrict__ gq,
                                                   const double *__restrict__ dgq, double far2,
                                                   double far1, const double *__restrict__ sigma,
                                                   int lds, int nc, double *__restrict__ R) {
  __shared__ double sdp[ND][SR_ROWS][SR_SC];
  __shared__ double sred[ND][256];
  __shared__ double sp[SR_ROWS][3];
  __shared__ int srow[SR_ROWS];
  const int k0 = blockIdx.x * SR_ROWS, tid = threadIdx.x;
  if (tid < SR_ROWS) {
    const int k = k0 + tid;
    srow[tid] = k < N && mv_of[k] < 0;
  }
  for (int e = tid; e < SR_ROWS * 3; e += 256) {
    const int r = e / 3, k = k0 + r;
    sp[r][e % 3] = k < N ? cen[3L * k + e % 3] : 0.0;
  }
  const int ng = max(1, 256 / nc), g = tid / nc, j = tid % nc;
  double acc[SR_ROWS][ND];
#pragma unroll
  for (int r = 0; r < SR_ROWS; ++r)
#pragma unroll
    for (int t = 0; t < ND; ++t) acc[r][t] = 0.0;
  const int q = tid % SR_SC, rb = tid / SR_SC;
  for (int s0 = 0; s0 < nm; s0 += SR_SC) {
    const int ns = min(SR_SC, nm - s0);
    __syncthreads();
    {
      const int mq = s0 + q;
      const int l = q < ns ? moved[mq] : 0;
#pragma unroll 1
      for (int r = rb; r < SR_ROWS; r += 256 / SR_SC) {
        double f[ND];
        if (q < ns && srow[r]) {
          sens_far_entry<ND>(l, mq, nm, &sp[r][0], nullptr, cen, dcen, darea, gq, dgq, far2, far1,
                             f);
        } else {
#pragma unroll
          for (int t = 0; t < ND; ++t) f[t] = 0.0;
        }
#pragma unroll
        for (int t = 0; t < ND; ++t) sdp[t][r][q] = f[t];
      }
    }
    __syncthreads();
    if (g < ng) {
      const int span = (ns + ng - 1) / ng, q1 = min(ns, (g + 1) * span);
      for (int qq = g * span; qq < q1; ++qq) {
        const double sg = sigma[(long)moved[s0 + qq] * lds + j];
#pragma unroll
        for (int r = 0; r < SR_ROWS; ++r)
#pragma unroll
          for (int t = 0; t < ND; ++t) acc[r][t] += sdp[t][r][qq] * sg;
      }
    }
  }
#pragma unroll
  for (int r = 0; r < SR_ROWS; ++r) {
    __syncthreads();
#pragma unroll
    for (int t = 0; t < ND; ++t) sred[t][tid] = g < ng ? acc[r][t] : 0.0;
    __syncthreads();
    if (tid < ND * nc && srow[r]) {
      const int t = tid / nc, jj = tid % nc;
      double s = 0.0;
      for (int gg = 0; gg < ng; ++gg) s += sred[t][gg * nc + jj];
      R[(long)t * N * nc + (long)(k0 + r) * nc + jj] = s;
    }
  }
}

constexpr int SC_ROWS = 256;  // rows per work-group of k_sens_contract

// J'[i][j] partial over the rows of block b:
//   -sum_k lambda[k][i] R[k][j] + sum_(k, cond k = i) da[k] sigma[k][j]
// (da[k] = 0 for panels that stay), partial[b][i][j].
__global__ __launch_bounds__(256) void k_sens_contract(int N, int nc, int rows,
                                                       const double *__restrict__ lambda, int ldl,
                                                       const double *__restrict__ R,
                                                       const double *__restrict__ sigma, int lds,
                                                       const double *__restrict__ da,
                                                       const uint32_t *__restrict__ cond,
                                                       int cond0, double *__restrict__ partial) {
  const int b = blockIdx.x;
  const int k0 = b * rows, k1 = min(N, k0 + rows);
  for (int e = threadIdx.x; e < nc * nc; e += blockDim.x) {
    const int i = e / nc, j = e % nc;
    double t = 0.0;
    for (int k = k0; k < k1; ++k) {
      t -= lambda[(long)k * ldl + i] * R[(long)k * nc + j];
      if (da[k] != 0.0 && (int)cond[k] - cond0 == i) t += da[k] * sigma[(long)k * lds + j];
    }
    partial[(long)b * nc * nc + e] = t;
  }
}

__global__ void k_sens_sum(int nblk, int n, const double *__restrict__ partial,
                           double *__restrict__ out) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double t = 0.0;
  for (int b = 0; b < nblk; ++b) t += partial[(long)b * n + e];
  out[e] = t;
}

__global__ void k_scatter_da(int nm, const int *__restrict__ moved, const double *__restrict__ darea,
                             double *__restrict__ da, int set) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nm) da[moved[i]] = set ? darea[i] : 0.0;
}

template <int ND>
void launch_sensitivity_nd(const SensArgs &a, hipStream_t s) {
  const int nm = a.nm, nc = a.nc, N = a.N;
  const int g = (nm + 255) / 256;
  hipLaunchKernelGGL(k_scatter_moved, dim3(g), dim3(256), 0, s, nm, a.moved, a.mv_of, 1);
  hipLaunchKernelGGL((k_moved_props<ND>), dim3(g), dim3(256), 0, s, nm, a.moved, a.verts, a.dv,
                     a.dstride, a.dcen, a.darea, a.dgq);
  const int nchunk = (N + 2047) / 2048;
  // chunk boxes for the closed-form scans, after the partials / contraction
  // blocks in the scratch buffer
  const size_t used = std::max((size_t)nchunk * nm * ND * nc,
                               (size_t)((N + SC_ROWS - 1) / SC_ROWS) * nc * nc);
  double *box_rows = a.scratch + used, *box_cols = box_rows + 8L * nchunk;
  hipLaunchKernelGGL(k_sens_boxes, dim3(nchunk), dim3(256), 0, s, N, nullptr, a.cen, a.gq, a.far2,
                     a.far1, box_rows);
  hipLaunchKernelGGL(k_sens_boxes, dim3((nm + 2047) / 2048), dim3(256), 0, s, nm, a.moved, a.cen,
                     a.gq, a.far2, a.far1, box_cols);
  hipLaunchKernelGGL((k_sens_rows<ND>), dim3((nm + SR_ROWS - 1) / SR_ROWS, nchunk), dim3(256), 0, s,
                     N, a.moved, a.mv_of, a.verts, a.cen, a.dv, a.dstride, a.dcen, a.darea, a.gq,
                     a.dgq, a.far2, a.far1, nm, a.sigma, a.lds, nc, a.scratch);
  hipLaunchKernelGGL((k_sens_near<ND>), dim3(nm, nchunk), dim3(256), 0, s, N, nm, 1, a.moved,
                     a.mv_of, a.verts, a.cen, a.dv, a.dstride, a.dcen, a.gq, a.far2, a.far1, a.sigma,
                     a.lds, nc, box_rows, a.scratch);
  hipLaunchKernelGGL(k_sens_rows_reduce, dim3((nm * ND * nc + 255) / 256), dim3(256), 0, s, N, nm,
                     nchunk, ND, nc, a.moved, a.scratch, a.R);
  hipLaunchKernelGGL((k_sens_cols<ND>), dim3((N + SR_ROWS - 1) / SR_ROWS), dim3(256), 0, s, N, nm,
                     a.moved, a.mv_of, a.verts, a.cen, a.dv, a.dstride, a.dcen, a.darea, a.gq,
                     a.dgq, a.far2, a.far1, a.sigma, a.lds, nc, a.R);
  hipLaunchKernelGGL((k_sens_near<ND>), dim3(N), dim3(256), 0, s, N, nm, 0, a.moved, a.mv_of,
                     a.verts, a.cen, a.dv, a.dstride, a.dcen, a.gq, a.far2, a.far1, a.sigma, a.lds, nc,
                     box_cols, a.R);
  const int rows = SC_ROWS, nblk = (N + rows - 1) / rows;  // enough work-groups to fill the chip
  for (int t = 0; t < ND; ++t) {
    hipLaunchKernelGGL(k_scatter_da, dim3(g), dim3(256), 0, s, nm, a.moved, a.darea + (long)t * nm,
                       a.da, 1);
    hipLaunchKernelGGL(k_sens_contract, dim3(nblk), dim3(256), 0, s, N, nc, rows, a.lambda, a.ldl,
                       a.R + (long)t * N * nc, a.sigma, a.lds, a.da, a.cond, a.cond0, a.scratch);
    hipLaunchKernelGGL(k_sens_sum, dim3((nc * nc + 255) / 256), dim3(256), 0, s, nblk, nc * nc,
                       a.scratch, a.J + (long)t * nc * nc);
    hipLaunchKernelGGL(k_scatter_da, dim3(g), dim3(256), 0, s, nm, a.moved, a.darea, a.da, 0);
  }
  hipLaunchKernelGGL(k_scatter_moved, dim3(g), dim3(256), 0, s, nm, a.moved, a.mv_of, 0);
  GP_LAUNCH_CHECK();
}

void launch_sensitivity(const SensArgs &a, hipStream_t s) {
  if (a.nm == 0) {
    hipLaunchKernelGGL(k_sens_sum, dim3((a.nd * a.nc * a.nc + 255) / 256), dim3(256), 0, s, 0,
                       a.nd * a.nc * a.nc, a.scratch, a.J);  // zero fill
    GP_LAUNCH_CHECK();
    return;
  }
  // algorithmic flops: 2 nm N dual entries (row and column of each moved
  // panel) at the 2x2 Gauss rule's ~60 real flops, times (1 + 2 nd) for the
  // dual arithmetic; the closed-form near entries cost more, so this is a
  // lower bound
  ProfScope ps("sens_wilton_dual", s, 2.0 * a.nm * (double)a.N * 60.0 * (1 + 2 * a.nd), 0.0);
  switch (a.nd) {
    case 1: launch_sensitivity_nd<1>(a, s); break;
    case 2: launch_sensitivity_nd<2>(a, s); break;
    default: throw Error(GP_ERR_INTERNAL, "sensitivity groups hold 1 or 2 directions");
  }
}

void launch_gauss_table(int n, const double *verts, double *gq, hipStream_t s) {
  hipLaunchKernelGGL(k_gauss_table, dim3((n + 255) / 256), dim3(256), 0, s, n, verts, gq);
  GP_LAUNCH_CHECK();
}

size_t sensitivity_scratch(int N, int nm_max, int nc, int nd) {
  const size_t rows = (size_t)((N + 2047) / 2048) * nm_max * nd * nc;
  const size_t con = (size_t)((N + SC_ROWS - 1) / SC_ROWS) * nc * nc;
  const size_t boxes = 8 * (size_t)((N + 2047) / 2048 + (nm_max + 2047) / 2048);
  return std::max(rows, con) + boxes;
}

}  // namespace gp
