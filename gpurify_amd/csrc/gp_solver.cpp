// Host orchestration of the device capacitance solve.
//
// Dense path (cap/solver.rs:100-134, Method::Direct): device Wilton assembly,
// single-work-group LU, device charge accumulation.
//
// FMM path (cap/fmm_solver.rs:147-172): the precorrected operator
// P sigma = (near exact-Wilton tiles) sigma + far(area * sigma), the far
// field by the black-box Chebyshev FMM (fmm/bbfmm.rs) with the uncompressed
// Chebyshev M2L matrix, right Jacobi preconditioning from the self terms and
// restarted GMRES(m) (krylov.rs:50-191) over a block of right-hand sides
// iterating in lockstep.  Mixed precision runs the inner GMRES on the FP32
// operator inside FP64 iterative refinement whose residual uses the FP64
// operator.
//
// Everything on the device is in scaled units: lengths divided by the root
// half-width L0 and centred on the root cube.  Then P' = P / (k_e L0),
// sigma' = k_e L0 sigma and C = 4 pi eps0 L0 C'.

#include "gp_solver.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>

#include "gp_kernels.h"

namespace gp {

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// Progress lines on stderr when GPURIFY_LOG is set (long solves).
bool log_enabled() {
  static const bool on = [] {
    const char *e = std::getenv("GPURIFY_LOG");
    return e && *e && *e != '0';
  }();
  return on;
}
#define GP_LOG(...)                          \
  do {                                       \
    if (log_enabled()) {                     \
      std::fprintf(stderr, "[gpurify] ");    \
      std::fprintf(stderr, __VA_ARGS__);     \
      std::fputc('\n', stderr);              \
      std::fflush(stderr);                   \
    }                                        \
  } while (0)

constexpr double kMinPositive = 2.2250738585072014e-308;

// Chebyshev roots (chebyshev.rs:11-17) and 1-D interpolation weights
// S_n(x, x_m) (chebyshev.rs:40-74).
std::vector<double> cheb_nodes(int n) {
  std::vector<double> v(n);
  for (int m = 0; m < n; ++m) v[m] = std::cos((2.0 * m + 1.0) * kPi / (2.0 * n));
  return v;
}
void interp_weights(double x, int n, const double *nodes, double *w) {
  double tx[32];
  tx[0] = 1.0;
  if (n > 1) {
    tx[1] = x;
    for (int k = 2; k < n; ++k) tx[k] = 2.0 * x * tx[k - 1] - tx[k - 2];
  }
  double inv_n = 1.0 / n;
  for (int m = 0; m < n; ++m) {
    double xm = nodes[m], sum = 0.5, tkm1 = 1.0, tk = xm;
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

template <typename T>
std::vector<T> cast_vec(const std::vector<double> &v) {
  return std::vector<T>(v.begin(), v.end());
}

// FP32 -> BF16 bits, round to nearest even (NaN stays NaN).
uint16_t bf16_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// Scaled SoA arrays of the panels in the order 'perm' (identity if null).
void scaled_arrays(const PanelGeom &g, const uint32_t *perm, const double c[3], double L0,
                   std::vector<double> &verts, std::vector<double> &cen,
                   std::vector<double> &area, std::vector<uint32_t> &cond) {
  int n = g.n;
  verts.resize((size_t)n * 12);
  cen.resize((size_t)n * 3);
  area.resize(n);
  cond.resize(n);
  double inv = 1.0 / L0;
  for (int p = 0; p < n; ++p) {
    int i = perm ? (int)perm[p] : p;
    for (int k = 0; k < 4; ++k)
      for (int d = 0; d < 3; ++d)
        verts[(size_t)p * 12 + 3 * k + d] = (g.verts[(size_t)i * 12 + 3 * k + d] - c[d]) * inv;
    cen[3 * p] = (g.cx[i] - c[0]) * inv;
    cen[3 * p + 1] = (g.cy[i] - c[1]) * inv;
    cen[3 * p + 2] = (g.cz[i] - c[2]) * inv;
    area[p] = g.area[i] * inv * inv;
    cond[p] = g.cond[i];
  }
}

// Panels of each conductor as positions in the device order, listed by
// ascending caller index (the summation order of accumulate_c,
// cap/solver.rs:67-71).
void cond_csr(const PanelGeom &g, const uint32_t *perm, std::vector<int> &off,
              std::vector<int> &idx) {
  std::vector<int> pos(g.n);
  for (int p = 0; p < g.n; ++p) pos[perm ? perm[p] : p] = p;
  off.assign(g.ncond + 1, 0);
  for (int i = 0; i < g.n; ++i) off[g.cond[i] + 1]++;
  for (int c = 0; c < g.ncond; ++c) off[c + 1] += off[c];
  idx.assign(g.n, 0);
  std::vector<int> fill(off.begin(), off.end() - 1);
  for (int i = 0; i < g.n; ++i) idx[fill[g.cond[i]]++] = pos[i];
}

// ------------------------------------------------------------------ dense
void solve_dense(gp_ctx *ctx, const PanelGeom &g, double *C, gp_cap_report *rep,
                 const double *eps_panel) {
  hipStream_t s = ctx->stream;
  double t0 = now_s();
  const int n = g.n, nc = g.ncond;
  double c[3], L0;
  root_cube(g, c, L0);
  std::vector<double> hv, hc, ha;
  std::vector<uint32_t> hcond;
  scaled_arrays(g, nullptr, c, L0, hv, hc, ha, hcond);
  std::vector<double> w(n);
  for (int i = 0; i < n; ++i) w[i] = ha[i] * (eps_panel ? eps_panel[i] : 1.0);
  std::vector<double> B((size_t)n * nc, 0.0);
  for (int i = 0; i < n; ++i) B[(size_t)i * nc + g.cond[i]] = 1.0;
  std::vector<int> coff, cidx;
  cond_csr(g, nullptr, coff, cidx);

  auto d_v = to_device(hv, s);
  auto d_c = to_device(hc, s);
  auto d_w = to_device(w, s);
  auto d_B = to_device(B, s);
  auto d_coff = to_device(coff, s);
  auto d_cidx = to_device(cidx, s);
  DBuf<double> d_P((size_t)n * n), d_work((size_t)n * nc), d_C((size_t)nc * nc);
  DBuf<int> d_piv(n), d_info(1);
  launch_dense_assemble(n, d_v.p, d_c.p, d_P.p, s);
  GP_HIP(hipStreamSynchronize(s));
  double t1 = now_s();
  launch_lu_solve(n, nc, d_P.p, d_piv.p, d_B.p, d_work.p, d_info.p, s);
  launch_cap_accumulate(nc, nc, nc, d_coff.p, d_cidx.p, d_w.p, d_B.p, d_C.p, s);
  int info = 0;
  GP_HIP(hipMemcpyAsync(&info, d_info.p, sizeof(int), hipMemcpyDeviceToHost, s));
  std::vector<double> hC((size_t)nc * nc);
  GP_HIP(hipMemcpyAsync(hC.data(), d_C.p, hC.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  GP_HIP(hipStreamSynchronize(s));
  if (info != 0) throw Error(GP_ERR_SINGULAR, "potential-coefficient matrix is singular");
  const double scale = 4.0 * kPi * kEps0 * L0;
  for (size_t e = 0; e < hC.size(); ++e) C[e] = hC[e] * scale;
  double t2 = now_s();
  if (rep) {
    rep->method_used = GP_METHOD_DENSE;
    rep->precision_used = GP_PREC_FP64;
    rep->converged = 1;
    rep->seconds_setup = t1 - t0;
    rep->seconds_solve = t2 - t1;
  }
}

// ------------------------------------------------------------------ FMM
struct FmmSystem {
  hipStream_t s = nullptr;
  int N = 0, n = 0, L = 0, ncond = 0;
  Tree t;
  double c[3] = {0, 0, 0}, L0 = 1.0;
  std::vector<int> lvl_off;
  int64_t nnz = 0;
  long pairs = 0;
  bool have32 = false;
  int n3p = 0;
  int m2l_inner = GP_M2L_BF16, m2l_outer = GP_M2L_FP64;
  DBuf<double> verts, cen, area, inv_diag;
  DBuf<uint32_t> cond;
  DBuf<int> leaf_lo, col_off, cols, panel_leaf;
  DBuf<int64_t> nf_off;
  DBuf<double> nf64;
  DBuf<float> nf32;
  DBuf<int> parent, octant, child_off, child_idx, targets, il_off, il_src, il_oidx, box_level;
  DBuf<double> inv_half;
  DBuf<double> w1d64, m2m64, G64;
  DBuf<float> w1d32, m2m32, G32, Gr32;
  DBuf<uint16_t> Gbf;
  DBuf<int> coff, cidx;
  NearTiles nt;
  FmmDev<double> f64;
  FmmDev<float> f32;
  int ld = 0;
  DBuf<double> z64, M64, Lc64;
  DBuf<float> z32, M32, Lc32;
  DBuf<uint16_t> Mbf;
  long matvecs = 0;

  void setup(const PanelGeom &g, const gp_cap_opts &o, bool need32, hipStream_t stream) {
    s = stream;
    N = g.n;
    ncond = g.ncond;
    n = (int)o.fmm_order;
    if (n < 2 || n > 6) throw Error(GP_ERR_INVALID_INPUT, "fmm_order must be in [2, 6]");
    m2l_inner = (int)o.m2l_inner;
    m2l_outer = (int)o.m2l_outer;
    if (m2l_inner != GP_M2L_BF16 && m2l_inner != GP_M2L_FP32 && m2l_inner != GP_M2L_BF16_VALU)
      throw Error(GP_ERR_INVALID_INPUT, "m2l_inner must be GP_M2L_BF16, _FP32 or _BF16_VALU");
    if (m2l_outer != GP_M2L_FP64 && m2l_outer != GP_M2L_FP64_VALU)
      throw Error(GP_ERR_INVALID_INPUT, "m2l_outer must be GP_M2L_FP64 or GP_M2L_FP64_VALU");
    L = o.max_level ? (int)o.max_level : choose_depth(g, (int)o.leaf_target);
    if (L < 2) L = 2;
    build_tree(g, L, t);
    L0 = t.half;
    std::memcpy(c, t.center, sizeof(c));
    lvl_off = t.lvl_off;
    std::vector<double> hv, hc, ha;
    std::vector<uint32_t> hcond;
    scaled_arrays(g, t.perm.data(), c, L0, hv, hc, ha, hcond);
    verts = to_device(hv, s);
    cen = to_device(hc, s);
    area = to_device(ha, s);
    cond = to_device(hcond, s);

    // Near-field tiles (cap/fmm_solver.rs:52-91 over bbfmm.rs:189-206).
    const int nl = t.n_leaves();
    std::vector<int> hcol_off(nl + 1, 0), hcols, hpl(N);
    std::vector<int64_t> hnf(nl + 1, 0);
    for (int T = 0; T < nl; ++T) {
      for (int q = t.near_off[T]; q < t.near_off[T + 1]; ++q) {
        int S = t.near_leaf[q];
        for (int p = t.leaf_lo[S]; p < t.leaf_lo[S + 1]; ++p) hcols.push_back(p);
      }
      hcol_off[T + 1] = (int)hcols.size();
      int nt_ = t.leaf_lo[T + 1] - t.leaf_lo[T];
      hnf[T + 1] = hnf[T] + (int64_t)nt_ * (hcol_off[T + 1] - hcol_off[T]);
      for (int p = t.leaf_lo[T]; p < t.leaf_lo[T + 1]; ++p) hpl[p] = T;
    }
    nnz = hnf[nl];
    leaf_lo = to_device(t.leaf_lo, s);
    col_off = to_device(hcol_off, s);
    cols = to_device(hcols, s);
    panel_leaf = to_device(hpl, s);
    nf_off = to_device(hnf, s);
    nf64.alloc((size_t)nnz);
    have32 = need32;
    if (need32) nf32.alloc((size_t)nnz);
    launch_near_assemble(nl, leaf_lo.p, nf_off.p, col_off.p, cols.p, verts.p, cen.p, nf64.p,
                         need32 ? nf32.p : nullptr, s);
    inv_diag.alloc(N);
    const double k_e = 1.0 / (4.0 * kPi * kEps0);
    launch_self_inv_diag(N, verts.p, cen.p, k_e * L0, inv_diag.p, s);
    nt.n = N;
    nt.nnz = (long)nnz;
    nt.n_leaves = nl;
    nt.leaf_lo = leaf_lo.p;
    nt.nf_off = nf_off.p;
    nt.col_off = col_off.p;
    nt.cols = cols.p;
    nt.panel_leaf = panel_leaf.p;

    // FMM tables.
    const int n3 = n * n * n;
    n3p = round_up(n3, 32);
    auto nodes = cheb_nodes(n);
    std::vector<double> hw1d((size_t)N * 3 * n);
    const double half_L = t.level_half(L);
    const double origin[3] = {t.center[0] - t.half, t.center[1] - t.half, t.center[2] - t.half};
    for (int T = 0; T < nl; ++T) {
      int b = t.leaf_box(T);
      double bc[3];
      for (int d = 0; d < 3; ++d) bc[d] = origin[d] + (t.cell[3 * b + d] + 0.5) * 2.0 * half_L;
      for (int p = t.leaf_lo[T]; p < t.leaf_lo[T + 1]; ++p) {
        int i = (int)t.perm[p];
        double a[3] = {g.cx[i], g.cy[i], g.cz[i]};
        for (int d = 0; d < 3; ++d)
          interp_weights((a[d] - bc[d]) / half_L, n, nodes.data(), &hw1d[(size_t)p * 3 * n + d * n]);
      }
    }
    // M2M 1-D matrices, bbfmm.rs:89-112: mat[m*n + k] = S_n(cc + x_k / 2, x_m).
    std::vector<double> hm2m((size_t)24 * n * n);
    for (int oz = 0; oz < 2; ++oz)
      for (int oy = 0; oy < 2; ++oy)
        for (int ox = 0; ox < 2; ++ox) {
          int o = ox + 2 * oy + 4 * oz;
          double cc[3] = {(2.0 * ox - 1.0) * 0.5, (2.0 * oy - 1.0) * 0.5, (2.0 * oz - 1.0) * 0.5};
          for (int axis = 0; axis < 3; ++axis)
            for (int k = 0; k < n; ++k) {
              double w[32];
              interp_weights(cc[axis] + 0.5 * nodes[k], n, nodes.data(), w);
              for (int m = 0; m < n; ++m) hm2m[(size_t)(o * 3 + axis) * n * n + m * n + k] = w[m];
            }
        }
    // Uncompressed M2L matrices on the reference box, bbfmm.rs:156-166.
    std::vector<double> hG((size_t)343 * n3p * n3p, 0.0);
    for (int o0 = -3; o0 <= 3; ++o0)
      for (int o1 = -3; o1 <= 3; ++o1)
        for (int o2 = -3; o2 <= 3; ++o2) {
          if (std::max(std::abs(o0), std::max(std::abs(o1), std::abs(o2))) <= 1) continue;
          int oi = (o0 + 3) * 49 + (o1 + 3) * 7 + (o2 + 3);
          double *G = &hG[(size_t)oi * n3p * n3p];
          for (int m = 0; m < n3; ++m) {
            int ma = m / (n * n), mb = (m / n) % n, mc = m % n;
            for (int l = 0; l < n3; ++l) {
              int la = l / (n * n), lb = (l / n) % n, lc = l % n;
              double dx = nodes[ma] - nodes[la] - 2.0 * o0;
              double dy = nodes[mb] - nodes[lb] - 2.0 * o1;
              double dz = nodes[mc] - nodes[lc] - 2.0 * o2;
              double r = std::sqrt(dx * dx + dy * dy + dz * dz);
              G[(size_t)m * n3p + l] = r > 0.0 ? 1.0 / r : 0.0;
            }
          }
        }
    int nb = t.n_boxes();
    std::vector<int> hlevel(nb), htargets;
    for (int l = 0; l <= L; ++l)
      for (int b = lvl_off[l]; b < lvl_off[l + 1]; ++b) hlevel[b] = l;
    for (int b = 0; b < nb; ++b)
      if (t.il_off[b + 1] > t.il_off[b]) htargets.push_back(b);
    pairs = (long)t.il_src.size();
    std::vector<double> hih(L + 1);
    for (int l = 0; l <= L; ++l) hih[l] = (double)(1u << l);  // 1/half_l in units of L0
    parent = to_device(t.parent, s);
    octant = to_device(t.octant, s);
    child_off = to_device(t.child_off, s);
    child_idx = to_device(t.child_idx, s);
    targets = to_device(htargets, s);
    il_off = to_device(t.il_off, s);
    il_src = to_device(t.il_src, s);
    il_oidx = to_device(t.il_oidx, s);
    box_level = to_device(hlevel, s);
    inv_half = to_device(hih, s);
    w1d64 = to_device(hw1d, s);
    m2m64 = to_device(hm2m, s);
    G64 = to_device(hG, s);
    if (need32) {
      w1d32 = to_device(cast_vec<float>(hw1d), s);
      m2m32 = to_device(cast_vec<float>(hm2m), s);
      G32 = to_device(cast_vec<float>(hG), s);
      std::vector<uint16_t> hb(hG.size());
      for (size_t e = 0; e < hG.size(); ++e) hb[e] = bf16_bits((float)hG[e]);
      Gbf = to_device(hb, s);
      if (m2l_inner == GP_M2L_BF16_VALU) {
        std::vector<float> hr(hG.size());
        for (size_t e = 0; e < hG.size(); ++e) {
          const uint32_t u = (uint32_t)hb[e] << 16;
          std::memcpy(&hr[e], &u, 4);
        }
        Gr32 = to_device(hr, s);
      }
    }
    auto fill = [&](auto &f) {
      f.n = n;
      f.L = L;
      f.n_boxes = nb;
      f.n_targets = (int)htargets.size();
      f.n_pairs = pairs;
      f.n_panels = N;
      f.lvl_off = lvl_off.data();
      f.leaf_lo = leaf_lo.p;
      f.parent = parent.p;
      f.octant = octant.p;
      f.child_off = child_off.p;
      f.child_idx = child_idx.p;
      f.targets = targets.p;
      f.il_off = il_off.p;
      f.il_src = il_src.p;
      f.il_oidx = il_oidx.p;
      f.box_level = box_level.p;
      f.inv_half = inv_half.p;
      f.area = area.p;
      f.n3p = n3p;
    };
    fill(f64);
    f64.w1d = w1d64.p;
    f64.m2m1d = m2m64.p;
    f64.G = G64.p;
    fill(f32);
    f32.w1d = w1d32.p;
    f32.m2m1d = m2m32.p;
    f32.G = G32.p;
    f32.Gbf = Gbf.p;
    f32.Gr = Gr32.p;
    std::vector<int> hco, hci;
    cond_csr(g, t.perm.data(), hco, hci);
    coff = to_device(hco, s);
    cidx = to_device(hci, s);
  }

  void ensure_work(int ld_) {
    if (ld_ == ld) return;
    ld = ld_;
    const size_t vec = (size_t)N * ld;
    const size_t mexp = (size_t)t.n_boxes() * n3p * ld;   // multipoles, padded
    const size_t lexp = (size_t)t.n_boxes() * n * n * n * ld;
    z64.alloc(vec);
    z64.zero(s);
    M64.alloc(mexp);
    M64.zero(s);
    Lc64.alloc(lexp);
    if (have32) {
      z32.alloc(vec);
      z32.zero(s);
      M32.alloc(mexp);
      M32.zero(s);
      Lc32.alloc(lexp);
      if (m2l_inner == GP_M2L_BF16 && ld % 16 == 0 && ld <= 128) {
        Mbf.alloc(mexp);
        Mbf.zero(s);
      }
    }
  }

  // y = P' M^-1 v (precond) or y = P' v, in FP64 (prec 0) or FP32 (prec 1).
  template <typename TV>
  void apply(const TV *v, bool precond, int nr, double *y, int prec) {
    const double *dinv = precond ? inv_diag.p : nullptr;
    ++matvecs;
    if (prec == 0) {
      if constexpr (sizeof(TV) != 8) throw Error(GP_ERR_INTERNAL, "FP64 operator on an FP32 basis");
      else kr_precond_cast(v, dinv, z64.p, N, ld, nr, s);
      launch_near_apply(nt, nf64.p, z64.p, ld, y, ld, nr, s);
      run_far_field(f64, z64.p, ld, nr, M64.p, Lc64.p, ld, m2l_outer, y, ld, s);
    } else {
      if (!have32) throw Error(GP_ERR_INTERNAL, "FP32 operator not assembled");
      kr_precond_cast(v, dinv, z32.p, N, ld, nr, s);
      launch_near_apply(nt, nf32.p, z32.p, ld, y, ld, nr, s);
      const int v = (m2l_inner == GP_M2L_BF16 && Mbf.p) ? GP_M2L_BF16
                    : m2l_inner == GP_M2L_BF16_VALU    ? GP_M2L_BF16_VALU
                                                       : GP_M2L_FP32;
      run_far_field(f32, z32.p, ld, nr, M32.p, Mbf.p, Lc32.p, ld, v, y, ld, s);
    }
  }
};

// Pinned host scratch for the per-iteration convergence flags.
struct Pinned {
  void *p = nullptr;
  explicit Pinned(size_t bytes) { GP_HIP(hipHostMalloc(&p, bytes, hipHostMallocDefault)); }
  ~Pinned() {
    if (p) (void)hipHostFree(p);
  }
  template <typename T>
  T *as() {
    return reinterpret_cast<T *>(p);
  }
};

struct GmresStats {
  std::vector<int> iters;
  std::vector<double> resid;
  std::vector<char> conv;
};

// Restarted GMRES(m) with right Jacobi preconditioning for nr columns of b
// (device, N x ld) starting from x = 0 (x must be zero on entry), per column
// exactly the control flow of krylov.rs:50-191: stop a cycle at the residual
// estimate, check the true residual when the estimate is within 10 tol.
// Orthogonalisation is classical Gram-Schmidt applied twice (CGS2, passes =
// 2), the block-friendly equivalent of the reference's modified Gram-Schmidt;
// the inner solves of the mixed solver (tolerance ~1e-2, corrected by the
// FP64 refinement) take one pass and keep the basis V in FP32 (TV = float).
template <typename TV>
void gmres_block(FmmSystem &sys, int prec, const double *b, double *x, int nr, double tol,
                 int m, int max_restarts, bool true_check, int passes,
                 const std::vector<int> &live_in, GmresStats &st) {
  hipStream_t s = sys.s;
  const int N = sys.N, ld = sys.ld;
  const size_t vec = (size_t)N * ld;
  int vcap = std::min(m, 32) + 1;
  DBuf<TV> V(vcap * vec);
  DBuf<double> w(vec), rv(vec), ax(vec);
  DBuf<double> partial(kr_partial_size(vcap, ld));
  DBuf<double> h1((size_t)(m + 1) * ld), h2((size_t)(m + 1) * ld), hn(ld);
  DBuf<double> H((size_t)(m + 1) * m * ld), cs((size_t)m * ld), sn((size_t)m * ld),
      g((size_t)(m + 1) * ld), Y((size_t)m * ld);
  DBuf<double> bnorm(ld), beta(ld), inv_beta(ld), est(ld), inv_h(ld);
  DBuf<int> live(ld), active(ld), kused(ld);
  Pinned pin(ld * (2 * sizeof(int) + 2 * sizeof(double)));
  int *h_active = pin.as<int>();
  int *h_kused = h_active + ld;
  double *h_est = reinterpret_cast<double *>(h_kused + ld);
  double *h_aux = h_est + ld;

  auto ensure_V = [&](int nv) {
    if (nv <= vcap) return;
    int nc = std::min(m + 1, std::max(nv, 2 * vcap));
    DBuf<TV> V2((size_t)nc * vec);
    GP_HIP(hipMemcpyAsync(V2.p, V.p, (size_t)vcap * vec * sizeof(TV),
                          hipMemcpyDeviceToDevice, s));
    GP_HIP(hipStreamSynchronize(s));  // the old basis is freed below
    V = std::move(V2);
    partial.alloc(kr_partial_size(nc, ld));
    vcap = nc;
  };

  st.iters.assign(nr, 0);
  st.resid.assign(nr, 0.0);
  st.conv.assign(nr, 0);
  std::vector<int> done(nr);
  for (int r = 0; r < nr; ++r) done[r] = !live_in[r];
  // bnorm = max(||b||, MIN_POSITIVE)   (krylov.rs:58)
  kr_dots(b, 0, 1, b, N, ld, nr, partial.p, bnorm.p, true, s);
  std::vector<double> hb(nr);
  GP_HIP(hipMemcpyAsync(hb.data(), bnorm.p, nr * sizeof(double), hipMemcpyDeviceToHost, s));
  GP_HIP(hipStreamSynchronize(s));
  for (auto &v : hb) v = std::max(v, kMinPositive);
  GP_HIP(hipMemcpyAsync(bnorm.p, hb.data(), nr * sizeof(double), hipMemcpyHostToDevice, s));

  bool x_zero = true;
  const size_t sv = vec;
  for (int cycle = 0; cycle < max_restarts; ++cycle) {
    if (x_zero) {
      GP_HIP(hipMemcpyAsync(rv.p, b, vec * sizeof(double), hipMemcpyDeviceToDevice, s));
    } else {
      sys.apply(x, false, nr, ax.p, prec);
      kr_sub(b, ax.p, rv.p, N, ld, nr, s);
    }
    kr_dots(rv.p, 0, 1, rv.p, N, ld, nr, partial.p, beta.p, true, s);
    std::vector<int> hl(ld, 0);
    for (int r = 0; r < nr; ++r) hl[r] = !done[r];
    GP_HIP(hipMemcpyAsync(live.p, hl.data(), ld * sizeof(int), hipMemcpyHostToDevice, s));
    kr_gmres_start(m, nr, ld, beta.p, bnorm.p, tol, live.p, active.p, kused.p, g.p, inv_beta.p,
                   est.p, s);
    GP_HIP(hipMemcpyAsync(h_active, active.p, nr * sizeof(int), hipMemcpyDeviceToHost, s));
    GP_HIP(hipMemcpyAsync(h_est, est.p, nr * sizeof(double), hipMemcpyDeviceToHost, s));
    GP_HIP(hipStreamSynchronize(s));
    std::vector<int> cyc(nr, 0);
    int nact = 0;
    for (int r = 0; r < nr; ++r) {
      if (!done[r] && !h_active[r]) {
        done[r] = 1;
        st.conv[r] = 1;
        st.resid[r] = h_est[r];
      }
      cyc[r] = h_active[r];
      nact += h_active[r];
    }
    if (nact == 0) break;
    kr_scale_cols(rv.p, V.p, inv_beta.p, active.p, N, ld, nr, s);
    int ksteps = 0;
    for (int k = 0; k < m; ++k) {
      ensure_V(k + 2);
      sys.apply(V.p + (size_t)k * sv, true, nr, w.p, prec);
      kr_dots(V.p, sv, k + 1, w.p, N, ld, nr, partial.p, h1.p, false, s);
      kr_axpy_multi(w.p, V.p, sv, h1.p, k + 1, N, ld, nr, -1.0, s);
      if (passes > 1) {
        kr_dots(V.p, sv, k + 1, w.p, N, ld, nr, partial.p, h2.p, false, s);
        kr_axpy_multi(w.p, V.p, sv, h2.p, k + 1, N, ld, nr, -1.0, s);
      }
      kr_dots(w.p, 0, 1, w.p, N, ld, nr, partial.p, hn.p, true, s);
      kr_gmres_givens(k, m, nr, ld, h1.p, passes > 1 ? h2.p : nullptr, hn.p, H.p, cs.p, sn.p, g.p,
                      bnorm.p, tol, active.p, kused.p, est.p, inv_h.p, s);
      kr_scale_cols(w.p, V.p + (size_t)(k + 1) * sv, inv_h.p, nullptr, N, ld, nr, s);
      ksteps = k + 1;
      GP_HIP(hipMemcpyAsync(h_active, active.p, nr * sizeof(int), hipMemcpyDeviceToHost, s));
      GP_HIP(hipStreamSynchronize(s));
      int any = 0;
      for (int r = 0; r < nr; ++r) any |= h_active[r];
      if (k % 25 == 24) {
        int na = 0;
        for (int r = 0; r < nr; ++r) na += h_active[r];
        GP_LOG("gmres prec %d cycle %d iteration %d: %d of %d columns active", prec, cycle, k + 1,
               na, nr);
      }
      if (!any) break;
    }
    kr_gmres_backsolve(m, nr, ld, H.p, g.p, kused.p, Y.p, s);
    kr_gmres_update_x(x, V.p, sv, Y.p, kused.p, ksteps, sys.inv_diag.p, N, ld, nr, s);
    x_zero = false;
    GP_HIP(hipMemcpyAsync(h_kused, kused.p, nr * sizeof(int), hipMemcpyDeviceToHost, s));
    GP_HIP(hipMemcpyAsync(h_est, est.p, nr * sizeof(double), hipMemcpyDeviceToHost, s));
    GP_HIP(hipStreamSynchronize(s));
    bool check = false;
    for (int r = 0; r < nr; ++r)
      if (cyc[r]) {
        st.iters[r] += h_kused[r];
        if (h_est[r] <= tol * 10.0) check = true;
      }
    if (true_check) {
      if (check) {
        sys.apply(x, false, nr, ax.p, prec);
        kr_sub(b, ax.p, rv.p, N, ld, nr, s);
        kr_dots(rv.p, 0, 1, rv.p, N, ld, nr, partial.p, beta.p, true, s);
        GP_HIP(hipMemcpyAsync(h_aux, beta.p, nr * sizeof(double), hipMemcpyDeviceToHost, s));
        GP_HIP(hipStreamSynchronize(s));
        for (int r = 0; r < nr; ++r)
          if (cyc[r] && h_est[r] <= tol * 10.0) {
            double rr = h_aux[r] / hb[r];
            if (rr <= tol) {
              done[r] = 1;
              st.conv[r] = 1;
              st.resid[r] = rr;
            }
          }
      }
    } else {
      for (int r = 0; r < nr; ++r)
        if (cyc[r] && h_est[r] <= tol) {
          done[r] = 1;
          st.conv[r] = 1;
          st.resid[r] = h_est[r];
        }
    }
    bool all = true;
    for (int r = 0; r < nr; ++r) all &= (bool)done[r];
    if (all) break;
  }
  // Columns left unconverged: report their true residual (krylov.rs:185-190).
  bool any_left = false;
  for (int r = 0; r < nr; ++r) any_left |= !done[r];
  if (any_left) {
    sys.apply(x, false, nr, ax.p, prec);
    kr_sub(b, ax.p, rv.p, N, ld, nr, s);
    kr_dots(rv.p, 0, 1, rv.p, N, ld, nr, partial.p, beta.p, true, s);
    GP_HIP(hipMemcpyAsync(h_aux, beta.p, nr * sizeof(double), hipMemcpyDeviceToHost, s));
    GP_HIP(hipStreamSynchronize(s));
    for (int r = 0; r < nr; ++r)
      if (!done[r]) st.resid[r] = h_aux[r] / hb[r];
  }
}

void solve_fmm(gp_ctx *ctx, const PanelGeom &g, const gp_cap_opts &o, double *C,
               gp_cap_report *rep) {
  hipStream_t s = ctx->stream;
  double t0 = now_s();
  const bool mixed = o.precision == GP_PREC_MIXED;
  FmmSystem sys;
  sys.setup(g, o, mixed, s);
  const int nc = g.ncond;
  const int jlo = std::min((int)o.rhs_first, nc);
  const int jhi = o.rhs_count ? std::min(nc, jlo + (int)o.rhs_count) : nc;
  const int nsolve = std::max(1, jhi - jlo);
  // Blocks of at most 128 columns: the matrix-core M2L holds one block.
  int blk = o.rhs_block ? (int)o.rhs_block : nsolve;
  blk = std::max(1, std::min(blk, std::min(nsolve, 128)));
  const int ld = blk == 1 ? 1 : round_up(blk, 16);
  sys.ensure_work(ld);
  GP_HIP(hipStreamSynchronize(s));
  double t1 = now_s();
  GP_LOG("setup %.3f s: %d panels, depth %d, %d boxes, near nnz %lld, %lld M2L pairs, ld %d",
         t1 - t0, sys.N, sys.L, sys.t.n_boxes(), (long long)sys.nnz, (long long)sys.pairs, ld);

  const size_t vec = (size_t)sys.N * ld;
  DBuf<double> b(vec), x(vec), dCb((size_t)nc * ld);
  DBuf<double> r(vec), dx(vec), ax(vec), part(kr_partial_size(1, ld)), rn(ld);
  std::vector<double> hC((size_t)nc * nc, 0.0), hblk((size_t)nc * ld);
  int max_iters = 0, refine_steps = 0;
  double worst = 0.0;
  bool conv_all = true;
  const int m = std::max(1, (int)o.restart), mr = std::max(1, (int)o.max_restarts);
  for (int j0 = jlo; j0 < jhi; j0 += blk) {
    const int nr = std::min(blk, jhi - j0);
    kr_unit_rhs(sys.cond.p, j0, sys.N, ld, nr, b.p, s);
    x.zero(s);
    std::vector<int> live(nr, 1);
    if (!mixed) {
      GmresStats st;
      gmres_block<double>(sys, 0, b.p, x.p, nr, o.tol, m, mr, true, 2, live, st);
      for (int q = 0; q < nr; ++q) {
        max_iters = std::max(max_iters, st.iters[q]);
        worst = std::max(worst, st.resid[q]);
        conv_all &= (bool)st.conv[q];
      }
    } else {
      // FP64 iterative refinement around the FP32 inner GMRES.
      GP_HIP(hipMemcpyAsync(r.p, b.p, vec * sizeof(double), hipMemcpyDeviceToDevice, s));
      std::vector<double> hbn(nr), hrn(nr);
      kr_dots(b.p, 0, 1, b.p, sys.N, ld, nr, part.p, rn.p, true, s);
      GP_HIP(hipMemcpyAsync(hbn.data(), rn.p, nr * sizeof(double), hipMemcpyDeviceToHost, s));
      GP_HIP(hipStreamSynchronize(s));
      for (auto &v : hbn) v = std::max(v, kMinPositive);
      std::vector<int> it_sum(nr, 0), done(nr, 0);
      int steps = 0;
      const int max_ref = std::max(1, (int)o.max_refine);
      const int passes = o.inner_passes >= 2 ? 2 : 1;
      for (;;) {
        kr_dots(r.p, 0, 1, r.p, sys.N, ld, nr, part.p, rn.p, true, s);
        GP_HIP(hipMemcpyAsync(hrn.data(), rn.p, nr * sizeof(double), hipMemcpyDeviceToHost, s));
        GP_HIP(hipStreamSynchronize(s));
        bool all = true;
        for (int q = 0; q < nr; ++q) {
          if (hrn[q] / hbn[q] <= o.tol) done[q] = 1;
          all &= (bool)done[q];
        }
        if (all || steps >= max_ref) break;
        std::vector<int> lv(nr);
        for (int q = 0; q < nr; ++q) lv[q] = !done[q];
        dx.zero(s);
        GmresStats st;
        if (o.inner_basis == GP_BASIS_FP64)
          gmres_block<double>(sys, 1, r.p, dx.p, nr, o.inner_tol, m, mr, false, passes, lv, st);
        else
          gmres_block<float>(sys, 1, r.p, dx.p, nr, o.inner_tol, m, mr, false, passes, lv, st);
        for (int q = 0; q < nr; ++q) it_sum[q] += st.iters[q];
        kr_axpy(x.p, dx.p, 1.0, sys.N, ld, nr, s);
        sys.apply(x.p, false, nr, ax.p, 0);
        kr_sub(b.p, ax.p, r.p, sys.N, ld, nr, s);
        ++steps;
        GP_LOG("refinement step %d: inner iterations up to %d", steps,
               *std::max_element(st.iters.begin(), st.iters.end()));
      }
      refine_steps = std::max(refine_steps, steps);
      for (int q = 0; q < nr; ++q) {
        max_iters = std::max(max_iters, it_sum[q]);
        worst = std::max(worst, hrn[q] / hbn[q]);
        conv_all &= (bool)done[q];
      }
    }
    launch_cap_accumulate(nc, nr, ld, sys.coff.p, sys.cidx.p, sys.area.p, x.p, dCb.p, s);
    GP_HIP(hipMemcpyAsync(hblk.data(), dCb.p, (size_t)nc * nr * sizeof(double),
                          hipMemcpyDeviceToHost, s));
    GP_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < nc; ++i)
      for (int q = 0; q < nr; ++q) hC[(size_t)i * nc + j0 + q] = hblk[(size_t)i * nr + q];
  }
  const double scale = 4.0 * kPi * kEps0 * sys.L0;
  for (size_t e = 0; e < hC.size(); ++e) C[e] = hC[e] * scale;
  double t2 = now_s();
  if (rep) {
    rep->method_used = GP_METHOD_FMM;
    rep->precision_used = mixed ? GP_PREC_MIXED : GP_PREC_FP64;
    rep->iterations = (uint32_t)max_iters;
    rep->matvecs = (uint32_t)sys.matvecs;
    rep->refine_steps = (uint32_t)refine_steps;
    rep->converged = conv_all ? 1u : 0u;
    rep->max_level = (uint32_t)sys.L;
    rep->n_leaves = (uint32_t)sys.t.n_leaves();
    rep->n_boxes = (uint32_t)sys.t.n_boxes();
    rep->fmm_order = (uint32_t)sys.n;
    rep->near_nnz = (uint64_t)sys.nnz;
    rep->m2l_pairs = (uint64_t)sys.pairs;
    rep->final_rel_residual = worst;
    rep->seconds_setup = t1 - t0;
    rep->seconds_solve = t2 - t1;
  }
  if (!conv_all && o.strict)
    throw Error(GP_ERR_NOT_CONVERGED, "GMRES did not reach tol (worst relative residual " +
                                          std::to_string(worst) + ")");
}

}  // namespace

void cap_solve(gp_ctx *ctx, const PanelGeom &g, const gp_cap_opts &o, double *C,
               gp_cap_report *rep, const double *eps_panel) {
  double t0 = now_s();
  if (!(o.tol > 0.0) || !std::isfinite(o.tol)) throw Error(GP_ERR_INVALID_INPUT, "tol must be > 0");
  if (o.precision != GP_PREC_FP64 && o.precision != GP_PREC_MIXED)
    throw Error(GP_ERR_INVALID_INPUT, "unknown precision");
  uint32_t method = o.method;
  uint32_t limit = o.dense_limit ? o.dense_limit : 500u;
  if (method == GP_METHOD_AUTO) method = ((uint32_t)g.n < limit) ? GP_METHOD_DENSE : GP_METHOD_FMM;
  if (eps_panel && method != GP_METHOD_DENSE)
    throw Error(GP_ERR_PANEL_CAPACITY, "inhomogeneous dielectric needs fewer than " +
                                           std::to_string(limit) + " panels (dense only)");
  if (rep) {
    std::memset(rep, 0, sizeof(*rep));
    rep->n_panels = (uint32_t)g.n;
    rep->n_conductors = (uint32_t)g.ncond;
  }
  if (method == GP_METHOD_DENSE)
    solve_dense(ctx, g, C, rep, eps_panel);
  else if (method == GP_METHOD_FMM)
    solve_fmm(ctx, g, o, C, rep);
  else
    throw Error(GP_ERR_INVALID_INPUT, "unknown method");
  for (size_t e = 0; e < (size_t)g.ncond * g.ncond; ++e)
    if (!std::isfinite(C[e])) throw Error(GP_ERR_NON_FINITE, "solver returned a non-finite capacitance");
  if (rep) rep->seconds_total = now_s() - t0;
}

void fmm_apply_host(gp_ctx *ctx, const PanelGeom &g, const gp_cap_opts &o, int nr,
                    const double *X, double *Y, int precision) {
  hipStream_t s = ctx->stream;
  FmmSystem sys;
  sys.setup(g, o, precision == 1, s);
  int ld = nr == 1 ? 1 : round_up(nr, 16);
  sys.ensure_work(ld);
  const int N = g.n;
  std::vector<double> hx((size_t)N * ld, 0.0), hy((size_t)N * ld);
  for (int p = 0; p < N; ++p)
    for (int r = 0; r < nr; ++r) hx[(size_t)p * ld + r] = X[(size_t)sys.t.perm[p] * nr + r];
  auto dxv = to_device(hx, s);
  DBuf<double> dy((size_t)N * ld);
  sys.apply(dxv.p, false, nr, dy.p, precision);
  GP_HIP(hipMemcpyAsync(hy.data(), dy.p, hy.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  GP_HIP(hipStreamSynchronize(s));
  // back to SI: P = k_e L0 P'
  const double sc = sys.L0 / (4.0 * kPi * kEps0);
  for (int p = 0; p < N; ++p)
    for (int r = 0; r < nr; ++r) Y[(size_t)sys.t.perm[p] * nr + r] = hy[(size_t)p * ld + r] * sc;
}

}  // namespace gp
