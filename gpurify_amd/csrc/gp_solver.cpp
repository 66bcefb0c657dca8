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

#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <climits>
#include <exception>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>

#include "gp_kernels.h"
#include "gp_parallel.h"
#include "gp_m2l_cpqr.h"
#include "gp_m2l_svd.h"

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

// Arnoldi steps per replayed graph of a batched GMRES (one host check of the
// convergence flags per chunk).
constexpr int kBatchGraphSteps = 4;

// Sensitivity parameters evaluated together when they move the same panels.
#ifndef GP_SENS_GROUP
#define GP_SENS_GROUP 2
#endif
constexpr int kSensGroup = GP_SENS_GROUP;

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
void scale_into(const PanelGeom &g, const uint32_t *perm, const double c[3], double L0,
                double *verts, double *cen, double *area, uint32_t *cond, uint32_t cond0) {
  const int n = g.n;
  const double inv = 1.0 / L0;
  for (int p = 0; p < n; ++p) {
    int i = perm ? (int)perm[p] : p;
    for (int k = 0; k < 4; ++k)
      for (int d = 0; d < 3; ++d)
        verts[(size_t)p * 12 + 3 * k + d] = (g.verts[(size_t)i * 12 + 3 * k + d] - c[d]) * inv;
    cen[3 * p] = (g.cx[i] - c[0]) * inv;
    cen[3 * p + 1] = (g.cy[i] - c[1]) * inv;
    cen[3 * p + 2] = (g.cz[i] - c[2]) * inv;
    area[p] = g.area[i] * inv * inv;
    cond[p] = g.cond[i] + cond0;
  }
}

void scaled_arrays(const PanelGeom &g, const uint32_t *perm, const double c[3], double L0,
                   std::vector<double> &verts, std::vector<double> &cen,
                   std::vector<double> &area, std::vector<uint32_t> &cond) {
  int n = g.n;
  verts.resize((size_t)n * 12);
  cen.resize((size_t)n * 3);
  area.resize(n);
  cond.resize(n);
  scale_into(g, perm, c, L0, verts.data(), cen.data(), area.data(), cond.data(), 0);
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

// ------------------------------------------------------------------ ground
// A grounded plane z = zg by the method of images (GP_GROUND_IMAGE; the
// reference meshes a finite substrate plate instead, bridge.rs:346-379):
// every vertex must lie above the plane.
void check_ground(const PanelGeom &g, double zg) {
  if (!std::isfinite(zg)) throw Error(GP_ERR_INVALID_INPUT, "ground_z must be finite");
  for (int i = 0; i < g.n; ++i)
    for (int k = 0; k < 4; ++k)
      if (!(g.verts[(size_t)i * 12 + 3 * k + 2] > zg))
        throw Error(GP_ERR_INVALID_INPUT, "panel " + std::to_string(i) +
                                              " has a vertex on or below the ground plane");
}

// The panels followed by their mirror images in z = zg (same conductors,
// sign -1): the image system as a signed panel set of 2n panels.
struct Mirrored {
  std::vector<double> v, sign;
  std::vector<uint32_t> c;
  PanelGeom g;
};
void mirror_geom(const PanelGeom &g, double zg, Mirrored &m) {
  const int n = g.n;
  m.v.resize((size_t)n * 24);
  std::copy(g.verts, g.verts + (size_t)n * 12, m.v.begin());
  m.c.assign(g.cond, g.cond + n);
  m.c.insert(m.c.end(), g.cond, g.cond + n);
  m.sign.assign((size_t)2 * n, 1.0);
  std::fill(m.sign.begin() + n, m.sign.end(), -1.0);
  m.g = PanelGeom();
  m.g.n = 2 * n;
  m.g.ncond = g.ncond;
  m.g.cx = g.cx, m.g.cy = g.cy, m.g.cz = g.cz, m.g.area = g.area;
  m.g.cx.resize(2 * n), m.g.cy.resize(2 * n), m.g.cz.resize(2 * n), m.g.area.resize(2 * n);
  parallel_for(n, [&](int i) {
    double *q = &m.v[(size_t)(n + i) * 12];
    for (int k = 0; k < 4; ++k) {
      q[3 * k] = g.verts[(size_t)i * 12 + 3 * k];
      q[3 * k + 1] = g.verts[(size_t)i * 12 + 3 * k + 1];
      q[3 * k + 2] = 2.0 * zg - g.verts[(size_t)i * 12 + 3 * k + 2];
    }
    double c[3], a;
    panel_props(q, c, a);
    m.g.cx[n + i] = c[0], m.g.cy[n + i] = c[1], m.g.cz[n + i] = c[2], m.g.area[n + i] = a;
  });
  m.g.verts = m.v.data();
  m.g.cond = m.c.data();
  m.g.sign = m.sign.data();
}

// ------------------------------------------------------------------ dense
// Above this size the one-work-group LU is slower than the host: dense
// systems then factor with rocSOLVER (solve_dense_roles).
constexpr int kOneGroupLU = 1024;

void solve_dense(gp_ctx *ctx, const PanelGeom &g, double *C, gp_cap_report *rep,
                 const double *eps_panel, const double *ground_z = nullptr) {
  if (g.n > kOneGroupLU) {
    std::vector<gp_panel_role> roles(g.n);
    for (int i = 0; i < g.n; ++i) {
      roles[i] = gp_panel_role{};
      roles[i].kind = GP_ROLE_CONDUCTOR;
      roles[i].conductor = g.cond[i];
      roles[i].eps_surrounding = eps_panel ? eps_panel[i] : 1.0;
    }
    solve_dense_roles(ctx, g, roles.data(), C, rep, ground_z);
    return;
  }
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

  auto d_v = upload_panels(hv, s);
  auto d_c = upload_panels(hc, s);
  auto d_w = upload_panels(w, s);
  auto d_B = to_device(B, s);
  auto d_coff = to_device(coff, s);
  auto d_cidx = to_device(cidx, s);
  DBuf<double> d_P((size_t)n * n), d_work((size_t)n * nc), d_C((size_t)nc * nc);
  DBuf<int> d_piv(n), d_info(1);
  launch_dense_assemble(n, d_v.p, d_c.p, d_P.p, s, ground_z != nullptr,
                        ground_z ? 2.0 * (*ground_z - c[2]) / L0 : 0.0);
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

// The dense problems idx of a batch in two launches (assembly, LU + solve;
// one work-group per problem), then C per problem on the host in the
// reference's summation order.
void solve_dense_batch(gp_ctx *ctx, const std::vector<PanelGeom> &gs, const std::vector<int> &idx,
                       const std::vector<double *> &C, gp_cap_report *reps,
                       const double *ground_z = nullptr) {
  hipStream_t s = ctx->stream;
  const double t0 = now_s();
  const int B = (int)idx.size();
  std::vector<int> hn(B), hrhs(B), hnoff(B + 1, 0);
  std::vector<long> haoff(B + 1, 0), hboff(B + 1, 0);
  std::vector<double> L0(B);
  for (int q = 0; q < B; ++q) {
    const PanelGeom &g = gs[idx[q]];
    hn[q] = g.n;
    hrhs[q] = g.ncond;
    hnoff[q + 1] = hnoff[q] + g.n;
    haoff[q + 1] = haoff[q] + (long)g.n * g.n;
    hboff[q + 1] = hboff[q] + (long)g.n * g.ncond;
  }
  std::vector<double> hv((size_t)hnoff[B] * 12), hc((size_t)hnoff[B] * 3), hB(hboff[B], 0.0);
  std::vector<std::vector<double>> area(B);
  std::vector<double> hzm2(B, 0.0);
  for (int q = 0; q < B; ++q) {
    const PanelGeom &g = gs[idx[q]];
    double c[3];
    root_cube(g, c, L0[q]);
    if (ground_z) hzm2[q] = 2.0 * (*ground_z - c[2]) / L0[q];
    std::vector<double> v, ce;
    std::vector<uint32_t> cd;
    scaled_arrays(g, nullptr, c, L0[q], v, ce, area[q], cd);
    std::copy(v.begin(), v.end(), hv.begin() + (size_t)hnoff[q] * 12);
    std::copy(ce.begin(), ce.end(), hc.begin() + (size_t)hnoff[q] * 3);
    for (int i = 0; i < g.n; ++i) hB[hboff[q] + (size_t)i * g.ncond + g.cond[i]] = 1.0;
  }
  auto d_n = to_device(hn, s);
  auto d_r = to_device(hrhs, s);
  auto d_noff = to_device(hnoff, s);
  auto d_aoff = to_device(haoff, s);
  auto d_boff = to_device(hboff, s);
  auto d_v = upload_panels(hv, s);
  auto d_c = upload_panels(hc, s);
  auto d_B = to_device(hB, s);
  DBuf<double> d_zm2;
  if (ground_z) d_zm2 = to_device(hzm2, s);
  DBuf<double> d_A(haoff[B]), d_work(hboff[B]);
  DBuf<int> d_piv(hnoff[B]), d_info(B);
  DenseBatch db;
  db.count = B;
  db.n = d_n.p;
  db.nrhs = d_r.p;
  db.n_off = d_noff.p;
  db.a_off = d_aoff.p;
  db.b_off = d_boff.p;
  db.verts = d_v.p;
  db.cen = d_c.p;
  db.A = d_A.p;
  db.B = d_B.p;
  db.work = d_work.p;
  db.piv = d_piv.p;
  db.info = d_info.p;
  db.zm2 = ground_z ? d_zm2.p : nullptr;
  launch_dense_batch(db, s);
  std::vector<int> info(B);
  GP_HIP(hipMemcpyAsync(info.data(), d_info.p, B * sizeof(int), hipMemcpyDeviceToHost, s));
  GP_HIP(hipMemcpyAsync(hB.data(), d_B.p, hB.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  GP_HIP(hipStreamSynchronize(s));
  const double t1 = now_s();
  for (int q = 0; q < B; ++q) {
    const PanelGeom &g = gs[idx[q]];
    if (info[q]) throw Error(GP_ERR_SINGULAR, "potential-coefficient matrix of problem " +
                                                  std::to_string(idx[q]) + " is singular");
    const int nc = g.ncond;
    const double scale = 4.0 * kPi * kEps0 * L0[q];
    double *Cp = C[idx[q]];
    std::fill(Cp, Cp + (size_t)nc * nc, 0.0);
    // accumulate_c (cap/solver.rs:67-71): panels in ascending order
    std::vector<double> acc((size_t)nc * nc, 0.0);
    for (int k = 0; k < g.n; ++k)
      for (int j = 0; j < nc; ++j)
        acc[(size_t)g.cond[k] * nc + j] += area[q][k] * hB[hboff[q] + (size_t)k * nc + j];
    for (size_t e = 0; e < acc.size(); ++e) Cp[e] = acc[e] * scale;
    if (reps) {
      gp_cap_report &r = reps[idx[q]];
      r.method_used = GP_METHOD_DENSE;
      r.precision_used = GP_PREC_FP64;
      r.converged = 1;
      r.seconds_setup = 0.0;
      r.seconds_solve = t1 - t0;
    }
  }
}

// ------------------------------------------------------------------ FMM
// Dielectric-interface rows of an FMM solve (caller order): the flag, the
// oriented normal and the unit-frame jump term 2 pi (eps_out + eps_in) /
// (eps_out - eps_in) of every interface panel.
struct DielRows {
  std::vector<unsigned char> iface;
  std::vector<double> nrm, diag;
};

// Low-rank M2L tables per (order, rank), computed once per process: they
// depend on the Chebyshev order only (reference box, unit half-width).
struct LowRank {
  std::vector<double> U, Gc;
};
std::shared_ptr<const LowRank> low_rank_tables(int n, int n3p, int kr, bool compressed,
                                               const std::vector<double> &hG) {
  static std::mutex mu;
  static std::map<std::pair<int, int>, std::shared_ptr<const LowRank>> cache;
  std::lock_guard<std::mutex> lock(mu);
  auto &e = cache[{n * 2 + (compressed ? 1 : 0), kr}];
  if (!e) {
    auto lr = std::make_shared<LowRank>();
    m2l_low_rank(n * n * n, n3p, hG.data(), kr, host_threads(), lr->U, lr->Gc);
    e = lr;
  }
  return e;
}

struct FmmSystem {
  hipStream_t s = nullptr;
  int N = 0, n = 0, L = 0, ncond = 0;
  // One tree per problem; a batch stacks independent problems (block-diagonal
  // operator) under one box numbering: per level the boxes with children
  // first, then the leaves of the problems whose depth is that level.
  std::vector<Tree> trees;
  std::vector<double> L0s;                // root half-width per problem, metres
  std::vector<int> panel_off, cond_off;   // per problem (+1): sorted panel / conductor ranges
  std::vector<int> lvl_off, npar;
  int n_leaves = 0;
  int64_t nnz = 0;
  long pairs = 0;
  bool have32 = false;
  bool m2l_compress = false;  // far-field M2L blocks hold the CPQR truncation
  bool skip_converged = true;  // a batch skips its converged problems' work
  int n3p = 0;
  // Chebyshev order of the FP32 / BF16 inner operator's far field (n_in <= n)
  int n_in = 0, n3p_in = 0;
  int m2l_inner = GP_M2L_BF16, m2l_outer = GP_M2L_FP64;
  DBuf<double> verts, cen, area, inv_diag;
  // Signed panel sets (PanelGeom::sign): the sign per sorted panel and the
  // charge weights 1/2 sign area; null otherwise (weights = area).
  DBuf<double> sgn, qw;
  DBuf<uint32_t> cond;
  DBuf<int> leaf_lo, leaf_box, col_off, cols, panel_leaf;
  DBuf<int> d_panel_off;    // problems' first panels in ascending order, then N
  DBuf<double> d_si_scale;  // k_e L0 of the same problems: the Jacobi threshold's SI scale
  DBuf<int64_t> nf_off;
  DBuf<double> nf64;
  DBuf<float> nf32;
  DBuf<int> packs;  // single-RHS near-field packs of leaves (NearTiles::pack)
  // single-RHS inner operator on FP16 tiles (near_inner): values and row scales
  bool near16 = false;
  DBuf<_Float16> nf16;
  DBuf<float> nscale;
  DBuf<int> parent, octant, child_off, child_idx, targets, il_off, il_src, il_oidx, box_level;
  // subtrees of the single-RHS tree passes (FmmDev::sub_root ..)
  int n_sub = 0;
  DBuf<int> sub_root, sub_leaf, sub_rng;
  DBuf<double> inv_half;
  DBuf<double> w1d64, m2m64, G64, lr_vt64, lr_up64, lr_u64;
  DBuf<int> lr_rank;
  DBuf<float> w1d32, m2m32, G32, Gr32, xl32, wc32;
  DBuf<uint16_t> Gbf;
  // Low-rank inner M2L: rank kr (0 = off), projection U and kr x kr tables.
  int kr = 0;
  DBuf<float> U32, Gc32, Gcr32;
  DBuf<uint16_t> Gcbf, Gcbf_lo;
  DBuf<int> coff, cidx;
  NearTiles nt;
  FmmDev<double> f64;
  FmmDev<float> f32;
  int ld = 0;
  DBuf<double> z64, M64, Lc64;
  DBuf<float> z32, M32, Lc32, Mt32, Lt32;
  DBuf<uint16_t> Mbf, Mtbf, Mtbf_lo;
  bool bf16_ok = false;  // the BF16 M2L runs at this ld
  long matvecs = 0;
  // Transposed operator P'^T (adjoint solves): set want_transpose before setup.
  bool want_transpose = false, transposed = false;
  // Single-RHS M2L grouped by offset: set want_by_offset before setup.
  bool want_by_offset = false;
  std::vector<int> po_off;
  DBuf<int> po_tgt, po_src;
  int n_tiles = 0, tile_size = 16;
  DBuf<int> tile_tgt, tile_eoff, ent_oidx, ent_src;
  int n_pairs2 = 0;
  DBuf<int> pair_tgt, pair_off, pe_src, pe_o1, pe_o2;
  DBuf<int> tr_off, tr_leaf, tr_stride;
  DBuf<int64_t> tr_base;
  DBuf<double> dinv_area, inv_area, ytmp;
  NearTranspose ntr;
  // The dielectric formulation (one problem, FP64): set diel before setup.
  const DielRows *diel = nullptr;
  DBuf<unsigned char> d_iface;
  DBuf<double> d_nrm, d_ddiag, dw1d64, yfar;
  // A batch: problem of each global box, skip flags, per-problem work
  DBuf<int> box_prob;
  DBuf<unsigned char> skip_box, skip_leaf;
  std::vector<double> p_panels, p_nnz, p_pairs, p_boxes;

  const Tree &t() const { return trees[0]; }
  double L0() const { return L0s[0]; }

  void setup(const PanelGeom &g, const gp_cap_opts &o, bool need32, hipStream_t stream) {
    std::vector<const PanelGeom *> one{&g};
    setup(one, o, need32, stream);
  }

  // sa: the stream of the near-field assembly and the kernels that read its
  // values (null: stream); the caller synchronises it before the solve
  void setup(const std::vector<const PanelGeom *> &probs, const gp_cap_opts &o, bool need32,
             hipStream_t stream, hipStream_t sa = nullptr) {
    s = stream;
    if (!sa) sa = stream;
    n = (int)o.fmm_order;
    if (n < 2 || n > 6) throw Error(GP_ERR_INVALID_INPUT, "fmm_order must be in [2, 6]");
    m2l_inner = (int)o.m2l_inner;
    m2l_outer = (int)o.m2l_outer;
    if (o.m2l_tile_order > 1)
      throw Error(GP_ERR_INVALID_INPUT, "m2l_tile_order must be 0 or 1");
    if (o.inner_order != 0 && (o.inner_order < 2 || (int)o.inner_order > n))
      throw Error(GP_ERR_INVALID_INPUT, "inner_order must be 0 or in [2, fmm_order]");
    n_in = o.inner_order ? (int)o.inner_order : n;
    if (m2l_inner != GP_M2L_BF16 && m2l_inner != GP_M2L_FP32 && m2l_inner != GP_M2L_BF16_VALU)
      throw Error(GP_ERR_INVALID_INPUT, "m2l_inner must be GP_M2L_BF16, _FP32 or _BF16_VALU");
    if (m2l_outer != GP_M2L_FP64 && m2l_outer != GP_M2L_FP64_VALU)
      throw Error(GP_ERR_INVALID_INPUT, "m2l_outer must be GP_M2L_FP64 or GP_M2L_FP64_VALU");
    const int P = (int)probs.size();
    double tph = now_s();
    auto phase = [&](const char *what) {
      const double t = now_s();
      GP_LOG("setup phase %-12s %.3f s", what, t - tph);
      tph = t;
    };
    trees.assign(P, Tree());
    L0s.assign(P, 1.0);
    L = 2;
    // Per-problem loops take problems one at a time, largest first, so the
    // threads finish together on batches whose sizes span two decades.
    std::vector<int> big_first(P);
    std::iota(big_first.begin(), big_first.end(), 0);
    std::stable_sort(big_first.begin(), big_first.end(),
                     [&](int a, int b) { return probs[a]->n > probs[b]->n; });
    parallel_for(
        P,
        [&](int i) {
          const int p = big_first[i];
          const PanelGeom &g = *probs[p];
          int Lp = o.max_level ? (int)o.max_level : choose_depth(g, (int)o.leaf_target);
          build_tree(g, std::max(Lp, 2), trees[p]);
          L0s[p] = trees[p].half;
        },
        1);
    for (int p = 0; p < P; ++p) L = std::max(L, trees[p].L);
    // Problems in order of depth: the global leaf order (and so the panel
    // order) lists shallower problems first.
    std::vector<int> order(P);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return trees[a].L < trees[b].L; });
    panel_off.assign(P + 1, 0);
    cond_off.assign(P + 1, 0);
    {
      std::vector<int> cnt(P);
      for (int p = 0; p < P; ++p) cnt[p] = probs[p]->n;
      int acc = 0;
      for (int p : order) {
        panel_off[p] = acc;
        acc += cnt[p];
      }
      N = acc;
      panel_off[P] = N;  // (per-problem ranges are panel_off[p] .. + n_p)
      for (int p = 0; p < P; ++p) cond_off[p + 1] = cond_off[p] + probs[p]->ncond;
      ncond = cond_off[P];
    }
    phase("trees");
    // Global box numbering.
    std::vector<std::vector<int>> gid(P);
    for (int p = 0; p < P; ++p) gid[p].assign(trees[p].n_boxes(), -1);
    lvl_off.assign(L + 2, 0);
    npar.assign(L + 1, 0);
    std::vector<int> gp_of, gb_of;  // global box -> (problem, local box)
    std::vector<int> hleaf_box, hleaf_lo{0};
    for (int l = 0; l <= L; ++l) {
      lvl_off[l] = (int)gp_of.size();
      for (int p : order) {
        const Tree &t = trees[p];
        if (t.L <= l) continue;
        for (int b = t.lvl_off[l]; b < t.lvl_off[l + 1]; ++b) {
          gid[p][b] = (int)gp_of.size();
          gp_of.push_back(p);
          gb_of.push_back(b);
        }
      }
      npar[l] = (int)gp_of.size() - lvl_off[l];
      for (int p : order) {
        const Tree &t = trees[p];
        if (t.L != l) continue;
        for (int T = 0; T < t.n_leaves(); ++T) {
          const int b = t.leaf_box(T);
          gid[p][b] = (int)gp_of.size();
          gp_of.push_back(p);
          gb_of.push_back(b);
          hleaf_box.push_back(gid[p][b]);
          hleaf_lo.push_back(panel_off[p] + t.leaf_lo[T + 1]);
        }
      }
    }
    lvl_off[L + 1] = (int)gp_of.size();
    const int nb = (int)gp_of.size();
    n_leaves = (int)hleaf_box.size();
    // Leaf index of each problem's first leaf, for the near-field tiles.
    std::vector<int> leaf_first(P, 0);
    {
      int acc = 0;
      for (int l = 0; l <= L; ++l)
        for (int p : order)
          if (trees[p].L == l) {
            leaf_first[p] = acc;
            acc += trees[p].n_leaves();
          }
    }

    phase("numbering");
    // Scaled panel arrays, per problem in its own unit frame.
    // (written in place by the problems' threads: no zero fill, no copies)
    std::unique_ptr<double[]> hv(new double[(size_t)N * 12]), hc(new double[(size_t)N * 3]),
        ha(new double[N]);
    std::unique_ptr<uint32_t[]> hcond(new uint32_t[N]);
    parallel_for(
        P,
        [&](int i) {
          const int p = big_first[i];
          const size_t o0 = panel_off[p];
          scale_into(*probs[p], trees[p].perm.data(), trees[p].center, L0s[p],
                     hv.get() + o0 * 12, hc.get() + o0 * 3, ha.get() + o0, hcond.get() + o0,
                     (uint32_t)cond_off[p]);
        },
        1);
    verts = upload_panels(hv.get(), (size_t)N * 12, s);
    cen = upload_panels(hc.get(), (size_t)N * 3, s);
    area = upload_panels(ha.get(), (size_t)N, s);
    cond = upload_panels(hcond.get(), (size_t)N, s);
    bool any_sign = false;
    for (int p = 0; p < P; ++p) any_sign |= probs[p]->sign != nullptr;
    if (any_sign) {
      std::vector<double> hs(N, 1.0), hq(N);
      for (int p = 0; p < P; ++p) {
        const double *sg = probs[p]->sign;
        for (int k = 0; k < probs[p]->n; ++k) {
          const size_t q = (size_t)panel_off[p] + k;
          hs[q] = sg ? sg[trees[p].perm[k]] : 1.0;
          hq[q] = sg ? 0.5 * hs[q] * ha[q] : ha[q];
        }
      }
      sgn = to_device(hs, s);
      qw = to_device(hq, s);
    }
    hv.reset();
    hc.reset();

    phase("panels");
    // Near-field tiles (cap/fmm_solver.rs:52-91 over bbfmm.rs:189-206), per
    // global leaf: its panels against the panels of its near leaves.
    std::vector<int> hcol_off(n_leaves + 1, 0), hpl(N);
    std::unique_ptr<int[]> hcols;  // every entry written by its leaf's thread
    std::vector<int64_t> hnf(n_leaves + 1, 0);
    std::vector<std::array<int, 3>> trent;  // (S, T, column start of S in T's tile)
    auto leaf_of = [&](int T, int &p, int &lt) {
      const int gbox = hleaf_box[T];
      p = gp_of[gbox];
      lt = gb_of[gbox] - trees[p].lvl_off[trees[p].L];
    };
    {
      std::vector<int64_t> ccount(n_leaves);
      parallel_for(n_leaves, [&](int T) {
        int p, lt;
        leaf_of(T, p, lt);
        const Tree &t = trees[p];
        int64_t c = 0;
        for (int q = t.near_off[lt]; q < t.near_off[lt + 1]; ++q)
          c += t.leaf_lo[t.near_leaf[q] + 1] - t.leaf_lo[t.near_leaf[q]];
        ccount[T] = c;
      });
      int64_t acc = 0;
      for (int T = 0; T < n_leaves; ++T) {
        acc += ccount[T];
        if (acc > INT_MAX) throw Error(GP_ERR_OOM, "near-field column lists exceed 2^31 entries");
        hcol_off[T + 1] = (int)acc;
        const int nt_ = hleaf_lo[T + 1] - hleaf_lo[T];
        hnf[T + 1] = hnf[T] + (int64_t)nt_ * ccount[T];
      }
    }
    phase("near count");
    hcols.reset(new int[std::max(1, hcol_off[n_leaves])]);
    parallel_for(n_leaves, [&](int T) {
      int p, lt;
      leaf_of(T, p, lt);
      const Tree &t = trees[p];
      int c = hcol_off[T];
      for (int q = t.near_off[lt]; q < t.near_off[lt + 1]; ++q) {
        const int S = t.near_leaf[q];
        for (int k = t.leaf_lo[S]; k < t.leaf_lo[S + 1]; ++k) hcols[c++] = panel_off[p] + k;
      }
      for (int k = hleaf_lo[T]; k < hleaf_lo[T + 1]; ++k) hpl[k] = T;
    });
    if (want_transpose)
      for (int T = 0; T < n_leaves; ++T) {
        int p, lt;
        leaf_of(T, p, lt);
        const Tree &t = trees[p];
        int c = 0;
        for (int q = t.near_off[lt]; q < t.near_off[lt + 1]; ++q) {
          const int S = t.near_leaf[q];
          trent.push_back({leaf_first[p] + S, T, c});
          c += t.leaf_lo[S + 1] - t.leaf_lo[S];
        }
      }
    nnz = hnf[n_leaves];
    if (want_transpose) {
      std::stable_sort(trent.begin(), trent.end(),
                       [](const std::array<int, 3> &a, const std::array<int, 3> &b) {
                         return a[0] < b[0];
                       });
      std::vector<int> ho(n_leaves + 1, 0), hl, hs;
      std::vector<int64_t> hbase;
      for (auto &e : trent) {
        ho[e[0] + 1]++;
        hl.push_back(e[1]);
        hbase.push_back(hnf[e[1]] + e[2]);
        hs.push_back(hcol_off[e[1] + 1] - hcol_off[e[1]]);
      }
      for (int T = 0; T < n_leaves; ++T) ho[T + 1] += ho[T];
      tr_off = to_device(ho, s);
      tr_leaf = to_device(hl, s);
      tr_base = to_device(hbase, s);
      tr_stride = to_device(hs, s);
      ntr.off = tr_off.p;
      ntr.leaf = tr_leaf.p;
      ntr.base = tr_base.p;
      ntr.stride = tr_stride.p;
    }
    phase("near cols");
    {
      // the problems' panel ranges in device order (ascending offsets) and
      // their threshold scales, for the one Jacobi launch below
      std::vector<int> ord(P), ho(P + 1);
      std::iota(ord.begin(), ord.end(), 0);
      std::stable_sort(ord.begin(), ord.end(),
                       [&](int a, int b) { return panel_off[a] < panel_off[b]; });
      std::vector<double> hsc(P);
      for (int k = 0; k < P; ++k) {
        ho[k] = panel_off[ord[k]];
        hsc[k] = L0s[ord[k]] / (4.0 * kPi * kEps0);
      }
      ho[P] = N;
      d_panel_off = to_device(ho, s);
      d_si_scale = to_device(hsc, s);
    }
    leaf_lo = to_device(hleaf_lo, s);
    leaf_box = to_device(hleaf_box, s);
    col_off = to_device(hcol_off, s);
    cols = to_device(hcols.get(), (size_t)hcol_off[n_leaves], s);
    panel_leaf = to_device(hpl, s);
    nf_off = to_device(hnf, s);
    phase("near upload");
    nf64.alloc((size_t)nnz);
    have32 = need32;
    if (o.near_inner != GP_NEAR_FP32 && o.near_inner != GP_NEAR_FP16)
      throw Error(GP_ERR_INVALID_INPUT, "near_inner must be GP_NEAR_FP32 or GP_NEAR_FP16");
    near16 = need32 && want_by_offset && o.near_inner == GP_NEAR_FP16 && !want_transpose && !diel;
    if (need32 && !near16) nf32.alloc((size_t)nnz);
    phase("near alloc");
    // (on sa: every input above is on the device, setup_stream synchronised
    // by its uploads; nothing below reads the values but on sa)
    if ((diel || want_transpose) && sa != s)
      throw Error(GP_ERR_INTERNAL, "dielectric and adjoint set-ups assemble on their own stream");
    launch_near_assemble(n_leaves, leaf_lo.p, nf_off.p, col_off.p, cols.p, verts.p, cen.p, nf64.p,
                         (need32 && !near16) ? nf32.p : nullptr, sa);
    phase("near launch");
    inv_diag.alloc(N);
    launch_self_inv_diag(N, P, d_panel_off.p, d_si_scale.p, verts.p, cen.p, inv_diag.p, sa);
    if (diel) {
      // interface rows of the near tiles and their Jacobi entries, in the
      // device (leaf-sorted) order
      if (P != 1 || need32)
        throw Error(GP_ERR_INVALID_INPUT, "the FMM dielectric solve is one problem in FP64");
      const auto &perm = trees[0].perm;
      std::vector<unsigned char> hi(N);
      std::vector<double> hn((size_t)N * 3), hd(N);
      for (int k = 0; k < N; ++k) {
        const int c = (int)perm[k];
        hi[k] = diel->iface[c];
        hd[k] = diel->diag[c];
        for (int d = 0; d < 3; ++d) hn[(size_t)k * 3 + d] = diel->nrm[(size_t)c * 3 + d];
      }
      d_iface = to_device(hi, s);
      d_nrm = to_device(hn, s);
      d_ddiag = to_device(hd, s);
      launch_near_dielectric_rows(n_leaves, N, leaf_lo.p, nf_off.p, col_off.p, cols.p, d_iface.p,
                                  d_nrm.p, d_ddiag.p, verts.p, cen.p, nf64.p, inv_diag.p, s);
    }
    if (want_transpose) {
      // far-field transpose: P'_far^T x = area * K (x / area), K symmetric
      std::vector<double> hdi(N);
      GP_HIP(hipMemcpyAsync(hdi.data(), inv_diag.p, N * sizeof(double), hipMemcpyDeviceToHost, s));
      GP_HIP(hipStreamSynchronize(s));
      std::vector<double> hia(N);
      for (int i = 0; i < N; ++i) {
        hdi[i] /= ha[i];
        hia[i] = 1.0 / ha[i];
      }
      dinv_area = to_device(hdi, s);
      inv_area = to_device(hia, s);
    }
    phase("near diag");
    nt.n = N;
    nt.nnz = (long)nnz;
    nt.n_leaves = n_leaves;
    nt.max_cols = 0;
    for (int T = 0; T < n_leaves; ++T)
      nt.max_cols = std::max(nt.max_cols, hcol_off[T + 1] - hcol_off[T]);
    if (want_by_offset) {
      // packs of consecutive leaves of one problem for the single-RHS
      // product: up to 32 leaves and 1024 columns (a wider leaf alone)
      const int lim = std::min(1024, nt.max_cols);
      std::vector<int> hp;
      for (int T = 0; T < n_leaves;) {
        hp.push_back(T);
        const int p = gp_of[hleaf_box[T]];
        int c = hcol_off[T + 1] - hcol_off[T], k = 1;
        for (++T; T < n_leaves && k < 32 && gp_of[hleaf_box[T]] == p &&
                  c + (hcol_off[T + 1] - hcol_off[T]) <= lim;
             ++T, ++k)
          c += hcol_off[T + 1] - hcol_off[T];
      }
      hp.push_back(n_leaves);
      packs = to_device(hp, s);
      nt.n_packs = (int)hp.size() - 1;
      nt.pack = packs.p;
    }
    nt.leaf_lo = leaf_lo.p;
    nt.nf_off = nf_off.p;
    nt.col_off = col_off.p;
    nt.cols = cols.p;
    nt.panel_leaf = panel_leaf.p;
    if (near16) {
      nf16.alloc((size_t)nnz);
      nscale.alloc(N);
      launch_near_to_f16(nt, nf64.p, nf16.p, nscale.p, sa);
    }

    phase("near");
    // FMM tables.
    const int n3 = n * n * n;
    n3p = round_up(n3, 32);
    n3p_in = round_up(n_in * n_in * n_in, 32);
    kr = 0;  // low-rank inner M2L: rank 32, 64 or 128 below n3p (blocks of RHS only:
             // single-RHS solves run plain BF16 target tiles, too coarse for it)
    if (need32 && o.m2l_rank > 0 && !want_by_offset) {
      kr = o.m2l_rank <= 32 ? 32 : o.m2l_rank <= 64 ? 64 : 128;
      if (kr >= n3p_in) kr = 0;
    }
    auto nodes = cheb_nodes(n);
    // Leaf boxes in each problem's unit frame (centre and half-width, exact
    // powers of two), for the interpolation weights computed on the device.
    std::vector<double> hgeo((size_t)n_leaves * 4);
    parallel_for(n_leaves, [&](int T) {
      int p, lt;
      leaf_of(T, p, lt);
      const Tree &t = trees[p];
      const int b = t.leaf_box(lt);
      const double hs = t.level_half(t.L) / t.half;
      for (int d = 0; d < 3; ++d) hgeo[(size_t)T * 4 + d] = -1.0 + (t.cell[3 * b + d] + 0.5) * 2.0 * hs;
      hgeo[(size_t)T * 4 + 3] = hs;
    });
    // M2M 1-D matrices, bbfmm.rs:89-112: mat[m*n + k] = S_n(cc + x_k / 2, x_m).
    auto m2m_table = [](int nn, const std::vector<double> &nd) {
      std::vector<double> t((size_t)24 * nn * nn);
      for (int oz = 0; oz < 2; ++oz)
        for (int oy = 0; oy < 2; ++oy)
          for (int ox = 0; ox < 2; ++ox) {
            int o = ox + 2 * oy + 4 * oz;
            double cc[3] = {(2.0 * ox - 1.0) * 0.5, (2.0 * oy - 1.0) * 0.5,
                            (2.0 * oz - 1.0) * 0.5};
            for (int axis = 0; axis < 3; ++axis)
              for (int k = 0; k < nn; ++k) {
                double w[32];
                interp_weights(cc[axis] + 0.5 * nd[k], nn, nd.data(), w);
                for (int m = 0; m < nn; ++m) t[(size_t)(o * 3 + axis) * nn * nn + m * nn + k] = w[m];
              }
          }
      return t;
    };
    const std::vector<double> hm2m = m2m_table(n, nodes);
    // M2L matrices on the reference box (bbfmm.rs:156-166), each far-field
    // block truncated by the reference's CPQR (bbfmm.rs:132-174) unless
    // m2l_compress is 0.  The transposed operator of the adjoint solves
    // relies on G_-o = G_o^T, which the truncation does not keep, so it
    // applies the full matrices.
    m2l_compress = o.m2l_compress != 0 && !want_transpose;
    skip_converged = o.skip_converged != 0;
    const auto m2l_tab = m2l_tables(n, n3p, m2l_compress);
    const std::vector<double> &hG = *m2l_tab;
    // the inner operator's tables at its own order (the same tables when
    // n_in == n)
    const auto nodes_in = n_in == n ? nodes : cheb_nodes(n_in);
    const std::vector<double> hm2m_in = n_in == n ? hm2m : m2m_table(n_in, nodes_in);
    const auto m2l_tab_in =
        (need32 && n_in != n) ? m2l_tables(n_in, n3p_in, m2l_compress) : m2l_tab;
    const std::vector<double> &hGi = *m2l_tab_in;
    phase("tables");
    // Box arrays in the global numbering.
    std::vector<int> hlevel(nb), hparent(nb), hoct(nb), hchild_off(nb + 1, 0), hchild, htargets;
    std::vector<int> hil_off(nb + 1, 0);
    std::unique_ptr<int[]> hil_src, hil_oidx;  // every entry written by its box's thread
    for (int l = 0; l <= L; ++l)
      for (int b = lvl_off[l]; b < lvl_off[l + 1]; ++b) hlevel[b] = l;
    parallel_for(nb, [&](int b) {
      const Tree &t = trees[gp_of[b]];
      const int lb = gb_of[b];
      hchild_off[b + 1] = t.child_off[lb + 1] - t.child_off[lb];
      hil_off[b + 1] = t.il_off[lb + 1] - t.il_off[lb];
    });
    for (int b = 0; b < nb; ++b) {
      if (hil_off[b + 1] > 0) htargets.push_back(b);
      hchild_off[b + 1] += hchild_off[b];
      hil_off[b + 1] += hil_off[b];
    }
    hchild.resize(hchild_off[nb]);
    const size_t n_il = (size_t)hil_off[nb];
    hil_src.reset(new int[std::max<size_t>(1, n_il)]);
    hil_oidx.reset(new int[std::max<size_t>(1, n_il)]);
    parallel_for(nb, [&](int b) {
      const int p = gp_of[b], lb = gb_of[b];
      const Tree &t = trees[p];
      hparent[b] = t.parent[lb] >= 0 ? gid[p][t.parent[lb]] : -1;
      hoct[b] = t.octant[lb];
      int c = hchild_off[b];
      for (int q = t.child_off[lb]; q < t.child_off[lb + 1]; ++q) hchild[c++] = gid[p][t.child_idx[q]];
      int e = hil_off[b];
      for (int q = t.il_off[lb]; q < t.il_off[lb + 1]; ++q, ++e) {
        hil_src[e] = gid[p][t.il_src[q]];
        hil_oidx[e] = t.il_oidx[q];
      }
    });
    phase("boxes");
    pairs = (long)n_il;
    if (need32 && m2l_inner == GP_M2L_BF16 && !want_by_offset) {
      // Sibling pairs for the multi-RHS BF16 M2L: consecutive targets with
      // one parent, the union of their lists ordered by source box.
      std::vector<int> gfirst;
      for (size_t i = 0; i < htargets.size();) {
        gfirst.push_back((int)i);
        i += (i + 1 < htargets.size() && hparent[htargets[i + 1]] == hparent[htargets[i]]) ? 2 : 1;
      }
      n_pairs2 = (int)gfirst.size();
      gfirst.push_back((int)htargets.size());
      std::vector<std::vector<std::array<int, 3>>> ents(n_pairs2);
      parallel_for(n_pairs2, [&](int gi) {
        auto &e = ents[gi];
        for (int i = gfirst[gi]; i < gfirst[gi + 1]; ++i) {
          const int b = htargets[i], side = i - gfirst[gi];
          for (int q = hil_off[b]; q < hil_off[b + 1]; ++q)
            e.push_back({hil_src[q], side ? -1 : hil_oidx[q], side ? hil_oidx[q] : -1});
        }
        std::sort(e.begin(), e.end());
        size_t w = 0;
        for (size_t r = 0; r < e.size(); ++r) {
          if (w > 0 && e[w - 1][0] == e[r][0]) {
            if (e[r][1] >= 0) e[w - 1][1] = e[r][1];
            if (e[r][2] >= 0) e[w - 1][2] = e[r][2];
          } else {
            e[w++] = e[r];
          }
        }
        e.resize(w);
      });
      std::vector<int> hpt((size_t)2 * n_pairs2, -1), hpo(n_pairs2 + 1, 0);
      for (int gi = 0; gi < n_pairs2; ++gi) {
        for (int i = gfirst[gi]; i < gfirst[gi + 1]; ++i) hpt[2 * gi + (i - gfirst[gi])] = htargets[i];
        hpo[gi + 1] = hpo[gi] + (int)ents[gi].size();
      }
      std::vector<int> hs(hpo[n_pairs2]), h1(hpo[n_pairs2]), h2(hpo[n_pairs2]);
      parallel_for(n_pairs2, [&](int gi) {
        int q = hpo[gi];
        for (const auto &e : ents[gi]) {
          hs[q] = e[0];
          h1[q] = e[1];
          h2[q] = e[2];
          ++q;
        }
      });
      pair_tgt = to_device(hpt, s);
      pair_off = to_device(hpo, s);
      pe_src = to_device(hs, s);
      pe_o1 = to_device(h1, s);
      pe_o2 = to_device(h2, s);
    }
    if (want_by_offset && n3p > 64) {
      // FP64 single-RHS M2L beyond the 64-target tiles: the pairs grouped by
      // offset (counting sort; targets stay ascending)
      po_off.assign(344, 0);
      for (size_t q = 0; q < n_il; ++q) po_off[hil_oidx[q] + 1]++;
      for (int o = 0; o < 343; ++o) po_off[o + 1] += po_off[o];
      std::vector<int> ht(pairs), hs(pairs), fillp(po_off.begin(), po_off.end() - 1);
      for (int b = 0; b < nb; ++b)
        for (int q = hil_off[b]; q < hil_off[b + 1]; ++q) {
          const int e = fillp[hil_oidx[q]]++;
          ht[e] = b;
          hs[e] = hil_src[q];
        }
      po_tgt = to_device(ht, s);
      po_src = to_device(hs, s);
    }
    if (want_by_offset) {
      // Target tiles for the single-RHS M2L (BF16, and FP64 up to n3p = 64):
      // TS targets of one octant, longest lists first (box order within; a
      // tile never spans octants), the union of their offsets.  TS = 64 where
      // the work-group stages G_o in LDS.
      const int mp = kr ? kr : n3p;
      tile_size = mp <= 128 ? 64 : 16;
      const int TS = tile_size;
      // (a stable counting sort on (octant, longest list first): the order of
      // a stable comparison sort, in linear time over millions of targets)
      std::vector<int> ord(htargets.size());
      {
        auto len = [&](int b) { return hil_off[b + 1] - hil_off[b]; };
        int maxlen = 0;
        for (int b : htargets) maxlen = std::max(maxlen, len(b));
        // m2l_tile_order 1: box order within the octant (neighbouring
        // targets, shared sources) instead of longest lists first
        const bool by_box = o.m2l_tile_order == 1;
        auto key = [&](int b) {
          return by_box ? hoct[b] * (maxlen + 1) : hoct[b] * (maxlen + 1) + (maxlen - len(b));
        };
        std::vector<int> start((size_t)8 * (maxlen + 1) + 1, 0);
        for (int b : htargets) start[key(b) + 1]++;
        for (size_t k = 0; k + 1 < start.size(); ++k) start[k + 1] += start[k];
        for (int b : htargets) ord[start[key(b)]++] = b;
      }
      std::vector<int> tfirst, tcount;
      for (size_t i = 0; i < ord.size();) {
        size_t j = i;
        while (j < ord.size() && hoct[ord[j]] == hoct[ord[i]]) ++j;
        for (size_t k = i; k < j; k += TS) {
          tfirst.push_back((int)k);
          tcount.push_back((int)std::min<size_t>(TS, j - k));
        }
        i = j;
      }
      n_tiles = (int)tfirst.size();
      std::vector<int> htt((size_t)n_tiles * TS, -1), heoff(n_tiles + 1, 0);
      parallel_for(n_tiles, [&](int T) {
        bool seen[343] = {};
        int cnt = 0;
        for (int j = 0; j < tcount[T]; ++j) {
          const int b = ord[tfirst[T] + j];
          htt[(size_t)T * TS + j] = b;
          for (int q = hil_off[b]; q < hil_off[b + 1]; ++q)
            if (!seen[hil_oidx[q]]) {
              seen[hil_oidx[q]] = true;
              ++cnt;
            }
        }
        heoff[T + 1] = cnt;
      });
      for (int T = 0; T < n_tiles; ++T) heoff[T + 1] += heoff[T];
      // (entries filled with -1 by each tile's own thread: no serial fill of
      // the whole array)
      std::vector<int> heo(heoff[n_tiles]);
      const size_t n_es = (size_t)heoff[n_tiles] * TS;
      std::unique_ptr<int[]> hes(new int[n_es]);
      parallel_for(n_tiles, [&](int T) {
        std::fill(hes.get() + (size_t)heoff[T] * TS, hes.get() + (size_t)heoff[T + 1] * TS, -1);
        int slot[343];
        std::fill(slot, slot + 343, -1);
        for (int j = 0; j < tcount[T]; ++j) {
          const int b = ord[tfirst[T] + j];
          for (int q = hil_off[b]; q < hil_off[b + 1]; ++q) slot[hil_oidx[q]] = 0;
        }
        int e = heoff[T];
        for (int o = 0; o < 343; ++o)
          if (slot[o] == 0) {
            slot[o] = e;
            heo[e++] = o;
          }
        for (int j = 0; j < tcount[T]; ++j) {
          const int b = ord[tfirst[T] + j];
          for (int q = hil_off[b]; q < hil_off[b + 1]; ++q)
            hes[(size_t)slot[hil_oidx[q]] * TS + j] = hil_src[q];
        }
      });
      tile_tgt = to_device(htt, s);
      tile_eoff = to_device(heoff, s);
      ent_oidx = to_device(heo, s);
      ent_src = to_device(hes.get(), n_es, s);
    }
    phase("m2l lists");
    // 1/half_l in units of each problem's own root half-width.
    std::vector<double> hih(L + 1);
    for (int l = 0; l <= L; ++l) hih[l] = (double)(1u << l);
    parent = to_device(hparent, s);
    octant = to_device(hoct, s);
    child_off = to_device(hchild_off, s);
    child_idx = to_device(hchild, s);
    targets = to_device(htargets, s);
    il_off = to_device(hil_off, s);
    il_src = to_device(hil_src.get(), n_il, s);
    il_oidx = to_device(hil_oidx.get(), n_il, s);
    box_level = to_device(hlevel, s);
    inv_half = to_device(hih, s);
    // Subtrees of the single-RHS tree passes: each level-2 box of a problem
    // with boxes below level 2, the level of its leaves, and per level the
    // global range of its descendants (Morton order keeps them contiguous;
    // checked here, else the passes keep their launch per level).  Largest
    // subtrees first.
    n_sub = 0;
    if (want_by_offset && L >= 3) {
      std::vector<int> sr, sl, sg, cnt;
      bool contiguous = true;
      for (int p : order) {
        const Tree &t = trees[p];
        if (t.L < 3) continue;
        std::vector<int> anc(t.n_boxes(), -1);
        for (int b = t.lvl_off[2]; b < t.lvl_off[3]; ++b) anc[b] = b;
        for (int l = 3; l <= t.L; ++l)
          for (int b = t.lvl_off[l]; b < t.lvl_off[l + 1]; ++b) anc[b] = anc[t.parent[b]];
        const int r0 = (int)sr.size();
        for (int b = t.lvl_off[2]; b < t.lvl_off[3]; ++b) {
          sr.push_back(gid[p][b]);
          sl.push_back(t.L);
          cnt.push_back(0);
        }
        sg.resize(sr.size() * 2 * (L + 1), 0);
        std::vector<int> lo(sr.size() - r0), hi(sr.size() - r0), nb(sr.size() - r0);
        for (int l = 2; l <= t.L; ++l) {
          std::fill(lo.begin(), lo.end(), INT_MAX);
          std::fill(hi.begin(), hi.end(), -1);
          std::fill(nb.begin(), nb.end(), 0);
          for (int b = t.lvl_off[l]; b < t.lvl_off[l + 1]; ++b) {
            const int k = anc[b] - t.lvl_off[2], g = gid[p][b];
            lo[k] = std::min(lo[k], g);
            hi[k] = std::max(hi[k], g + 1);
            ++nb[k];
          }
          for (int k = 0; k < (int)lo.size(); ++k) {
            int *rg = &sg[(size_t)(r0 + k) * 2 * (L + 1) + 2 * l];
            if (nb[k] == 0) continue;
            contiguous &= hi[k] - lo[k] == nb[k];
            rg[0] = lo[k];
            rg[1] = hi[k];
            cnt[r0 + k] += nb[k];
          }
        }
      }
      if (!contiguous) {
        GP_LOG("tree passes by level: a subtree's boxes are not contiguous");
      } else if (!sr.empty()) {
        std::vector<int> ord(sr.size());
        std::iota(ord.begin(), ord.end(), 0);
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return cnt[a] > cnt[b]; });
        std::vector<int> r2(sr.size()), l2(sr.size()), g2(sg.size());
        for (size_t i = 0; i < ord.size(); ++i) {
          r2[i] = sr[ord[i]];
          l2[i] = sl[ord[i]];
          std::copy_n(&sg[(size_t)ord[i] * 2 * (L + 1)], 2 * (L + 1), &g2[i * 2 * (L + 1)]);
        }
        sub_root = to_device(r2, s);
        sub_leaf = to_device(l2, s);
        sub_rng = to_device(g2, s);
        n_sub = (int)r2.size();
      }
    }
    {
      DBuf<double> dgeo = to_device(hgeo, s), dnodes = to_device(nodes, s);
      w1d64.alloc((size_t)N * 3 * n);
      if (need32) w1d32.alloc((size_t)N * 3 * n_in);
      launch_interp_w1d(N, n, panel_leaf.p, dgeo.p, cen.p, dnodes.p, w1d64.p,
                        (need32 && n_in == n) ? w1d32.p : nullptr, s);
      DBuf<double> dnodes_in;
      if (need32 && n_in != n) {
        dnodes_in = to_device(nodes_in, s);
        launch_interp_w1d(N, n_in, panel_leaf.p, dgeo.p, cen.p, dnodes_in.p, nullptr, w1d32.p, s);
      }
      if (diel) {
        dw1d64.alloc((size_t)N * 3 * n);
        launch_interp_dw1d(N, n, panel_leaf.p, dgeo.p, cen.p, dnodes.p, dw1d64.p, s);
      }
      // single RHS: the FP32 operator's P2M and L2P form the weights from
      // the panels' box coordinates (16 bytes a panel instead of 3 n weights
      // and the FP64 area)
      if (need32 && want_by_offset && n_in >= 2 && n_in <= 6) {
        xl32.alloc((size_t)N * 4);
        launch_local_coords(N, panel_leaf.p, dgeo.p, cen.p, area.p, xl32.p, s);
        std::vector<float> hwc((size_t)n_in * n_in);
        for (int m = 0; m < n_in; ++m) {
          const double xm = nodes_in[m];
          double tkm1 = 1.0, tk = xm;
          for (int k = 0; k < n_in; ++k) {
            const double t = k == 0 ? 1.0 : k == 1 ? xm : 2.0 * xm * tk - tkm1;
            if (k >= 2) tkm1 = tk, tk = t;
            hwc[(size_t)m * n_in + k] = (float)(2.0 / n_in * (k == 0 ? 0.5 : 1.0) * t);
          }
        }
        wc32 = to_device(hwc, s);
      }
      GP_HIP(hipStreamSynchronize(s));  // dgeo and dnodes are freed below
    }
    m2m64 = to_device(hm2m, s);
    G64 = to_device(hG, s);
    // the truncated blocks' factors for the factored single-RHS tile kernel
    // (64-target tiles of the FP64 operator)
    if (m2l_compress && want_by_offset && tile_size == 64 && n3p <= 64) {
      const auto lr = m2l_factors(n, n3p);
      lr_vt64 = to_device(lr->Vt, s);
      lr_up64 = to_device(lr->Up, s);
      lr_u64 = to_device(lr->U, s);
      lr_rank = to_device(lr->rank, s);
      f64.lr_kp = lr->kp;
      f64.lr_kmax = lr->kmax;
    }
    if (need32) {
      m2m32 = to_device(cast_vec<float>(hm2m_in), s);
      G32 = to_device(cast_vec<float>(hGi), s);
      std::vector<uint16_t> hb(hGi.size());
      for (size_t e = 0; e < hGi.size(); ++e) hb[e] = bf16_bits((float)hGi[e]);
      Gbf = to_device(hb, s);
      if (m2l_inner == GP_M2L_BF16_VALU) {
        std::vector<float> hr(hGi.size());
        for (size_t e = 0; e < hGi.size(); ++e) {
          const uint32_t u = (uint32_t)hb[e] << 16;
          std::memcpy(&hr[e], &u, 4);
        }
        Gr32 = to_device(hr, s);
      }
      if (kr) {
        auto lr = low_rank_tables(n_in, n3p_in, kr, m2l_compress, hGi);
        U32 = to_device(cast_vec<float>(lr->U), s);
        Gc32 = to_device(cast_vec<float>(lr->Gc), s);
        std::vector<uint16_t> hcb(lr->Gc.size()), hcl(lr->Gc.size());
        for (size_t e = 0; e < hcb.size(); ++e) {
          const float v = (float)lr->Gc[e];
          hcb[e] = bf16_bits(v);
          const uint32_t u = (uint32_t)hcb[e] << 16;
          float h;
          std::memcpy(&h, &u, 4);
          hcl[e] = bf16_bits(v - h);  // the pair hi + lo for the split-BF16 products
        }
        Gcbf = to_device(hcb, s);
        Gcbf_lo = to_device(hcl, s);
        if (m2l_inner == GP_M2L_BF16_VALU) {
          std::vector<float> hr(hcb.size());
          for (size_t e = 0; e < hcb.size(); ++e) {
            const uint32_t u = (uint32_t)hcb[e] << 16;
            std::memcpy(&hr[e], &u, 4);
          }
          Gcr32 = to_device(hr, s);
        }
      }
    }
    phase("uploads");
    const bool batch = P > 1;
    auto fill = [&](auto &f) {
      f.n = n;
      f.fuse_leaf = o.fuse_leaf != 0;
      f.L = L;
      f.n_boxes = nb;
      f.n_targets = (int)htargets.size();
      f.n_pairs = pairs;
      f.n_panels = N;
      f.lvl_off = lvl_off.data();
      f.npar = batch ? npar.data() : nullptr;
      f.leaf_box = batch ? leaf_box.p : nullptr;
      f.n_leaves = n_leaves;
      f.leaf_lo = leaf_lo.p;
      f.panel_leaf = panel_leaf.p;
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
      f.po_off = (want_by_offset && !po_off.empty()) ? po_off.data() : nullptr;
      f.po_tgt = po_tgt.p;
      f.po_src = po_src.p;
      f.n_tiles = n_tiles;
      f.tile_size = tile_size;
      f.n_pairs2 = n_pairs2;
      f.pair_tgt = n_pairs2 ? pair_tgt.p : nullptr;
      f.pair_off = pair_off.p;
      f.pe_src = pe_src.p;
      f.pe_o1 = pe_o1.p;
      f.pe_o2 = pe_o2.p;
      f.tile_tgt = want_by_offset ? tile_tgt.p : nullptr;
      f.tile_eoff = tile_eoff.p;
      f.ent_oidx = ent_oidx.p;
      f.ent_src = ent_src.p;
      f.n_sub = n_sub;
      f.sub_root = sub_root.p;
      f.sub_leaf = sub_leaf.p;
      f.sub_rng = sub_rng.p;
    };
    fill(f64);
    f64.w1d = w1d64.p;
    f64.m2m1d = m2m64.p;
    f64.G = G64.p;
    f64.lr_vt = lr_vt64.p;
    f64.lr_up = lr_up64.p;
    f64.lr_u = lr_u64.p;
    f64.lr_rank = lr_rank.p;
    fill(f32);
    f32.n = n_in;
    f32.n3p = n3p_in;
    f32.w1d = w1d32.p;
    f32.xl = xl32.p;
    f32.wc = wc32.p;
    f32.m2m1d = m2m32.p;
    f32.G = G32.p;
    f32.Gbf = Gbf.p;
    f32.Gr = Gr32.p;
    f32.kr = kr;
    f32.U = U32.p;
    f32.Gc = Gc32.p;
    f32.Gcbf = Gcbf.p;
    f32.Gcbf_lo = Gcbf_lo.p;
    f32.Gcr = Gcr32.p;
    // Panels of each (batch-global) conductor, ascending caller order.
    std::vector<int> hco{0}, hci;
    for (int p = 0; p < P; ++p) {
      std::vector<int> co, ci;
      cond_csr(*probs[p], trees[p].perm.data(), co, ci);
      for (int c = 0; c < probs[p]->ncond; ++c) {
        for (int q = co[c]; q < co[c + 1]; ++q) hci.push_back(panel_off[p] + ci[q]);
        hco.push_back((int)hci.size());
      }
    }
    coff = to_device(hco, s);
    cidx = to_device(hci, s);
    // A batch: the problem of every box, for skipping converged problems,
    // and each problem's share of the work, for the profiler's accounting.
    if (P > 1) {
      box_prob = to_device(gp_of, s);
      skip_box.alloc(nb);
      skip_leaf.alloc(n_leaves);
      p_panels.assign(P, 0.0);
      p_nnz.assign(P, 0.0);
      p_pairs.assign(P, 0.0);
      p_boxes.assign(P, 0.0);
      for (int p = 0; p < P; ++p) p_panels[p] = probs[p]->n;
      for (int T = 0; T < n_leaves; ++T) p_nnz[gp_of[hleaf_box[T]]] += (double)(hnf[T + 1] - hnf[T]);
      for (int b = 0; b < nb; ++b) {
        p_pairs[gp_of[b]] += hil_off[b + 1] - hil_off[b];
        p_boxes[gp_of[b]] += 1.0;
      }
    }
  }

  // The operator applies that follow skip the work of the batch's problems
  // whose system is inactive (d_active, one flag per problem, read on the
  // device when the applies run); h_active, when given, is its host copy for
  // the profiler's accounting.  No effect on one problem.
  void mask(const int *d_active, const int *h_active) {
    if (!box_prob.p) return;
    kr_skip_flags(f64.n_boxes, n_leaves, box_prob.p, leaf_box.p, d_active, skip_box.p,
                  skip_leaf.p, s);
    double fr[4] = {1.0, 1.0, 1.0, 1.0};  // panels, nnz, pairs, boxes
    if (h_active) {
      double a[4] = {0, 0, 0, 0}, t[4] = {0, 0, 0, 0};
      for (size_t p = 0; p < p_panels.size(); ++p) {
        const double *v[4] = {&p_panels[p], &p_nnz[p], &p_pairs[p], &p_boxes[p]};
        for (int q = 0; q < 4; ++q) {
          t[q] += *v[q];
          if (h_active[p]) a[q] += *v[q];
        }
      }
      for (int q = 0; q < 4; ++q) fr[q] = t[q] > 0 ? a[q] / t[q] : 1.0;
    }
    auto set = [&](auto &f) {
      f.skip_box = skip_box.p;
      f.skip_leaf = skip_leaf.p;
      f.frac_panels = fr[0];
      f.frac_pairs = fr[2];
      f.frac_boxes = fr[3];
    };
    set(f32);
    set(f64);
    nt.skip_leaf = skip_leaf.p;
    nt.frac_nnz = fr[1];
    nt.frac_rows = fr[0];
  }
  void unmask() {
    auto clr = [](auto &f) {
      f.skip_box = f.skip_leaf = nullptr;
      f.frac_panels = f.frac_pairs = f.frac_boxes = 1.0;
    };
    clr(f32);
    clr(f64);
    nt.skip_leaf = nullptr;
    nt.frac_nnz = nt.frac_rows = 1.0;
  }

  void ensure_work(int ld_) {
    if (ld_ == ld) return;
    ld = ld_;
    const size_t vec = (size_t)N * ld;
    const size_t mexp = (size_t)f64.n_boxes * n3p * ld;   // multipoles, padded
    const size_t lexp = (size_t)f64.n_boxes * n * n * n * ld;
    z64.alloc(vec);
    z64.zero(s);
    if (diel) yfar.alloc(vec);
    if (want_transpose) {
      ytmp.alloc(vec);
      ytmp.zero(s);
    }
    M64.alloc(mexp);
    M64.zero(s);
    Lc64.alloc(lexp);
    if (have32) {
      z32.alloc(vec);
      z32.zero(s);
      M32.alloc(mexp);
      M32.zero(s);
      Lc32.alloc(lexp);
      bf16_ok = m2l_inner == GP_M2L_BF16 &&
                ((ld % 16 == 0 && ld <= 128) || (ld == 1 && want_by_offset));
      if (kr) {
        const size_t texp = (size_t)f64.n_boxes * kr * ld;
        Mt32.alloc(texp);
        Mt32.zero(s);
        Lt32.alloc(texp);
        if (bf16_ok) {
          Mtbf.alloc(texp);
          Mtbf.zero(s);
          if (ld > 1) {  // blocks of RHS take the split-BF16 products
            Mtbf_lo.alloc(texp);
            Mtbf_lo.zero(s);
          }
        }
      } else if (bf16_ok) {
        Mbf.alloc(mexp);
        Mbf.zero(s);
      }
    }
  }

  // M2L kernel of the FP32 operator at the current ld
  int inner_m2l() const {
    return bf16_ok ? GP_M2L_BF16 : m2l_inner == GP_M2L_BF16_VALU ? GP_M2L_BF16_VALU : GP_M2L_FP32;
  }

  // y = P' M^-1 v (precond) or y = P' v, in FP64 (prec 0) or FP32 (prec 1).
  template <typename TV>
  void apply(const TV *v, bool precond, int nr, double *y, int prec) {
    const double *dinv = precond ? inv_diag.p : nullptr;
    ++matvecs;
    if (transposed) {
      apply_transposed(v, precond, nr, y, prec);
      return;
    }
    if (prec == 0) {
      if constexpr (sizeof(TV) != 8) throw Error(GP_ERR_INTERNAL, "FP64 operator on an FP32 basis");
      else kr_precond_cast(v, dinv, z64.p, N, ld, nr, s);
      launch_near_apply(nt, nf64.p, z64.p, ld, y, ld, nr, s);
      if (diel) {
        // potentials of all panels into yfar; conductor rows take them,
        // interface rows the normal derivative of the same local expansions
        yfar.zero(s);
        run_far_field(f64, z64.p, ld, nr, M64.p, Lc64.p, ld, m2l_outer, yfar.p, ld, s);
        launch_far_dielectric(f64, nr, ld, ld, dw1d64.p, d_nrm.p, d_iface.p, Lc64.p, yfar.p, y, s);
      } else {
        run_far_field(f64, z64.p, ld, nr, M64.p, Lc64.p, ld, m2l_outer, y, ld, s);
      }
    } else {
      if (!have32) throw Error(GP_ERR_INTERNAL, "FP32 operator not assembled");
      kr_precond_cast(v, dinv, z32.p, N, ld, nr, s);
      if (near16) {
        if (ld != 1 || nr != 1) throw Error(GP_ERR_INTERNAL, "FP16 near tiles take one RHS");
        launch_near_apply_f16(nt, nf16.p, nscale.p, z32.p, y, s);
      } else {
        launch_near_apply(nt, nf32.p, z32.p, ld, y, ld, nr, s);
      }
      run_far_field(f32, z32.p, ld, nr, M32.p, Mbf.p, Lc32.p, ld, inner_m2l(), y, ld, s, Mt32.p,
                    Mtbf.p, Lt32.p, Mtbf_lo.p);
    }
  }

  // y = P'^T M^-1 v: transposed near tiles, and the symmetric far kernel K
  // between area scalings (P'_far = K diag(area)).  The Jacobi preconditioner
  // is the same diagonal.
  template <typename TV>
  void apply_transposed(const TV *v, bool precond, int nr, double *y, int prec) {
    const double *dinv = precond ? inv_diag.p : nullptr;
    // dinv / area, or 1 / area without the preconditioner
    const double *dfar = precond ? dinv_area.p : inv_area.p;
    if (prec == 0) {
      if constexpr (sizeof(TV) != 8) throw Error(GP_ERR_INTERNAL, "FP64 operator on an FP32 basis");
      else {
        kr_precond_cast(v, dinv, z64.p, N, ld, nr, s);
        launch_near_apply_tr(nt, ntr, nf64.p, z64.p, ld, y, ld, nr, s);
        kr_precond_cast(v, dfar, z64.p, N, ld, nr, s);
      }
      run_far_field(f64, z64.p, ld, nr, M64.p, Lc64.p, ld, m2l_outer, ytmp.p, ld, s);
    } else {
      kr_precond_cast(v, dinv, z32.p, N, ld, nr, s);
      launch_near_apply_tr(nt, ntr, nf32.p, z32.p, ld, y, ld, nr, s);
      kr_precond_cast(v, dfar, z32.p, N, ld, nr, s);
      run_far_field(f32, z32.p, ld, nr, M32.p, Mbf.p, Lc32.p, ld, inner_m2l(), ytmp.p, ld, s,
                    Mt32.p, Mtbf.p, Lt32.p, Mtbf_lo.p);
    }
    kr_rows_axpy_clear(y, ytmp.p, area.p, N, ld, nr, s);
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

// Vector operations of the Krylov drivers over nsys() independent systems.
// BlockOps: nr right-hand sides of one problem, N x ld row-major, per-system
// state stride ld.  SegOps: a batch of problems stacked in one column, each
// system a row segment, per-system state stride nsys.
struct BlockOps {
  // a block of RHS of one problem: the operator applies every column
  static constexpr bool maskable = false;
  int N = 0, ld = 0, nr = 0;
  hipStream_t s = nullptr;
  int nsys() const { return nr; }
  int S() const { return ld; }
  int cols() const { return nr; }
  size_t vec() const { return (size_t)N * ld; }
  size_t partial_size(int nv) const { return kr_partial_size(nv, ld); }
  template <typename TX>
  void dots(const TX *X, size_t sx, int nv, const double *Y, double *part, double *out,
            bool sq, const int * = nullptr) const {
    kr_dots(X, sx, nv, Y, N, ld, nr, part, out, sq, s);
  }
  template <typename TV>
  void axpy_multi(double *w, const TV *V, size_t sv, const double *coef, int nv, double a,
                  const int * = nullptr) const {
    kr_axpy_multi(w, V, sv, coef, nv, N, ld, nr, a, s);
  }
  template <typename TO>
  void scale(const double *in, TO *out, const double *sc, const int *mask) const {
    kr_scale_cols(in, out, sc, mask, N, ld, nr, s);
  }
  void sub(const double *a, const double *b, double *out) const { kr_sub(a, b, out, N, ld, nr, s); }
  void axpy(double *x, const double *y, double a) const { kr_axpy(x, y, a, N, ld, nr, s); }
  template <typename TV>
  void update_x(double *x, const TV *V, size_t sv, const double *Y, const int *ku, int kmax,
                const double *dinv) const {
    kr_gmres_update_x(x, V, sv, Y, ku, kmax, dinv, N, ld, nr, s);
  }
};

struct SegOps {
  // a batch of problems, one system each: converged problems are skipped
  static constexpr bool maskable = true;
  SegMap sm;
  hipStream_t s = nullptr;
  int nsys() const { return sm.nsys; }
  int S() const { return sm.nsys; }
  int cols() const { return 1; }
  size_t vec() const { return (size_t)sm.N; }
  size_t partial_size(int nv) const { return seg_partial_size(sm, nv); }
  template <typename TX>
  void dots(const TX *X, size_t sx, int nv, const double *Y, double *part, double *out,
            bool sq, const int *act = nullptr) const {
    seg_dots(sm, X, sx, nv, Y, part, out, sq, s, act);
  }
  template <typename TV>
  void axpy_multi(double *w, const TV *V, size_t sv, const double *coef, int nv, double a,
                  const int *act = nullptr) const {
    seg_axpy_multi(sm, w, V, sv, coef, nv, a, s, act);
  }
  template <typename TO>
  void scale(const double *in, TO *out, const double *sc, const int *mask) const {
    seg_scale(sm, in, out, sc, mask, s);
  }
  void sub(const double *a, const double *b, double *out) const { kr_sub(a, b, out, sm.N, 1, 1, s); }
  void axpy(double *x, const double *y, double a) const { kr_axpy(x, y, a, sm.N, 1, 1, s); }
  template <typename TV>
  void update_x(double *x, const TV *V, size_t sv, const double *Y, const int *ku, int kmax,
                const double *dinv) const {
    seg_update_x(sm, x, V, sv, Y, ku, kmax, dinv, s);
  }
};

// GMRES workspace, kept across the solves of one system (the refinement
// steps of the mixed solver), with the captured Arnoldi steps that replay on
// it: graph k0 holds the steps k0 .. k0 + g_chunk - 1 of a cycle (each: the
// operator apply, Gram-Schmidt passes, norm, Givens update and the scaled new
// basis vector; then the copy of the active flags to the host), valid for
// the current basis allocation and the scalars it was captured with.
template <typename TV>
struct GmresWork {
  int m = 0, S = 0, vcap = 0;
  size_t vec = 0;
  DBuf<TV> V;
  DBuf<double> w, rv, ax, partial, h1, h2, hn, H, cs, sn, g, Y, bnorm, beta, inv_beta, est, inv_h;
  DBuf<int> live, active, kused;
  std::unique_ptr<Pinned> pin;
  std::vector<hipGraphExec_t> graphs;
  double g_tol = -1.0;
  int g_passes = 0, g_prec = -1, g_nc = 0, g_nr = 0, g_chunk = 0;
  void clear_graphs() {
    for (auto e : graphs)
      if (e) (void)hipGraphExecDestroy(e);
    graphs.clear();
  }
  ~GmresWork() { clear_graphs(); }
  template <typename Ops>
  void prepare(int m_, int S_, size_t vec_, const Ops &ops) {
    if (m_ == m && S_ == S && vec_ == vec) return;
    clear_graphs();
    m = m_;
    S = S_;
    vec = vec_;
    vcap = std::min(m, 32) + 1;
    V.alloc(vcap * vec);
    w.alloc(vec);
    rv.alloc(vec);
    ax.alloc(vec);
    partial.alloc(ops.partial_size(vcap));
    h1.alloc((size_t)(m + 1) * S);
    h2.alloc((size_t)(m + 1) * S);
    hn.alloc(S);
    H.alloc((size_t)(m + 1) * m * S);
    cs.alloc((size_t)m * S);
    sn.alloc((size_t)m * S);
    g.alloc((size_t)(m + 1) * S);
    Y.alloc((size_t)m * S);
    for (auto *b : {&bnorm, &beta, &inv_beta, &est, &inv_h}) b->alloc(S);
    for (auto *b : {&live, &active, &kused}) b->alloc(S);
    pin = std::make_unique<Pinned>((size_t)S * (2 * sizeof(int) + 2 * sizeof(double)));
  }
};

// Restarted GMRES(m) with right Jacobi preconditioning for the systems of
// ops (right-hand sides b, device) starting from x = 0 (x must be zero on
// entry), per system exactly the control flow of krylov.rs:50-191: stop a
// cycle at the residual estimate, check the true residual when the estimate
// is within 10 tol.  Orthogonalisation is classical Gram-Schmidt applied
// twice (CGS2, passes = 2), the block-friendly equivalent of the reference's
// modified Gram-Schmidt; the inner solves of the mixed solver (tolerance
// ~1e-2, corrected by the FP64 refinement) take one pass and keep the basis V
// in FP32 (TV = float).  work (optional) keeps the workspace across calls;
// with capture set, each Arnoldi step k runs as a HIP graph captured on its
// first use and replayed afterwards (never while the kernel profiler runs).
template <typename TV, typename Ops>
void gmres_run(FmmSystem &sys, const Ops &ops, int prec, const double *b, double *x, double tol,
               int m, int max_restarts, bool true_check, int passes,
               const std::vector<int> &live_in, GmresStats &st, GmresWork<TV> *work = nullptr,
               bool capture = false) {
  hipStream_t s = sys.s;
  const int nr = ops.nsys(), S = ops.S(), nc = ops.cols();
  const size_t vec = ops.vec();
  GmresWork<TV> local;
  GmresWork<TV> &W = work ? *work : local;
  W.prepare(m, S, vec, ops);
  int &vcap = W.vcap;
  DBuf<TV> &V = W.V;
  DBuf<double> &w = W.w, &rv = W.rv, &ax = W.ax, &partial = W.partial, &h1 = W.h1, &h2 = W.h2,
               &hn = W.hn, &H = W.H, &cs = W.cs, &sn = W.sn, &g = W.g, &Y = W.Y,
               &bnorm = W.bnorm, &beta = W.beta, &inv_beta = W.inv_beta, &est = W.est,
               &inv_h = W.inv_h;
  DBuf<int> &live = W.live, &active = W.active, &kused = W.kused;
  int *h_active = W.pin->template as<int>();
  int *h_kused = h_active + S;
  double *h_est = reinterpret_cast<double *>(h_kused + S);
  double *h_aux = h_est + S;
  const Prof *prof = active_prof();
  const bool graphs = capture && !(prof && prof->on);
  if (graphs && (W.g_tol != tol || W.g_passes != passes || W.g_prec != prec || W.g_nc != nc ||
                 W.g_nr != nr)) {
    W.clear_graphs();
    W.g_tol = tol;
    W.g_passes = passes;
    W.g_prec = prec;
    W.g_nc = nc;
    W.g_nr = nr;
  }

  auto ensure_V = [&](int nv) {
    if (nv <= vcap) return;
    int ncap = std::min(m + 1, std::max(nv, 2 * vcap));
    DBuf<TV> V2((size_t)ncap * vec);
    GP_HIP(hipMemcpyAsync(V2.p, V.p, (size_t)vcap * vec * sizeof(TV), hipMemcpyDeviceToDevice, s));
    GP_HIP(hipStreamSynchronize(s));  // the old basis is freed below
    V = std::move(V2);
    partial.alloc(ops.partial_size(ncap));
    vcap = ncap;
    W.clear_graphs();  // captured with the old basis
  };

  st.iters.assign(nr, 0);
  st.resid.assign(nr, 0.0);
  st.conv.assign(nr, 0);
  std::vector<int> done(nr);
  for (int r = 0; r < nr; ++r) done[r] = !live_in[r];
  // bnorm = max(||b||, MIN_POSITIVE)   (krylov.rs:58)
  ops.dots(b, 0, 1, b, partial.p, bnorm.p, true);
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
      sys.apply(x, false, nc, ax.p, prec);
      ops.sub(b, ax.p, rv.p);
    }
    ops.dots(rv.p, 0, 1, rv.p, partial.p, beta.p, true);
    std::vector<int> hl(S, 0);
    for (int r = 0; r < nr; ++r) hl[r] = !done[r];
    GP_HIP(hipMemcpyAsync(live.p, hl.data(), S * sizeof(int), hipMemcpyHostToDevice, s));
    kr_gmres_start(m, nr, S, beta.p, bnorm.p, tol, live.p, active.p, kused.p, g.p, inv_beta.p,
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
    ops.scale(rv.p, V.p, inv_beta.p, active.p);
    int ksteps = 0;
    // Steps per host check: with graphs, a batch (whose converged systems'
    // work the kernels skip) replays chunks of kBatchGraphSteps steps and
    // reads the active flags once per chunk; steps after a system converges
    // leave its state unchanged (inv_h = 0, kused fixed), so the answer is
    // the one-step-at-a-time answer.
    const int chunk = (graphs && Ops::maskable && sys.skip_converged) ? kBatchGraphSteps : 1;
    if (graphs && W.g_chunk != chunk) {
      W.clear_graphs();
      W.g_chunk = chunk;
    }
    for (int k0 = 0, len = 0; k0 < m; k0 += len) {
      len = std::min(chunk, m - k0);
      ensure_V(k0 + len + 1);
      // a batch skips the operator and Gram-Schmidt work of its inactive
      // systems (read from the device flags, so graph replays see the
      // current ones): their new basis vectors are zero (inv_h = 0) and
      // never enter x, and their coefficients are not used
      const int *act = (Ops::maskable && sys.skip_converged) ? active.p : nullptr;
      auto step = [&](int k) {
        if constexpr (Ops::maskable)
          if (sys.skip_converged) sys.mask(active.p, graphs ? nullptr : h_active);
        sys.apply(V.p + (size_t)k * sv, true, nc, w.p, prec);
        if constexpr (Ops::maskable) sys.unmask();
        ops.dots(V.p, sv, k + 1, w.p, partial.p, h1.p, false, act);
        ops.axpy_multi(w.p, V.p, sv, h1.p, k + 1, -1.0, act);
        if (passes > 1) {
          ops.dots(V.p, sv, k + 1, w.p, partial.p, h2.p, false, act);
          ops.axpy_multi(w.p, V.p, sv, h2.p, k + 1, -1.0, act);
        }
        ops.dots(w.p, 0, 1, w.p, partial.p, hn.p, true, act);
        kr_gmres_givens(k, m, nr, S, h1.p, passes > 1 ? h2.p : nullptr, hn.p, H.p, cs.p, sn.p,
                        g.p, bnorm.p, tol, active.p, kused.p, est.p, inv_h.p, s);
        ops.scale(w.p, V.p + (size_t)(k + 1) * sv, inv_h.p, act);
      };
      auto steps = [&] {
        for (int k = k0; k < k0 + len; ++k) step(k);
        GP_HIP(hipMemcpyAsync(h_active, active.p, nr * sizeof(int), hipMemcpyDeviceToHost, s));
      };
      if (graphs) {
        if ((int)W.graphs.size() <= k0) W.graphs.resize(k0 + 1, nullptr);
        if (!W.graphs[k0]) {
          // (no device allocation or copy of another thread inside the capture)
          std::lock_guard<std::recursive_mutex> lk(device_mem_mutex());
          hipGraph_t gr = nullptr;
          GP_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
          try {
            steps();
          } catch (...) {
            (void)hipStreamEndCapture(s, &gr);
            if (gr) (void)hipGraphDestroy(gr);
            throw;
          }
          GP_HIP(hipStreamEndCapture(s, &gr));
          hipGraphExec_t ex = nullptr;
          const hipError_t e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
          (void)hipGraphDestroy(gr);
          GP_HIP(e);
          W.graphs[k0] = ex;
        }
        GP_HIP(hipGraphLaunch(W.graphs[k0], s));
      } else {
        steps();
      }
      ksteps = k0 + len;
      GP_HIP(hipStreamSynchronize(s));
      int na = 0;
      for (int r = 0; r < nr; ++r) na += h_active[r];
      if (ksteps / 25 != k0 / 25)
        GP_LOG("gmres prec %d cycle %d iteration %d: %d of %d systems active", prec, cycle, ksteps,
               na, nr);
      if (!na) break;
    }
    kr_gmres_backsolve(m, nr, S, H.p, g.p, kused.p, Y.p, s);
    ops.update_x(x, V.p, sv, Y.p, kused.p, ksteps, sys.inv_diag.p);
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
        sys.apply(x, false, nc, ax.p, prec);
        ops.sub(b, ax.p, rv.p);
        ops.dots(rv.p, 0, 1, rv.p, partial.p, beta.p, true);
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
  // Systems left unconverged: report their true residual (krylov.rs:185-190).
  bool any_left = false;
  for (int r = 0; r < nr; ++r) any_left |= !done[r];
  if (any_left) {
    sys.apply(x, false, nc, ax.p, prec);
    ops.sub(b, ax.p, rv.p);
    ops.dots(rv.p, 0, 1, rv.p, partial.p, beta.p, true);
    GP_HIP(hipMemcpyAsync(h_aux, beta.p, nr * sizeof(double), hipMemcpyDeviceToHost, s));
    GP_HIP(hipStreamSynchronize(s));
    for (int r = 0; r < nr; ++r)
      if (!done[r]) st.resid[r] = h_aux[r] / hb[r];
  }
}

// Per-system outcome of one solve of the systems of ops.
struct SolveStats {
  std::vector<int> iters;       // GMRES iterations (summed over refinement steps)
  std::vector<double> resid;    // final relative residual
  std::vector<char> conv;
  int refine_steps = 0;
};

// Solves sys x = b for the live systems of ops: FP64 GMRES, or (mixed) FP64
// iterative refinement around the FP32 inner GMRES.  x must be zero on entry.
template <typename Ops>
void solve_systems(FmmSystem &sys, const Ops &ops, const gp_cap_opts &o, bool mixed,
                   const double *b, double *x, const std::vector<int> &live, SolveStats &out) {
  hipStream_t s = sys.s;
  const int nr = ops.nsys(), S = ops.S(), nc = ops.cols();
  const size_t vec = ops.vec();
  const int m = std::max(1, (int)o.restart), mr = std::max(1, (int)o.max_restarts);
  out.iters.assign(nr, 0);
  out.resid.assign(nr, 0.0);
  out.conv.assign(nr, 0);
  out.refine_steps = 0;
  if (!mixed) {
    GmresStats st;
    gmres_run<double>(sys, ops, 0, b, x, o.tol, m, mr, true, 2, live, st);
    out.iters = st.iters;
    out.resid = st.resid;
    out.conv = st.conv;
    return;
  }
  // FP64 iterative refinement around the FP32 inner GMRES.
  DBuf<double> r(vec), dx(vec), ax(vec), part(ops.partial_size(1)), rn(S);
  GP_HIP(hipMemcpyAsync(r.p, b, vec * sizeof(double), hipMemcpyDeviceToDevice, s));
  std::vector<double> hbn(nr), hrn(nr);
  ops.dots(b, 0, 1, b, part.p, rn.p, true);
  GP_HIP(hipMemcpyAsync(hbn.data(), rn.p, nr * sizeof(double), hipMemcpyDeviceToHost, s));
  GP_HIP(hipStreamSynchronize(s));
  for (auto &v : hbn) v = std::max(v, kMinPositive);
  std::vector<int> done(nr, 0);
  for (int q = 0; q < nr; ++q) done[q] = !live[q];
  const int max_ref = std::max(1, (int)o.max_refine);
  const int passes = o.inner_passes >= 2 ? 2 : 1;
  // inner workspace and its captured Arnoldi steps, shared by the refinement
  // steps; capture starts with the second step (the first one also settles
  // the autotuned kernel variants)
  GmresWork<float> work32;
  GmresWork<double> work64;
  // A batch recomputes the outer residual only for the problems still being
  // refined: a finished problem's x no longer changes, so its rows of A x
  // (kept in ax from the step that finished it; zero for problems with no
  // column here) and of r stay what a new apply would give, bit for bit.
  bool mask_outer = false;
  DBuf<int> d_lv;
  std::vector<int> h_lv(nr);  // (the source of d_lv's copy outlives it)
  if constexpr (Ops::maskable) {
    mask_outer = sys.skip_converged && sys.box_prob.p;
    if (mask_outer) {
      ax.zero(s);
      d_lv.alloc(nr);
    }
  }
  for (;;) {
    ops.dots(r.p, 0, 1, r.p, part.p, rn.p, true);
    GP_HIP(hipMemcpyAsync(hrn.data(), rn.p, nr * sizeof(double), hipMemcpyDeviceToHost, s));
    GP_HIP(hipStreamSynchronize(s));
    bool all = true;
    for (int q = 0; q < nr; ++q) {
      if (live[q] && hrn[q] / hbn[q] <= o.tol) done[q] = 1;
      all &= (bool)done[q];
    }
    if (all || out.refine_steps >= max_ref) break;
    std::vector<int> lv(nr);
    for (int q = 0; q < nr; ++q) lv[q] = !done[q];
    dx.zero(s);
    GmresStats st;
    const bool capture = o.step_graphs != 0 && out.refine_steps >= 1;
    if (o.inner_basis == GP_BASIS_FP64)
      gmres_run<double>(sys, ops, 1, r.p, dx.p, o.inner_tol, m, mr, false, passes, lv, st,
                        &work64, capture);
    else
      gmres_run<float>(sys, ops, 1, r.p, dx.p, o.inner_tol, m, mr, false, passes, lv, st, &work32,
                       capture);
    for (int q = 0; q < nr; ++q) out.iters[q] += st.iters[q];
    ops.axpy(x, dx.p, 1.0);
    if (mask_outer) {
      h_lv = lv;
      GP_HIP(hipMemcpyAsync(d_lv.p, h_lv.data(), nr * sizeof(int), hipMemcpyHostToDevice, s));
      sys.mask(d_lv.p, h_lv.data());
    }
    sys.apply(x, false, nc, ax.p, 0);
    if (mask_outer) sys.unmask();
    ops.sub(b, ax.p, r.p);
    ++out.refine_steps;
    GP_LOG("refinement step %d: inner iterations up to %d", out.refine_steps,
           *std::max_element(st.iters.begin(), st.iters.end()));
  }
  for (int q = 0; q < nr; ++q) {
    out.resid[q] = live[q] ? hrn[q] / hbn[q] : 0.0;
    out.conv[q] = live[q] ? (char)done[q] : 1;
  }
}

void solve_fmm(gp_ctx *ctx, const PanelGeom &g, const gp_cap_opts &o, double *C,
               gp_cap_report *rep, const double *eps_panel, const DielRows *diel = nullptr) {
  hipStream_t s = ctx->stream;
  double t0 = now_s();
  const bool mixed = o.precision == GP_PREC_MIXED;
  const int nc = g.ncond;
  const int jlo = std::min((int)o.rhs_first, nc);
  const int jhi = o.rhs_count ? std::min(nc, jlo + (int)o.rhs_count) : nc;
  const int nsolve = std::max(1, jhi - jlo);
  // Blocks of at most 128 columns: the matrix-core M2L holds one block.
  int blk = o.rhs_block ? (int)o.rhs_block : nsolve;
  blk = std::max(1, std::min(blk, std::min(nsolve, 128)));
  const int ld = blk == 1 ? 1 : round_up(blk, 16);
  FmmSystem sys;
  sys.want_by_offset = ld == 1;
  sys.diel = diel;
  sys.setup(g, o, mixed, s);
  sys.ensure_work(ld);
  GP_HIP(hipStreamSynchronize(s));
  double t1 = now_s();
  GP_LOG("setup %.3f s: %d panels, depth %d, %d boxes, near nnz %lld, %lld M2L pairs, ld %d",
         t1 - t0, sys.N, sys.L, sys.f64.n_boxes, (long long)sys.nnz, (long long)sys.pairs, ld);

  const size_t vec = (size_t)sys.N * ld;
  DBuf<double> b(vec), x(vec), dCb((size_t)nc * ld);
  std::vector<double> hC((size_t)nc * nc, 0.0), hblk((size_t)nc * ld);
  // Charge weights: the panel areas, times the surrounding permittivity of
  // each panel for the conductor-role dielectric solve (free charge
  // eps_k * sigma_k * a_k, cap/dielectric.rs:141-149).
  DBuf<double> wq;
  const double *w = sys.qw.p ? sys.qw.p : sys.area.p;
  if (eps_panel) {
    std::vector<double> ha(sys.N);
    GP_HIP(hipMemcpyAsync(ha.data(), w, sys.N * sizeof(double), hipMemcpyDeviceToHost, s));
    GP_HIP(hipStreamSynchronize(s));
    for (int k = 0; k < sys.N; ++k) ha[k] *= eps_panel[sys.t().perm[k]];
    wq = to_device(ha, s);
    w = wq.p;
  }
  int max_iters = 0, refine_steps = 0;
  double worst = 0.0;
  bool conv_all = true;
  for (int j0 = jlo; j0 < jhi; j0 += blk) {
    const int nr = std::min(blk, jhi - j0);
    kr_unit_rhs(sys.cond.p, j0, sys.N, ld, nr, b.p, s, sys.sgn.p);
    x.zero(s);
    BlockOps ops;
    ops.N = sys.N;
    ops.ld = ld;
    ops.nr = nr;
    ops.s = s;
    SolveStats st;
    solve_systems(sys, ops, o, mixed, b.p, x.p, std::vector<int>(nr, 1), st);
    refine_steps = std::max(refine_steps, st.refine_steps);
    for (int q = 0; q < nr; ++q) {
      max_iters = std::max(max_iters, st.iters[q]);
      worst = std::max(worst, st.resid[q]);
      conv_all &= (bool)st.conv[q];
    }
    launch_cap_accumulate(nc, nr, ld, sys.coff.p, sys.cidx.p, w, x.p, dCb.p, s);
    GP_HIP(hipMemcpyAsync(hblk.data(), dCb.p, (size_t)nc * nr * sizeof(double),
                          hipMemcpyDeviceToHost, s));
    GP_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < nc; ++i)
      for (int q = 0; q < nr; ++q) hC[(size_t)i * nc + j0 + q] = hblk[(size_t)i * nr + q];
  }
  const double scale = 4.0 * kPi * kEps0 * sys.L0();
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
    rep->n_leaves = (uint32_t)sys.n_leaves;
    rep->n_boxes = (uint32_t)sys.f64.n_boxes;
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

// A batch of independent FMM problems solved together: one stacked,
// block-diagonal operator (every kernel launch covers all problems) and one
// GMRES per problem running in lockstep with per-problem reductions.  Each
// right-hand-side column j of [rhs_first, rhs_first + rhs_count) is one pass
// over the batch (problems with fewer conductors sit it out).
// Splits a batch's FMM problems into k groups, group j holding a share of
// the panels proportional to kGroupGrowth^j (largest problems first, each
// into the group furthest below its share); problems keep their order within
// a group.  The host set-up of a group (its near field assembled on a stream
// of its own) takes about half the device solve of the same panels, so a
// set-up 1.8 times the previous group's stays hidden behind that group's
// solve, and the first group, whose set-up nothing hides, is small (nets4096:
// 8.96M unknowns/s with growth 1.5, 9.17M with 1.8, first set-up 0.12 ->
// 0.09 s).
constexpr double kGroupGrowth = 1.8;
std::vector<std::vector<int>> batch_groups(const std::vector<const PanelGeom *> &probs, int k) {
  const int P = (int)probs.size();
  k = std::max(1, std::min({k, P, 16}));
  double total = 0.0;
  for (auto *g : probs) total += g->n;
  std::vector<double> share(k), fill(k, 0.0);
  double wsum = 0.0;
  for (int j = 0; j < k; ++j) wsum += std::pow(kGroupGrowth, j);
  for (int j = 0; j < k; ++j) share[j] = total * std::pow(kGroupGrowth, j) / wsum;
  std::vector<int> by_size(P);
  std::iota(by_size.begin(), by_size.end(), 0);
  std::stable_sort(by_size.begin(), by_size.end(),
                   [&](int a, int b) { return probs[a]->n > probs[b]->n; });
  std::vector<std::vector<int>> grp(k);
  for (int p : by_size) {
    int best = 0;
    for (int j = 1; j < k; ++j)
      if (share[j] - fill[j] > share[best] - fill[best]) best = j;
    grp[best].push_back(p);
    fill[best] += probs[p]->n;
  }
  std::vector<std::vector<int>> out;
  for (auto &g : grp)
    if (!g.empty()) {
      std::sort(g.begin(), g.end());
      out.push_back(std::move(g));
    }
  return out;
}

// One group of a batch, set up: its stacked operator and the set-up's
// timing, profile records and first error.
struct StagedGroup {
  std::vector<int> idx;  // positions in the batch
  std::vector<const PanelGeom *> probs;
  std::unique_ptr<FmmSystem> sys;
  double t0 = 0.0, t1 = 0.0, upload = 0.0;
  Prof prof;
  std::exception_ptr err;
};

// FmmSystem::setup of one group on ctx->setup_stream (device current,
// profiling into the group's own records, uploads timed apart), from a
// thread of its own or from the calling thread, whose per-thread profiler
// and upload time are restored.
void stage_group(gp_ctx *ctx, const gp_cap_opts &o, bool mixed, bool profile, StagedGroup &st) {
  st.t0 = now_s();
  Prof *const prev_prof = active_prof();
  const double prev_upload = upload_seconds();
  try {
    GP_HIP(hipSetDevice(ctx->device));
    st.prof.on = profile;
    active_prof() = &st.prof;
    upload_seconds() = 0.0;
    st.sys.reset(new FmmSystem());
    st.sys->want_by_offset = true;
    st.sys->setup(st.probs, o, mixed, ctx->setup_stream, ctx->asm_stream);
    GP_HIP(hipStreamSynchronize(ctx->setup_stream));
    GP_HIP(hipStreamSynchronize(ctx->asm_stream));
  } catch (...) {
    st.err = std::current_exception();
  }
  active_prof() = prev_prof;
  st.upload = upload_seconds();
  upload_seconds() = prev_upload;
  st.t1 = now_s();
}

// The GMRES solves of one set-up group, on ctx->stream: every conductor
// column of every problem, C and the reports by group position.
void solve_group(gp_ctx *ctx, FmmSystem &sys, const std::vector<const PanelGeom *> &probs,
                 const int *batch_idx, const gp_cap_opts &o, bool mixed, double t_setup,
                 double *const *C, gp_cap_report *reps) {
  hipStream_t s = ctx->stream;
  sys.s = s;
  const int P = (int)probs.size();
  const double t_start = now_s();
  sys.ensure_work(1);
  // Row segments and the reduction block map (at most 2048 rows a block).
  std::vector<int> hseg(sys.N), blo, bhi, sblk{0};
  for (int p = 0; p < P; ++p) {
    const int r0 = sys.panel_off[p], r1 = r0 + probs[p]->n;
    for (int r = r0; r < r1; ++r) hseg[r] = p;
    for (int r = r0; r < r1; r += 2048) {
      blo.push_back(r);
      bhi.push_back(std::min(r1, r + 2048));
    }
    sblk.push_back((int)blo.size());
  }
  DBuf<int> dseg = to_device(hseg, s), dblo = to_device(blo, s), dbhi = to_device(bhi, s),
            dsblk = to_device(sblk, s);
  SegOps ops;
  ops.sm.N = sys.N;
  ops.sm.nsys = P;
  ops.sm.nblk = (int)blo.size();
  ops.sm.seg = dseg.p;
  ops.sm.blk_lo = dblo.p;
  ops.sm.blk_hi = dbhi.p;
  ops.sm.sys_blk = dsblk.p;
  ops.s = s;
  GP_HIP(hipStreamSynchronize(s));
  GP_LOG("group set-up %.3f s: %d problems, %d panels, max depth %d, %d boxes, near nnz %lld, "
         "%lld M2L pairs", t_setup, P, sys.N, sys.L, sys.f64.n_boxes, (long long)sys.nnz,
         (long long)sys.pairs);
  const double t1 = now_s();

  int maxc = 0;
  for (auto *g : probs) maxc = std::max(maxc, g->ncond);
  const int jlo = std::min((int)o.rhs_first, maxc);
  const int jhi = o.rhs_count ? std::min(maxc, jlo + (int)o.rhs_count) : maxc;
  DBuf<double> b(sys.N), x(sys.N), dC(sys.ncond);
  DBuf<int> drhs(P);
  std::vector<double> hdC(sys.ncond);
  std::vector<int> iters(P, 0), refine(P, 0), conv(P, 1);
  std::vector<double> resid(P, 0.0);
  for (int p = 0; p < P; ++p)
    std::fill(C[p], C[p] + (size_t)probs[p]->ncond * probs[p]->ncond, 0.0);
  for (int j = jlo; j < jhi; ++j) {
    std::vector<int> hr(P), live(P);
    for (int p = 0; p < P; ++p) {
      live[p] = j < probs[p]->ncond;
      hr[p] = live[p] ? sys.cond_off[p] + j : -1;
    }
    GP_HIP(hipMemcpyAsync(drhs.p, hr.data(), P * sizeof(int), hipMemcpyHostToDevice, s));
    seg_unit_rhs(ops.sm, sys.cond.p, drhs.p, b.p, s, sys.sgn.p);
    x.zero(s);
    SolveStats st;
    solve_systems(sys, ops, o, mixed, b.p, x.p, live, st);
    launch_cap_accumulate(sys.ncond, 1, 1, sys.coff.p, sys.cidx.p,
                          sys.qw.p ? sys.qw.p : sys.area.p, x.p, dC.p, s);
    GP_HIP(hipMemcpyAsync(hdC.data(), dC.p, sys.ncond * sizeof(double), hipMemcpyDeviceToHost, s));
    GP_HIP(hipStreamSynchronize(s));
    for (int p = 0; p < P; ++p) {
      if (!live[p]) continue;
      const int nc = probs[p]->ncond;
      const double scale = 4.0 * kPi * kEps0 * sys.L0s[p];
      for (int i = 0; i < nc; ++i) C[p][(size_t)i * nc + j] = hdC[sys.cond_off[p] + i] * scale;
      iters[p] = std::max(iters[p], st.iters[p]);
      refine[p] = std::max(refine[p], st.refine_steps);
      resid[p] = std::max(resid[p], st.resid[p]);
      conv[p] &= (int)st.conv[p];
    }
  }
  const double t2 = now_s();
  if (reps)
    for (int p = 0; p < P; ++p) {
      gp_cap_report &r = reps[p];
      r.method_used = GP_METHOD_FMM;
      r.precision_used = mixed ? GP_PREC_MIXED : GP_PREC_FP64;
      r.iterations = (uint32_t)iters[p];
      r.matvecs = (uint32_t)sys.matvecs;
      r.refine_steps = (uint32_t)refine[p];
      r.converged = (uint32_t)conv[p];
      r.max_level = (uint32_t)sys.trees[p].L;
      r.n_leaves = (uint32_t)sys.trees[p].n_leaves();
      r.n_boxes = (uint32_t)sys.trees[p].n_boxes();
      r.fmm_order = (uint32_t)sys.n;
      // the problem's own share of the stacked operator
      r.near_nnz = P > 1 ? (uint64_t)sys.p_nnz[p] : (uint64_t)sys.nnz;
      r.m2l_pairs = P > 1 ? (uint64_t)sys.p_pairs[p] : (uint64_t)sys.pairs;
      r.final_rel_residual = resid[p];
      r.seconds_setup = t_setup + (t1 - t_start);
      r.seconds_solve = t2 - t1;
    }
  if (o.strict)
    for (int p = 0; p < P; ++p)
      if (!conv[p])
        throw Error(GP_ERR_NOT_CONVERGED, "GMRES did not reach tol for problem " +
                                              std::to_string(batch_idx[p]) + " (relative residual " +
                                              std::to_string(resid[p]) + ")");
}

// The FMM problems of a batched call, in groups (gp_cap_opts::batch_groups):
// a second host thread sets up group k + 1 on ctx->setup_stream while this
// thread solves group k on ctx->stream, and `before` (the dense problems'
// solve) runs while group 0 is set up.  Only group 0's uploads count in
// upload_seconds(): the later groups' copies overlap a solve.
void solve_fmm_batch(gp_ctx *ctx, const std::vector<const PanelGeom *> &probs,
                     const gp_cap_opts &o, const std::vector<double *> &C, gp_cap_report *reps,
                     const std::function<void()> &before) {
  const bool mixed = o.precision == GP_PREC_MIXED;
  double total = 0.0;
  for (auto *g : probs) total += g->n;
  const int k = o.batch_groups ? (int)o.batch_groups : (total < 2.0e6 ? 1 : 4);
  const auto groups = batch_groups(probs, k);
  const bool profile = active_prof() && active_prof()->on;
  auto make = [&](int j) {
    auto st = std::make_unique<StagedGroup>();
    st->idx = groups[j];
    for (int p : st->idx) st->probs.push_back(probs[p]);
    return st;
  };
  // the set-up thread's profile records move to this thread's profile
  auto absorb = [&](StagedGroup &st) {
    Prof *mine = active_prof();
    if (mine) {
      mine->recs.insert(mine->recs.end(), st.prof.recs.begin(), st.prof.recs.end());
      mine->pool.insert(mine->pool.end(), st.prof.pool.begin(), st.prof.pool.end());
      st.prof.recs.clear();
      st.prof.pool.clear();
    }
  };
  std::unique_ptr<StagedGroup> cur = make(0), next;
  if (groups.size() == 1) {
    // one group: nothing to overlap but the dense problems; every launch
    // from this thread (the set-up runs here, on the set-up stream)
    if (before) before();
    stage_group(ctx, o, mixed, profile, *cur);
  } else {
    StagedGroup *c = cur.get();
    std::thread th([&, c] { stage_group(ctx, o, mixed, profile, *c); });
    try {
      if (before) before();
    } catch (...) {
      th.join();
      throw;
    }
    th.join();
  }
  absorb(*cur);
  if (cur->err) std::rethrow_exception(cur->err);
  upload_seconds() += cur->upload;
  for (size_t j = 0; j < groups.size(); ++j) {
    std::thread th;
    if (j + 1 < groups.size()) {
      next = make((int)j + 1);
      StagedGroup *nx = next.get();
      th = std::thread([&, nx] { stage_group(ctx, o, mixed, profile, *nx); });
    }
    const double ts0 = now_s();
    try {
      const int G = (int)cur->idx.size();
      std::vector<double *> Cg(G);
      std::vector<gp_cap_report> rg(G);
      for (int q = 0; q < G; ++q) {
        Cg[q] = C[cur->idx[q]];
        if (reps) rg[q] = reps[cur->idx[q]];
      }
      solve_group(ctx, *cur->sys, cur->probs, cur->idx.data(), o, mixed, cur->t1 - cur->t0,
                  Cg.data(), reps ? rg.data() : nullptr);
      if (reps)
        for (int q = 0; q < G; ++q) reps[cur->idx[q]] = rg[q];
      cur->sys.reset();  // this group's device memory
    } catch (...) {
      if (th.joinable()) th.join();
      throw;
    }
    const double ts1 = now_s();
    if (th.joinable()) {
      th.join();
      GP_LOG("group %zu: solve %.3f s; then waited %.3f s for group %zu's set-up (%.3f s)", j,
             ts1 - ts0, now_s() - ts1, j + 1, next->t1 - next->t0);
      absorb(*next);
      if (next->err) std::rethrow_exception(next->err);
      cur = std::move(next);
    }
  }
}

}  // namespace

// cap::solve_dielectric (cap/dielectric.rs:47-151): conductor rows of
// potential coefficients, dielectric-interface rows of normal fields with the
// jump term on the diagonal, one LU, one solve per conductor, and the charge
// of each conductor's panels weighted by the permittivity around them.  In
// the unit frame (sigma' = k_e L0 sigma) the interface rows are scaled by L0:
// -F(l, c_k, n_k) off the diagonal (scale-free) and
// 2 pi (eps_out + eps_in) / (eps_out - eps_in) on it.
void solve_dense_roles(gp_ctx *ctx, const PanelGeom &g, const gp_panel_role *roles, double *C,
                       gp_cap_report *rep, const double *ground_z) {
  hipStream_t s = ctx->stream;
  double t0 = now_s();
  const int n = g.n, nc = g.ncond;
  double c[3], L0;
  root_cube(g, c, L0);
  std::vector<double> hv, hc, ha;
  std::vector<uint32_t> hcond;
  scaled_arrays(g, nullptr, c, L0, hv, hc, ha, hcond);
  std::vector<double> w(n, 0.0), B((size_t)n * nc, 0.0), hn, hd;
  std::vector<int> drow, coff(nc + 1, 0), cidx;
  for (int k = 0; k < n; ++k) {
    const gp_panel_role &r = roles[k];
    if (r.kind == GP_ROLE_CONDUCTOR) {
      if (r.conductor >= (uint32_t)nc)
        throw Error(GP_ERR_INVALID_INPUT, "panel " + std::to_string(k) + " conductor out of range");
      if (!std::isfinite(r.eps_surrounding))
        throw Error(GP_ERR_INVALID_INPUT, "panel " + std::to_string(k) + " has a non-finite eps");
      w[k] = ha[k] * r.eps_surrounding;
      B[(size_t)k * nc + r.conductor] = 1.0;
      coff[r.conductor + 1]++;
    } else if (r.kind == GP_ROLE_DIELECTRIC) {
      if (!(std::isfinite(r.eps_out) && std::isfinite(r.eps_in)) || r.eps_out == r.eps_in)
        throw Error(GP_ERR_INVALID_INPUT, "panel " + std::to_string(k) +
                                              ": interface permittivities must be finite and differ");
      // Panel::normal (panel.rs:50-55) toward the outer side (dielectric.rs:39-46)
      const double *q = &g.verts[(size_t)k * 12];
      const double ux = q[3] - q[0], uy = q[4] - q[1], uz = q[5] - q[2];
      const double vx = q[6] - q[0], vy = q[7] - q[1], vz = q[8] - q[2];
      double nx = uy * vz - uz * vy, ny = uz * vx - ux * vz, nz = ux * vy - uy * vx;
      const double len = std::sqrt(nx * nx + ny * ny + nz * nz);
      if (!(len > 0.0)) throw Error(GP_ERR_INVALID_INPUT, "panel " + std::to_string(k) + " has no normal");
      nx /= len, ny /= len, nz /= len;
      if (nx * (g.cx[k] - r.reference[0]) + ny * (g.cy[k] - r.reference[1]) +
              nz * (g.cz[k] - r.reference[2]) < 0.0)
        nx = -nx, ny = -ny, nz = -nz;
      drow.push_back(k);
      hn.insert(hn.end(), {nx, ny, nz});
      hd.push_back(2.0 * kPi * (r.eps_out + r.eps_in) / (r.eps_out - r.eps_in));
    } else {
      throw Error(GP_ERR_INVALID_INPUT, "panel " + std::to_string(k) + " has an unknown role");
    }
  }
  // conductor panels by conductor, ascending panel index (the summation order)
  for (int j = 0; j < nc; ++j) coff[j + 1] += coff[j];
  cidx.resize(coff[nc]);
  {
    std::vector<int> fill(coff.begin(), coff.end() - 1);
    for (int k = 0; k < n; ++k)
      if (roles[k].kind == GP_ROLE_CONDUCTOR) cidx[fill[roles[k].conductor]++] = k;
  }
  if (ground_z && !drow.empty())
    throw Error(GP_ERR_INVALID_INPUT, "the ground plane by images takes conductor panels only");
  auto d_v = to_device(hv, s);
  auto d_c = to_device(hc, s);
  DBuf<double> d_P((size_t)n * n);
  launch_dense_assemble(n, d_v.p, d_c.p, d_P.p, s, ground_z != nullptr,
                        ground_z ? 2.0 * (*ground_z - c[2]) / L0 : 0.0);
  if (!drow.empty()) {
    auto d_dr = to_device(drow, s);
    auto d_n = to_device(hn, s);
    auto d_d = to_device(hd, s);
    launch_dielectric_rows(n, (int)drow.size(), d_dr.p, d_n.p, d_d.p, d_v.p, d_c.p, d_P.p, s);
    GP_HIP(hipStreamSynchronize(s));
  }
  GP_HIP(hipStreamSynchronize(s));
  double t1 = now_s();
  std::vector<double> hC((size_t)nc * nc, 0.0);
  if (n <= kOneGroupLU) {
    // one work-group: the LU of the dense conductor path
    auto d_w = to_device(w, s);
    auto d_B = to_device(B, s);
    auto d_coff = to_device(coff, s);
    auto d_cidx = to_device(cidx, s);
    DBuf<double> d_work((size_t)n * nc), d_C((size_t)nc * nc);
    DBuf<int> d_piv(n), d_info(1);
    launch_lu_solve(n, nc, d_P.p, d_piv.p, d_B.p, d_work.p, d_info.p, s);
    launch_cap_accumulate(nc, nc, nc, d_coff.p, d_cidx.p, d_w.p, d_B.p, d_C.p, s);
    int info = 0;
    GP_HIP(hipMemcpyAsync(&info, d_info.p, sizeof(int), hipMemcpyDeviceToHost, s));
    GP_HIP(hipMemcpyAsync(hC.data(), d_C.p, hC.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    GP_HIP(hipStreamSynchronize(s));
    if (info != 0) throw Error(GP_ERR_SINGULAR, "dielectric system is singular");
  } else {
    // larger systems: rocSOLVER's blocked LU over the whole device.  The
    // row-major matrix is the column-major A^T, so getrs solves with the
    // transpose of that factorization.
    std::vector<double> Bc((size_t)n * nc);
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < nc; ++j) Bc[(size_t)j * n + k] = B[(size_t)k * nc + j];
    auto d_B = to_device(Bc, s);
    DBuf<int> d_piv(n), d_info(1);
    if (!ctx->blas) {
      rocblas_handle h0;
      if (rocblas_create_handle(&h0) != rocblas_status_success)
        throw Error(GP_ERR_INTERNAL, "rocblas_create_handle failed");
      ctx->blas = std::shared_ptr<void>(h0, [](void *p) {
        (void)rocblas_destroy_handle(static_cast<rocblas_handle>(p));
      });
    }
    rocblas_handle h = static_cast<rocblas_handle>(ctx->blas.get());
    rocblas_status st = rocblas_set_stream(h, s);
    if (st == rocblas_status_success) st = rocsolver_dgetrf(h, n, n, d_P.p, n, d_piv.p, d_info.p);
    if (st == rocblas_status_success)
      st = rocsolver_dgetrs(h, rocblas_operation_transpose, n, nc, d_P.p, n, d_piv.p, d_B.p, n);
    int info = 0;
    if (st == rocblas_status_success) {
      GP_HIP(hipMemcpyAsync(&info, d_info.p, sizeof(int), hipMemcpyDeviceToHost, s));
      GP_HIP(hipMemcpyAsync(Bc.data(), d_B.p, Bc.size() * sizeof(double), hipMemcpyDeviceToHost, s));
      GP_HIP(hipStreamSynchronize(s));
    }
    if (st != rocblas_status_success)
      throw Error(GP_ERR_INTERNAL, "rocSOLVER getrf/getrs failed (" + std::to_string((int)st) + ")");
    if (info != 0) throw Error(GP_ERR_SINGULAR, "dielectric system is singular");
    // charges of each conductor's panels in ascending panel order
    for (int i = 0; i < nc; ++i)
      for (int q = coff[i]; q < coff[i + 1]; ++q) {
        const int k = cidx[q];
        for (int j = 0; j < nc; ++j) hC[(size_t)i * nc + j] += Bc[(size_t)j * n + k] * w[k];
      }
  }
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


void solve_dielectric_fmm(gp_ctx *ctx, const PanelGeom &g, const gp_panel_role *roles,
                          const gp_cap_opts &opts, double *C, gp_cap_report *rep) {
  const double t0 = now_s();
  const int n = g.n, nc = g.ncond;
  if (opts.precision != GP_PREC_FP64)
    throw Error(GP_ERR_INVALID_INPUT, "the FMM dielectric solve runs in FP64 only");
  if (!(opts.tol > 0.0) || !std::isfinite(opts.tol))
    throw Error(GP_ERR_INVALID_INPUT, "tol must be > 0");
  DielRows dr;
  dr.iface.assign(n, 0);
  dr.nrm.assign((size_t)n * 3, 0.0);
  dr.diag.assign(n, 0.0);
  // conductor ids for the solve: interface panels form one extra
  // pseudo-conductor nc whose row of C is dropped; charge weights
  // eps_surrounding * area on conductor panels only (dielectric.rs:141-149)
  std::vector<uint32_t> cond(n);
  std::vector<double> w(n, 0.0);
  for (int k = 0; k < n; ++k) {
    const gp_panel_role &r = roles[k];
    if (r.kind == GP_ROLE_CONDUCTOR) {
      if (r.conductor >= (uint32_t)nc)
        throw Error(GP_ERR_INVALID_INPUT, "panel " + std::to_string(k) + " conductor out of range");
      if (!std::isfinite(r.eps_surrounding))
        throw Error(GP_ERR_INVALID_INPUT, "panel " + std::to_string(k) + " has a non-finite eps");
      cond[k] = r.conductor;
      w[k] = r.eps_surrounding;
    } else if (r.kind == GP_ROLE_DIELECTRIC) {
      if (!(std::isfinite(r.eps_out) && std::isfinite(r.eps_in)) || r.eps_out == r.eps_in)
        throw Error(GP_ERR_INVALID_INPUT, "panel " + std::to_string(k) +
                                              ": interface permittivities must be finite and differ");
      // Panel::normal (panel.rs:50-55) toward the outer side (dielectric.rs:39-46)
      const double *q = &g.verts[(size_t)k * 12];
      const double ux = q[3] - q[0], uy = q[4] - q[1], uz = q[5] - q[2];
      const double vx = q[6] - q[0], vy = q[7] - q[1], vz = q[8] - q[2];
      double nx = uy * vz - uz * vy, ny = uz * vx - ux * vz, nz = ux * vy - uy * vx;
      const double len = std::sqrt(nx * nx + ny * ny + nz * nz);
      if (!(len > 0.0)) throw Error(GP_ERR_INVALID_INPUT, "panel " + std::to_string(k) + " has no normal");
      nx /= len, ny /= len, nz /= len;
      if (nx * (g.cx[k] - r.reference[0]) + ny * (g.cy[k] - r.reference[1]) +
              nz * (g.cz[k] - r.reference[2]) < 0.0)
        nx = -nx, ny = -ny, nz = -nz;
      dr.iface[k] = 1;
      dr.nrm[(size_t)k * 3] = nx;
      dr.nrm[(size_t)k * 3 + 1] = ny;
      dr.nrm[(size_t)k * 3 + 2] = nz;
      dr.diag[k] = 2.0 * kPi * (r.eps_out + r.eps_in) / (r.eps_out - r.eps_in);
      cond[k] = (uint32_t)nc;
    } else {
      throw Error(GP_ERR_INVALID_INPUT, "panel " + std::to_string(k) + " has an unknown role");
    }
  }
  PanelGeom ga = g;
  ga.cond = cond.data();
  ga.ncond = nc + 1;
  gp_cap_opts o = opts;
  o.method = GP_METHOD_FMM;
  o.rhs_first = 0;
  o.rhs_count = (uint32_t)nc;
  if (rep) {
    std::memset(rep, 0, sizeof(*rep));
    rep->n_panels = (uint32_t)n;
    rep->n_conductors = (uint32_t)nc;
  }
  std::vector<double> Cf((size_t)(nc + 1) * (nc + 1));
  solve_fmm(ctx, ga, o, Cf.data(), rep, w.data(), &dr);
  for (int i = 0; i < nc; ++i)
    for (int j = 0; j < nc; ++j) C[(size_t)i * nc + j] = Cf[(size_t)i * (nc + 1) + j];
  for (int e = 0; e < nc * nc; ++e)
    if (!std::isfinite(C[e])) throw Error(GP_ERR_NON_FINITE, "solver returned a non-finite capacitance");
  if (rep) {
    rep->n_conductors = (uint32_t)nc;
    rep->seconds_total = now_s() - t0;
  }
}

void cap_solve(gp_ctx *ctx, const PanelGeom &g, const gp_cap_opts &o, double *C,
               gp_cap_report *rep, const double *eps_panel) {
  double t0 = now_s();
  upload_seconds() = 0.0;
  if (!(o.tol > 0.0) || !std::isfinite(o.tol)) throw Error(GP_ERR_INVALID_INPUT, "tol must be > 0");
  if (o.precision != GP_PREC_FP64 && o.precision != GP_PREC_MIXED)
    throw Error(GP_ERR_INVALID_INPUT, "unknown precision");
  uint32_t method = o.method;
  uint32_t limit = o.dense_limit ? o.dense_limit : 500u;
  if (method == GP_METHOD_AUTO) method = ((uint32_t)g.n < limit) ? GP_METHOD_DENSE : GP_METHOD_FMM;
  if (rep) {
    std::memset(rep, 0, sizeof(*rep));
    rep->n_panels = (uint32_t)g.n;
    rep->n_conductors = (uint32_t)g.ncond;
  }
  if (o.ground != GP_GROUND_NONE && o.ground != GP_GROUND_IMAGE)
    throw Error(GP_ERR_INVALID_INPUT, "unknown ground model");
  const bool image = o.ground == GP_GROUND_IMAGE;
  if (image) check_ground(g, o.ground_z);
  if (method == GP_METHOD_DENSE) {
    solve_dense(ctx, g, C, rep, eps_panel, image ? &o.ground_z : nullptr);
  } else if (method == GP_METHOD_FMM) {
    if (image) {
      // the panels and their mirrors in one octree, antisymmetric RHS
      Mirrored m;
      mirror_geom(g, o.ground_z, m);
      std::vector<double> eps2;
      if (eps_panel) {
        eps2.assign(eps_panel, eps_panel + g.n);
        eps2.insert(eps2.end(), eps_panel, eps_panel + g.n);
      }
      solve_fmm(ctx, m.g, o, C, rep, eps_panel ? eps2.data() : nullptr);
    } else {
      solve_fmm(ctx, g, o, C, rep, eps_panel);
    }
  } else {
    throw Error(GP_ERR_INVALID_INPUT, "unknown method");
  }
  for (size_t e = 0; e < (size_t)g.ncond * g.ncond; ++e)
    if (!std::isfinite(C[e])) throw Error(GP_ERR_NON_FINITE, "solver returned a non-finite capacitance");
  if (rep) {
    rep->seconds_total = now_s() - t0;
    rep->seconds_upload = upload_seconds();
  }
}

void cap_jacobian(gp_ctx *ctx, const PanelGeom &g, const gp_cap_opts &o, const ParamMap &pm,
                  double *C, double *J, gp_cap_report *rep) {
  const double t0 = now_s();
  hipStream_t s = ctx->stream;
  if (!(o.tol > 0.0) || !std::isfinite(o.tol)) throw Error(GP_ERR_INVALID_INPUT, "tol must be > 0");
  if (o.precision != GP_PREC_FP64 && o.precision != GP_PREC_MIXED)
    throw Error(GP_ERR_INVALID_INPUT, "unknown precision");
  const int N = g.n, nc = g.ncond;
  if (nc > 128) throw Error(GP_ERR_INVALID_INPUT, "sensitivities support at most 128 conductors");
  // Parameter entries: panels in range, no panel twice within a parameter.
  const int np = pm.n;
  const uint32_t ne = np ? pm.off[np] : 0;
  int nm_max = 0;
  {
    std::vector<int> seen(N, -1);
    for (int p = 0; p < np; ++p) {
      if (pm.off[p + 1] < pm.off[p]) throw Error(GP_ERR_INVALID_INPUT, "param_off must not decrease");
      nm_max = std::max(nm_max, (int)(pm.off[p + 1] - pm.off[p]));
      for (uint32_t e = pm.off[p]; e < pm.off[p + 1]; ++e) {
        const uint32_t k = pm.panel[e];
        if (k >= (uint32_t)N) throw Error(GP_ERR_INVALID_INPUT, "parameter panel index out of range");
        if (seen[k] == p) throw Error(GP_ERR_INVALID_INPUT, "panel listed twice in one parameter");
        seen[k] = p;
        for (int c = 0; c < 12; ++c)
          if (!std::isfinite(pm.dv[12 * (size_t)e + c]))
            throw Error(GP_ERR_INVALID_INPUT, "non-finite vertex velocity");
      }
    }
  }
  const uint32_t limit = o.dense_limit ? o.dense_limit : 500u;
  uint32_t method = o.method;
  if (method == GP_METHOD_AUTO) method = ((uint32_t)N < limit) ? GP_METHOD_DENSE : GP_METHOD_FMM;
  if (method != GP_METHOD_DENSE && method != GP_METHOD_FMM)
    throw Error(GP_ERR_INVALID_INPUT, "unknown method");
  if (rep) {
    std::memset(rep, 0, sizeof(*rep));
    rep->n_panels = (uint32_t)N;
    rep->n_conductors = (uint32_t)nc;
    rep->method_used = method;
    rep->precision_used = method == GP_METHOD_FMM ? o.precision : GP_PREC_FP64;
  }
  std::vector<uint32_t> hmoved(ne);
  double L0 = 1.0;
  std::vector<double> hC((size_t)nc * nc);
  // Device data the sensitivity kernels read, set by either path.
  const double *dverts = nullptr, *dcen = nullptr, *dsig = nullptr, *dlam = nullptr;
  const uint32_t *dcond = nullptr;
  int lds = 0;
  // Dense path storage.
  DBuf<double> dv_, dc_, dB, dA;
  DBuf<uint32_t> dcond_;
  // FMM path storage.
  FmmSystem sys;
  DBuf<double> xf, xa;
  if (method == GP_METHOD_DENSE) {
    double c[3];
    root_cube(g, c, L0);
    std::vector<double> hv, hc, ha;
    std::vector<uint32_t> hcond;
    scaled_arrays(g, nullptr, c, L0, hv, hc, ha, hcond);
    std::vector<double> B((size_t)N * nc, 0.0), A((size_t)N * nc, 0.0);
    for (int i = 0; i < N; ++i) {
      B[(size_t)i * nc + g.cond[i]] = 1.0;
      A[(size_t)i * nc + g.cond[i]] = ha[i];
    }
    dv_ = to_device(hv, s);
    dc_ = to_device(hc, s);
    dcond_ = to_device(hcond, s);
    dB = to_device(B, s);
    dA = to_device(A, s);
    auto d_w = to_device(ha, s);
    std::vector<int> coff, cidx;
    cond_csr(g, nullptr, coff, cidx);
    auto d_coff = to_device(coff, s);
    auto d_cidx = to_device(cidx, s);
    DBuf<double> d_P((size_t)N * N), d_PT((size_t)N * N), d_work((size_t)N * nc), d_C((size_t)nc * nc);
    DBuf<int> d_piv(N), d_info(2);
    launch_dense_assemble(N, dv_.p, dc_.p, d_P.p, s);
    launch_transpose(N, d_P.p, d_PT.p, s);
    launch_lu_solve(N, nc, d_P.p, d_piv.p, dB.p, d_work.p, d_info.p, s);
    launch_lu_solve(N, nc, d_PT.p, d_piv.p, dA.p, d_work.p, d_info.p + 1, s);
    launch_cap_accumulate(nc, nc, nc, d_coff.p, d_cidx.p, d_w.p, dB.p, d_C.p, s);
    int info[2] = {0, 0};
    GP_HIP(hipMemcpyAsync(info, d_info.p, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    GP_HIP(hipMemcpyAsync(hC.data(), d_C.p, hC.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    GP_HIP(hipStreamSynchronize(s));
    if (info[0] || info[1]) throw Error(GP_ERR_SINGULAR, "potential-coefficient matrix is singular");
    for (uint32_t e = 0; e < ne; ++e) hmoved[e] = pm.panel[e];
    dverts = dv_.p;
    dcen = dc_.p;
    dcond = dcond_.p;
    dsig = dB.p;
    dlam = dA.p;
    lds = nc;
    if (rep) rep->converged = 1;
  } else {
    const bool mixed = o.precision == GP_PREC_MIXED;
    sys.want_transpose = true;
    sys.setup(g, o, mixed, s);
    L0 = sys.L0();
    const int ld = nc == 1 ? 1 : round_up(nc, 16);
    sys.ensure_work(ld);
    const size_t vec = (size_t)N * ld;
    DBuf<double> b(vec), dCb((size_t)nc * ld);
    xf.alloc(vec);
    xa.alloc(vec);
    xf.zero(s);
    xa.zero(s);
    BlockOps ops;
    ops.N = N;
    ops.ld = ld;
    ops.nr = nc;
    ops.s = s;
    const std::vector<int> live(nc, 1);
    SolveStats sf, sa;
    kr_unit_rhs(sys.cond.p, 0, N, ld, nc, b.p, s);
    solve_systems(sys, ops, o, mixed, b.p, xf.p, live, sf);
    // adjoint: P'^T lambda_i = scaled areas of conductor i
    sys.transposed = true;
    kr_weight_rhs(sys.cond.p, sys.area.p, 0, N, ld, nc, b.p, s);
    solve_systems(sys, ops, o, mixed, b.p, xa.p, live, sa);
    sys.transposed = false;
    launch_cap_accumulate(nc, nc, ld, sys.coff.p, sys.cidx.p, sys.area.p, xf.p, dCb.p, s);
    GP_HIP(hipMemcpyAsync(hC.data(), dCb.p, hC.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    GP_HIP(hipStreamSynchronize(s));
    std::vector<int> pos(N);
    for (int k = 0; k < N; ++k) pos[sys.t().perm[k]] = k;
    for (uint32_t e = 0; e < ne; ++e) hmoved[e] = (uint32_t)pos[pm.panel[e]];
    dverts = sys.verts.p;
    dcen = sys.cen.p;
    dcond = sys.cond.p;
    dsig = xf.p;
    dlam = xa.p;
    lds = ld;
    if (rep) {
      bool conv = true;
      double worst = 0.0;
      int its = 0;
      for (int q = 0; q < nc; ++q) {
        conv &= sf.conv[q] && sa.conv[q];
        worst = std::max(worst, std::max(sf.resid[q], sa.resid[q]));
        its = std::max(its, std::max(sf.iters[q], sa.iters[q]));
      }
      rep->converged = conv;
      rep->final_rel_residual = worst;
      rep->iterations = (uint32_t)its;
      rep->refine_steps = (uint32_t)std::max(sf.refine_steps, sa.refine_steps);
      rep->matvecs = (uint32_t)sys.matvecs;
      rep->max_level = (uint32_t)sys.L;
      rep->n_leaves = (uint32_t)sys.n_leaves;
      rep->n_boxes = (uint32_t)sys.f64.n_boxes;
      rep->fmm_order = (uint32_t)sys.n;
      rep->near_nnz = (uint64_t)sys.nnz;
      rep->m2l_pairs = (uint64_t)sys.pairs;
    }
  }
  const double t1 = now_s();
  for (size_t e = 0; e < hC.size(); ++e) C[e] = hC[e] * 4.0 * kPi * kEps0 * L0;
  if (np > 0) {
    std::vector<int> hm(hmoved.begin(), hmoved.end());
    std::vector<double> hdv(pm.dv, pm.dv + 12 * (size_t)ne);
    auto d_moved = to_device(hm, s);
    auto d_dv = to_device(hdv, s);
    DBuf<int> mv_of(N);
    const int nmx = std::max(nm_max, 1);
    DBuf<double> da(N), R((size_t)2 * N * nc), dcn((size_t)6 * nmx), dar((size_t)2 * nmx),
        scratch(sensitivity_scratch(N, nmx, nc, 2)), dJ((size_t)np * nc * nc), gq((size_t)20 * N),
        dgq((size_t)32 * nmx);
    launch_gauss_table(N, dverts, gq.p, s);
    // The dense path differentiates every entry in closed form; the FMM path
    // takes the Gauss rule beyond 6 source diagonals (relative entry error
    // below 4e-6) and the centroid rule beyond 30 (below 8e-5), both under
    // the FMM's own far-field error.
    const double far2 = method == GP_METHOD_FMM ? 36.0 : 0.0;
    const double far1 = method == GP_METHOD_FMM ? 900.0 : 0.0;
    GP_HIP(hipMemsetAsync(mv_of.p, 0xff, N * sizeof(int), s));
    da.zero(s);
    R.zero(s);
    // Parameters that move the same panels (consecutive, e.g. the width,
    // spacing and height of one wire) share their integral evaluations.
    for (int p = 0; p < np;) {
      const uint32_t o0 = pm.off[p], nm = pm.off[p + 1] - o0;
      int nd = 1;
      while (nd < kSensGroup && p + nd < np && pm.off[p + nd + 1] - pm.off[p + nd] == nm &&
             std::equal(pm.panel + o0, pm.panel + o0 + nm, pm.panel + pm.off[p + nd]))
        ++nd;
      SensArgs a;
      a.N = N;
      a.nm = (int)nm;
      a.nc = nc;
      a.nd = nd;
      a.lds = lds;
      a.ldl = lds;
      a.cond0 = 0;
      a.moved = d_moved.p + o0;
      a.dv = d_dv.p + 12 * (size_t)o0;
      a.dstride = 12L * nm;
      a.verts = dverts;
      a.cen = dcen;
      a.sigma = dsig;
      a.lambda = dlam;
      a.cond = dcond;
      a.mv_of = mv_of.p;
      a.da = da.p;
      a.dcen = dcn.p;
      a.darea = dar.p;
      a.R = R.p;
      a.scratch = scratch.p;
      a.gq = gq.p;
      a.dgq = dgq.p;
      a.far2 = far2;
      a.far1 = far1;
      a.J = dJ.p + (size_t)p * nc * nc;
      launch_sensitivity(a, s);
      p += nd;
    }
    std::vector<double> hJ((size_t)np * nc * nc);
    GP_HIP(hipMemcpyAsync(hJ.data(), dJ.p, hJ.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    GP_HIP(hipStreamSynchronize(s));
    // dC/dtheta = 4 pi eps0 dC'/dtheta' (theta' = theta / L0)
    for (size_t e = 0; e < hJ.size(); ++e) J[e] = hJ[e] * 4.0 * kPi * kEps0;
  }
  const double t2 = now_s();
  for (size_t e = 0; e < hC.size(); ++e)
    if (!std::isfinite(C[e])) throw Error(GP_ERR_NON_FINITE, "solver returned a non-finite capacitance");
  for (size_t e = 0; e < (size_t)np * nc * nc; ++e)
    if (!std::isfinite(J[e])) throw Error(GP_ERR_NON_FINITE, "non-finite sensitivity");
  if (rep) {
    rep->seconds_total = t2 - t0;
    rep->seconds_solve = t1 - t0;
    rep->seconds_setup = 0.0;
  }
  if (rep && !rep->converged && o.strict)
    throw Error(GP_ERR_NOT_CONVERGED, "GMRES did not reach tol in the forward or adjoint solve");
}

void cap_solve_batched(gp_ctx *ctx, const std::vector<PanelGeom> &gs, const gp_cap_opts &o,
                       const std::vector<double *> &C, gp_cap_report *reps) {
  const double t0 = now_s();
  upload_seconds() = 0.0;
  if (!(o.tol > 0.0) || !std::isfinite(o.tol)) throw Error(GP_ERR_INVALID_INPUT, "tol must be > 0");
  if (o.precision != GP_PREC_FP64 && o.precision != GP_PREC_MIXED)
    throw Error(GP_ERR_INVALID_INPUT, "unknown precision");
  if (o.method != GP_METHOD_AUTO && o.method != GP_METHOD_DENSE && o.method != GP_METHOD_FMM)
    throw Error(GP_ERR_INVALID_INPUT, "unknown method");
  if (o.ground != GP_GROUND_NONE && o.ground != GP_GROUND_IMAGE)
    throw Error(GP_ERR_INVALID_INPUT, "unknown ground model");
  const bool image = o.ground == GP_GROUND_IMAGE;
  const uint32_t limit = o.dense_limit ? o.dense_limit : 500u;
  std::vector<const PanelGeom *> fmm;
  std::vector<double *> Cf;
  std::vector<int> fidx, dprobs;
  const int P = (int)gs.size();
  if (image)
    for (const PanelGeom &g : gs) check_ground(g, o.ground_z);
  // with a ground plane the FMM problems are their signed mirrored sets
  std::vector<Mirrored> mir;
  mir.reserve(P);
  if (reps) std::memset(reps, 0, sizeof(gp_cap_report) * P);
  for (int p = 0; p < P; ++p) {
    const PanelGeom &g = gs[p];
    if (reps) {
      reps[p].n_panels = (uint32_t)g.n;
      reps[p].n_conductors = (uint32_t)g.ncond;
    }
    const bool dense = o.method == GP_METHOD_DENSE ||
                       (o.method == GP_METHOD_AUTO && (uint32_t)g.n < limit);
    if (dense) {
      dprobs.push_back(p);
    } else if (image) {
      mir.emplace_back();
      mirror_geom(g, o.ground_z, mir.back());
      fmm.push_back(&mir.back().g);
      Cf.push_back(C[p]);
      fidx.push_back(p);
    } else {
      fmm.push_back(&g);
      Cf.push_back(C[p]);
      fidx.push_back(p);
    }
  }
  auto dense = [&] {
    if (!dprobs.empty()) solve_dense_batch(ctx, gs, dprobs, C, reps, image ? &o.ground_z : nullptr);
  };
  if (fmm.empty()) {
    dense();
  } else {
    // (the dense problems solve while the first group of FMM problems is set up)
    std::vector<gp_cap_report> fr(fmm.size());
    for (size_t q = 0; q < fmm.size(); ++q) fr[q] = reps ? reps[fidx[q]] : gp_cap_report{};
    solve_fmm_batch(ctx, fmm, o, Cf, fr.data(), dense);
    if (reps)
      for (size_t q = 0; q < fmm.size(); ++q) reps[fidx[q]] = fr[q];
  }
  const double t1 = now_s();
  for (int p = 0; p < P; ++p) {
    for (size_t e = 0; e < (size_t)gs[p].ncond * gs[p].ncond; ++e)
      if (!std::isfinite(C[p][e]))
        throw Error(GP_ERR_NON_FINITE, "solver returned a non-finite capacitance for problem " +
                                           std::to_string(p));
    if (reps) {
      reps[p].seconds_total = t1 - t0;
      reps[p].seconds_upload = upload_seconds();
    }
  }
}

void fmm_apply_host(gp_ctx *ctx, const PanelGeom &g, const gp_cap_opts &o, int nr,
                    const double *X, double *Y, int precision) {
  hipStream_t s = ctx->stream;
  FmmSystem sys;
  sys.want_by_offset = nr == 1;
  sys.setup(g, o, precision == 1, s);
  int ld = nr == 1 ? 1 : round_up(nr, 16);
  sys.ensure_work(ld);
  const int N = g.n;
  std::vector<double> hx((size_t)N * ld, 0.0), hy((size_t)N * ld);
  for (int p = 0; p < N; ++p)
    for (int r = 0; r < nr; ++r) hx[(size_t)p * ld + r] = X[(size_t)sys.t().perm[p] * nr + r];
  auto dxv = to_device(hx, s);
  DBuf<double> dy((size_t)N * ld);
  sys.apply(dxv.p, false, nr, dy.p, precision);
  GP_HIP(hipMemcpyAsync(hy.data(), dy.p, hy.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  GP_HIP(hipStreamSynchronize(s));
  // back to SI: P = k_e L0 P'
  const double sc = sys.L0() / (4.0 * kPi * kEps0);
  for (int p = 0; p < N; ++p)
    for (int r = 0; r < nr; ++r) Y[(size_t)sys.t().perm[p] * nr + r] = hy[(size_t)p * ld + r] * sc;
}

}  // namespace gp
