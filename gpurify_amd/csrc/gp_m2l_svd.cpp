// Low-rank M2L tables (host).  The projection is the SVD compression of
// Fong & Darve, "The black-box fast multipole method" (J. Comput. Phys. 228,
// 2009), sec. 3.2, here without the Chebyshev weights; the eigenvectors come
// from the classical cyclic Jacobi method (Golub & Van Loan, Matrix
// Computations, 4th ed., alg. 8.5.2-8.5.3).
#include "gp_m2l_svd.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <thread>

namespace gp {

namespace {

template <typename F>
void split(int n, int nthreads, F f) {
  const int nt = std::max(1, std::min(nthreads, n));
  std::vector<std::thread> th;
  for (int t = 1; t < nt; ++t) th.emplace_back([&, t] {
    for (int i = t; i < n; i += nt) f(i);
  });
  for (int i = 0; i < n; i += nt) f(i);
  for (auto &x : th) x.join();
}

}  // namespace

void sym_eig(int n, std::vector<double> &A, std::vector<double> &w, std::vector<double> &V) {
  V.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
  auto a = [&](int i, int j) -> double & { return A[(size_t)i * n + j]; };
  double total = 0.0;
  for (size_t e = 0; e < A.size(); ++e) total += A[e] * A[e];
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) off += a(p, q) * a(p, q);
    if (off <= 1e-30 * total) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = a(p, q);
        if (apq == 0.0) continue;
        // symmetric 2x2 Schur decomposition: J^T [[app, apq], [apq, aqq]] J diagonal
        const double tau = (a(q, q) - a(p, p)) / (2.0 * apq);
        const double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
        const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
        for (int k = 0; k < n; ++k) {  // A <- A J
          const double kp = a(k, p), kq = a(k, q);
          a(k, p) = c * kp - s * kq;
          a(k, q) = s * kp + c * kq;
        }
        for (int k = 0; k < n; ++k) {  // A <- J^T A
          const double pk = a(p, k), qk = a(q, k);
          a(p, k) = c * pk - s * qk;
          a(q, k) = s * pk + c * qk;
        }
        a(p, q) = a(q, p) = 0.0;
        for (int k = 0; k < n; ++k) {  // V <- V J
          double &vp = V[(size_t)k * n + p], &vq = V[(size_t)k * n + q];
          const double x = vp, y = vq;
          vp = c * x - s * y;
          vq = s * x + c * y;
        }
      }
  }
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int i, int j) { return a(i, i) > a(j, j); });
  w.resize(n);
  std::vector<double> Vs((size_t)n * n);
  for (int c = 0; c < n; ++c) {
    w[c] = a(idx[c], idx[c]);
    for (int r = 0; r < n; ++r) Vs[(size_t)r * n + c] = V[(size_t)r * n + idx[c]];
  }
  V.swap(Vs);
}

void m2l_low_rank(int n3, int ld, const double *G, int kr, int nthreads, std::vector<double> &U,
                  std::vector<double> &Gc) {
  auto nonzero = [&](int o) {
    const double *g = G + (size_t)o * ld * ld;
    for (int m = 0; m < n3; ++m)
      for (int l = 0; l < n3; ++l)
        if (g[(size_t)m * ld + l] != 0.0) return true;
    return false;
  };
  std::vector<int> far;
  for (int o = 0; o < 343; ++o)
    if (nonzero(o)) far.push_back(o);
  // S = sum_o G_o G_o^T: the upper triangle one row per task, then mirrored
  std::vector<double> S((size_t)n3 * n3, 0.0);
  split(n3, nthreads, [&](int m) {
    double *row = &S[(size_t)m * n3];
    for (int o : far) {
      const double *g = G + (size_t)o * ld * ld;
      const double *gm = g + (size_t)m * ld;
      for (int mp = m; mp < n3; ++mp) {
        const double *gq = g + (size_t)mp * ld;
        double acc = 0.0;
        for (int l = 0; l < n3; ++l) acc += gm[l] * gq[l];
        row[mp] += acc;
      }
    }
  });
  for (int m = 0; m < n3; ++m)
    for (int mp = 0; mp < m; ++mp) S[(size_t)m * n3 + mp] = S[(size_t)mp * n3 + m];
  std::vector<double> w, V;
  sym_eig(n3, S, w, V);
  U.assign((size_t)n3 * kr, 0.0);
  for (int m = 0; m < n3; ++m)
    for (int j = 0; j < kr; ++j) U[(size_t)m * kr + j] = V[(size_t)m * n3 + j];
  // Gc_o = U^T G_o U, one offset per task
  Gc.assign((size_t)343 * kr * kr, 0.0);
  split((int)far.size(), nthreads, [&](int fi) {
    const int o = far[fi];
    const double *g = G + (size_t)o * ld * ld;
    std::vector<double> GU((size_t)n3 * kr, 0.0);  // G_o U
    for (int m = 0; m < n3; ++m)
      for (int l = 0; l < n3; ++l) {
        const double v = g[(size_t)m * ld + l];
        const double *ul = &U[(size_t)l * kr];
        double *out = &GU[(size_t)m * kr];
        for (int j = 0; j < kr; ++j) out[j] += v * ul[j];
      }
    double *c = &Gc[(size_t)o * kr * kr];
    for (int m = 0; m < n3; ++m) {
      const double *um = &U[(size_t)m * kr];
      const double *gu = &GU[(size_t)m * kr];
      for (int i = 0; i < kr; ++i)
        for (int j = 0; j < kr; ++j) c[(size_t)i * kr + j] += um[i] * gu[j];
    }
  });
}

}  // namespace gp
