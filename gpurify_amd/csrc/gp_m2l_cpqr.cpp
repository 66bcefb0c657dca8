// Truncated M2L operators of the reference (host): CPQR compression of each
// far-field interaction matrix, applied on the device as one dense matrix.
//
// Working layout: the matrix being factored is held column-major (column j
// contiguous), so the pivot swaps, norms, dot products and updates of the
// modified Gram-Schmidt sweep each stream one or two contiguous columns.
// Every sum runs over the row index in ascending order, and the pivot search
// keeps the first maximum, as linalg.rs:277-381 does: with the same IEEE
// arithmetic (no fused multiply-add on either side) the truncation rank and
// pivots equal the reference's.
#include "gp_m2l_cpqr.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <map>
#include <mutex>
#include <numeric>
#include <thread>

namespace gp {

double m2l_cpqr_eps(int n) { return std::max(1e-3 * std::pow(0.1, n), 1e-14); }

int cpqr_truncate(int m, int n, double *a, int lda, double eps, double *U, int ldu, double *Vt,
                  int ldv) {
  std::vector<double> q((size_t)m * n);      // column j at q[j * m]
  std::vector<double> r((size_t)n * n, 0.0);  // R row-major, rows = pivot steps
  std::vector<int> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  std::vector<double> cn(n);  // remaining squared column norms
  for (int j = 0; j < n; ++j) {
    double *qj = &q[(size_t)j * m];
    double s = 0.0;
    for (int i = 0; i < m; ++i) {
      const double v = a[(size_t)i * lda + j];
      qj[i] = v;
      s += v * v;
    }
    cn[j] = s;
  }
  double mx = 0.0;
  for (double c : cn) mx = std::max(mx, c);
  const double thr = eps * mx;
  int k = 0;
  const int kmax = std::min(m, n);
  for (int step = 0; step < kmax; ++step) {
    int best = step;
    for (int j = step + 1; j < n; ++j)
      if (cn[j] > cn[best]) best = j;
    if (cn[best] < thr) break;
    double *qs = &q[(size_t)step * m];
    if (best != step) {
      std::swap(perm[step], perm[best]);
      std::swap(cn[step], cn[best]);
      std::swap_ranges(qs, qs + m, &q[(size_t)best * m]);
      for (int row = 0; row < step; ++row)
        std::swap(r[(size_t)row * n + step], r[(size_t)row * n + best]);
    }
    double nrm = 0.0;
    for (int i = 0; i < m; ++i) nrm += qs[i] * qs[i];
    nrm = std::sqrt(nrm);
    if (nrm < 1e-15) break;
    r[(size_t)step * n + step] = nrm;
    const double inv = 1.0 / nrm;
    for (int i = 0; i < m; ++i) qs[i] *= inv;
    for (int j = step + 1; j < n; ++j) {
      double *qj = &q[(size_t)j * m];
      double d = 0.0;
      for (int i = 0; i < m; ++i) d += qs[i] * qj[i];
      r[(size_t)step * n + j] = d;
      for (int i = 0; i < m; ++i) qj[i] -= d * qs[i];
      cn[j] -= d * d;
      if (cn[j] < 0.0) cn[j] = 0.0;
    }
    k = step + 1;
  }
  // a = U V^T with U = Q[:, :k] and V^T[s][perm[j]] = R[s][j].
  if (U && Vt)
    for (int s = 0; s < k; ++s) {
      for (int i = 0; i < m; ++i) U[(size_t)i * ldu + s] = q[(size_t)s * m + i];
      for (int j = 0; j < n; ++j) Vt[(size_t)s * ldv + perm[j]] = r[(size_t)s * n + j];
    }
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < n; ++j) {
      double v = 0.0;
      for (int s = 0; s < k; ++s) v += q[(size_t)s * m + i] * r[(size_t)s * n + j];
      a[(size_t)i * lda + perm[j]] = v;
    }
  return k;
}

void m2l_matrix(int n, int o0, int o1, int o2, const double *nodes, double *G, int ld) {
  const int n3 = n * n * n;
  for (int m = 0; m < n3; ++m) {
    const int ma = m / (n * n), mb = (m / n) % n, mc = m % n;
    for (int l = 0; l < n3; ++l) {
      const int la = l / (n * n), lb = (l / n) % n, lc = l % n;
      const double dx = nodes[ma] - nodes[la] - 2.0 * o0;
      const double dy = nodes[mb] - nodes[lb] - 2.0 * o1;
      const double dz = nodes[mc] - nodes[lc] - 2.0 * o2;
      const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
      G[(size_t)m * ld + l] = r > 0.0 ? 1.0 / r : 0.0;
    }
  }
}

namespace {

struct Tables {
  std::shared_ptr<const std::vector<double>> G;
  std::vector<int> ranks;
};

}  // namespace

std::shared_ptr<const std::vector<double>> m2l_tables(int n, int n3p, bool compress,
                                                      std::vector<int> *ranks) {
  static std::mutex mu;
  static std::map<std::pair<int, int>, Tables> cache;
  std::lock_guard<std::mutex> lock(mu);
  Tables &e = cache[{n * 1024 + n3p, compress ? 1 : 0}];
  if (!e.G) {
    const int n3 = n * n * n;
    std::vector<double> nodes(n);
    // Chebyshev roots cos((2m+1) pi / 2n) (chebyshev.rs:11-17)
    for (int m = 0; m < n; ++m) nodes[m] = std::cos((2.0 * m + 1.0) * M_PI / (2.0 * n));
    auto G = std::make_shared<std::vector<double>>((size_t)343 * n3p * n3p, 0.0);
    e.ranks.assign(343, 0);
    std::vector<int> far;
    for (int oi = 0; oi < 343; ++oi) {
      const int o0 = oi / 49 - 3, o1 = (oi / 7) % 7 - 3, o2 = oi % 7 - 3;
      if (std::max(std::abs(o0), std::max(std::abs(o1), std::abs(o2))) > 1) far.push_back(oi);
    }
    const double eps = m2l_cpqr_eps(n);
    std::atomic<int> next(0);
    auto work = [&] {
      for (int t; (t = next.fetch_add(1)) < (int)far.size();) {
        const int oi = far[t];
        double *blk = G->data() + (size_t)oi * n3p * n3p;
        m2l_matrix(n, oi / 49 - 3, (oi / 7) % 7 - 3, oi % 7 - 3, nodes.data(), blk, n3p);
        e.ranks[oi] = compress ? cpqr_truncate(n3, n3, blk, n3p, eps) : n3;
      }
    };
    const int nt = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) th.emplace_back(work);
    work();
    for (auto &t : th) t.join();
    e.G = G;
  }
  if (ranks) *ranks = e.ranks;
  return e.G;
}

std::shared_ptr<const M2lFactors> m2l_factors(int n, int n3p) {
  static std::mutex mu;
  static std::map<int, std::shared_ptr<const M2lFactors>> cache;
  std::lock_guard<std::mutex> lock(mu);
  auto &slot = cache[n * 1024 + n3p];
  if (slot) return slot;
  const int n3 = n * n * n;
  std::vector<double> nodes(n);
  for (int m = 0; m < n; ++m) nodes[m] = std::cos((2.0 * m + 1.0) * M_PI / (2.0 * n));
  const double eps = m2l_cpqr_eps(n);
  // factors at full width first (rank <= n3), then packed to kp
  std::vector<int> rank(343, 0);
  std::vector<double> U((size_t)343 * n3 * n3, 0.0), V((size_t)343 * n3 * n3, 0.0);
  std::vector<int> far;
  for (int oi = 0; oi < 343; ++oi) {
    const int o0 = oi / 49 - 3, o1 = (oi / 7) % 7 - 3, o2 = oi % 7 - 3;
    if (std::max(std::abs(o0), std::max(std::abs(o1), std::abs(o2))) > 1) far.push_back(oi);
  }
  std::atomic<int> next(0);
  auto work = [&] {
    std::vector<double> blk((size_t)n3 * n3);
    for (int t; (t = next.fetch_add(1)) < (int)far.size();) {
      const int oi = far[t];
      m2l_matrix(n, oi / 49 - 3, (oi / 7) % 7 - 3, oi % 7 - 3, nodes.data(), blk.data(), n3);
      rank[oi] = cpqr_truncate(n3, n3, blk.data(), n3, eps, &U[(size_t)oi * n3 * n3], n3,
                               &V[(size_t)oi * n3 * n3], n3);
    }
  };
  const int nt = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  for (int t = 1; t < nt; ++t) th.emplace_back(work);
  work();
  for (auto &t : th) t.join();
  auto f = std::make_shared<M2lFactors>();
  const int kmax = *std::max_element(rank.begin(), rank.end());
  f->kp = std::max(16, (kmax + 15) / 16 * 16);
  const int kp = f->kp, k4 = kp / 4;
  f->kmax = kmax;
  f->rank = rank;
  f->Vt.assign((size_t)343 * kp * n3p, 0.0);
  f->Up.assign((size_t)343 * n3p * kp, 0.0);
  f->U.assign((size_t)343 * n3p * kp, 0.0);
  for (int oi : far)
    for (int s = 0; s < rank[oi]; ++s) {
      for (int l = 0; l < n3; ++l)
        f->Vt[((size_t)oi * kp + s) * n3p + l] = V[(size_t)oi * n3 * n3 + (size_t)s * n3 + l];
      for (int m = 0; m < n3; ++m) {
        const double u = U[(size_t)oi * n3 * n3 + (size_t)m * n3 + s];
        f->Up[(((size_t)oi * n3p + m) * 4 + s % 4) * k4 + s / 4] = u;
        f->U[((size_t)oi * n3p + m) * kp + s] = u;
      }
    }
  slot = f;
  return slot;
}

}  // namespace gp
