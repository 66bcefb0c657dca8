// The reference's truncated M2L operators (host, plain C++): each far-field
// interaction matrix G_o of the Chebyshev bbFMM (bbfmm.rs:132-174) is
// compressed by LowRankApprox::compress (linalg.rs:277-381), a column-pivoted
// QR by modified Gram-Schmidt that stops at the first pivot whose remaining
// squared column norm is below eps times the largest squared column norm of
// G_o, with eps = max(1e-3 * 0.1^n, 1e-14) for Chebyshev order n
// (bbfmm.rs:150).  The reference applies y += U (V^T x) per pair; the device
// applies the same truncated operator as one dense matrix U V^T per offset,
// so the FP64 M2L kernels serve both the truncated and the full operator.
#pragma once
#include <memory>
#include <vector>

namespace gp {

// Truncation threshold of order n (bbfmm.rs:150).
double m2l_cpqr_eps(int n);

// Truncated column-pivoted QR of the m x n row-major matrix a (row stride
// lda), pivot rule, threshold and arithmetic order of linalg.rs:277-381;
// overwrites a with the rank-k product U V^T and returns k.  With U and Vt
// given, also writes the factors: U[i * ldu + s] (m x k) and Vt[s * ldv + j]
// (k x n), the u and vt of LowRankApprox (linalg.rs:277-381).
int cpqr_truncate(int m, int n, double *a, int lda, double eps, double *U = nullptr,
                  int ldu = 0, double *Vt = nullptr, int ldv = 0);

// Dense M2L interaction matrix of offset (o0, o1, o2) (box units) on the
// reference box of order n: G[m * ld + l] = 1 / |x_m - x_l - 2 o| over the
// tensor Chebyshev nodes (bbfmm.rs:156-166), zero for r = 0.
void m2l_matrix(int n, int o0, int o1, int o2, const double *nodes, double *G, int ld);

// The 343 offset blocks of [-3, 3]^3 (index (o0+3)*49 + (o1+3)*7 + o2+3),
// each n3p x n3p zero padded, near-field offsets (max |o_i| <= 1) all zero;
// with compress set every far-field block holds its CPQR truncation.  Built
// once per (order, compress) and process.  ranks (optional) receives the 343
// truncation ranks (0 for the near-field offsets).
std::shared_ptr<const std::vector<double>> m2l_tables(int n, int n3p, bool compress,
                                                      std::vector<int> *ranks = nullptr);

// The factors of the truncated far-field blocks of order n, for applying
// them as the reference does, y += U (V^T x) (bbfmm.rs:132-174): kp = the
// largest rank rounded up to a multiple of 16; per offset index o (343, the
// near-field ones zero) rank[o], V^T as Vt[(o * kp + s) * n3p + l] and U
// with its rank index split in four, Up[((o * n3p + m) * 4 + q) * (kp / 4) +
// s'] = U[m][4 s' + q] (so that a lane holding rank entries q, q + 4, ...
// reads one contiguous run), and plain row-major as U; zero beyond the rank
// and n^3.
struct M2lFactors {
  int kp = 0, kmax = 0;  // padded and largest rank
  std::vector<int> rank;
  std::vector<double> Vt, Up;
  std::vector<double> U;  // U[(o * n3p + m) * kp + s] = U[m][s] (row-major)
};
std::shared_ptr<const M2lFactors> m2l_factors(int n, int n3p);

}  // namespace gp
