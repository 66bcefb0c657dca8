// Low-rank form of the far-field M2L operator of the Chebyshev bbFMM (host,
// plain C++).  The 316 far-field interaction matrices G_o of one order share
// their leading singular vectors: with S = sum_o G_o G_o^T (the kernel 1/r is
// symmetric and the offset set is closed under o -> -o, so G_-o = G_o^T and
// the row and column spaces coincide), U = the kr leading eigenvectors of S
// and G_o ~ U (U^T G_o U) U^T.  The M2L then runs on kr-vectors: multipoles
// projected by U^T, locals expanded by U.
#pragma once
#include <vector>

namespace gp {

// Cyclic Jacobi eigen-decomposition of a symmetric matrix A (n x n,
// row-major, overwritten): eigenvalues w in descending order and the matching
// unit eigenvectors as the columns of V (n x n, row-major).
void sym_eig(int n, std::vector<double> &A, std::vector<double> &w, std::vector<double> &V);

// G: 343 offset blocks of ld x ld (row m, column l, the first n3 rows and
// columns used; near-field offsets all zero).  Outputs U (n3 x kr, row-major)
// and Gc (343 x kr x kr, row-major: Gc_o = U^T G_o U).  Uses up to nthreads
// threads.
void m2l_low_rank(int n3, int ld, const double *G, int kr, int nthreads, std::vector<double> &U,
                  std::vector<double> &Gc);

}  // namespace gp
