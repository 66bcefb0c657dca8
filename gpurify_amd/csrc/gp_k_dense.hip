// Dense direct solve on the device for the small-panel rule of
// bridge.rs:399-401 (cap::solve with Method::Direct): partial-pivot LU
// (linalg.rs:172-224) and the multi-RHS triangular solves (linalg.rs:227-258),
// one work-group per matrix so a batch of small nets runs as one launch.
// The matrix stays in global memory (L2-resident below ~700 panels); each
// elimination step is a pivot reduction, a row swap and a rank-1 update with
// a barrier between phases.

#include "gp_common.h"
#include "gp_kernels.h"

namespace gp {

constexpr int kLuThreads = 1024;

// A: n x n row-major (overwritten by L\U), piv: n (row permutation, as the
// reference's piv vector), B: n x nrhs row-major (overwritten by the solution
// of A X = B).  info[0] = 0, or k+1 when column k has no non-zero pivot
// (LinalgError::Singular, linalg.rs:191-193).
__device__ void lu_solve_block(int n, int nrhs, double *__restrict__ A, int *__restrict__ piv,
                               double *__restrict__ B, double *__restrict__ work,
                               int *__restrict__ info) {
  __shared__ double red_v[kLuThreads / 64];
  __shared__ int red_i[kLuThreads / 64];
  __shared__ int s_p;
  __shared__ double s_piv;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  for (int i = tid; i < n; i += blockDim.x) piv[i] = i;
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    // Pivot: largest |a_ik| for i >= k; the first (smallest i) wins ties, as
    // the strict comparison of linalg.rs:184-189.
    double bv = -1.0;
    int bi = k;
    for (int i = k + tid; i < n; i += blockDim.x) {
      double m = fabs(A[(long)i * n + k]);
      if (m > bv) {
        bv = m;
        bi = i;
      }
    }
    for (int off = 32; off > 0; off >>= 1) {
      double ov = __shfl_down(bv, off);
      int oi = __shfl_down(bi, off);
      if (ov > bv || (ov == bv && oi < bi)) {
        bv = ov;
        bi = oi;
      }
    }
    if (lane == 0) {
      red_v[wid] = bv;
      red_i[wid] = bi;
    }
    __syncthreads();
    if (tid == 0) {
      double v = red_v[0];
      int ix = red_i[0];
      for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
        if (red_v[w] > v || (red_v[w] == v && red_i[w] < ix)) {
          v = red_v[w];
          ix = red_i[w];
        }
      s_p = (v > 0.0) ? ix : -1;
      if (v > 0.0 && ix != k) {
        int t = piv[k];
        piv[k] = piv[ix];
        piv[ix] = t;
      }
    }
    __syncthreads();
    int p = s_p;
    if (p < 0) {
      if (tid == 0) info[0] = k + 1;
      return;
    }
    if (p != k)
      for (int j = tid; j < n; j += blockDim.x) {
        double t = A[(long)k * n + j];
        A[(long)k * n + j] = A[(long)p * n + j];
        A[(long)p * n + j] = t;
      }
    __syncthreads();
    if (tid == 0) s_piv = A[(long)k * n + k];
    __syncthreads();
    const double pivot = s_piv;
    for (int i = k + 1 + tid; i < n; i += blockDim.x) A[(long)i * n + k] = A[(long)i * n + k] / pivot;
    __syncthreads();
    // rank-1 update, a wavefront per row, lanes along the row (no index
    // division; every element the same expression as before)
    const int nw = blockDim.x >> 6;
    const double *Ak = A + (long)k * n;
    for (int i = k + 1 + wid; i < n; i += nw) {
      double *Ai = A + (long)i * n;
      const double l = Ai[k];
      for (int j = k + 1 + lane; j < n; j += 64) Ai[j] = Ai[j] - l * Ak[j];
    }
    __syncthreads();
  }
  // Permute: work[i][r] = B[piv[i]][r]  (linalg.rs:231-234).
  for (int i = tid; i < n; i += blockDim.x)
    for (int r = 0; r < nrhs; ++r) work[(long)i * nrhs + r] = B[(long)piv[i] * nrhs + r];
  __syncthreads();
  // Forward substitution, unit-diagonal L (column-oriented; a thread per
  // row, every element the same expression as before).
  for (int j = 0; j < n; ++j) {
    const double *wj = work + (long)j * nrhs;
    for (int i = j + 1 + tid; i < n; i += blockDim.x) {
      const double a = A[(long)i * n + j];
      double *wi = work + (long)i * nrhs;
      for (int r = 0; r < nrhs; ++r) wi[r] -= a * wj[r];
    }
    __syncthreads();
  }
  // Back substitution with U.
  for (int j = n - 1; j >= 0; --j) {
    for (int r = tid; r < nrhs; r += blockDim.x) work[(long)j * nrhs + r] /= A[(long)j * n + j];
    __syncthreads();
    const double *wj = work + (long)j * nrhs;
    for (int i = tid; i < j; i += blockDim.x) {
      const double a = A[(long)i * n + j];
      double *wi = work + (long)i * nrhs;
      for (int r = 0; r < nrhs; ++r) wi[r] -= a * wj[r];
    }
    __syncthreads();
  }
  for (long e = tid; e < (long)n * nrhs; e += blockDim.x) B[e] = work[e];
  if (tid == 0) info[0] = 0;
}

__global__ __launch_bounds__(kLuThreads) void k_lu_solve(int n, int nrhs, double *__restrict__ A,
                                                         int *__restrict__ piv,
                                                         double *__restrict__ B,
                                                         double *__restrict__ work,
                                                         int *__restrict__ info) {
  lu_solve_block(n, nrhs, A, piv, B, work, info);
}

// A batch of independent dense solves, one work-group per problem p:
// n[p] x n[p] at A + a_off[p], n[p] x nrhs[p] at B + b_off[p] (work alike),
// pivots at piv + n_off[p], status info[p].
__global__ __launch_bounds__(kLuThreads) void k_lu_solve_batched(
    const int *__restrict__ n, const int *__restrict__ nrhs, const long *__restrict__ a_off,
    const long *__restrict__ b_off, const int *__restrict__ n_off, double *__restrict__ A,
    int *__restrict__ piv, double *__restrict__ B, double *__restrict__ work,
    int *__restrict__ info) {
  const int p = blockIdx.x;
  lu_solve_block(n[p], nrhs[p], A + a_off[p], piv + n_off[p], B + b_off[p], work + b_off[p],
                 info + p);
}

// C'[i][j] = sum over panels k of conductor i of w_k * X[k][j] (accumulate_c,
// cap/solver.rs:67-71; with per-panel eps weights for the conductor-role
// dielectric solve, cap/dielectric.rs:139-144).  One block per conductor i;
// panels of conductor i listed in cidx[coff[i]..coff[i+1]) in ascending
// order, so the sum order is fixed and the result deterministic.
__global__ __launch_bounds__(256) void k_cap_accumulate(int ncond, int nrhs, int ldx,
                                                        const int *__restrict__ coff,
                                                        const int *__restrict__ cidx,
                                                        const double *__restrict__ w,
                                                        const double *__restrict__ X,
                                                        double *__restrict__ C) {
  int i = blockIdx.x;
  for (int j = threadIdx.x; j < nrhs; j += blockDim.x) {
    double acc = 0.0;
    for (int q = coff[i]; q < coff[i + 1]; ++q) {
      int k = cidx[q];
      acc += X[(long)k * ldx + j] * w[k];
    }
    C[(long)i * nrhs + j] = acc;
  }
}

void launch_dense_batch(const DenseBatch &b, hipStream_t s) {
  if (b.count == 0) return;
  launch_dense_assemble_batched(b, s);
  ProfScope ps("dense_lu_solve", s, 0.0, 0.0);
  hipLaunchKernelGGL(k_lu_solve_batched, dim3(b.count), dim3(kLuThreads), 0, s, b.n, b.nrhs,
                     b.a_off, b.b_off, b.n_off, b.A, b.piv, b.B, b.work, b.info);
  GP_LAUNCH_CHECK();
}

void launch_lu_solve(int n, int nrhs, double *A, int *piv, double *B, double *work, int *info,
                     hipStream_t s) {
  ProfScope ps("dense_lu_solve", s, 2.0 / 3.0 * n * (double)n * n + 2.0 * n * (double)n * nrhs, 0.0);
  hipLaunchKernelGGL(k_lu_solve, dim3(1), dim3(kLuThreads), 0, s, n, nrhs, A, piv, B, work, info);
  GP_LAUNCH_CHECK();
}

void launch_cap_accumulate(int ncond, int nrhs, int ldx, const int *coff, const int *cidx,
                           const double *w, const double *X, double *C, hipStream_t s) {
  hipLaunchKernelGGL(k_cap_accumulate, dim3(ncond), dim3(256), 0, s, ncond, nrhs, ldx, coff, cidx,
                     w, X, C);
  GP_LAUNCH_CHECK();
}

}  // namespace gp
