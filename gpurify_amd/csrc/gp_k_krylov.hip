// GMRES building blocks on the device (krylov.rs:50-207) for a block of
// independent right-hand sides that iterate in lockstep.  Arithmetic and
// every reduction are FP64; the Krylov basis is stored in FP64 or, for the
// inner solves of the mixed-precision solver, in FP32 (template TV).
//
// Vectors are N x ld row-major (RHS fastest); nr <= ld columns are live.
// Reductions are two-stage with a fixed order (per-block partials, then one
// thread per column summing the partials in block order), so every run gives
// the same bits.  Per-column GMRES state (Hessenberg column, Givens
// rotations, residual estimate, iteration count) lives in small device
// arrays updated by one thread per column.

#include <type_traits>

#include "gp_common.h"
#include "gp_kernels.h"

namespace gp {

constexpr int KR_BLOCKS = 512;
constexpr int KR_VB = 8;  // basis vectors per block: Y is read once per 8 dots

// partial[(v * nblk + b) * ld + r] = sum over rows of block b of X_v[row][r] * Y[row][r]
// for the KR_VB vectors v0 .. v0 + 7 of this block (grid.y = ceil(nvec / 8)).
template <typename TX>
__global__ __launch_bounds__(256) void k_dot_partial(const TX *__restrict__ X,
                                                     size_t strideX, int nvec,
                                                     const double *__restrict__ Y, int N, int ld,
                                                     int nr, double *__restrict__ partial) {
  __shared__ double red[KR_VB][256];
  const int v0 = blockIdx.y * KR_VB, b = blockIdx.x, nblk = gridDim.x;
  const int nv = min(KR_VB, nvec - v0);
  const int rpp = 256 / nr;  // rows per pass
  const int tid = threadIdx.x;
  const int r = tid % nr, ro = tid / nr;
  const int rows_per_block = (N + nblk - 1) / nblk;
  const int row0 = b * rows_per_block, row1 = min(N, row0 + rows_per_block);
  double acc[KR_VB];
#pragma unroll
  for (int v = 0; v < KR_VB; ++v) acc[v] = 0.0;
  if (ro < rpp)
    for (int row = row0 + ro; row < row1; row += rpp) {
      const size_t idx = (size_t)row * ld + r;
      const double y = Y[idx];
#pragma unroll
      for (int v = 0; v < KR_VB; ++v)
        if (v < nv) acc[v] += (double)X[(size_t)(v0 + v) * strideX + idx] * y;
    }
#pragma unroll
  for (int v = 0; v < KR_VB; ++v) red[v][tid] = acc[v];
  __syncthreads();
  if (tid < nr)
    for (int v = 0; v < nv; ++v) {
      double s = 0.0;
      for (int j = 0; j < rpp; ++j) s += red[v][j * nr + tid];
      partial[((size_t)(v0 + v) * nblk + b) * ld + tid] = s;
    }
}

// The same dots for an FP32 basis with ld % 4 == 0: thread (ro, cq) takes the
// four columns 4 cq .. 4 cq + 3 (16-byte loads of the basis, two of y) of
// rows row0 + ro, + rpp, ... with rpp = 256 / (ld / 4), two rows an
// iteration; the column groups' sums go through LDS one column at a time.
__global__ __launch_bounds__(256) void k_dot_partial4(const float *__restrict__ X,
                                                      size_t strideX, int nvec,
                                                      const double *__restrict__ Y, int N, int ld,
                                                      double *__restrict__ partial) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  typedef double d2 __attribute__((ext_vector_type(2)));
  __shared__ double red[KR_VB][256];
  const int v0 = blockIdx.y * KR_VB, b = blockIdx.x, nblk = gridDim.x;
  const int nv = min(KR_VB, nvec - v0);
  const int qpr = ld >> 2, rpp = 256 / qpr;
  const int tid = threadIdx.x, cq = tid % qpr, ro = tid / qpr;
  const int rows_per_block = (N + nblk - 1) / nblk;
  const int row0 = b * rows_per_block, row1 = min(N, row0 + rows_per_block);
  double acc[KR_VB][4];
#pragma unroll
  for (int v = 0; v < KR_VB; ++v)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[v][c] = 0.0;
  if (ro < rpp) {
    auto step = [&](int row) {
      const size_t idx = (size_t)row * ld + 4 * cq;
      const d2 ya = *reinterpret_cast<const d2 *>(Y + idx), yb = *reinterpret_cast<const d2 *>(Y + idx + 2);
#pragma unroll
      for (int v = 0; v < KR_VB; ++v)
        if (v < nv) {
          const f4 x = *reinterpret_cast<const f4 *>(X + (size_t)(v0 + v) * strideX + idx);
          acc[v][0] += (double)x[0] * ya[0];
          acc[v][1] += (double)x[1] * ya[1];
          acc[v][2] += (double)x[2] * yb[0];
          acc[v][3] += (double)x[3] * yb[1];
        }
    };
    int row = row0 + ro;
    for (; row + rpp < row1; row += 2 * rpp) {
      step(row);
      step(row + rpp);
    }
    if (row < row1) step(row);
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    __syncthreads();
#pragma unroll
    for (int v = 0; v < KR_VB; ++v) red[v][tid] = ro < rpp ? acc[v][c] : 0.0;
    __syncthreads();
    if (tid < qpr)
      for (int v = 0; v < nv; ++v) {
        double s = 0.0;
        for (int j = 0; j < rpp; ++j) s += red[v][j * qpr + tid];
        partial[((size_t)(v0 + v) * nblk + b) * ld + 4 * tid + c] = s;
      }
  }
}

// out[v * ld + r] = sum_b partial[(v * nblk + b) * ld + r]  (optionally sqrt):
// one wavefront per output, lane l summing blocks l, l + 64, ... in order,
// then a fixed shuffle tree, so every run gives the same bits.
__global__ __launch_bounds__(256) void k_reduce_partial(const double *__restrict__ partial,
                                                        int nblk, int nvec, int ld, int nr,
                                                        double *__restrict__ out, int take_sqrt) {
  const int e = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (e >= nvec * nr) return;
  const int v = e / nr, r = e % nr;
  double s = 0.0;
  for (int b = lane; b < nblk; b += 64) s += partial[((size_t)v * nblk + b) * ld + r];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
  if (lane == 0) out[(size_t)v * ld + r] = take_sqrt ? sqrt(s) : s;
}

// w[row][r] += alpha * sum_{v < nvec} coef[v][r] * V_v[row][r]
template <typename TV>
__global__ __launch_bounds__(256) void k_axpy_multi(double *__restrict__ w,
                                                    const TV *__restrict__ V, size_t strideV,
                                                    const double *__restrict__ coef, int nvec,
                                                    int N, int ld, int nr, double alpha) {
  size_t total = (size_t)N * nr;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    int row = (int)(e / nr), r = (int)(e % nr);
    size_t idx = (size_t)row * ld + r;
    double s = 0.0;
    for (int v = 0; v < nvec; ++v)
      s += coef[(size_t)v * ld + r] * (double)V[(size_t)v * strideV + idx];
    w[idx] += alpha * s;
  }
}

// The same update for an FP32 basis with ld % 4 == 0 and ld <= 128: one
// thread per (row, group of 4 columns), 16-byte basis loads issued four
// vectors at a time; the coefficients go through LDS 16 vectors at a time
// (one 16-byte basis load per vector and thread instead of three loads);
// columns r >= nr keep their values.
__global__ __launch_bounds__(256) void k_axpy_multi4(double *__restrict__ w,
                                                     const float *__restrict__ V, size_t strideV,
                                                     const double *__restrict__ coef, int nvec,
                                                     int N, int ld, int nr, double alpha) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  typedef double d2 __attribute__((ext_vector_type(2)));
  constexpr int VC = 16;
  __shared__ __attribute__((aligned(16))) double sc[VC * 128];
  const int qpr = ld >> 2;
  const size_t total = (size_t)N * qpr;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool ok = e < total;
  const int row = ok ? (int)(e / qpr) : 0, r = ok ? 4 * (int)(e % qpr) : 0;
  const size_t idx = (size_t)row * ld + r;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  for (int v0 = 0; v0 < nvec; v0 += VC) {
    const int vn = min(VC, nvec - v0);
    __syncthreads();
    for (int i = threadIdx.x; i < vn * ld; i += 256) sc[i] = coef[(size_t)v0 * ld + i];
    __syncthreads();
    if (ok) {
#pragma unroll 4
      for (int u = 0; u < vn; ++u) {
        const f4 x = *reinterpret_cast<const f4 *>(V + (size_t)(v0 + u) * strideV + idx);
        const d2 ca = *reinterpret_cast<const d2 *>(sc + u * ld + r);
        const d2 cb = *reinterpret_cast<const d2 *>(sc + u * ld + r + 2);
        s0 += ca[0] * (double)x[0];
        s1 += ca[1] * (double)x[1];
        s2 += cb[0] * (double)x[2];
        s3 += cb[1] * (double)x[3];
      }
    }
  }
  if (ok) {
    d2 wa = *reinterpret_cast<d2 *>(w + idx), wb = *reinterpret_cast<d2 *>(w + idx + 2);
    if (r < nr) wa[0] += alpha * s0;
    if (r + 1 < nr) wa[1] += alpha * s1;
    if (r + 2 < nr) wb[0] += alpha * s2;
    if (r + 3 < nr) wb[1] += alpha * s3;
    *reinterpret_cast<d2 *>(w + idx) = wa;
    *reinterpret_cast<d2 *>(w + idx + 2) = wb;
  }
}

// out[row][r] = scale[r] * in[row][r] (0 where mask[r] == 0); mask may be null.
template <typename TO>
__global__ __launch_bounds__(256) void k_scale_cols(const double *__restrict__ in,
                                                    TO *__restrict__ out,
                                                    const double *__restrict__ scale,
                                                    const int *__restrict__ mask, int N, int ld,
                                                    int nr) {
  size_t total = (size_t)N * nr;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    int row = (int)(e / nr), r = (int)(e % nr);
    size_t idx = (size_t)row * ld + r;
    out[idx] = (TO)((mask && !mask[r]) ? 0.0 : scale[r] * in[idx]);
  }
}

// out = a - b (elementwise over live columns)
__global__ __launch_bounds__(256) void k_sub(const double *__restrict__ a,
                                             const double *__restrict__ b,
                                             double *__restrict__ out, int N, int ld, int nr) {
  size_t total = (size_t)N * nr;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    int row = (int)(e / nr), r = (int)(e % nr);
    size_t idx = (size_t)row * ld + r;
    out[idx] = a[idx] - b[idx];
  }
}

// z[row][r] = (T)(v[row][r] * (dinv ? dinv[row] : 1)): right Jacobi
// preconditioner (cap/fmm_solver.rs:133-144) fused with the cast to the
// operator's working precision.
template <typename TI, typename T>
__global__ __launch_bounds__(256) void k_precond_cast(const TI *__restrict__ v,
                                                      const double *__restrict__ dinv,
                                                      T *__restrict__ z, int N, int ld, int nr) {
  size_t total = (size_t)N * nr;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    int row = (int)(e / nr), r = (int)(e % nr);
    size_t idx = (size_t)row * ld + r;
    double s = dinv ? dinv[row] : 1.0;
    z[idx] = (T)((double)v[idx] * s);
  }
}

// x += alpha * y over live columns.
__global__ __launch_bounds__(256) void k_axpy(double *__restrict__ x, const double *__restrict__ y,
                                              double alpha, int N, int ld, int nr) {
  size_t total = (size_t)N * nr;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    int row = (int)(e / nr), r = (int)(e % nr);
    size_t idx = (size_t)row * ld + r;
    x[idx] += alpha * y[idx];
  }
}

// Unit-conductor right-hand sides (rhs_unit_conductor, cap/solver.rs:55-63):
// b[p][r] = 1 where panel p belongs to conductor j0 + r, else 0 (all ld columns).
__global__ __launch_bounds__(256) void k_unit_rhs(const uint32_t *__restrict__ cond, int j0, int N,
                                                  int ld, int nr, const double *__restrict__ sign,
                                                  double *__restrict__ b) {
  size_t total = (size_t)N * ld;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    int row = (int)(e / ld), r = (int)(e % ld);
    b[e] = (r < nr && (int)cond[row] == j0 + r) ? (sign ? sign[row] : 1.0) : 0.0;
  }
}

// Adjoint right-hand sides: b[p][r] = w[p] where panel p belongs to
// conductor j0 + r (its scaled area), else 0 (all ld columns).
__global__ __launch_bounds__(256) void k_weight_rhs(const uint32_t *__restrict__ cond,
                                                    const double *__restrict__ w, int j0, int N,
                                                    int ld, int nr, double *__restrict__ b) {
  size_t total = (size_t)N * ld;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    int row = (int)(e / ld), r = (int)(e % ld);
    b[e] = (r < nr && (int)cond[row] == j0 + r) ? w[row] : 0.0;
  }
}

// y[row][r] += w[row] * t[row][r] over live columns, then t = 0 there.
__global__ __launch_bounds__(256) void k_rows_axpy_clear(double *__restrict__ y,
                                                         double *__restrict__ t,
                                                         const double *__restrict__ w, int N, int ld,
                                                         int nr) {
  size_t total = (size_t)N * nr;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    int row = (int)(e / nr), r = (int)(e % nr);
    size_t idx = (size_t)row * ld + r;
    y[idx] += w[row] * t[idx];
    t[idx] = 0.0;
  }
}

// Givens rotation eliminating b against a (krylov.rs:195-207).
__device__ inline void givens(double a, double b, double &c, double &s) {
  if (b == 0.0) {
    c = 1.0;
    s = 0.0;
  } else if (fabs(b) > fabs(a)) {
    double t = a / b;
    s = 1.0 / sqrt(1.0 + t * t);
    c = t * s;
  } else {
    double t = b / a;
    c = 1.0 / sqrt(1.0 + t * t);
    s = t * c;
  }
}

// Cycle start: beta = ||r||, g[0] = beta, active = beta/bnorm > tol.
// Zeroes g[1..m]; inv_beta scales the first basis vector.
__global__ void k_gmres_start(int m, int nr, int ld, const double *__restrict__ beta,
                              const double *__restrict__ bnorm, double tol,
                              const int *__restrict__ live, int *__restrict__ active,
                              int *__restrict__ kused, double *__restrict__ g,
                              double *__restrict__ inv_beta, double *__restrict__ est) {
  int r = threadIdx.x + blockIdx.x * blockDim.x;
  if (r >= nr) return;
  double rel = beta[r] / bnorm[r];
  est[r] = rel;
  int a = live[r] && (rel > tol);
  active[r] = a;
  kused[r] = 0;
  g[r] = beta[r];
  for (int i = 1; i <= m; ++i) g[(size_t)i * ld + r] = 0.0;
  inv_beta[r] = (a && beta[r] > 0.0) ? 1.0 / beta[r] : 0.0;
}

// One Arnoldi step's Hessenberg update for every active column: store the
// new column h[0..k+1] (from Gram-Schmidt), apply the previous rotations,
// form and apply the new one, update g and the residual estimate
// (krylov.rs:119-141).  Columns reaching tol stop (kused frozen).
// Also sets inv_h[r] = 1/h[k+1] (or 1 when h[k+1] <= 1e-14, krylov.rs:112-116).
__global__ void k_gmres_givens(int k, int m, int nr, int ld, const double *__restrict__ h1,
                               const double *__restrict__ h2, const double *__restrict__ hn,
                               double *__restrict__ H, double *__restrict__ cs,
                               double *__restrict__ sn, double *__restrict__ g,
                               const double *__restrict__ bnorm, double tol,
                               int *__restrict__ active, int *__restrict__ kused,
                               double *__restrict__ est, double *__restrict__ inv_h) {
  int r = threadIdx.x + blockIdx.x * blockDim.x;
  if (r >= nr) return;
  if (!active[r]) {
    inv_h[r] = 0.0;
    return;
  }
  double hk1 = hn[r];
  inv_h[r] = hk1 > 1e-14 ? 1.0 / hk1 : 1.0;
  // H[i][k] stored at ((i * m) + k) * ld + r; Gram-Schmidt coefficients of
  // both passes summed.
  for (int i = 0; i <= k; ++i)
    H[((size_t)i * m + k) * ld + r] = h1[(size_t)i * ld + r] + (h2 ? h2[(size_t)i * ld + r] : 0.0);
  for (int i = 0; i < k; ++i) {
    double hi = H[((size_t)i * m + k) * ld + r], hi1 = H[((size_t)(i + 1) * m + k) * ld + r];
    double c = cs[(size_t)i * ld + r], s = sn[(size_t)i * ld + r];
    H[((size_t)i * m + k) * ld + r] = c * hi + s * hi1;
    H[((size_t)(i + 1) * m + k) * ld + r] = -s * hi + c * hi1;
  }
  double hkk = H[((size_t)k * m + k) * ld + r];
  double c, s;
  givens(hkk, hk1, c, s);
  cs[(size_t)k * ld + r] = c;
  sn[(size_t)k * ld + r] = s;
  H[((size_t)k * m + k) * ld + r] = c * hkk + s * hk1;
  double gk = g[(size_t)k * ld + r], gk1 = g[(size_t)(k + 1) * ld + r];
  g[(size_t)k * ld + r] = c * gk + s * gk1;
  g[(size_t)(k + 1) * ld + r] = -s * gk + c * gk1;
  kused[r] = k + 1;
  double e = fabs(g[(size_t)(k + 1) * ld + r]) / bnorm[r];
  est[r] = e;
  if (e <= tol) active[r] = 0;
}

// y = H^-1 g over the kused[r] leading rows (krylov.rs:144-152).
__global__ void k_gmres_backsolve(int m, int nr, int ld, const double *__restrict__ H,
                                  const double *__restrict__ g, const int *__restrict__ kused,
                                  double *__restrict__ Y) {
  int r = threadIdx.x + blockIdx.x * blockDim.x;
  if (r >= nr) return;
  int ku = kused[r];
  for (int i = ku - 1; i >= 0; --i) {
    double s = g[(size_t)i * ld + r];
    for (int j = i + 1; j < ku; ++j) s -= H[((size_t)i * m + j) * ld + r] * Y[(size_t)j * ld + r];
    Y[(size_t)i * ld + r] = s / H[((size_t)i * m + i) * ld + r];
  }
  for (int i = ku; i < m; ++i) Y[(size_t)i * ld + r] = 0.0;
}

// x[row][r] += dinv[row] * sum_{i < kused[r]} Y[i][r] V_i[row][r]  (krylov.rs:154-170)
template <typename TV>
__global__ __launch_bounds__(256) void k_gmres_update_x(double *__restrict__ x,
                                                        const TV *__restrict__ V,
                                                        size_t strideV,
                                                        const double *__restrict__ Y,
                                                        const int *__restrict__ kused, int kmax,
                                                        const double *__restrict__ dinv, int N,
                                                        int ld, int nr) {
  size_t total = (size_t)N * nr;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    int row = (int)(e / nr), r = (int)(e % nr);
    size_t idx = (size_t)row * ld + r;
    int ku = min(kused[r], kmax);
    double s = 0.0;
    for (int i = 0; i < ku; ++i) s += Y[(size_t)i * ld + r] * (double)V[(size_t)i * strideV + idx];
    x[idx] += (dinv ? dinv[row] : 1.0) * s;
  }
}

// ---------------------------------------------------------------- segmented
// A batch of independent systems stacked in one vector of N rows (one
// column): rows [seg_lo[p], seg_lo[p + 1]) belong to system p = seg[row].
// Per-system state (coefficients, scales, Hessenberg columns) has stride
// nsys.  Dot products reduce per system over a fixed block map: block b
// covers rows [blk_lo[b], blk_hi[b]) of system blk_sys[b], and the blocks of
// system p are [sys_blk[p], sys_blk[p + 1]) (fixed order, same bits every run).

// partial[v * nblk + b] = sum over the rows of block b of X_v[row] * Y[row];
// with active given, 0 for the blocks of inactive systems (none of their
// rows read)
template <typename TX>
__global__ __launch_bounds__(256) void k_seg_dot_partial(const TX *__restrict__ X, size_t strideX,
                                                         int nvec, const double *__restrict__ Y,
                                                         const int *__restrict__ blk_lo,
                                                         const int *__restrict__ blk_hi,
                                                         const int *__restrict__ seg,
                                                         const int *__restrict__ active,
                                                         double *__restrict__ partial) {
  __shared__ double red[KR_VB][256 / 64];
  const int b = blockIdx.x, nblk = gridDim.x, v0 = blockIdx.y * KR_VB;
  const int nv = min(KR_VB, nvec - v0);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (active && (blk_lo[b] >= blk_hi[b] || !active[seg[blk_lo[b]]])) {  // work-group-uniform
    if (tid < nv) partial[(size_t)(v0 + tid) * nblk + b] = 0.0;
    return;
  }
  double acc[KR_VB];
#pragma unroll
  for (int v = 0; v < KR_VB; ++v) acc[v] = 0.0;
  for (int row = blk_lo[b] + tid; row < blk_hi[b]; row += 256) {
    const double y = Y[row];
#pragma unroll
    for (int v = 0; v < KR_VB; ++v)
      if (v < nv) acc[v] += (double)X[(size_t)(v0 + v) * strideX + row] * y;
  }
#pragma unroll
  for (int v = 0; v < KR_VB; ++v) {
    double a = acc[v];
    for (int off = 32; off > 0; off >>= 1) a += __shfl_down(a, off, 64);
    if (lane == 0) red[v][wave] = a;
  }
  __syncthreads();
  if (tid < nv) {
    double t = 0.0;
    for (int w = 0; w < 4; ++w) t += red[tid][w];
    partial[(size_t)(v0 + tid) * nblk + b] = t;
  }
}

// out[v * nsys + p] = sum of the partials of system p's blocks (optionally sqrt)
__global__ void k_seg_reduce(const double *__restrict__ partial, int nblk, int nvec,
                             const int *__restrict__ sys_blk, int nsys, double *__restrict__ out,
                             int take_sqrt) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nvec * nsys) return;
  int v = e / nsys, p = e % nsys;
  double t = 0.0;
  for (int b = sys_blk[p]; b < sys_blk[p + 1]; ++b) t += partial[(size_t)v * nblk + b];
  out[e] = take_sqrt ? sqrt(t) : t;
}

// w[row] += alpha * sum_{v < nvec} coef[v][seg[row]] * V_v[row] (rows of
// inactive systems untouched when active is given)
template <typename TV>
__global__ __launch_bounds__(256) void k_seg_axpy_multi(double *__restrict__ w,
                                                        const TV *__restrict__ V, size_t strideV,
                                                        const double *__restrict__ coef, int nvec,
                                                        int N, const int *__restrict__ seg, int nsys,
                                                        double alpha,
                                                        const int *__restrict__ active) {
  for (int row = blockIdx.x * blockDim.x + threadIdx.x; row < N; row += gridDim.x * blockDim.x) {
    const int p = seg[row];
    if (active && !active[p]) continue;
    double t = 0.0;
    for (int v = 0; v < nvec; ++v) t += coef[(size_t)v * nsys + p] * (double)V[(size_t)v * strideV + row];
    w[row] += alpha * t;
  }
}

// out[row] = scale[seg[row]] * in[row] (0 where mask[seg[row]] == 0)
template <typename TO>
__global__ __launch_bounds__(256) void k_seg_scale(const double *__restrict__ in, TO *__restrict__ out,
                                                   const double *__restrict__ scale,
                                                   const int *__restrict__ mask, int N,
                                                   const int *__restrict__ seg) {
  for (int row = blockIdx.x * blockDim.x + threadIdx.x; row < N; row += gridDim.x * blockDim.x) {
    const int p = seg[row];
    out[row] = (TO)((mask && !mask[p]) ? 0.0 : scale[p] * in[row]);
  }
}

// x[row] += dinv[row] * sum_{i < kused[p]} Y[i][p] V_i[row], p = seg[row]
template <typename TV>
__global__ __launch_bounds__(256) void k_seg_update_x(double *__restrict__ x,
                                                      const TV *__restrict__ V, size_t strideV,
                                                      const double *__restrict__ Y,
                                                      const int *__restrict__ kused, int kmax,
                                                      const double *__restrict__ dinv, int N,
                                                      const int *__restrict__ seg, int nsys) {
  for (int row = blockIdx.x * blockDim.x + threadIdx.x; row < N; row += gridDim.x * blockDim.x) {
    const int p = seg[row];
    const int ku = min(kused[p], kmax);
    double t = 0.0;
    for (int i = 0; i < ku; ++i) t += Y[(size_t)i * nsys + p] * (double)V[(size_t)i * strideV + row];
    x[row] += (dinv ? dinv[row] : 1.0) * t;
  }
}

// b[row] = 1 (sign[row] when given) where the panel's (batch-global)
// conductor is system p's right-hand-side conductor, else 0.
__global__ __launch_bounds__(256) void k_seg_unit_rhs(const uint32_t *__restrict__ cond, int N,
                                                      const int *__restrict__ seg,
                                                      const int *__restrict__ rhs_cond,
                                                      const double *__restrict__ sign,
                                                      double *__restrict__ b) {
  for (int row = blockIdx.x * blockDim.x + threadIdx.x; row < N; row += gridDim.x * blockDim.x)
    b[row] = ((int)cond[row] == rhs_cond[seg[row]]) ? (sign ? sign[row] : 1.0) : 0.0;
}

namespace {
int grid_for(size_t total) { return (int)std::min<size_t>((total + 255) / 256, 8192); }
}  // namespace

template <typename TX>
void seg_dots(const SegMap &sm, const TX *X, size_t strideX, int nvec, const double *Y,
              double *partial, double *out, bool take_sqrt, hipStream_t s, const int *active) {
  if (nvec <= 0) return;
  ProfScope ps("gmres_dots", s, 2.0 * nvec * sm.N, (sizeof(TX) * nvec + 8.0) * sm.N);
  hipLaunchKernelGGL((k_seg_dot_partial<TX>), dim3(sm.nblk, (nvec + KR_VB - 1) / KR_VB), dim3(256), 0,
                     s, X, strideX, nvec, Y, sm.blk_lo, sm.blk_hi, sm.seg, active, partial);
  GP_LAUNCH_CHECK();
  const int tot = nvec * sm.nsys;
  hipLaunchKernelGGL(k_seg_reduce, dim3((tot + 255) / 256), dim3(256), 0, s, partial, sm.nblk, nvec,
                     sm.sys_blk, sm.nsys, out, take_sqrt ? 1 : 0);
  GP_LAUNCH_CHECK();
}
size_t seg_partial_size(const SegMap &sm, int nvec_max) {
  return (size_t)sm.nblk * ((nvec_max + KR_VB - 1) / KR_VB) * KR_VB;
}
template <typename TV>
void seg_axpy_multi(const SegMap &sm, double *w, const TV *V, size_t strideV, const double *coef,
                    int nvec, double alpha, hipStream_t s, const int *active) {
  if (nvec <= 0) return;
  ProfScope ps("gmres_axpy", s, 2.0 * nvec * sm.N, (sizeof(TV) * nvec + 16.0) * sm.N);
  hipLaunchKernelGGL((k_seg_axpy_multi<TV>), dim3(grid_for(sm.N)), dim3(256), 0, s, w, V, strideV,
                     coef, nvec, sm.N, sm.seg, sm.nsys, alpha, active);
  GP_LAUNCH_CHECK();
}
template <typename TO>
void seg_scale(const SegMap &sm, const double *in, TO *out, const double *scale, const int *mask,
               hipStream_t s) {
  hipLaunchKernelGGL((k_seg_scale<TO>), dim3(grid_for(sm.N)), dim3(256), 0, s, in, out, scale, mask,
                     sm.N, sm.seg);
  GP_LAUNCH_CHECK();
}
template <typename TV>
void seg_update_x(const SegMap &sm, double *x, const TV *V, size_t strideV, const double *Y,
                  const int *kused, int kmax, const double *dinv, hipStream_t s) {
  hipLaunchKernelGGL((k_seg_update_x<TV>), dim3(grid_for(sm.N)), dim3(256), 0, s, x, V, strideV, Y,
                     kused, kmax, dinv, sm.N, sm.seg, sm.nsys);
  GP_LAUNCH_CHECK();
}
void seg_unit_rhs(const SegMap &sm, const uint32_t *cond, const int *rhs_cond, double *b,
                  hipStream_t s, const double *sign) {
  hipLaunchKernelGGL(k_seg_unit_rhs, dim3(grid_for(sm.N)), dim3(256), 0, s, cond, sm.N, sm.seg,
                     rhs_cond, sign, b);
  GP_LAUNCH_CHECK();
}
#define GP_SEG_BASIS(TV)                                                                        \
  template void seg_dots<TV>(const SegMap &, const TV *, size_t, int, const double *, double *,  \
                             double *, bool, hipStream_t, const int *);                         \
  template void seg_axpy_multi<TV>(const SegMap &, double *, const TV *, size_t, const double *, \
                                   int, double, hipStream_t, const int *);                      \
  template void seg_scale<TV>(const SegMap &, const double *, TV *, const double *, const int *, \
                              hipStream_t);                                                     \
  template void seg_update_x<TV>(const SegMap &, double *, const TV *, size_t, const double *,   \
                                 const int *, int, const double *, hipStream_t);
GP_SEG_BASIS(double)
GP_SEG_BASIS(float)
#undef GP_SEG_BASIS

template <typename TX>
void kr_dots(const TX *X, size_t strideX, int nvec, const double *Y, int N, int ld, int nr,
             double *partial, double *out, bool take_sqrt, hipStream_t s) {
  if (nvec <= 0) return;
  ProfScope ps("gmres_dots", s, 2.0 * nvec * N * nr, (sizeof(TX) * nvec + 8.0) * N * nr);
  int nblk = std::min(KR_BLOCKS, std::max(1, (N + 255) / 256));
  if (std::is_same<TX, float>::value && ld % 4 == 0 && ld >= 16 && ld <= 1024 && strideX % 4 == 0)
    hipLaunchKernelGGL(k_dot_partial4, dim3(nblk, (nvec + KR_VB - 1) / KR_VB), dim3(256), 0, s,
                       reinterpret_cast<const float *>(X), strideX, nvec, Y, N, ld, partial);
  else
    hipLaunchKernelGGL((k_dot_partial<TX>), dim3(nblk, (nvec + KR_VB - 1) / KR_VB), dim3(256), 0, s,
                       X, strideX, nvec, Y, N, ld, nr, partial);
  GP_LAUNCH_CHECK();
  int tot = nvec * nr;
  hipLaunchKernelGGL(k_reduce_partial, dim3((tot + 3) / 4), dim3(256), 0, s, partial, nblk, nvec,
                     ld, nr, out, take_sqrt ? 1 : 0);
  GP_LAUNCH_CHECK();
}

size_t kr_partial_size(int nvec_max, int ld) { return (size_t)KR_BLOCKS * nvec_max * ld; }

template <typename TV>
void kr_axpy_multi(double *w, const TV *V, size_t strideV, const double *coef, int nvec,
                   int N, int ld, int nr, double alpha, hipStream_t s) {
  if (nvec <= 0) return;
  ProfScope ps("gmres_axpy", s, 2.0 * nvec * N * nr, (sizeof(TV) * nvec + 16.0) * N * nr);
  if (std::is_same<TV, float>::value && ld % 4 == 0 && ld >= 16 && ld <= 128 && strideV % 4 == 0)
    hipLaunchKernelGGL(k_axpy_multi4, dim3((unsigned)(((size_t)N * (ld / 4) + 255) / 256)), dim3(256),
                       0, s, w, reinterpret_cast<const float *>(V), strideV, coef, nvec, N, ld, nr,
                       alpha);
  else
    hipLaunchKernelGGL((k_axpy_multi<TV>), dim3(grid_for((size_t)N * nr)), dim3(256), 0, s, w, V,
                       strideV, coef, nvec, N, ld, nr, alpha);
  GP_LAUNCH_CHECK();
}

template <typename TO>
void kr_scale_cols(const double *in, TO *out, const double *scale, const int *mask, int N,
                   int ld, int nr, hipStream_t s) {
  hipLaunchKernelGGL((k_scale_cols<TO>), dim3(grid_for((size_t)N * nr)), dim3(256), 0, s, in, out,
                     scale, mask, N, ld, nr);
  GP_LAUNCH_CHECK();
}

void kr_sub(const double *a, const double *b, double *out, int N, int ld, int nr, hipStream_t s) {
  hipLaunchKernelGGL(k_sub, dim3(grid_for((size_t)N * nr)), dim3(256), 0, s, a, b, out, N, ld, nr);
  GP_LAUNCH_CHECK();
}

template <typename TI, typename T>
void kr_precond_cast(const TI *v, const double *dinv, T *z, int N, int ld, int nr, hipStream_t s) {
  hipLaunchKernelGGL((k_precond_cast<TI, T>), dim3(grid_for((size_t)N * nr)), dim3(256), 0, s, v,
                     dinv, z, N, ld, nr);
  GP_LAUNCH_CHECK();
}

void kr_gmres_start(int m, int nr, int ld, const double *beta, const double *bnorm, double tol,
                    const int *live, int *active, int *kused, double *g, double *inv_beta,
                    double *est, hipStream_t s) {
  hipLaunchKernelGGL(k_gmres_start, dim3((nr + 255) / 256), dim3(256), 0, s, m, nr, ld, beta, bnorm,
                     tol, live, active, kused, g, inv_beta, est);
  GP_LAUNCH_CHECK();
}

void kr_gmres_givens(int k, int m, int nr, int ld, const double *h1, const double *h2,
                     const double *hn, double *H, double *cs, double *sn, double *g,
                     const double *bnorm, double tol, int *active, int *kused, double *est,
                     double *inv_h, hipStream_t s) {
  hipLaunchKernelGGL(k_gmres_givens, dim3((nr + 255) / 256), dim3(256), 0, s, k, m, nr, ld, h1, h2,
                     hn, H, cs, sn, g, bnorm, tol, active, kused, est, inv_h);
  GP_LAUNCH_CHECK();
}

void kr_gmres_backsolve(int m, int nr, int ld, const double *H, const double *g, const int *kused,
                        double *Y, hipStream_t s) {
  hipLaunchKernelGGL(k_gmres_backsolve, dim3((nr + 255) / 256), dim3(256), 0, s, m, nr, ld, H, g,
                     kused, Y);
  GP_LAUNCH_CHECK();
}

template <typename TV>
void kr_gmres_update_x(double *x, const TV *V, size_t strideV, const double *Y,
                       const int *kused, int kmax, const double *dinv, int N, int ld, int nr,
                       hipStream_t s) {
  hipLaunchKernelGGL((k_gmres_update_x<TV>), dim3(grid_for((size_t)N * nr)), dim3(256), 0, s, x, V,
                     strideV, Y, kused, kmax, dinv, N, ld, nr);
  GP_LAUNCH_CHECK();
}

void kr_axpy(double *x, const double *y, double alpha, int N, int ld, int nr, hipStream_t s) {
  hipLaunchKernelGGL(k_axpy, dim3(grid_for((size_t)N * nr)), dim3(256), 0, s, x, y, alpha, N, ld,
                     nr);
  GP_LAUNCH_CHECK();
}

void kr_weight_rhs(const uint32_t *cond, const double *w, int j0, int N, int ld, int nr, double *b,
                   hipStream_t s) {
  hipLaunchKernelGGL(k_weight_rhs, dim3(grid_for((size_t)N * ld)), dim3(256), 0, s, cond, w, j0, N,
                     ld, nr, b);
  GP_LAUNCH_CHECK();
}

void kr_rows_axpy_clear(double *y, double *t, const double *w, int N, int ld, int nr,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_rows_axpy_clear, dim3(grid_for((size_t)N * nr)), dim3(256), 0, s, y, t, w,
                     N, ld, nr);
  GP_LAUNCH_CHECK();
}

// Skip flags of a batch from its systems' active flags: box b belongs to
// system box_prob[b], leaf T is box leaf_box[T].  1 = the system has
// converged (its kernels' work is skipped).
__global__ __launch_bounds__(256) void k_skip_flags(int nb, int nl, const int *__restrict__ box_prob,
                                                    const int *__restrict__ leaf_box,
                                                    const int *__restrict__ active,
                                                    unsigned char *__restrict__ skip_box,
                                                    unsigned char *__restrict__ skip_leaf) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nb + nl; i += gridDim.x * blockDim.x) {
    if (i < nb)
      skip_box[i] = active[box_prob[i]] == 0;
    else
      skip_leaf[i - nb] = active[box_prob[leaf_box[i - nb]]] == 0;
  }
}

void kr_skip_flags(int nb, int nl, const int *box_prob, const int *leaf_box, const int *active,
                   unsigned char *skip_box, unsigned char *skip_leaf, hipStream_t s) {
  hipLaunchKernelGGL(k_skip_flags, dim3(grid_for((size_t)nb + nl)), dim3(256), 0, s, nb, nl,
                     box_prob, leaf_box, active, skip_box, skip_leaf);
  GP_LAUNCH_CHECK();
}

void kr_unit_rhs(const uint32_t *cond, int j0, int N, int ld, int nr, double *b, hipStream_t s,
                 const double *sign) {
  hipLaunchKernelGGL(k_unit_rhs, dim3(grid_for((size_t)N * ld)), dim3(256), 0, s, cond, j0, N, ld,
                     nr, sign, b);
  GP_LAUNCH_CHECK();
}

#define GP_KR_BASIS(TV)                                                                        \
  template void kr_dots<TV>(const TV *, size_t, int, const double *, int, int, int, double *,      \
                            double *, bool, hipStream_t);                                          \
  template void kr_axpy_multi<TV>(double *, const TV *, size_t, const double *, int, int, int, int, \
                                  double, hipStream_t);                                            \
  template void kr_scale_cols<TV>(const double *, TV *, const double *, const int *, int, int, int, \
                                  hipStream_t);                                                    \
  template void kr_gmres_update_x<TV>(double *, const TV *, size_t, const double *, const int *,   \
                                      int, const double *, int, int, int, hipStream_t);
GP_KR_BASIS(double)
GP_KR_BASIS(float)
#undef GP_KR_BASIS
template void kr_precond_cast<double, double>(const double *, const double *, double *, int, int,
                                              int, hipStream_t);
template void kr_precond_cast<double, float>(const double *, const double *, float *, int, int,
                                             int, hipStream_t);
template void kr_precond_cast<float, float>(const float *, const double *, float *, int, int, int,
                                            hipStream_t);

}  // namespace gp
