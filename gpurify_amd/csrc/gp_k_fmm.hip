// Black-box Chebyshev FMM passes on the device (fmm/bbfmm.rs:209-448), for a
// block of right-hand sides, in FP64 or FP32 (template T).
//
// Expansions are stored per box as n^3 x ld (Chebyshev node m = a n^2 + b n + c
// major, RHS fastest).  One work-group handles one box and a chunk of up to
// 16 RHS columns; each thread owns a fixed set of (node, column) outputs in
// registers, inputs are staged through LDS.  The far-field potential of
// panel charges q = area * sigma is accumulated into the FP64 output y, which
// already holds the near field.

#include "gp_common.h"
#include "gp_kernels.h"

namespace gp {

constexpr int FM_RC = 16;   // RHS columns per work-group
constexpr int FM_OPT = 16;  // max outputs per thread: n^3 * 16 / 256 for n <= 6

// P2M (bbfmm.rs:215-228): M[leaf][m][r] = sum_p S(p, m) area_p z[p][r], with
// the tensor weights S(p, (a,b,c)) = wx[a] wy[b] wz[c] of the point's
// 1-D interpolation weights w1d[p][axis][k] (chebyshev.rs:77-94).
template <typename T>
__global__ __launch_bounds__(256) void k_p2m(int n, const int *__restrict__ leaf_lo,
                                             int leaf_box0, const T *__restrict__ w1d,
                                             const double *__restrict__ area,
                                             const T *__restrict__ z, int ldz, int nr,
                                             T *__restrict__ M, int ldm) {
  const int li = blockIdx.x, r0 = blockIdx.y * FM_RC;
  const int rc = min(FM_RC, nr - r0);
  const int n3 = n * n * n, nout = n3 * rc;
  const int p0 = leaf_lo[li], p1 = leaf_lo[li + 1];
  T acc[FM_OPT];
#pragma unroll
  for (int k = 0; k < FM_OPT; ++k) acc[k] = 0;
  for (int p = p0; p < p1; ++p) {
    const T *w = w1d + (size_t)p * 3 * n;
    const T ap = (T)area[p];
#pragma unroll
    for (int k = 0; k < FM_OPT; ++k) {
      int o = threadIdx.x + k * 256;
      if (o < nout) {
        int m = o / rc, r = r0 + o % rc;
        int a = m / (n * n), b = (m / n) % n, c = m % n;
        T q = ap * z[(size_t)p * ldz + r];
        acc[k] += w[a] * w[n + b] * w[2 * n + c] * q;
      }
    }
  }
  T *out = M + (size_t)(leaf_box0 + li) * n3 * ldm;
#pragma unroll
  for (int k = 0; k < FM_OPT; ++k) {
    int o = threadIdx.x + k * 256;
    if (o < nout) {
      int m = o / rc, r = r0 + o % rc;
      out[(size_t)m * ldm + r] = acc[k];
    }
  }
}

// M2M (bbfmm.rs:230-294): for every box of one level with children,
// M[parent] = sum over children of (X (x) Y (x) Z)_octant M[child], applied as
// three 1-D contractions (z, then y, then x) through two LDS buffers.
template <typename T>
__global__ __launch_bounds__(256) void k_m2m(int n, int box0, const int *__restrict__ child_off,
                                             const int *__restrict__ child_idx,
                                             const int *__restrict__ octant,
                                             const T *__restrict__ m2m1d, int nr,
                                             T *__restrict__ M, int ldm) {
  extern __shared__ unsigned char smem_raw[];
  T *s0 = reinterpret_cast<T *>(smem_raw);
  const int n3 = n * n * n, nn = n * n;
  T *s1 = s0 + n3 * FM_RC;
  const int b = box0 + blockIdx.x, r0 = blockIdx.y * FM_RC;
  const int rc = min(FM_RC, nr - r0);
  const int nout = n3 * rc;
  T acc[FM_OPT];
#pragma unroll
  for (int k = 0; k < FM_OPT; ++k) acc[k] = 0;
  for (int q = child_off[b]; q < child_off[b + 1]; ++q) {
    const int ch = child_idx[q];
    const int o = octant[ch];
    const T *mx = m2m1d + (o * 3 + 0) * nn, *my = m2m1d + (o * 3 + 1) * nn,
            *mz = m2m1d + (o * 3 + 2) * nn;
    const T *wc = M + (size_t)ch * n3 * ldm;
    for (int e = threadIdx.x; e < nout; e += 256) s0[e] = wc[(size_t)(e / rc) * ldm + r0 + e % rc];
    __syncthreads();
    // z: s1[a][b][cp] = sum_c mz[cp][c] s0[a][b][c]
    for (int e = threadIdx.x; e < nout; e += 256) {
      int m = e / rc, r = e % rc;
      int ab = m / n, cp = m % n;
      T s = 0;
      for (int c = 0; c < n; ++c) s += mz[cp * n + c] * s0[(ab * n + c) * rc + r];
      s1[e] = s;
    }
    __syncthreads();
    // y: s0[a][bp][cp] = sum_b my[bp][b] s1[a][b][cp]
    for (int e = threadIdx.x; e < nout; e += 256) {
      int m = e / rc, r = e % rc;
      int a = m / nn, bp = (m / n) % n, cp = m % n;
      T s = 0;
      for (int bb = 0; bb < n; ++bb) s += my[bp * n + bb] * s1[((a * n + bb) * n + cp) * rc + r];
      s0[e] = s;
    }
    __syncthreads();
    // x: acc[ap][bp][cp] += sum_a mx[ap][a] s0[a][bp][cp]
#pragma unroll
    for (int k = 0; k < FM_OPT; ++k) {
      int e = threadIdx.x + k * 256;
      if (e < nout) {
        int m = e / rc, r = e % rc;
        int ap = m / nn, bc = m % nn;
        T s = 0;
        for (int a = 0; a < n; ++a) s += mx[ap * n + a] * s0[(a * nn + bc) * rc + r];
        acc[k] += s;
      }
    }
    __syncthreads();
  }
  T *out = M + (size_t)b * n3 * ldm;
#pragma unroll
  for (int k = 0; k < FM_OPT; ++k) {
    int e = threadIdx.x + k * 256;
    if (e < nout) out[(size_t)(e / rc) * ldm + r0 + e % rc] = acc[k];
  }
}

// M2L (bbfmm.rs:296-335) with the uncompressed Chebyshev interaction matrix
// G_off[m][l] = 1/|x_m - x_l - 2 off| on the reference box (bbfmm.rs:156-166),
// scaled by 1/half of the level: Lc[t] = (1/half) sum_(s,off) G_off M[s].
// targets: box ids with a non-empty interaction list.
template <typename T>
__global__ __launch_bounds__(256) void k_m2l(int n, const int *__restrict__ targets,
                                             const int *__restrict__ il_off,
                                             const int *__restrict__ il_src,
                                             const int *__restrict__ il_oidx,
                                             const int *__restrict__ box_level,
                                             const double *__restrict__ inv_half,
                                             const T *__restrict__ G, int nr,
                                             const T *__restrict__ M, T *__restrict__ Lc,
                                             int ldm) {
  extern __shared__ unsigned char smem_raw[];
  T *s0 = reinterpret_cast<T *>(smem_raw);
  const int n3 = n * n * n;
  const int t = targets[blockIdx.x], r0 = blockIdx.y * FM_RC;
  const int rc = min(FM_RC, nr - r0);
  const int nout = n3 * rc;
  T acc[FM_OPT];
#pragma unroll
  for (int k = 0; k < FM_OPT; ++k) acc[k] = 0;
  for (int q = il_off[t]; q < il_off[t + 1]; ++q) {
    const T *ws = M + (size_t)il_src[q] * n3 * ldm;
    const T *g = G + (size_t)il_oidx[q] * n3 * n3;
    for (int e = threadIdx.x; e < nout; e += 256) s0[e] = ws[(size_t)(e / rc) * ldm + r0 + e % rc];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < FM_OPT; ++k) {
      int e = threadIdx.x + k * 256;
      if (e < nout) {
        int m = e / rc, r = e % rc;
        const T *gr = g + (size_t)m * n3;
        T s = 0;
        for (int l = 0; l < n3; ++l) s += gr[l] * s0[l * rc + r];
        acc[k] += s;
      }
    }
    __syncthreads();
  }
  const T ih = (T)inv_half[box_level[t]];
  T *out = Lc + (size_t)t * n3 * ldm;
#pragma unroll
  for (int k = 0; k < FM_OPT; ++k) {
    int e = threadIdx.x + k * 256;
    if (e < nout) out[(size_t)(e / rc) * ldm + r0 + e % rc] = acc[k] * ih;
  }
}

// L2L (bbfmm.rs:337-398): for every box of one level,
// Lc[child] += (X^T (x) Y^T (x) Z^T)_octant Lc[parent].
template <typename T>
__global__ __launch_bounds__(256) void k_l2l(int n, int box0, const int *__restrict__ parent,
                                             const int *__restrict__ octant,
                                             const T *__restrict__ m2m1d, int nr,
                                             T *__restrict__ Lc, int ldm) {
  extern __shared__ unsigned char smem_raw[];
  T *s0 = reinterpret_cast<T *>(smem_raw);
  const int n3 = n * n * n, nn = n * n;
  T *s1 = s0 + n3 * FM_RC;
  const int ch = box0 + blockIdx.x, r0 = blockIdx.y * FM_RC;
  const int rc = min(FM_RC, nr - r0);
  const int nout = n3 * rc;
  const int p = parent[ch], o = octant[ch];
  const T *mx = m2m1d + (o * 3 + 0) * nn, *my = m2m1d + (o * 3 + 1) * nn,
          *mz = m2m1d + (o * 3 + 2) * nn;
  const T *lp = Lc + (size_t)p * n3 * ldm;
  for (int e = threadIdx.x; e < nout; e += 256) s0[e] = lp[(size_t)(e / rc) * ldm + r0 + e % rc];
  __syncthreads();
  // x^T: s1[ap][b][c] = sum_a mx[a][ap] s0[a][b][c]
  for (int e = threadIdx.x; e < nout; e += 256) {
    int m = e / rc, r = e % rc;
    int ap = m / nn, bc = m % nn;
    T s = 0;
    for (int a = 0; a < n; ++a) s += mx[a * n + ap] * s0[(a * nn + bc) * rc + r];
    s1[e] = s;
  }
  __syncthreads();
  // y^T: s0[a][bp][c] = sum_b my[b][bp] s1[a][b][c]
  for (int e = threadIdx.x; e < nout; e += 256) {
    int m = e / rc, r = e % rc;
    int a = m / nn, bp = (m / n) % n, c = m % n;
    T s = 0;
    for (int bb = 0; bb < n; ++bb) s += my[bb * n + bp] * s1[((a * n + bb) * n + c) * rc + r];
    s0[e] = s;
  }
  __syncthreads();
  // z^T: Lc[child][a][b][cp] += sum_c mz[c][cp] s0[a][b][c]
  T *lc = Lc + (size_t)ch * n3 * ldm;
  for (int e = threadIdx.x; e < nout; e += 256) {
    int m = e / rc, r = e % rc;
    int ab = m / n, cp = m % n;
    T s = 0;
    for (int c = 0; c < n; ++c) s += mz[c * n + cp] * s0[(ab * n + c) * rc + r];
    lc[(size_t)m * ldm + r0 + r] += s;
  }
}

// L2P (bbfmm.rs:400-417): y[p][r] += sum_m S(p, m) Lc[leaf][m][r].
template <typename T>
__global__ __launch_bounds__(256) void k_l2p(int n, const int *__restrict__ leaf_lo, int leaf_box0,
                                             const T *__restrict__ w1d, int nr,
                                             const T *__restrict__ Lc, int ldm,
                                             double *__restrict__ y, int ldy) {
  extern __shared__ unsigned char smem_raw[];
  T *s0 = reinterpret_cast<T *>(smem_raw);
  const int li = blockIdx.x, r0 = blockIdx.y * FM_RC;
  const int rc = min(FM_RC, nr - r0);
  const int n3 = n * n * n, nn = n * n;
  const T *ll = Lc + (size_t)(leaf_box0 + li) * n3 * ldm;
  for (int e = threadIdx.x; e < n3 * rc; e += 256) s0[e] = ll[(size_t)(e / rc) * ldm + r0 + e % rc];
  __syncthreads();
  const int p0 = leaf_lo[li], p1 = leaf_lo[li + 1];
  const int cnt = (p1 - p0) * rc;
  for (int e = threadIdx.x; e < cnt; e += 256) {
    int p = p0 + e / rc, r = e % rc;
    const T *w = w1d + (size_t)p * 3 * n;
    T acc = 0;
    for (int a = 0; a < n; ++a) {
      T sa = 0;
      for (int bb = 0; bb < n; ++bb) {
        T sb = 0;
        for (int c = 0; c < n; ++c) sb += w[2 * n + c] * s0[(a * nn + bb * n + c) * rc + r];
        sa += w[n + bb] * sb;
      }
      acc += w[a] * sa;
    }
    y[(size_t)p * ldy + r0 + r] += (double)acc;
  }
}

template <typename T>
void run_far_field_t(const FmmDev<T> &f, const T *z, int ldz, int nr, T *M, T *Lc, int ldm,
                     double *y, int ldy, hipStream_t s) {
  const int n = f.n, n3 = n * n * n;
  const int nrc = (nr + FM_RC - 1) / FM_RC;
  const size_t lds1 = (size_t)n3 * FM_RC * sizeof(T);
  const int L = f.L;
  const int leaf0 = f.lvl_off[L];
  const int nleaf = f.lvl_off[L + 1] - leaf0;
  const bool f32 = sizeof(T) == 4;
  const double exp_bytes = (double)n3 * nr * sizeof(T);
  // P2M
  ProfScope *ps = new ProfScope(f32 ? "p2m_f32" : "p2m_f64", s, 2.0 * f.n_panels * n3 * nr,
                                (double)f.n_panels * nr * sizeof(T) + nleaf * exp_bytes);
  hipLaunchKernelGGL((k_p2m<T>), dim3(nleaf, nrc), dim3(256), 0, s, n, f.leaf_lo, leaf0, f.w1d,
                     f.area, z, ldz, nr, M, ldm);
  GP_LAUNCH_CHECK();
  delete ps;
  // M2M, finest to level 2 (multipoles above level 2 are never used).
  for (int l = L - 1; l >= 2; --l) {
    int nb = f.lvl_off[l + 1] - f.lvl_off[l];
    if (nb == 0) continue;
    int nch = f.lvl_off[l + 2] - f.lvl_off[l + 1];
    ProfScope pm(f32 ? "m2m_f32" : "m2m_f64", s, 6.0 * n * n3 * nr * nch, (nb + nch) * exp_bytes);
    hipLaunchKernelGGL((k_m2m<T>), dim3(nb, nrc), dim3(256), 2 * lds1, s, n, f.lvl_off[l],
                       f.child_off, f.child_idx, f.octant, f.m2m1d, nr, M, ldm);
    GP_LAUNCH_CHECK();
  }
  // M2L over every box of level >= 2 with a non-empty list; boxes without
  // one get a zero local expansion.
  GP_HIP(hipMemsetAsync(Lc, 0, (size_t)f.n_boxes * n3 * ldm * sizeof(T), s));
  if (f.n_targets > 0) {
    ProfScope pl(f32 ? "m2l_f32" : "m2l_f64", s, 2.0 * (double)f.n_pairs * n3 * n3 * nr,
                 (double)(f.n_boxes + f.n_targets) * exp_bytes);
    hipLaunchKernelGGL((k_m2l<T>), dim3(f.n_targets, nrc), dim3(256), lds1, s, n, f.targets,
                       f.il_off, f.il_src, f.il_oidx, f.box_level, f.inv_half, f.G, nr, M, Lc,
                       ldm);
    GP_LAUNCH_CHECK();
  }
  // L2L, level 2 to the leaves.
  for (int l = 2; l < L; ++l) {
    int nb = f.lvl_off[l + 2] - f.lvl_off[l + 1];
    if (nb == 0) continue;
    ProfScope pm(f32 ? "l2l_f32" : "l2l_f64", s, 6.0 * n * n3 * nr * nb, 2.0 * nb * exp_bytes);
    hipLaunchKernelGGL((k_l2l<T>), dim3(nb, nrc), dim3(256), 2 * lds1, s, n, f.lvl_off[l + 1],
                       f.parent, f.octant, f.m2m1d, nr, Lc, ldm);
    GP_LAUNCH_CHECK();
  }
  ProfScope pp(f32 ? "l2p_f32" : "l2p_f64", s, 2.0 * f.n_panels * n3 * nr,
               (double)f.n_panels * nr * 16.0 + nleaf * exp_bytes);
  hipLaunchKernelGGL((k_l2p<T>), dim3(nleaf, nrc), dim3(256), lds1, s, n, f.leaf_lo, leaf0, f.w1d,
                     nr, Lc, ldm, y, ldy);
  GP_LAUNCH_CHECK();
}

void run_far_field(const FmmDev<double> &f, const double *z, int ldz, int nr, double *M,
                   double *Lc, int ldm, double *y, int ldy, hipStream_t s) {
  run_far_field_t<double>(f, z, ldz, nr, M, Lc, ldm, y, ldy, s);
}
void run_far_field(const FmmDev<float> &f, const float *z, int ldz, int nr, float *M, float *Lc,
                   int ldm, double *y, int ldy, hipStream_t s) {
  run_far_field_t<float>(f, z, ldz, nr, M, Lc, ldm, y, ldy, s);
}

}  // namespace gp
