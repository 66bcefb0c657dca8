// Black-box Chebyshev FMM passes on the device (fmm/bbfmm.rs:209-448), for a
// block of right-hand sides, in FP64 or FP32 (template T).
//
// Layouts (per box, m = Chebyshev node a n^2 + b n + c, r = RHS column):
//   multipoles M[box][r][m]  (m fastest, padded to n3p, a multiple of 32),
//   locals    Lc[box][m][r]  (r fastest).
// The multipole layout makes both operands of the M2L matrix-core product
// contiguous along the contraction index.  P2M and M2M also write a BF16
// copy of the multipoles when the M2L runs on BF16 MFMA.
//
// Generic kernels: one work-group per box and 16-column RHS chunk, each
// thread owns fixed (node, column) outputs in registers, inputs staged in LDS.
// MFMA M2L: one work-group (4 wavefronts) per target box and all RHS columns
// (<= 128); the accumulator is a grid of 16x16 tiles (RHS x node), the four
// wavefronts split the node dimension, and every interaction pair streams its
// source multipole and its interaction matrix through 16x16x32 BF16 (or
// 16x16x4 FP64) matrix-core products.

#include <cstdlib>
#include <type_traits>

#include "gp_common.h"
#include "gp_kernels.h"

namespace gp {

constexpr int FM_RC = 16;   // RHS columns per work-group (generic kernels)
constexpr int FM_OPT = 16;  // max outputs per thread: n^3 * 16 / 256 for n <= 6

template <typename T>
__device__ inline void store_mult(T *M, __bf16 *Mb, size_t idx, T v) {
  M[idx] = v;
  if (Mb) Mb[idx] = (__bf16)(float)v;
}

// P2M (bbfmm.rs:215-228): M[leaf][r][m] = sum_p S(p, m) area_p z[p][r], with
// the tensor weights S(p, (a,b,c)) = wx[a] wy[b] wz[c] of the point's
// 1-D interpolation weights w1d[p][axis][k] (chebyshev.rs:77-94).
template <typename T>
__global__ __launch_bounds__(256) void k_p2m(int n, int n3p, const int *__restrict__ leaf_lo,
                                             const int *__restrict__ leaf_box, int leaf_box0, const T *__restrict__ w1d,
                                             const double *__restrict__ area,
                                             const T *__restrict__ z, int ldz, int nr,
                                             T *__restrict__ M, __bf16 *__restrict__ Mb, int ldm) {
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
        int r = r0 + o / n3, m = o % n3;
        int a = m / (n * n), b = (m / n) % n, c = m % n;
        T q = ap * z[(size_t)p * ldz + r];
        acc[k] += w[a] * w[n + b] * w[2 * n + c] * q;
      }
    }
  }
  const size_t base = (size_t)(leaf_box ? leaf_box[li] : leaf_box0 + li) * ldm * n3p;
#pragma unroll
  for (int k = 0; k < FM_OPT; ++k) {
    int o = threadIdx.x + k * 256;
    if (o < nout) {
      int r = r0 + o / n3, m = o % n3;
      store_mult(M, Mb, base + (size_t)r * n3p + m, acc[k]);
    }
  }
}

// M2M (bbfmm.rs:230-294): for every box of one level with children,
// M[parent] = sum over children of (X (x) Y (x) Z)_octant M[child], applied as
// three 1-D contractions (z, then y, then x) through two LDS buffers laid out
// [m][r] within the chunk.
template <typename T>
__global__ __launch_bounds__(256) void k_m2m(int n, int n3p, int box0,
                                             const int *__restrict__ child_off,
                                             const int *__restrict__ child_idx,
                                             const int *__restrict__ octant,
                                             const T *__restrict__ m2m1d, int nr,
                                             T *__restrict__ M, __bf16 *__restrict__ Mb, int ldm) {
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
    const T *wc = M + (size_t)ch * ldm * n3p;
    for (int e = threadIdx.x; e < nout; e += 256) {
      int r = e / n3, m = e % n3;
      s0[m * rc + r] = wc[(size_t)(r0 + r) * n3p + m];
    }
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
    // x: acc[ap][bp][cp] += sum_a mx[ap][a] s0[a][bp][cp]; thread output o maps
    // to (r = o / n3, m = o % n3) so the final store is contiguous in m.
#pragma unroll
    for (int k = 0; k < FM_OPT; ++k) {
      int o = threadIdx.x + k * 256;
      if (o < nout) {
        int r = o / n3, m = o % n3;
        int ap = m / nn, bc = m % nn;
        T s = 0;
        for (int a = 0; a < n; ++a) s += mx[ap * n + a] * s0[(a * nn + bc) * rc + r];
        acc[k] += s;
      }
    }
    __syncthreads();
  }
  const size_t base = (size_t)b * ldm * n3p;
#pragma unroll
  for (int k = 0; k < FM_OPT; ++k) {
    int o = threadIdx.x + k * 256;
    if (o < nout) {
      int r = o / n3, m = o % n3;
      store_mult(M, Mb, base + (size_t)(r0 + r) * n3p + m, acc[k]);
    }
  }
}

// Generic M2L (bbfmm.rs:296-335) with the uncompressed Chebyshev interaction
// matrix G_off[m][l] = 1/|x_m - x_l - 2 off| on the reference box
// (bbfmm.rs:156-166, stored n3p x n3p), scaled by 1/half of the level:
// Lc[t] = (1/half) sum_(s,off) G_off M[s].  targets: boxes with a non-empty
// interaction list.
template <typename T>
__global__ __launch_bounds__(256) void k_m2l(int n3, int n3p, const int *__restrict__ targets,
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
  const int t = targets[blockIdx.x], r0 = blockIdx.y * FM_RC;
  const int rc = min(FM_RC, nr - r0);
  const int nout = n3 * rc;
  T acc[FM_OPT];
#pragma unroll
  for (int k = 0; k < FM_OPT; ++k) acc[k] = 0;
  for (int q = il_off[t]; q < il_off[t + 1]; ++q) {
    const T *ws = M + (size_t)il_src[q] * ldm * n3p;
    const T *g = G + (size_t)il_oidx[q] * n3p * n3p;
    for (int e = threadIdx.x; e < nout; e += 256) {
      int r = e / n3, m = e % n3;
      s0[m * rc + r] = ws[(size_t)(r0 + r) * n3p + m];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < FM_OPT; ++k) {
      int e = threadIdx.x + k * 256;
      if (e < nout) {
        int m = e / rc, r = e % rc;
        const T *gr = g + (size_t)m * n3p;
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

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

// Matrix-core M2L.  Product per pair: D (RHS x node) += A (RHS x l) * B
// (l x node) with A = the source's multipole rows and B[l][m] = G_off[m][l];
// both operands are contiguous along the contraction index l.
//   v_mfma_f32_16x16x32_bf16: lane holds A[i = lane % 16][k = 8 (lane / 16) + 0..7],
//     B[k = 8 (lane / 16) + 0..7][j = lane % 16], D[i = 4 (lane / 16) + v][j = lane % 16];
//   v_mfma_f64_16x16x4f64 interleaves its output rows: D[i][j] sits in lane
//     j + 16 (i % 4), register i / 4.
// One work-group (four wavefronts) per target box and all RHS columns
// (RT <= RTM tiles of 16, ldm <= 128); n3p = 16 MT and wavefront w owns node
// tiles [w MTW, (w + 1) MTW), MTW = ceil(MT / 4).

// BF16 inputs, FP32 accumulation.  Per pair the source multipole rows
// (RT * 16 rows of n3p) are staged once per work-group in LDS, rows padded by
// 8 elements so that the 8 lanes of each 128-byte read phase hit distinct
// banks, and read by all four wavefronts as the A operand; each wavefront
// keeps its G_off fragments (B operand) in registers.  The next pair's
// multipole rows are prefetched into registers while the current pair runs
// on the matrix cores, and so are its G_off fragments when they fit
// (MTW * n3p / 32 <= 8 fragments, orders 2 to 5).  DEPTH 2 keeps the source
// rows of the next two pairs in flight (two register buffers).
template <int MT, int RTM, int DEPTH = 1>
__global__ __launch_bounds__(256) void k_m2l_bf16_pipe(int n3, int nr, int ldm, int xrun,
                                                       const int *__restrict__ targets,
                                                       const int *__restrict__ il_off,
                                                       const int *__restrict__ il_src,
                                                       const int *__restrict__ il_oidx,
                                                       const int *__restrict__ box_level,
                                                       const double *__restrict__ inv_half,
                                                       const __bf16 *__restrict__ G,
                                                       const __bf16 *__restrict__ Mb,
                                                       float *__restrict__ Lc) {
  constexpr int NP = 16 * MT;      // n3p
  constexpr int KS = NP / 32;      // contraction steps of 32
  constexpr int MTW = (MT + 3) / 4;
  constexpr int LDA = NP + 8;      // padded LDS row (elements)
  constexpr int VPR = NP / 8;      // 16-byte vectors per multipole row
  constexpr int APT = (RTM * 16 * VPR + 255) / 256;
  constexpr bool BPRE = MTW * KS <= 8;
  __shared__ __attribute__((aligned(16))) __bf16 sA[RTM * 16 * LDA];
  const int t = targets[xcd_block(blockIdx.x, gridDim.x, xrun)];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int li = lane & 15, lk = (lane >> 4) * 8;
  const int rt0 = blockIdx.y * RTM;  // RHS tiles [rt0, rt0 + RT) of this work-group
  const int RT = min((nr + 15) >> 4, rt0 + RTM) - rt0;
  const int AV = RT * 16 * VPR;    // vectors of one pair's A block
  const size_t MS = (size_t)ldm * NP;
  f32x4 acc[RTM][MTW];
#pragma unroll
  for (int i = 0; i < RTM; ++i)
#pragma unroll
    for (int j = 0; j < MTW; ++j) acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
  const int q0 = il_off[t], q1 = il_off[t + 1];
  bf16x8 apre[APT], apre2[DEPTH > 1 ? APT : 1];
  bf16x8 bcur[MTW][KS];
  bf16x8 bnext[BPRE ? MTW : 1][BPRE ? KS : 1];
  auto load_a = [&](int q, bf16x8 (&dst)[APT]) {
    const bf16x8 *src =
        reinterpret_cast<const bf16x8 *>(Mb + (size_t)il_src[q] * MS + (size_t)rt0 * 16 * NP);
#pragma unroll
    for (int i = 0; i < APT; ++i) {
      const int v = tid + 256 * i;
      if (v < AV) dst[i] = src[v];
    }
  };
  auto load_b = [&](int q, bf16x8 (&dst)[MTW][KS]) {
    const __bf16 *Go = G + (size_t)il_oidx[q] * NP * NP;
#pragma unroll
    for (int j = 0; j < MTW; ++j) {
      const int mt = wave * MTW + j;
      if (mt < MT)
#pragma unroll
        for (int k = 0; k < KS; ++k)
          dst[j][k] = *(const bf16x8 *)(Go + (size_t)(mt * 16 + li) * NP + k * 32 + lk);
    }
  };
  if (q0 < q1) {
    load_a(q0, apre);
    if constexpr (DEPTH > 1)
      if (q0 + 1 < q1) load_a(q0 + 1, apre2);
    if constexpr (BPRE) load_b(q0, bcur);
  }
  // one pair: its source rows (ab) into LDS, the rows DEPTH pairs ahead into
  // ab, the next G_off fragments, the matrix-core products
  auto step = [&](int q, bf16x8 (&ab)[APT]) {
    __syncthreads();  // previous pair's LDS reads are done
#pragma unroll
    for (int i = 0; i < APT; ++i) {
      const int v = tid + 256 * i;
      if (v < AV) *(bf16x8 *)(sA + (v / VPR) * LDA + (v % VPR) * 8) = ab[i];
    }
    __syncthreads();
    if constexpr (!BPRE) load_b(q, bcur);
    if (q + DEPTH < q1) load_a(q + DEPTH, ab);
    if constexpr (BPRE)
      if (q + 1 < q1) load_b(q + 1, bnext);
#pragma unroll
    for (int k = 0; k < KS; ++k) {
#pragma unroll
      for (int rt = 0; rt < RTM; ++rt) {
        if (rt < RT) {
          const bf16x8 a = *(const bf16x8 *)(sA + (rt * 16 + li) * LDA + k * 32 + lk);
#pragma unroll
          for (int j = 0; j < MTW; ++j)
            if (wave * MTW + j < MT)
              acc[rt][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bcur[j][k], acc[rt][j], 0, 0, 0);
        }
      }
    }
    if constexpr (BPRE) {
#pragma unroll
      for (int j = 0; j < MTW; ++j)
#pragma unroll
        for (int k = 0; k < KS; ++k) bcur[j][k] = bnext[j][k];
    }
  };
  if constexpr (DEPTH > 1) {
    for (int q = q0; q < q1; q += 2) {
      step(q, apre);
      if (q + 1 < q1) step(q + 1, apre2);
    }
  } else {
    for (int q = q0; q < q1; ++q) step(q, apre);
  }
  const float ih = (float)inv_half[box_level[t]];
  float *out = Lc + (size_t)t * n3 * ldm;
#pragma unroll
  for (int rt = 0; rt < RTM; ++rt)
#pragma unroll
    for (int j = 0; j < MTW; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = (rt0 + rt) * 16 + (lane >> 4) * 4 + v, m = (wave * MTW + j) * 16 + li;
        if (rt < RT && r < nr && m < n3) out[(size_t)m * ldm + r] = acc[rt][j][v] * ih;
      }
}

// FP64 (v_mfma_f64_16x16x4f64).  The contraction runs in steps of 16 over
// every (pair, k-slab); the source multipole slab (RT * 16 rows x 16) is
// staged in a double-buffered LDS tile shared by the four wavefronts, each
// wavefront keeps its node tiles' G_off fragments in registers, and the next
// step's slab and fragments are loaded while the current step runs on the
// matrix cores (one barrier per step).  Each lane loads 4 consecutive
// contraction entries (32 bytes) of both operands and feeds them to 4
// successive MFMAs: in MFMA kk lane group g = lane / 16 supplies contraction
// index k0 + 4 g + kk for both operands, so the product sums every index once.
template <int MT, int RTM>
__global__ __launch_bounds__(256) void k_m2l_f64_pipe(int n3, int nr, int ldm, int xrun,
                                                      const int *__restrict__ targets,
                                                      const int *__restrict__ il_off,
                                                      const int *__restrict__ il_src,
                                                      const int *__restrict__ il_oidx,
                                                      const int *__restrict__ box_level,
                                                      const double *__restrict__ inv_half,
                                                      const double *__restrict__ G,
                                                      const double *__restrict__ M,
                                                      double *__restrict__ Lc) {
  constexpr int NP = 16 * MT;   // n3p
  constexpr int MTW = (MT + 3) / 4;
  constexpr int NKS = NP / 16;  // k-slabs per pair
  constexpr int KP = 20;        // padded LDS row (doubles)
  constexpr int APT = RTM;      // slab elements per thread (RTM * 16 rows x 16)
  __shared__ __attribute__((aligned(32))) double sA[2][RTM * 16 * KP];
  const int t = targets[xcd_block(blockIdx.x, gridDim.x, xrun)];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int li = lane & 15, lk = (lane >> 4) * 4;
  const int RT = min((nr + 15) >> 4, RTM);
  const int AE = RT * 16 * 16;
  const size_t MS = (size_t)ldm * NP;
  f64x4 acc[RTM][MTW];
#pragma unroll
  for (int i = 0; i < RTM; ++i)
#pragma unroll
    for (int j = 0; j < MTW; ++j) acc[i][j] = (f64x4){0.0, 0.0, 0.0, 0.0};
  const int q0 = il_off[t], q1 = il_off[t + 1];
  const int nsteps = (q1 - q0) * NKS;
  double apre[APT];
  f64x4 bcur[MTW], bnext[MTW];
  auto load_step = [&](int st, f64x4 (&bd)[MTW]) {
    const int q = q0 + st / NKS, k0 = (st % NKS) * 16;
    const double *Ms = M + (size_t)il_src[q] * MS + k0;
#pragma unroll
    for (int i = 0; i < APT; ++i) {
      const int e = tid + 256 * i;
      if (e < AE) apre[i] = Ms[(size_t)(e >> 4) * NP + (e & 15)];
    }
    const double *Go = G + (size_t)il_oidx[q] * NP * NP + k0 + lk;
#pragma unroll
    for (int j = 0; j < MTW; ++j) {
      const int mt = wave * MTW + j;
      if (mt < MT) bd[j] = *(const f64x4 *)(Go + (size_t)(mt * 16 + li) * NP);
    }
  };
  auto store_a = [&](int buf) {
#pragma unroll
    for (int i = 0; i < APT; ++i) {
      const int e = tid + 256 * i;
      if (e < AE) sA[buf][(e >> 4) * KP + (e & 15)] = apre[i];
    }
  };
  if (nsteps > 0) {
    load_step(0, bcur);
    store_a(0);
  }
  __syncthreads();
  for (int st = 0; st < nsteps; ++st) {
    const int buf = st & 1;
    if (st + 1 < nsteps) load_step(st + 1, bnext);
#pragma unroll
    for (int rt = 0; rt < RTM; ++rt) {
      if (rt < RT) {
        const f64x4 a = *(const f64x4 *)(&sA[buf][(rt * 16 + li) * KP + lk]);
#pragma unroll
        for (int j = 0; j < MTW; ++j)
          if (wave * MTW + j < MT)
#pragma unroll
            for (int kk = 0; kk < 4; ++kk)
              acc[rt][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kk], bcur[j][kk], acc[rt][j], 0, 0, 0);
      }
    }
    if (st + 1 < nsteps) {
      store_a(buf ^ 1);
#pragma unroll
      for (int j = 0; j < MTW; ++j) bcur[j] = bnext[j];
    }
    __syncthreads();
  }
  const double ih = inv_half[box_level[t]];
  double *out = Lc + (size_t)t * n3 * ldm;
#pragma unroll
  for (int rt = 0; rt < RTM; ++rt)
#pragma unroll
    for (int j = 0; j < MTW; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = rt * 16 + (lane >> 4) + 4 * v, m = (wave * MTW + j) * 16 + li;
        if (rt < RT && r < nr && m < n3) out[(size_t)m * ldm + r] = acc[rt][j][v] * ih;
      }
}

// Single right-hand side L2P for n^3 <= 64: one wavefront per leaf, lane m
// holding local coefficient m (one coalesced load); per panel of the leaf the
// lanes form w_x[a] w_y[b] w_z[c] L[m] and a shuffle tree sums them (the
// panel rows of the leaf are read as broadcasts, four panels in flight).
template <typename T>
__global__ __launch_bounds__(256) void k_l2p1w(int n, int nleaf, const int *__restrict__ leaf_lo,
                                               const int *__restrict__ leaf_box, int leaf_box0,
                                               const T *__restrict__ w1d,
                                               const T *__restrict__ Lc, double *__restrict__ y,
                                               const unsigned char *__restrict__ skip) {
  const int li = blockIdx.x * 4 + (threadIdx.x >> 6), m = threadIdx.x & 63;
  if (li >= nleaf || (skip && skip[li])) return;  // wavefront-uniform
  const int n3 = n * n * n;
  const int mm = m < n3 ? m : 0;  // idle lanes read node 0
  const int a = mm / (n * n), b = (mm / n) % n, c = mm % n;
  const int box = leaf_box ? leaf_box[li] : leaf_box0 + li;
  const T l = m < n3 ? Lc[(size_t)box * n3 + m] : (T)0;
  const int p0 = leaf_lo[li], p1 = leaf_lo[li + 1];
  for (int p = p0; p < p1; p += 4) {
    T t[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const T *w = w1d + (size_t)min(p + u, p1 - 1) * 3 * n;  // in range even past the leaf
      const T v = w[a] * w[n + b] * w[2 * n + c] * l;
      t[u] = (p + u < p1 && m < n3) ? v : (T)0;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
      for (int u = 0; u < 4; ++u) t[u] += __shfl_xor(t[u], off);
    if (m == 0)
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (p + u < p1) y[p + u] += (double)t[u];
  }
}

// Single right-hand side (ldm = 1) P2M: several leaves per work-group, one
// thread per (leaf, node), each summing its leaf's panels in order (the
// same expression and order as k_p2m).
template <typename T, int NF = 0>  // NF > 0: the order fixed at compile time
__global__ __launch_bounds__(256) void k_p2m1(int n_rt, int n3p, int tpl, int nleaf,
                                              const int *__restrict__ leaf_lo,
                                              const int *__restrict__ leaf_box, int leaf_box0,
                                              const T *__restrict__ w1d,
                                              const double *__restrict__ area,
                                              const T *__restrict__ z, T *__restrict__ M,
                                              __bf16 *__restrict__ Mb,
                                              const unsigned char *__restrict__ skip) {
  const int n = NF > 0 ? NF : n_rt;
  const int li = blockIdx.x * (256 / tpl) + threadIdx.x / tpl, m = threadIdx.x % tpl;
  const int n3 = n * n * n;
  if (li >= nleaf || m >= n3 || (skip && skip[li])) return;
  const int a = m / (n * n), b = (m / n) % n, c = m % n;
  T acc = 0;
#pragma unroll 4
  for (int p = leaf_lo[li]; p < leaf_lo[li + 1]; ++p) {
    const T *w = w1d + (size_t)p * 3 * n;
    const T q = (T)area[p] * z[p];
    acc += w[a] * w[n + b] * w[2 * n + c] * q;
  }
  const int box = leaf_box ? leaf_box[li] : leaf_box0 + li;
  store_mult(M, Mb, (size_t)box * n3p + m, acc);
}

// Single-RHS P2M with the leaf's panels staged: one work-group of
// round_up(n3, 64) threads per leaf loads the 1-D weights and charges of up to
// 64 panels at a time as one contiguous block into LDS, then thread m sums
// them for node m (the expression and order of k_p2m1).  k_p2m1 issues one
// round of dependent loads per panel; leaves of a few panels are all latency.
template <typename T, int NF>
__global__ __launch_bounds__(256) void k_p2m1s(int n3p, int nleaf, const int *__restrict__ leaf_lo,
                                               const int *__restrict__ leaf_box, int leaf_box0,
                                               const T *__restrict__ w1d,
                                               const double *__restrict__ area,
                                               const T *__restrict__ z, T *__restrict__ M,
                                               __bf16 *__restrict__ Mb,
                                               const unsigned char *__restrict__ skip) {
  constexpr int n = NF, n3 = n * n * n, W3 = 3 * n, CH = 64;
  __shared__ T sw[CH * W3];
  __shared__ T sq[CH];
  const int li = blockIdx.x, m = threadIdx.x;
  if (skip && skip[li]) return;  // work-group-uniform, before any barrier
  const int lo = leaf_lo[li], hi = leaf_lo[li + 1];
  const int a = m / (n * n), b = (m / n) % n, c = m % n;
  T acc = 0;
  for (int p0 = lo; p0 < hi; p0 += CH) {
    const int cnt = min(CH, hi - p0);
    __syncthreads();  // the previous chunk has been read
    for (int e = m; e < cnt * W3; e += blockDim.x) sw[e] = w1d[(size_t)p0 * W3 + e];
    for (int j = m; j < cnt; j += blockDim.x) sq[j] = (T)area[p0 + j] * z[p0 + j];
    __syncthreads();
    if (m < n3)
      for (int j = 0; j < cnt; ++j) {
        const T *w = sw + j * W3;
        acc += w[a] * w[n + b] * w[2 * n + c] * sq[j];
      }
  }
  if (m < n3) {
    const int box = leaf_box ? leaf_box[li] : leaf_box0 + li;
    store_mult(M, Mb, (size_t)box * n3p + m, acc);
  }
}

// Single-RHS P2M over several consecutive leaves per work-group (LPB = 256 /
// tpl leaf slots of tpl threads): the 1-D weights and charges of the
// group's panels (contiguous: its leaves are consecutive) are staged in LDS
// 256 panels at a time by coalesced loads of the whole block, then thread
// (slot, m) sums the panels of its leaf in that chunk for node m (the
// expression and order of k_p2m1).  With leaves of a few panels k_p2m1
// waits on one small load per panel and k_p2m1s runs a work-group per leaf;
// here every load instruction moves a full block.  A work-group whose leaves
// are all skipped leaves at once; otherwise panels of skipped leaves are
// staged but not summed.
template <typename T, int NF>
__global__ __launch_bounds__(256) void k_p2m1m(int n3p, int nleaf, const int *__restrict__ leaf_lo,
                                               const int *__restrict__ leaf_box, int leaf_box0,
                                               const T *__restrict__ w1d,
                                               const double *__restrict__ area,
                                               const T *__restrict__ z, T *__restrict__ M,
                                               __bf16 *__restrict__ Mb,
                                               const unsigned char *__restrict__ skip) {
  constexpr int n = NF, n3 = n * n * n, W3 = 3 * n, CH = 256;
  constexpr int TPL = n3 <= 32 ? 32 : n3 <= 64 ? 64 : n3 <= 128 ? 128 : 256, LPB = 256 / TPL;
  __shared__ T sw[CH * W3];
  __shared__ T sq[CH];
  const int l0 = blockIdx.x * LPB, l1 = min(l0 + LPB, nleaf);
  const int slot = threadIdx.x / TPL, m = threadIdx.x % TPL, li = l0 + slot;
  const bool mine = li < l1 && m < n3 && !(skip && skip[li]);
  if (!__syncthreads_or(mine)) return;  // every leaf of the group skipped
  const int P0 = leaf_lo[l0], P1 = leaf_lo[l1];
  const int lo = li < l1 ? leaf_lo[li] : P1, hi = li < l1 ? leaf_lo[li + 1] : P1;
  const int a = m / (n * n), b = (m / n) % n, c = m % n;
  T acc = 0;
  for (int c0 = P0; c0 < P1; c0 += CH) {
    const int cnt = min(CH, P1 - c0);
    __syncthreads();  // the previous chunk has been read
    for (int e = threadIdx.x; e < cnt * W3; e += 256) sw[e] = w1d[(size_t)c0 * W3 + e];
    if ((int)threadIdx.x < cnt) sq[threadIdx.x] = (T)area[c0 + threadIdx.x] * z[c0 + threadIdx.x];
    __syncthreads();
    if (mine) {
      const int j1 = min(hi, c0 + cnt) - c0;
      for (int j = max(lo, c0) - c0; j < j1; ++j) {
        const T *w = sw + j * W3;
        acc += w[a] * w[n + b] * w[2 * n + c] * sq[j];
      }
    }
  }
  if (mine) {
    const int box = leaf_box ? leaf_box[li] : leaf_box0 + li;
    store_mult(M, Mb, (size_t)box * n3p + m, acc);
  }
}

// ---------------------------------------------------------------- weights on the fly
// The FP32 single-RHS operator's P2M and L2P can form each panel's 1-D
// interpolation weights from its local coordinates (k_interp_w1d's sums,
// w_m(x) = sum_k wc[m n + k] T_k(x) with wc = (2 / n) T_k(x_m), the k = 0
// term halved), reading 16 bytes a panel (x, y, z in the leaf's box frame
// and the area, FP32) instead of the 3 n stored weights and the FP64 area
// (48 bytes at order 3).  The weights then differ from the stored ones (FP64
// sums rounded to FP32) by FP32 rounding.

// w_m(x) for one node row c of wc (n coefficients), by the Chebyshev
// recurrence.
template <int n>
__device__ __forceinline__ float cheb_w(float x, const float (&c)[n]) {
  float tkm1 = 1.0f, tk = x, s = c[0];
  if (n > 1) s += c[1] * x;
#pragma unroll
  for (int k = 2; k < n; ++k) {
    const float t = 2.0f * x * tk - tkm1;
    tkm1 = tk;
    tk = t;
    s += c[k] * t;
  }
  return s;
}

// xl[p] = (x, y, z, area): panel p's centroid in its leaf box's frame
// ((cen - centre) / half width, as k_interp_w1d) and its area, FP32.
__global__ __launch_bounds__(256) void k_local_coords(int N, const int *__restrict__ panel_leaf,
                                                      const double *__restrict__ geo,
                                                      const double *__restrict__ cen,
                                                      const double *__restrict__ area,
                                                      float4 *__restrict__ xl) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double *g = geo + (size_t)panel_leaf[i] * 4;
  xl[i] = make_float4((float)((cen[(size_t)i * 3] - g[0]) / g[3]),
                      (float)((cen[(size_t)i * 3 + 1] - g[1]) / g[3]),
                      (float)((cen[(size_t)i * 3 + 2] - g[2]) / g[3]), (float)area[i]);
}

void launch_local_coords(int N, const int *panel_leaf, const double *geo, const double *cen,
                         const double *area, float *xl, hipStream_t s) {
  if (N == 0) return;
  hipLaunchKernelGGL(k_local_coords, dim3((N + 255) / 256), dim3(256), 0, s, N, panel_leaf, geo, cen,
                     area, reinterpret_cast<float4 *>(xl));
  GP_LAUNCH_CHECK();
}

// k_p2m1 with the weights formed from xl: thread (leaf, node m = (a, b, c))
// evaluates w_a(x) w_b(y) w_c(z) per panel.
template <int NF>
__global__ __launch_bounds__(256) void k_p2m1x(int n3p, int tpl, int nleaf,
                                               const int *__restrict__ leaf_lo,
                                               const int *__restrict__ leaf_box, int leaf_box0,
                                               const float4 *__restrict__ xl,
                                               const float *__restrict__ wc,
                                               const float *__restrict__ z, float *__restrict__ M,
                                               __bf16 *__restrict__ Mb,
                                               const unsigned char *__restrict__ skip) {
  constexpr int n = NF, n3 = n * n * n;
  const int li = blockIdx.x * (256 / tpl) + threadIdx.x / tpl, m = threadIdx.x % tpl;
  if (li >= nleaf || m >= n3 || (skip && skip[li])) return;
  const int a = m / (n * n), b = (m / n) % n, c = m % n;
  float ca[n], cb[n], cc[n];
#pragma unroll
  for (int k = 0; k < n; ++k) ca[k] = wc[a * n + k], cb[k] = wc[b * n + k], cc[k] = wc[c * n + k];
  float acc = 0;
#pragma unroll 4
  for (int p = leaf_lo[li]; p < leaf_lo[li + 1]; ++p) {
    const float4 v = xl[p];
    const float q = v.w * z[p];
    acc += cheb_w<n>(v.x, ca) * cheb_w<n>(v.y, cb) * cheb_w<n>(v.z, cc) * q;
  }
  const int box = leaf_box ? leaf_box[li] : leaf_box0 + li;
  store_mult(M, Mb, (size_t)box * n3p + m, acc);
}

// k_p2m1m with the weights formed from xl: the group's panels' (x, y, z,
// area z) staged in LDS 256 at a time, thread (slot, m) evaluating its
// node's three weights per panel.
template <int NF>
__global__ __launch_bounds__(256) void k_p2m1mx(int n3p, int nleaf, const int *__restrict__ leaf_lo,
                                                const int *__restrict__ leaf_box, int leaf_box0,
                                                const float4 *__restrict__ xl,
                                                const float *__restrict__ wc,
                                                const float *__restrict__ z, float *__restrict__ M,
                                                __bf16 *__restrict__ Mb,
                                                const unsigned char *__restrict__ skip) {
  constexpr int n = NF, n3 = n * n * n, CH = 256;
  constexpr int TPL = n3 <= 32 ? 32 : n3 <= 64 ? 64 : n3 <= 128 ? 128 : 256, LPB = 256 / TPL;
  __shared__ float4 sx[CH];
  const int l0 = blockIdx.x * LPB, l1 = min(l0 + LPB, nleaf);
  const int slot = threadIdx.x / TPL, m = threadIdx.x % TPL, li = l0 + slot;
  const bool mine = li < l1 && m < n3 && !(skip && skip[li]);
  if (!__syncthreads_or(mine)) return;  // every leaf of the group skipped
  const int P0 = leaf_lo[l0], P1 = leaf_lo[l1];
  const int lo = li < l1 ? leaf_lo[li] : P1, hi = li < l1 ? leaf_lo[li + 1] : P1;
  const int mm = min(m, n3 - 1), a = mm / (n * n), b = (mm / n) % n, c = mm % n;
  float ca[n], cb[n], cc[n];
#pragma unroll
  for (int k = 0; k < n; ++k) ca[k] = wc[a * n + k], cb[k] = wc[b * n + k], cc[k] = wc[c * n + k];
  float acc = 0;
  for (int c0 = P0; c0 < P1; c0 += CH) {
    const int cnt = min(CH, P1 - c0);
    __syncthreads();  // the previous chunk has been read
    if ((int)threadIdx.x < cnt) {
      float4 v = xl[c0 + threadIdx.x];
      v.w *= z[c0 + threadIdx.x];  // area * z, the charge
      sx[threadIdx.x] = v;
    }
    __syncthreads();
    if (mine) {
      const int j1 = min(hi, c0 + cnt) - c0;
      for (int j = max(lo, c0) - c0; j < j1; ++j) {
        const float4 v = sx[j];
        acc += cheb_w<n>(v.x, ca) * cheb_w<n>(v.y, cb) * cheb_w<n>(v.z, cc) * v.w;
      }
    }
  }
  if (mine) {
    const int box = leaf_box ? leaf_box[li] : leaf_box0 + li;
    store_mult(M, Mb, (size_t)box * n3p + m, acc);
  }
}

// k_l2p1r with the weights formed from xl.
template <int NF>
__global__ __launch_bounds__(256) void k_l2p1rx(int N, const int *__restrict__ panel_leaf,
                                                const int *__restrict__ leaf_box, int leaf_box0,
                                                const float4 *__restrict__ xl,
                                                const float *__restrict__ wc,
                                                const float *__restrict__ Lc,
                                                double *__restrict__ y,
                                                const unsigned char *__restrict__ skip) {
  constexpr int n = NF, nn = n * n;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const int li = panel_leaf[p];
  if (skip && skip[li]) return;
  const float *ll = Lc + (size_t)(leaf_box ? leaf_box[li] : leaf_box0 + li) * n * nn;
  const float4 v = xl[p];
  float wx[n], wy[n], wz[n];
#pragma unroll
  for (int i = 0; i < n; ++i) {
    float c[n];
#pragma unroll
    for (int k = 0; k < n; ++k) c[k] = wc[i * n + k];
    wx[i] = cheb_w<n>(v.x, c);
    wy[i] = cheb_w<n>(v.y, c);
    wz[i] = cheb_w<n>(v.z, c);
  }
  float acc = 0;
#pragma unroll
  for (int a = 0; a < n; ++a) {
    float sa = 0;
#pragma unroll
    for (int bb = 0; bb < n; ++bb) {
      float sb = 0;
#pragma unroll
      for (int c = 0; c < n; ++c) sb += wz[c] * ll[a * nn + bb * n + c];
      sa += wy[bb] * sb;
    }
    acc += wx[a] * sa;
  }
  y[p] += (double)acc;
}

// Single right-hand side L2P: one thread per panel, y[p] += sum over nodes of
// S(p, node) Lc[leaf][node] (the same expression and order as k_l2p).
template <typename T>
__global__ __launch_bounds__(256) void k_l2p1(int n, int N, const int *__restrict__ panel_leaf,
                                              const int *__restrict__ leaf_box, int leaf_box0,
                                              const T *__restrict__ w1d, const T *__restrict__ Lc,
                                              double *__restrict__ y,
                                              const unsigned char *__restrict__ skip) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const int li = panel_leaf[p], nn = n * n;
  if (skip && skip[li]) return;
  const T *ll = Lc + (size_t)(leaf_box ? leaf_box[li] : leaf_box0 + li) * n * nn;
  const T *w = w1d + (size_t)p * 3 * n;
  T acc = 0;
  for (int a = 0; a < n; ++a) {
    T sa = 0;
    for (int bb = 0; bb < n; ++bb) {
      T sb = 0;
      for (int c = 0; c < n; ++c) sb += w[2 * n + c] * ll[a * nn + bb * n + c];
      sa += w[n + bb] * sb;
    }
    acc += w[a] * sa;
  }
  y[p] += (double)acc;
}

// Single right-hand side L2P with the order fixed at compile time: one
// thread per panel (neighbouring threads are the panels of one leaf and share
// its locals' cache lines), the weights in registers, the same expression and
// order as k_l2p1.  A leaf of a few panels keeps most lanes of the
// wavefront-per-leaf form idle; this form keeps every lane on a panel.
template <typename T, int NF>
__global__ __launch_bounds__(256) void k_l2p1r(int N, const int *__restrict__ panel_leaf,
                                               const int *__restrict__ leaf_box, int leaf_box0,
                                               const T *__restrict__ w1d, const T *__restrict__ Lc,
                                               double *__restrict__ y,
                                               const unsigned char *__restrict__ skip) {
  constexpr int n = NF, nn = n * n;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const int li = panel_leaf[p];
  if (skip && skip[li]) return;
  const T *ll = Lc + (size_t)(leaf_box ? leaf_box[li] : leaf_box0 + li) * n * nn;
  const T *w = w1d + (size_t)p * 3 * n;
  T wx[n], wy[n], wz[n];
#pragma unroll
  for (int i = 0; i < n; ++i) {
    wx[i] = w[i];
    wy[i] = w[n + i];
    wz[i] = w[2 * n + i];
  }
  T acc = 0;
#pragma unroll
  for (int a = 0; a < n; ++a) {
    T sa = 0;
#pragma unroll
    for (int bb = 0; bb < n; ++bb) {
      T sb = 0;
#pragma unroll
      for (int c = 0; c < n; ++c) sb += wz[c] * ll[a * nn + bb * n + c];
      sa += wy[bb] * sb;
    }
    acc += wx[a] * sa;
  }
  y[p] += (double)acc;
}

// Single right-hand side L2P, order 4, FP32: one thread per panel, its
// leaf's 64 local coefficients as 16 aligned 16-byte loads (the panels of a
// leaf are neighbouring threads and share those lines), the tensor-product
// sum sum_a wx[a] sum_b wy[b] sum_c wz[c] L[a][b][c] in registers.
__global__ __launch_bounds__(256) void k_l2p1v4(int N, const int *__restrict__ panel_leaf,
                                                const int *__restrict__ leaf_box, int leaf_box0,
                                                const float *__restrict__ w1d,
                                                const float *__restrict__ Lc,
                                                double *__restrict__ y,
                                                const unsigned char *__restrict__ skip) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const int li = panel_leaf[p];
  if (skip && skip[li]) return;
  const f4 *l4 = reinterpret_cast<const f4 *>(Lc + (size_t)(leaf_box ? leaf_box[li] : leaf_box0 + li) * 64);
  const f4 *w4 = reinterpret_cast<const f4 *>(w1d + (size_t)p * 12);
  f4 L[16];
#pragma unroll
  for (int e = 0; e < 16; ++e) L[e] = l4[e];
  const f4 wx = w4[0], wy = w4[1], wz = w4[2];
  float acc = 0.f;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    float sa = 0.f;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const f4 l = L[a * 4 + b];
      sa += wy[b] * (wz[0] * l[0] + wz[1] * l[1] + wz[2] * l[2] + wz[3] * l[3]);
    }
    acc += wx[a] * sa;
  }
  y[p] += (double)acc;
}

// Single right-hand side M2M of parent box b by the tpb threads of one LDS
// slot of the work-group (thread m: node m; on = false: loads and stores
// nothing but passes the barriers).  All children are loaded into the slot
// at once (8 loads in flight), then the three 1-D contractions (z, then y,
// then x, as k_m2m) run for every child between three barriers, the x pass
// summing over the children in order.
template <typename T, int NF>
__device__ __forceinline__ void m2m1_box(int n_rt, int n3p, int tpb, int b, bool on, int slot,
                                         int m, unsigned char *smem_raw,
                                         const int *__restrict__ child_off,
                                         const int *__restrict__ child_idx,
                                         const int *__restrict__ octant,
                                         const T *__restrict__ m2m1d, T *__restrict__ M,
                                         __bf16 *__restrict__ Mb) {
  const int n = NF > 0 ? NF : n_rt;
  const int n3 = n * n * n, nn = n * n;
  T *s0 = reinterpret_cast<T *>(smem_raw) + (size_t)slot * 16 * tpb, *s1 = s0 + 8 * tpb;
  const bool live = on && m < n3;
  const int c0 = child_off[b], nch = on ? child_off[b + 1] - c0 : 0;
  const int a = m / nn, bq = (m / n) % n, cq = m % n;
  int oc[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const bool valid = q < nch;
    const int ch = valid ? child_idx[c0 + q] : 0;
    oc[q] = valid ? octant[ch] : 0;
    if (m < n3) s0[q * tpb + m] = valid ? M[(size_t)ch * n3p + m] : (T)0;
  }
  __syncthreads();
  if (m < n3)
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (q < nch) {  // z: s1[a][b][cp] = sum_c mz[cp][c] s0[a][b][c]
        const T *mz = m2m1d + (oc[q] * 3 + 2) * nn;
        const T *x0 = s0 + q * tpb + (a * n + bq) * n;
        T t = 0;
        for (int c = 0; c < n; ++c) t += mz[cq * n + c] * x0[c];
        s1[q * tpb + m] = t;
      }
  __syncthreads();
  if (m < n3)
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (q < nch) {  // y: s0[a][bp][cp] = sum_b my[bp][b] s1[a][b][cp]
        const T *my = m2m1d + (oc[q] * 3 + 1) * nn;
        const T *x1 = s1 + q * tpb + a * nn + cq;
        T t = 0;
        for (int bb = 0; bb < n; ++bb) t += my[bq * n + bb] * x1[bb * n];
        s0[q * tpb + m] = t;
      }
  __syncthreads();
  T acc = 0;
  if (live)
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (q < nch) {  // x: acc[ap][bp][cp] += sum_a mx[ap][a] s0[a][bp][cp]
        const T *mx = m2m1d + (oc[q] * 3 + 0) * nn;
        const T *x2 = s0 + q * tpb + bq * n + cq;
        T t = 0;
        for (int aa = 0; aa < n; ++aa) t += mx[a * n + aa] * x2[aa * nn];
        acc += t;
      }
  if (live) store_mult(M, Mb, (size_t)b * n3p + m, acc);
}

// Single right-hand side M2M of one level: 256 / tpb parent boxes per
// work-group, one thread per (parent, node).
template <typename T, int NF = 0>  // NF > 0: the order fixed at compile time (loops unrolled)
__global__ __launch_bounds__(256) void k_m2m1(int n_rt, int n3p, int tpb, int box0, int nb,
                                              const int *__restrict__ child_off,
                                              const int *__restrict__ child_idx,
                                              const int *__restrict__ octant,
                                              const T *__restrict__ m2m1d, T *__restrict__ M,
                                              __bf16 *__restrict__ Mb,
                                              const unsigned char *__restrict__ skip) {
  extern __shared__ unsigned char smem_raw[];
  const int slot = threadIdx.x / tpb, m = threadIdx.x % tpb;
  const int bi = blockIdx.x * (256 / tpb) + slot;
  const int b = box0 + min(bi, nb - 1);
  // a skipped box loads no child and stores nothing, but its threads pass
  // every barrier with the others
  const bool on = bi < nb && !(skip && skip[b]);
  m2m1_box<T, NF>(n_rt, n3p, tpb, b, on, slot, m, smem_raw, child_off, child_idx, octant, m2m1d,
                  M, Mb);
}

// The whole single-RHS upward pass (M2M from the finest parents up to level
// 2) of one subtree per work-group: the descendants of level-2 box
// sub_root[r] at level l are the global boxes [rng[2 l], rng[2 l + 1]),
// rng = sub_rng + 2 (L + 1) r (contiguous: boxes are Morton-ordered within a
// problem's level, and a problem's boxes of one level are contiguous), and
// sub_leaf[r] is the level of the subtree's leaves.  Level by level, deepest
// parents first, 256 / tpb parents at a time, with a barrier after each
// batch (the slots' LDS is reused) and between levels (a level reads the
// multipoles the level below wrote).  Subtrees are independent, so one
// launch replaces a launch per level, most of which held a few hundred boxes.
// Per box the arithmetic of k_m2m1.
template <typename T, int NF = 0>
__global__ __launch_bounds__(256) void k_m2m1_sub(int n_rt, int n3p, int tpb, int L,
                                                  const int *__restrict__ sub_root,
                                                  const int *__restrict__ sub_leaf,
                                                  const int *__restrict__ sub_rng,
                                                  const int *__restrict__ child_off,
                                                  const int *__restrict__ child_idx,
                                                  const int *__restrict__ octant,
                                                  const T *__restrict__ m2m1d, T *__restrict__ M,
                                                  __bf16 *__restrict__ Mb,
                                                  const unsigned char *__restrict__ skip) {
  extern __shared__ unsigned char smem_raw[];
  const int r = blockIdx.x;
  if (skip && skip[sub_root[r]]) return;  // every box of a problem shares its flag
  const int slot = threadIdx.x / tpb, m = threadIdx.x % tpb, bpb = 256 / tpb;
  const int *rng = sub_rng + (size_t)r * 2 * (L + 1);
  for (int l = sub_leaf[r] - 1; l >= 2; --l) {
    const int lo = rng[2 * l], hi = rng[2 * l + 1];
    for (int b0 = lo; b0 < hi; b0 += bpb) {
      const int b = b0 + slot;
      m2m1_box<T, NF>(n_rt, n3p, tpb, b < hi ? b : lo, b < hi, slot, m, smem_raw, child_off,
                      child_idx, octant, m2m1d, M, Mb);
      __syncthreads();
    }
  }
}

// Single right-hand side L2L of child box ch by the tpb threads of one LDS
// slot (live = false: loads and stores nothing, passes the barriers): the
// transposed contractions x^T, y^T, z^T of k_l2l through the slot.
template <typename T, int NF>
__device__ __forceinline__ void l2l1_box(int n_rt, int tpb, int ch, bool on, int slot, int m,
                                         unsigned char *smem_raw,
                                         const int *__restrict__ parent,
                                         const int *__restrict__ octant,
                                         const T *__restrict__ m2m1d, T *__restrict__ Lc) {
  const int n = NF > 0 ? NF : n_rt;
  const int n3 = n * n * n, nn = n * n;
  T *s0 = reinterpret_cast<T *>(smem_raw) + slot * 2 * tpb, *s1 = s0 + tpb;
  const bool live = on && m < n3;
  const int p = parent[ch], o = octant[ch];
  const T *mx = m2m1d + (o * 3 + 0) * nn, *my = m2m1d + (o * 3 + 1) * nn,
          *mz = m2m1d + (o * 3 + 2) * nn;
  const int a = m / nn, bq = (m / n) % n, cq = m % n;
  const T own = live ? Lc[(size_t)ch * n3 + m] : (T)0;  // the child's M2L part, loaded early
  if (m < n3) s0[m] = on ? Lc[(size_t)p * n3 + m] : (T)0;
  __syncthreads();
  if (m < n3) {  // x^T: s1[ap][b][c] = sum_a mx[a][ap] s0[a][b][c]
    T t = 0;
    for (int aa = 0; aa < n; ++aa) t += mx[aa * n + a] * s0[aa * nn + bq * n + cq];
    s1[m] = t;
  }
  __syncthreads();
  if (m < n3) {  // y^T: s0[a][bp][c] = sum_b my[b][bp] s1[a][b][c]
    T t = 0;
    for (int bb = 0; bb < n; ++bb) t += my[bb * n + bq] * s1[(a * n + bb) * n + cq];
    s0[m] = t;
  }
  __syncthreads();
  if (live) {  // z^T: Lc[child][a][b][cp] += sum_c mz[c][cp] s0[a][b][c]
    T t = 0;
    for (int c = 0; c < n; ++c) t += mz[c * n + cq] * s0[(a * n + bq) * n + c];
    Lc[(size_t)ch * n3 + m] = own + t;
  }
}

// Single right-hand side L2L of one level: one thread per (child, node).
template <typename T, int NF = 0>  // NF > 0: the order fixed at compile time (loops unrolled)
__global__ __launch_bounds__(256) void k_l2l1(int n_rt, int tpb, int box0, int nb,
                                              const int *__restrict__ parent,
                                              const int *__restrict__ octant,
                                              const T *__restrict__ m2m1d, T *__restrict__ Lc,
                                              const unsigned char *__restrict__ skip) {
  extern __shared__ unsigned char smem_raw[];
  const int slot = threadIdx.x / tpb, m = threadIdx.x % tpb;
  const int bi = blockIdx.x * (256 / tpb) + slot;
  const int ch = box0 + min(bi, nb - 1);
  // a skipped child loads and stores nothing; its threads pass every barrier
  const bool on = bi < nb && !(skip && skip[ch]);
  l2l1_box<T, NF>(n_rt, tpb, ch, on, slot, m, smem_raw, parent, octant, m2m1d, Lc);
}

// The whole single-RHS downward pass (L2L from level 2 to the leaves) of one
// subtree per work-group, the subtrees and ranges of k_m2m1_sub: level by
// level from the children of the root down, a barrier after each batch and
// between levels (a level reads the locals the level above completed).  Per
// box the arithmetic of k_l2l1.
template <typename T, int NF = 0>
__global__ __launch_bounds__(256) void k_l2l1_sub(int n_rt, int tpb, int L,
                                                  const int *__restrict__ sub_root,
                                                  const int *__restrict__ sub_leaf,
                                                  const int *__restrict__ sub_rng,
                                                  const int *__restrict__ parent,
                                                  const int *__restrict__ octant,
                                                  const T *__restrict__ m2m1d, T *__restrict__ Lc,
                                                  const unsigned char *__restrict__ skip) {
  extern __shared__ unsigned char smem_raw[];
  const int r = blockIdx.x;
  if (skip && skip[sub_root[r]]) return;  // every box of a problem shares its flag
  const int slot = threadIdx.x / tpb, m = threadIdx.x % tpb, bpb = 256 / tpb;
  const int *rng = sub_rng + (size_t)r * 2 * (L + 1);
  for (int l = 3; l <= sub_leaf[r]; ++l) {
    const int lo = rng[2 * l], hi = rng[2 * l + 1];
    for (int b0 = lo; b0 < hi; b0 += bpb) {
      const int ch = b0 + slot;
      l2l1_box<T, NF>(n_rt, tpb, ch < hi ? ch : lo, ch < hi, slot, m, smem_raw, parent, octant,
                      m2m1d, Lc);
      __syncthreads();
    }
  }
}

// Split-BF16 products for the low-rank M2L: every projected multipole and
// every kr x kr table entry is a BF16 pair hi + lo (lo = the rounding
// residual of hi), and the matrix cores form hi*hi + hi*lo + lo*hi with FP32
// accumulation: products to ~2^-16 relative (the dropped lo*lo term and the
// pair's mantissa) instead of BF16's 2^-8, at three BF16 operations.  The
// projected basis has mixed signs and a wide range, where plain BF16 rounding
// slows the refinement.  Otherwise as k_m2l_bf16_pipe, with DEPTH pairs of
// prefetch: DEPTH = 2 or 4 keeps that many register sets, the source rows of
// pair q + DEPTH issued as soon as pair q's rows are in LDS and its G_off
// fragments right after pair q's products, so DEPTH times the bytes are in
// flight per work-group.
template <int MT, int RTM, int DEPTH = 1>
__global__ __launch_bounds__(256) void k_m2l_bf16x3(int n3, int nr, int ldm, int xrun,
                                                    const int *__restrict__ targets,
                                                    const int *__restrict__ il_off,
                                                    const int *__restrict__ il_src,
                                                    const int *__restrict__ il_oidx,
                                                    const int *__restrict__ box_level,
                                                    const double *__restrict__ inv_half,
                                                    const __bf16 *__restrict__ G,
                                                    const __bf16 *__restrict__ Gl,
                                                    const __bf16 *__restrict__ Mb,
                                                    const __bf16 *__restrict__ Mbl,
                                                    float *__restrict__ Lc) {
  constexpr int NP = 16 * MT, KS = NP / 32, MTW = (MT + 3) / 4, LDA = NP + 8, VPR = NP / 8;
  constexpr int APT = (RTM * 16 * VPR + 255) / 256;
  __shared__ __attribute__((aligned(16))) __bf16 sA[RTM * 16 * LDA], sAl[RTM * 16 * LDA];
  const int t = targets[xcd_block(blockIdx.x, gridDim.x, xrun)];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int li = lane & 15, lk = (lane >> 4) * 8;
  const int rt0 = blockIdx.y * RTM;
  const int RT = min((nr + 15) >> 4, rt0 + RTM) - rt0;
  const int AV = RT * 16 * VPR;
  const size_t MS = (size_t)ldm * NP;
  f32x4 acc[RTM][MTW];
#pragma unroll
  for (int i = 0; i < RTM; ++i)
#pragma unroll
    for (int j = 0; j < MTW; ++j) acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
  const int q0 = il_off[t], q1 = il_off[t + 1];
  if constexpr (DEPTH == 1) {
    bf16x8 ah[APT], al[APT];
    bf16x8 bh[MTW][KS], bl[MTW][KS], nh[MTW][KS], nl[MTW][KS];
    auto load_a = [&](int q) {
      const size_t off = (size_t)il_src[q] * MS + (size_t)rt0 * 16 * NP;
      const bf16x8 *sh = reinterpret_cast<const bf16x8 *>(Mb + off);
      const bf16x8 *sl = reinterpret_cast<const bf16x8 *>(Mbl + off);
  #pragma unroll
      for (int i = 0; i < APT; ++i) {
        const int v = tid + 256 * i;
        if (v < AV) {
          ah[i] = sh[v];
          al[i] = sl[v];
        }
      }
    };
    auto load_b = [&](int q, bf16x8 (&dh)[MTW][KS], bf16x8 (&dl)[MTW][KS]) {
      const size_t off = (size_t)il_oidx[q] * NP * NP;
  #pragma unroll
      for (int j = 0; j < MTW; ++j) {
        const int mt = wave * MTW + j;
        if (mt < MT)
  #pragma unroll
          for (int k = 0; k < KS; ++k) {
            const size_t e = off + (size_t)(mt * 16 + li) * NP + k * 32 + lk;
            dh[j][k] = *(const bf16x8 *)(G + e);
            dl[j][k] = *(const bf16x8 *)(Gl + e);
          }
      }
    };
    if (q0 < q1) {
      load_a(q0);
      load_b(q0, bh, bl);
    }
    for (int q = q0; q < q1; ++q) {
      __syncthreads();
  #pragma unroll
      for (int i = 0; i < APT; ++i) {
        const int v = tid + 256 * i;
        if (v < AV) {
          const int o = (v / VPR) * LDA + (v % VPR) * 8;
          *(bf16x8 *)(sA + o) = ah[i];
          *(bf16x8 *)(sAl + o) = al[i];
        }
      }
      __syncthreads();
      if (q + 1 < q1) {
        load_a(q + 1);
        load_b(q + 1, nh, nl);
      }
  #pragma unroll
      for (int k = 0; k < KS; ++k) {
  #pragma unroll
        for (int rt = 0; rt < RTM; ++rt) {
          if (rt < RT) {
            const int o = (rt * 16 + li) * LDA + k * 32 + lk;
            const bf16x8 a = *(const bf16x8 *)(sA + o), a2 = *(const bf16x8 *)(sAl + o);
  #pragma unroll
            for (int j = 0; j < MTW; ++j)
              if (wave * MTW + j < MT) {
                f32x4 c = acc[rt][j];
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2, bh[j][k], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bl[j][k], c, 0, 0, 0);
                acc[rt][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bh[j][k], c, 0, 0, 0);
              }
          }
        }
      }
  #pragma unroll
      for (int j = 0; j < MTW; ++j)
  #pragma unroll
        for (int k = 0; k < KS; ++k) {
          bh[j][k] = nh[j][k];
          bl[j][k] = nl[j][k];
        }
    }
  } else {
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using I3 = std::integral_constant<int, 3>;
    bf16x8 ah[DEPTH][APT], al[DEPTH][APT];
    bf16x8 bh[DEPTH][MTW][KS], bl[DEPTH][MTW][KS];
    auto load_a = [&](int q, auto bc) {
      constexpr int b = decltype(bc)::value;
      const size_t off = (size_t)il_src[q] * MS + (size_t)rt0 * 16 * NP;
      const bf16x8 *sh = reinterpret_cast<const bf16x8 *>(Mb + off);
      const bf16x8 *sl = reinterpret_cast<const bf16x8 *>(Mbl + off);
#pragma unroll
      for (int i = 0; i < APT; ++i) {
        const int v = tid + 256 * i;
        if (v < AV) {
          ah[b][i] = sh[v];
          al[b][i] = sl[v];
        }
      }
    };
    auto load_b = [&](int q, auto bc) {
      constexpr int b = decltype(bc)::value;
      const size_t off = (size_t)il_oidx[q] * NP * NP;
#pragma unroll
      for (int j = 0; j < MTW; ++j) {
        const int mt = wave * MTW + j;
        if (mt < MT)
#pragma unroll
          for (int k = 0; k < KS; ++k) {
            const size_t e = off + (size_t)(mt * 16 + li) * NP + k * 32 + lk;
            bh[b][j][k] = *(const bf16x8 *)(G + e);
            bl[b][j][k] = *(const bf16x8 *)(Gl + e);
          }
      }
    };
    auto step = [&](int q, auto bc) {
      constexpr int b = decltype(bc)::value;
      __syncthreads();
#pragma unroll
      for (int i = 0; i < APT; ++i) {
        const int v = tid + 256 * i;
        if (v < AV) {
          const int o = (v / VPR) * LDA + (v % VPR) * 8;
          *(bf16x8 *)(sA + o) = ah[b][i];
          *(bf16x8 *)(sAl + o) = al[b][i];
        }
      }
      __syncthreads();
      if (q + DEPTH < q1) load_a(q + DEPTH, bc);
#pragma unroll
      for (int k = 0; k < KS; ++k) {
#pragma unroll
        for (int rt = 0; rt < RTM; ++rt) {
          if (rt < RT) {
            const int o = (rt * 16 + li) * LDA + k * 32 + lk;
            const bf16x8 a = *(const bf16x8 *)(sA + o), a2 = *(const bf16x8 *)(sAl + o);
#pragma unroll
            for (int j = 0; j < MTW; ++j)
              if (wave * MTW + j < MT) {
                f32x4 c = acc[rt][j];
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2, bh[b][j][k], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bl[b][j][k], c, 0, 0, 0);
                acc[rt][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bh[b][j][k], c, 0, 0, 0);
              }
          }
        }
      }
      if (q + DEPTH < q1) load_b(q + DEPTH, bc);
    };
    auto prime = [&](int q, auto bc) {
      if (q < q1) {
        load_a(q, bc);
        load_b(q, bc);
      }
    };
    prime(q0, I0{});
    prime(q0 + 1, I1{});
    if constexpr (DEPTH == 4) {
      prime(q0 + 2, I2{});
      prime(q0 + 3, I3{});
    }
    int q = q0;
    for (; q + DEPTH - 1 < q1; q += DEPTH) {
      step(q, I0{});
      step(q + 1, I1{});
      if constexpr (DEPTH == 4) {
        step(q + 2, I2{});
        step(q + 3, I3{});
      }
    }
    // the last pairs, in the register sets they were loaded into
    if (q < q1) step(q, I0{});
    if (q + 1 < q1) step(q + 1, I1{});
    if constexpr (DEPTH == 4)
      if (q + 2 < q1) step(q + 2, I2{});
  }
  const float ih = (float)inv_half[box_level[t]];
  float *out = Lc + (size_t)t * n3 * ldm;
#pragma unroll
  for (int rt = 0; rt < RTM; ++rt)
#pragma unroll
    for (int j = 0; j < MTW; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = (rt0 + rt) * 16 + (lane >> 4) * 4 + v, m = (wave * MTW + j) * 16 + li;
        if (rt < RT && r < nr && m < n3) out[(size_t)m * ldm + r] = acc[rt][j][v] * ih;
      }
}

// Sibling pairs: one work-group per (pair of sibling targets, RTM RHS tiles
// from tile rt0 = RTM blockIdx.y).  The pair's entries are the union of the
// two interaction lists ordered by source box, each with the source's offset
// for either target (-1: not in that list).  Every source block is staged in
// LDS once and feeds the matrix-core products of both targets (two
// accumulator sets), so a source shared by the siblings is read once: the
// per-target kernel reads a source block from HBM for every pair.  The next
// entry's source rows and both G_off fragment sets are prefetched into
// registers during the current products.
template <int MT, int RTM>
__global__ __launch_bounds__(256) void k_m2l_bf16_pair(int n3, int nr, int ldm, int xrun,
                                                       const int *__restrict__ pair_tgt,
                                                       const int *__restrict__ pair_off,
                                                       const int *__restrict__ pe_src,
                                                       const int *__restrict__ pe_o1,
                                                       const int *__restrict__ pe_o2,
                                                       const int *__restrict__ box_level,
                                                       const double *__restrict__ inv_half,
                                                       const __bf16 *__restrict__ G,
                                                       const __bf16 *__restrict__ Mb,
                                                       float *__restrict__ Lc) {
  constexpr int NP = 16 * MT, KS = NP / 32, MTW = (MT + 3) / 4, LDA = NP + 8, VPR = NP / 8;
  constexpr int APT = (RTM * 16 * VPR + 255) / 256;
  __shared__ __attribute__((aligned(16))) __bf16 sA[RTM * 16 * LDA];
  const int grp = xcd_block(blockIdx.x, gridDim.x, xrun);
  const int rt0 = blockIdx.y * RTM;
  const int RT = min((nr + 15) >> 4, rt0 + RTM) - rt0;
  const int t1 = pair_tgt[2 * grp], t2 = pair_tgt[2 * grp + 1];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int li = lane & 15, lk = (lane >> 4) * 8;
  const int AV = RT * 16 * VPR;
  const size_t MS = (size_t)ldm * NP;
  f32x4 acc1[RTM][MTW], acc2[RTM][MTW];
#pragma unroll
  for (int i = 0; i < RTM; ++i)
#pragma unroll
    for (int j = 0; j < MTW; ++j) acc1[i][j] = acc2[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
  const int q0 = pair_off[grp], q1 = pair_off[grp + 1];
  bf16x8 apre[APT];
  bf16x8 b1[MTW][KS], b2[MTW][KS], n1[MTW][KS], n2[MTW][KS];
  auto load_a = [&](int q) {
    const bf16x8 *src =
        reinterpret_cast<const bf16x8 *>(Mb + (size_t)pe_src[q] * MS + (size_t)rt0 * 16 * NP);
#pragma unroll
    for (int i = 0; i < APT; ++i) {
      const int v = tid + 256 * i;
      if (v < AV) apre[i] = src[v];
    }
  };
  auto load_b = [&](int o, bf16x8 (&dst)[MTW][KS]) {
    if (o < 0) return;
    const __bf16 *Go = G + (size_t)o * NP * NP;
#pragma unroll
    for (int j = 0; j < MTW; ++j) {
      const int mt = wave * MTW + j;
      if (mt < MT)
#pragma unroll
        for (int k = 0; k < KS; ++k)
          dst[j][k] = *(const bf16x8 *)(Go + (size_t)(mt * 16 + li) * NP + k * 32 + lk);
    }
  };
  int o1 = -1, o2 = -1;
  if (q0 < q1) {
    load_a(q0);
    o1 = pe_o1[q0];
    o2 = pe_o2[q0];
    load_b(o1, b1);
    load_b(o2, b2);
  }
  for (int q = q0; q < q1; ++q) {
    __syncthreads();  // previous entry's LDS reads are done
#pragma unroll
    for (int i = 0; i < APT; ++i) {
      const int v = tid + 256 * i;
      if (v < AV) *(bf16x8 *)(sA + (v / VPR) * LDA + (v % VPR) * 8) = apre[i];
    }
    __syncthreads();
    int p1 = -1, p2 = -1;
    if (q + 1 < q1) {
      load_a(q + 1);
      p1 = pe_o1[q + 1];
      p2 = pe_o2[q + 1];
      load_b(p1, n1);
      load_b(p2, n2);
    }
#pragma unroll
    for (int k = 0; k < KS; ++k) {
#pragma unroll
      for (int rt = 0; rt < RTM; ++rt) {
        if (rt < RT) {
          const bf16x8 a = *(const bf16x8 *)(sA + (rt * 16 + li) * LDA + k * 32 + lk);
#pragma unroll
          for (int j = 0; j < MTW; ++j)
            if (wave * MTW + j < MT) {
              if (o1 >= 0)
                acc1[rt][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b1[j][k], acc1[rt][j], 0, 0, 0);
              if (o2 >= 0)
                acc2[rt][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b2[j][k], acc2[rt][j], 0, 0, 0);
            }
        }
      }
    }
    o1 = p1;
    o2 = p2;
#pragma unroll
    for (int j = 0; j < MTW; ++j)
#pragma unroll
      for (int k = 0; k < KS; ++k) {
        b1[j][k] = n1[j][k];
        b2[j][k] = n2[j][k];
      }
  }
#pragma unroll
  for (int s2 = 0; s2 < 2; ++s2) {
    const int t = s2 ? t2 : t1;
    if (t < 0) continue;
    const float ih = (float)inv_half[box_level[t]];
    float *out = Lc + (size_t)t * n3 * ldm;
#pragma unroll
    for (int rt = 0; rt < RTM; ++rt)
#pragma unroll
      for (int j = 0; j < MTW; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int r = (rt0 + rt) * 16 + (lane >> 4) * 4 + v, m = (wave * MTW + j) * 16 + li;
          if (rt < RT && r < nr && m < n3)
            out[(size_t)m * ldm + r] = (s2 ? acc2[rt][j][v] : acc1[rt][j][v]) * ih;
        }
  }
}

// Single right-hand side (ldm = 1), target tiles: one wavefront per tile of
// 16 targets of one octant, whose interaction lists share most offsets.  Per
// entry (an offset o of the tile's union), D (node x target) += G_o (node x l)
// * [M_s1 .. M_s16] (l x target) on the matrix cores, zero columns for
// targets without that offset; the next entry's source multipoles (and, when
// they fit in registers, its G_o fragments) load during the current products.
// Each target's local expansion is complete at the end and written once,
// scaled by 1/half of its level: no read-modify-write of the locals.
template <int MT>
__global__ __launch_bounds__(256) void k_m2l1t_bf16(int n3, int n_tiles,
                                                    const int *__restrict__ tile_tgt,
                                                    const int *__restrict__ tile_eoff,
                                                    const int *__restrict__ ent_oidx,
                                                    const int *__restrict__ ent_src,
                                                    const int *__restrict__ box_level,
                                                    const double *__restrict__ inv_half,
                                                    const __bf16 *__restrict__ G,
                                                    const __bf16 *__restrict__ Mb,
                                                    float *__restrict__ Lc,
                                                    const unsigned char *__restrict__ skip) {
  constexpr int NP = 16 * MT, KS = NP / 32;
  constexpr bool APRE = MT * KS <= 8;  // G_o fragments of the next entry in registers
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tile = blockIdx.x * 4 + wave;
  if (tile >= n_tiles) return;
  const int li = lane & 15, lk = (lane >> 4) * 8;
  // a skipped target column reads no source and is not written
  const int t = tile_tgt[tile * 16 + li];
  const bool off = t < 0 || (skip && skip[t]);
  if (__all(off)) return;  // wavefront-uniform
  const int e0 = tile_eoff[tile], e1 = tile_eoff[tile + 1];
  f32x4 acc[MT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) acc[mt] = (f32x4){0.f, 0.f, 0.f, 0.f};
  const bf16x8 zero = {};
  bf16x8 b[KS], bn[KS];
  bf16x8 a[APRE ? MT : 1][APRE ? KS : 1], an[APRE ? MT : 1][APRE ? KS : 1];
  auto load_b = [&](int e, bf16x8 (&B)[KS]) {
    const int sb = off ? -1 : ent_src[(size_t)e * 16 + li];
    const __bf16 *Ms = Mb + (size_t)(sb < 0 ? 0 : sb) * NP;
#pragma unroll
    for (int k = 0; k < KS; ++k) B[k] = sb >= 0 ? *(const bf16x8 *)(Ms + k * 32 + lk) : zero;
  };
  auto load_a = [&](int e, bf16x8 (&A)[APRE ? MT : 1][APRE ? KS : 1]) {
    const __bf16 *Go = G + (size_t)ent_oidx[e] * NP * NP;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
      for (int k = 0; k < KS; ++k)
        if constexpr (APRE) A[mt][k] = *(const bf16x8 *)(Go + (size_t)(mt * 16 + li) * NP + k * 32 + lk);
  };
  if (e0 < e1) {
    load_b(e0, b);
    if constexpr (APRE) load_a(e0, a);
  }
  for (int e = e0; e < e1; ++e) {
    const __bf16 *Go = G + (size_t)ent_oidx[e] * NP * NP;
    if (e + 1 < e1) {
      load_b(e + 1, bn);
      if constexpr (APRE) load_a(e + 1, an);
    }
#pragma unroll
    for (int k = 0; k < KS; ++k)
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) {
        bf16x8 g;
        if constexpr (APRE) g = a[mt][k];
        else g = *(const bf16x8 *)(Go + (size_t)(mt * 16 + li) * NP + k * 32 + lk);
        acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(g, b[k], acc[mt], 0, 0, 0);
      }
    if (e + 1 < e1) {
#pragma unroll
      for (int k = 0; k < KS; ++k) b[k] = bn[k];
      if constexpr (APRE) {
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
          for (int k = 0; k < KS; ++k) a[mt][k] = an[mt][k];
      }
    }
  }
  // D[i = node][j = target]: lane holds nodes 4 (lane / 16) + v of target lane % 16
  if (off) return;
  const float ih = (float)inv_half[box_level[t]];
  float *out = Lc + (size_t)t * n3;
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) {
    const int m0 = mt * 16 + (lane >> 4) * 4;
    if ((n3 & 3) == 0) {
      if (m0 < n3)
        *(f32x4 *)(out + m0) = (f32x4){acc[mt][0] * ih, acc[mt][1] * ih, acc[mt][2] * ih,
                                       acc[mt][3] * ih};
    } else {
#pragma unroll
      for (int v = 0; v < 4; ++v)
        if (m0 + v < n3) out[m0 + v] = acc[mt][v] * ih;
    }
  }
}

// The same over tiles of 64 targets (one per wavefront quarter): the four
// wavefronts walk one entry list, G_o is staged once per entry in LDS (double
// buffered, rows padded by 8 BF16) by the whole work-group, each wavefront
// streams the source multipoles of its 16 targets.  One barrier per entry.
template <int MT>
__global__ __launch_bounds__(256) void k_m2l1t64_bf16(int n3, int n_tiles, int xrun,
                                                      const int *__restrict__ tile_tgt,
                                                      const int *__restrict__ tile_eoff,
                                                      const int *__restrict__ ent_oidx,
                                                      const int *__restrict__ ent_src,
                                                      const int *__restrict__ box_level,
                                                      const double *__restrict__ inv_half,
                                                      const __bf16 *__restrict__ G,
                                                      const __bf16 *__restrict__ Mb,
                                                      float *__restrict__ Lc,
                                                      const unsigned char *__restrict__ skip) {
  constexpr int NP = 16 * MT, KS = NP / 32, LDG = NP + 8;
  constexpr int VPR = NP / 8;                  // 16-byte vectors per G_o row
  constexpr int GPT = (NP * VPR + 255) / 256;  // G_o vectors per thread
  __shared__ __attribute__((aligned(16))) __bf16 sG[2][NP * LDG];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int tile = xcd_block(blockIdx.x, gridDim.x, xrun);
  const int li = lane & 15, lk = (lane >> 4) * 8, col = wave * 16 + li;
  const int e0 = tile_eoff[tile], e1 = tile_eoff[tile + 1];
  // a skipped target column reads no source and is not written; a tile
  // whose 64 columns are all skipped leaves at once (work-group-uniform,
  // every thread votes before the first barrier)
  const int t = tile_tgt[tile * 64 + col];
  const bool off = t < 0 || (skip && skip[t]);
  if (!__syncthreads_or(!off)) return;
  f32x4 acc[MT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) acc[mt] = (f32x4){0.f, 0.f, 0.f, 0.f};
  const bf16x8 zero = {};
  bf16x8 b[KS], bn[KS], gp[GPT];
  auto load_b = [&](int e, bf16x8 (&B)[KS]) {
    const int sb = off ? -1 : ent_src[(size_t)e * 64 + col];
    const __bf16 *Ms = Mb + (size_t)(sb < 0 ? 0 : sb) * NP;
#pragma unroll
    for (int k = 0; k < KS; ++k) B[k] = sb >= 0 ? *(const bf16x8 *)(Ms + k * 32 + lk) : zero;
  };
  auto load_g = [&](int e) {
    const bf16x8 *Go = reinterpret_cast<const bf16x8 *>(G + (size_t)ent_oidx[e] * NP * NP);
#pragma unroll
    for (int i = 0; i < GPT; ++i) {
      const int v = tid + 256 * i;
      if (v < NP * VPR) gp[i] = Go[v];
    }
  };
  auto store_g = [&](int buf) {
#pragma unroll
    for (int i = 0; i < GPT; ++i) {
      const int v = tid + 256 * i;
      if (v < NP * VPR) *(bf16x8 *)(&sG[buf][(v / VPR) * LDG + (v % VPR) * 8]) = gp[i];
    }
  };
  if (e0 < e1) {
    load_g(e0);
    load_b(e0, b);
    store_g(0);
  }
  __syncthreads();
  for (int e = e0; e < e1; ++e) {
    const int buf = (e - e0) & 1;
    const bool more = e + 1 < e1;
    if (more) {
      load_g(e + 1);
      load_b(e + 1, bn);
    }
    const __bf16 *sg = sG[buf];
#pragma unroll
    for (int k = 0; k < KS; ++k)
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) {
        const bf16x8 g = *(const bf16x8 *)(sg + (mt * 16 + li) * LDG + k * 32 + lk);
        acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(g, b[k], acc[mt], 0, 0, 0);
      }
    if (more) {
      store_g(buf ^ 1);
#pragma unroll
      for (int k = 0; k < KS; ++k) b[k] = bn[k];
    }
    __syncthreads();
  }
  if (off) return;
  const float ih = (float)inv_half[box_level[t]];
  float *out = Lc + (size_t)t * n3;
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) {
    const int m0 = mt * 16 + (lane >> 4) * 4;
    if ((n3 & 3) == 0) {
      if (m0 < n3)
        *(f32x4 *)(out + m0) = (f32x4){acc[mt][0] * ih, acc[mt][1] * ih, acc[mt][2] * ih,
                                       acc[mt][3] * ih};
    } else {
#pragma unroll
      for (int v = 0; v < 4; ++v)
        if (m0 + v < n3) out[m0 + v] = acc[mt][v] * ih;
    }
  }
}

// FP64 form of the 64-target tiles (orders <= 4: G_o of 64 x 64 doubles, two
// LDS buffers of 33 KB): 16x16x4 FP64 matrix-core products, G_o rows padded
// by 2 doubles in LDS.
template <int MT>
__global__ __launch_bounds__(256) void k_m2l1t64_f64(int n3, int n_tiles, int xrun,
                                                     const int *__restrict__ tile_tgt,
                                                     const int *__restrict__ tile_eoff,
                                                     const int *__restrict__ ent_oidx,
                                                     const int *__restrict__ ent_src,
                                                     const int *__restrict__ box_level,
                                                     const double *__restrict__ inv_half,
                                                     const double *__restrict__ G,
                                                     const double *__restrict__ M,
                                                     double *__restrict__ Lc,
                                                     const unsigned char *__restrict__ skip) {
  constexpr int NP = 16 * MT, K4 = NP / 4, LDG = NP + 2;
  constexpr int VPR = NP / 2;                  // double2 vectors per G_o row
  constexpr int GPT = (NP * VPR + 255) / 256;  // G_o vectors per thread
  typedef double f64x2 __attribute__((ext_vector_type(2)));
  __shared__ __attribute__((aligned(16))) double sG[2][NP * LDG];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int tile = xcd_block(blockIdx.x, gridDim.x, xrun);
  const int li = lane & 15, lq = lane >> 4, col = wave * 16 + li;
  const int e0 = tile_eoff[tile], e1 = tile_eoff[tile + 1];
  // skipped target columns as in k_m2l1t64_bf16
  const int t = tile_tgt[tile * 64 + col];
  const bool off = t < 0 || (skip && skip[t]);
  if (!__syncthreads_or(!off)) return;
  f64x4 acc[MT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) acc[mt] = (f64x4){0.0, 0.0, 0.0, 0.0};
  double b[K4], bn[K4];
  f64x2 gp[GPT];
  auto load_b = [&](int e, double (&B)[K4]) {
    const int sb = off ? -1 : ent_src[(size_t)e * 64 + col];
    const double *Ms = M + (size_t)(sb < 0 ? 0 : sb) * NP;
#pragma unroll
    for (int k = 0; k < K4; ++k) B[k] = sb >= 0 ? Ms[4 * k + lq] : 0.0;
  };
  auto load_g = [&](int e) {
    const f64x2 *Go = reinterpret_cast<const f64x2 *>(G + (size_t)ent_oidx[e] * NP * NP);
#pragma unroll
    for (int i = 0; i < GPT; ++i) {
      const int v = tid + 256 * i;
      if (v < NP * VPR) gp[i] = Go[v];
    }
  };
  auto store_g = [&](int buf) {
#pragma unroll
    for (int i = 0; i < GPT; ++i) {
      const int v = tid + 256 * i;
      if (v < NP * VPR) *(f64x2 *)(&sG[buf][(v / VPR) * LDG + (v % VPR) * 2]) = gp[i];
    }
  };
  if (e0 < e1) {
    load_g(e0);
    load_b(e0, b);
    store_g(0);
  }
  __syncthreads();
  for (int e = e0; e < e1; ++e) {
    const int buf = (e - e0) & 1;
    const bool more = e + 1 < e1;
    if (more) {
      load_g(e + 1);
      load_b(e + 1, bn);
    }
    const double *sg = sG[buf];
#pragma unroll
    for (int k = 0; k < K4; ++k)
#pragma unroll
      for (int mt = 0; mt < MT; ++mt)
        acc[mt] = __builtin_amdgcn_mfma_f64_16x16x4f64(sg[(mt * 16 + li) * LDG + 4 * k + lq], b[k],
                                                       acc[mt], 0, 0, 0);
    if (more) {
      store_g(buf ^ 1);
#pragma unroll
      for (int k = 0; k < K4; ++k) b[k] = bn[k];
    }
    __syncthreads();
  }
  // D[i][j]: lane holds nodes lq + 4 v of target column li
  if (off) return;
  const double ih = inv_half[box_level[t]];
  double *out = Lc + (size_t)t * n3;
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int m = mt * 16 + lq + 4 * v;
      if (m < n3) out[m] = acc[mt][v] * ih;
    }
}


// The 64-target tiles of the FP64 operator with truncated M2L blocks,
// applied in factored form as the reference applies them (y += U (V^T x),
// bbfmm.rs:132-174): per entry (offset o of rank k), T = V^T_o [M_s ...]
// (k x 16 per wavefront, in ceil(k / 16) blocks of 16 rows, 16 K-steps of 4
// over the n^3 nodes) and then acc += U_o T over ceil(k / 4) K-steps of 4:
// at k = 8 that is 24 matrix-core products of 16x16x4 a wavefront instead of
// the 64 of the dense block.  The D layout of the first product (row 4 r + q
// of a block in lane group q, register r) is the B layout of the second
// (K-step r, row q), so T never leaves its registers.  The nodes are summed
// in the order lane group q takes nodes q K4 .. q K4 + K4 - 1, so each lane
// reads contiguous runs of V^T_o and of its source's multipole; U_o rows are
// stored with the rank index split in four (M2lFactors::Up) for the same
// reason.  Wavefronts are independent (no LDS, no barrier): the factors come
// through the caches, and a wavefront whose 16 targets are all skipped
// leaves at once.
template <int MT, int KB>
__global__ __launch_bounds__(256) void k_m2l1t64_f64lr(int n3, int n_tiles, int xrun,
                                                       const int *__restrict__ tile_tgt,
                                                       const int *__restrict__ tile_eoff,
                                                       const int *__restrict__ ent_oidx,
                                                       const int *__restrict__ ent_src,
                                                       const int *__restrict__ box_level,
                                                       const double *__restrict__ inv_half,
                                                       const double *__restrict__ Vt,
                                                       const double *__restrict__ Up,
                                                       const int *__restrict__ rank, int kp,
                                                       const double *__restrict__ M,
                                                       double *__restrict__ Lc,
                                                       const unsigned char *__restrict__ skip) {
  constexpr int NP = 16 * MT, K4 = NP / 4;
  typedef double f64x2 __attribute__((ext_vector_type(2)));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tile = xcd_block(blockIdx.x, gridDim.x, xrun);
  const int li = lane & 15, lq = lane >> 4, col = wave * 16 + li;
  const int e0 = tile_eoff[tile], e1 = tile_eoff[tile + 1];
  const int t = tile_tgt[tile * 64 + col];
  const bool off = t < 0 || (skip && skip[t]);
  if (__all(off)) return;  // wavefront-uniform; no barrier follows
  const int k4 = kp / 4;
  f64x4 acc[MT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) acc[mt] = (f64x4){0.0, 0.0, 0.0, 0.0};
  double b[K4], bn[K4];
  // nodes lq K4 .. lq K4 + K4 - 1 of the target's source (zeros: none)
  auto load_b = [&](int e, double (&B)[K4]) {
    const int sb = off ? -1 : ent_src[(size_t)e * 64 + col];
    const f64x2 *Ms = reinterpret_cast<const f64x2 *>(M + (size_t)(sb < 0 ? 0 : sb) * NP + lq * K4);
#pragma unroll
    for (int c = 0; c < K4 / 2; ++c) {
      const f64x2 v = sb >= 0 ? Ms[c] : (f64x2){0.0, 0.0};
      B[2 * c] = v[0];
      B[2 * c + 1] = v[1];
    }
  };
  if (e0 < e1) load_b(e0, b);
  for (int e = e0; e < e1; ++e) {
    if (e + 1 < e1) load_b(e + 1, bn);
    const int o = ent_oidx[e], k = rank[o];
    f64x4 tt[KB];
#pragma unroll
    for (int tb = 0; tb < KB; ++tb) {
      tt[tb] = (f64x4){0.0, 0.0, 0.0, 0.0};
      if (tb * 16 < k) {  // wavefront-uniform
        const f64x2 *va =
            reinterpret_cast<const f64x2 *>(Vt + ((size_t)o * kp + tb * 16 + li) * NP + lq * K4);
#pragma unroll
        for (int c = 0; c < K4 / 2; ++c) {
          const f64x2 a = va[c];
          tt[tb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[0], b[2 * c], tt[tb], 0, 0, 0);
          tt[tb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[1], b[2 * c + 1], tt[tb], 0, 0, 0);
        }
      }
    }
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      const double *ua = Up + (((size_t)o * NP + mt * 16 + li) * 4 + lq) * k4;
#pragma unroll
      for (int st = 0; st < 4 * KB; ++st)
        if (st * 4 < k)  // wavefront-uniform
          acc[mt] = __builtin_amdgcn_mfma_f64_16x16x4f64(ua[st], tt[st / 4][st % 4], acc[mt], 0,
                                                         0, 0);
    }
    if (e + 1 < e1) {
#pragma unroll
      for (int c = 0; c < K4; ++c) b[c] = bn[c];
    }
  }
  // D[i][j]: lane holds nodes lq + 4 v of target column li
  if (off) return;
  const double ih = inv_half[box_level[t]];
  double *out = Lc + (size_t)t * n3;
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int m = mt * 16 + lq + 4 * v;
      if (m < n3) out[m] = acc[mt][v] * ih;
    }
}

// The factored FP64 tiles with the factors staged in LDS by the work-group,
// as k_m2l1t64_f64 stages G_o: per entry only the k rows of V^T_o and the k
// columns of U_o (2 k n3 doubles: 17 KB at order 4, k <= 17, against 32 KB
// for G_o), loaded into registers one entry ahead and stored into the other
// of two LDS buffers (dynamic LDS of 2 (kr (n3p + 2) + n3p (kr + 1))
// doubles, kr the largest rank), one barrier per entry.  The products and
// their order are those of k_m2l1t64_f64lr; rows and columns beyond the
// entry's rank read as zero.  PB: the next entry's source multipoles are
// loaded before this entry's products (a second register set); otherwise
// into the same registers once the first product has read them, so the
// loads overlap the second product (fewer registers: three waves a SIMD at
// order 4).
template <int MT, int KRB, bool PB>  // KRB: a bound on the largest rank (kr <= KRB)
__global__ __launch_bounds__(256) void k_m2l1t64_f64lrs(int n3, int n_tiles, int xrun,
                                                        const int *__restrict__ tile_tgt,
                                                        const int *__restrict__ tile_eoff,
                                                        const int *__restrict__ ent_oidx,
                                                        const int *__restrict__ ent_src,
                                                        const int *__restrict__ box_level,
                                                        const double *__restrict__ inv_half,
                                                        const double *__restrict__ Vt,
                                                        const double *__restrict__ U,
                                                        const int *__restrict__ rank, int kp,
                                                        int kr, const double *__restrict__ M,
                                                        double *__restrict__ Lc,
                                                        const unsigned char *__restrict__ skip) {
  constexpr int NP = 16 * MT, K4 = NP / 4, LDV = NP + 2, KB = (KRB + 15) / 16;
  constexpr int GPT = (2 * KRB * NP + 255) / 256;  // staged doubles per thread, at most
  typedef double f64x2 __attribute__((ext_vector_type(2)));
  extern __shared__ double lds_f[];
  const int LDU = kr + 1, bufsz = kr * LDV + NP * LDU;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int tile = xcd_block(blockIdx.x, gridDim.x, xrun);
  const int li = lane & 15, lq = lane >> 4, col = wave * 16 + li;
  const int e0 = tile_eoff[tile], e1 = tile_eoff[tile + 1];
  const int t = tile_tgt[tile * 64 + col];
  const bool off = t < 0 || (skip && skip[t]);
  if (!__syncthreads_or(!off)) return;
  f64x4 acc[MT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) acc[mt] = (f64x4){0.0, 0.0, 0.0, 0.0};
  double b[K4], bn[PB ? K4 : 1], gp[GPT];
  auto load_b = [&](int e, double *B) {
    const int sb = off ? -1 : ent_src[(size_t)e * 64 + col];
    const f64x2 *Ms = reinterpret_cast<const f64x2 *>(M + (size_t)(sb < 0 ? 0 : sb) * NP + lq * K4);
#pragma unroll
    for (int c = 0; c < K4 / 2; ++c) {
      const f64x2 v = sb >= 0 ? Ms[c] : (f64x2){0.0, 0.0};
      B[2 * c] = v[0];
      B[2 * c + 1] = v[1];
    }
  };
  // V^T_o rows 0 .. k-1 (k NP contiguous doubles), then U_o columns 0 .. k-1
  auto load_g = [&](int o, int k) {
    const double *Vo = Vt + (size_t)o * kp * NP, *Uo = U + (size_t)o * NP * kp;
    const int nv = k * NP;
#pragma unroll
    for (int i = 0; i < GPT; ++i) {
      const int v = tid + 256 * i;
      if (v < nv) gp[i] = Vo[v];
      else if (v < 2 * nv) gp[i] = Uo[(size_t)((v - nv) / k) * kp + (v - nv) % k];
    }
  };
  auto store_g = [&](int buf, int k) {
    double *sV = lds_f + (size_t)buf * bufsz, *sU = sV + kr * LDV;
    const int nv = k * NP;
#pragma unroll
    for (int i = 0; i < GPT; ++i) {
      const int v = tid + 256 * i;
      if (v < nv) sV[(v / NP) * LDV + v % NP] = gp[i];
      else if (v < 2 * nv) sU[((v - nv) / k) * LDU + (v - nv) % k] = gp[i];
    }
  };
  int kc = 0;
  if (e0 < e1) {
    kc = rank[ent_oidx[e0]];
    load_g(ent_oidx[e0], kc);
    load_b(e0, b);
    store_g(0, kc);
  }
  __syncthreads();
  for (int e = e0; e < e1; ++e) {
    const int buf = (e - e0) & 1;
    const bool more = e + 1 < e1;
    int kn = 0;
    if (more) {
      const int on = ent_oidx[e + 1];
      kn = rank[on];
      load_g(on, kn);
      if constexpr (PB) load_b(e + 1, bn);
    }
    const double *sV = lds_f + (size_t)buf * bufsz, *sU = sV + kr * LDV;
    f64x4 tt[KB];
#pragma unroll
    for (int tb = 0; tb < KB; ++tb) {
      tt[tb] = (f64x4){0.0, 0.0, 0.0, 0.0};
      if (tb * 16 < kc) {  // work-group-uniform
        const int r = tb * 16 + li;
        const bool rv = r < kc;
        const double *va = sV + (rv ? r : 0) * LDV + lq * K4;
#pragma unroll
        for (int j = 0; j < K4; ++j)
          tt[tb] = __builtin_amdgcn_mfma_f64_16x16x4f64(rv ? va[j] : 0.0, b[j], tt[tb], 0, 0, 0);
      }
    }
    if constexpr (!PB)
      if (more) load_b(e + 1, b);  // b is read; the loads overlap U_o T
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      const double *ua = sU + (mt * 16 + li) * LDU;
#pragma unroll
      for (int st = 0; st < 4 * KB; ++st)
        if (st * 4 < kc) {  // work-group-uniform
          const int si = 4 * st + lq;
          acc[mt] = __builtin_amdgcn_mfma_f64_16x16x4f64(si < kc ? ua[si] : 0.0, tt[st / 4][st % 4],
                                                         acc[mt], 0, 0, 0);
        }
    }
    if (more) {
      store_g(buf ^ 1, kn);
      if constexpr (PB) {
#pragma unroll
        for (int c = 0; c < K4; ++c) b[c] = bn[c];
      }
    }
    kc = kn;
    __syncthreads();
  }
  if (off) return;
  const double ih = inv_half[box_level[t]];
  double *out = Lc + (size_t)t * n3;
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int m = mt * 16 + lq + 4 * v;
      if (m < n3) out[m] = acc[mt][v] * ih;
    }
}

// Single right-hand side (ldm = 1), FP64 beyond the 64-target tiles: M2L
// grouped by offset.  One launch per M2L offset o covers every (target,
// source) pair with that offset (at most one per target, so the updates of one
// launch never collide and the sum order over offsets is fixed): D (node x
// pair) = G_o (node x l) * [M_s1 .. M_s16] (l x pair) on the FP64 matrix
// cores (v_mfma_f64_16x16x4f64: A[i][k] in lane i + 16 k, B[k][j] in lane
// j + 16 k, D[i][j] in lane j + 16 (i % 4), register i / 4), 16 pairs per
// wavefront, then Lc[t] += D[:, pair] / half.
template <int MT>
__global__ __launch_bounds__(256) void k_m2l1_f64(int n3, int npairs, const int *__restrict__ tgt,
                                                  const int *__restrict__ src,
                                                  const int *__restrict__ box_level,
                                                  const double *__restrict__ inv_half,
                                                  const double *__restrict__ Go,
                                                  const double *__restrict__ M,
                                                  double *__restrict__ Lc,
                                                  const unsigned char *__restrict__ skip) {
  constexpr int NP = 16 * MT;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j0 = (blockIdx.x * 4 + wave) * 16;
  if (j0 >= npairs) return;
  const int li = lane & 15, lq = lane >> 4;
  const int jp = min(j0 + li, npairs - 1);
  // a pair of a skipped target reads zeros and is not written
  const bool off = skip && skip[tgt[jp]];
  const double *Ms = M + (size_t)src[jp] * NP;
  f64x4 acc[MT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) acc[mt] = (f64x4){0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < NP; k0 += 4) {
    const double b = off ? 0.0 : Ms[k0 + lq];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      const double a = Go[(size_t)(mt * 16 + li) * NP + k0 + lq];
      acc[mt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[mt], 0, 0, 0);
    }
  }
  const int j = j0 + li;
  if (j < npairs && !off) {
    const int t = tgt[j];
    const double ih = inv_half[box_level[t]];
    double *out = Lc + (size_t)t * n3;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int m = mt * 16 + lq + 4 * v;
        if (m < n3) out[m] += acc[mt][v] * ih;
      }
  }
}

void launch_m2l_by_offset(const FmmDev<double> &f, int n3, const double *M, double *Lc,
                          hipStream_t s) {
  const int n3p = f.n3p, MT = n3p / 16;
  for (int o = 0; o < 343; ++o) {
    const int p0 = f.po_off[o], np = f.po_off[o + 1] - p0;
    if (np == 0) continue;
    const dim3 grid((np + 63) / 64);
    const int *tg = f.po_tgt + p0, *sr = f.po_src + p0;
    const double *G = f.G + (size_t)o * n3p * n3p;
    switch (MT) {
      case 2: hipLaunchKernelGGL((k_m2l1_f64<2>), grid, dim3(256), 0, s, n3, np, tg, sr, f.box_level, f.inv_half, G, M, Lc, f.skip_box); break;
      case 4: hipLaunchKernelGGL((k_m2l1_f64<4>), grid, dim3(256), 0, s, n3, np, tg, sr, f.box_level, f.inv_half, G, M, Lc, f.skip_box); break;
      case 8: hipLaunchKernelGGL((k_m2l1_f64<8>), grid, dim3(256), 0, s, n3, np, tg, sr, f.box_level, f.inv_half, G, M, Lc, f.skip_box); break;
      case 14: hipLaunchKernelGGL((k_m2l1_f64<14>), grid, dim3(256), 0, s, n3, np, tg, sr, f.box_level, f.inv_half, G, M, Lc, f.skip_box); break;
      default: throw Error(GP_ERR_INTERNAL, "no single-RHS M2L for this n3p");
    }
    GP_LAUNCH_CHECK();
  }
}

void launch_m2l_tiles(const FmmDev<float> &f, int n3, const __bf16 *Mb, float *Lc, int xrun,
                      hipStream_t s) {
  const __bf16 *G = reinterpret_cast<const __bf16 *>(f.Gbf);
  if (f.tile_size == 64) {
    const dim3 grid(f.n_tiles);
    switch (f.n3p / 16) {
      case 2: hipLaunchKernelGGL((k_m2l1t64_bf16<2>), grid, dim3(256), 0, s, n3, f.n_tiles, xrun, f.tile_tgt, f.tile_eoff, f.ent_oidx, f.ent_src, f.box_level, f.inv_half, G, Mb, Lc, f.skip_box); break;
      case 4: hipLaunchKernelGGL((k_m2l1t64_bf16<4>), grid, dim3(256), 0, s, n3, f.n_tiles, xrun, f.tile_tgt, f.tile_eoff, f.ent_oidx, f.ent_src, f.box_level, f.inv_half, G, Mb, Lc, f.skip_box); break;
      case 8: hipLaunchKernelGGL((k_m2l1t64_bf16<8>), grid, dim3(256), 0, s, n3, f.n_tiles, xrun, f.tile_tgt, f.tile_eoff, f.ent_oidx, f.ent_src, f.box_level, f.inv_half, G, Mb, Lc, f.skip_box); break;
      default: throw Error(GP_ERR_INTERNAL, "no 64-target M2L tiles for this n3p");
    }
    GP_LAUNCH_CHECK();
    return;
  }
  if (f.tile_size != 16) throw Error(GP_ERR_INTERNAL, "M2L target tiles of 16 or 64");
  const dim3 grid((f.n_tiles + 3) / 4);
  switch (f.n3p / 16) {
    case 2: hipLaunchKernelGGL((k_m2l1t_bf16<2>), grid, dim3(256), 0, s, n3, f.n_tiles, f.tile_tgt, f.tile_eoff, f.ent_oidx, f.ent_src, f.box_level, f.inv_half, G, Mb, Lc, f.skip_box); break;
    case 4: hipLaunchKernelGGL((k_m2l1t_bf16<4>), grid, dim3(256), 0, s, n3, f.n_tiles, f.tile_tgt, f.tile_eoff, f.ent_oidx, f.ent_src, f.box_level, f.inv_half, G, Mb, Lc, f.skip_box); break;
    case 8: hipLaunchKernelGGL((k_m2l1t_bf16<8>), grid, dim3(256), 0, s, n3, f.n_tiles, f.tile_tgt, f.tile_eoff, f.ent_oidx, f.ent_src, f.box_level, f.inv_half, G, Mb, Lc, f.skip_box); break;
    case 14: hipLaunchKernelGGL((k_m2l1t_bf16<14>), grid, dim3(256), 0, s, n3, f.n_tiles, f.tile_tgt, f.tile_eoff, f.ent_oidx, f.ent_src, f.box_level, f.inv_half, G, Mb, Lc, f.skip_box); break;
    default: throw Error(GP_ERR_INTERNAL, "no single-RHS M2L for this n3p");
  }
  GP_LAUNCH_CHECK();
}

void launch_m2l_tiles_f64lr(const FmmDev<double> &f, int n3, const double *M, double *Lc,
                            int xrun, hipStream_t s) {
  const dim3 grid(f.n_tiles);
  if (f.lr_kp != 16 && f.lr_kp != 32)
    throw Error(GP_ERR_INTERNAL, "factored FP64 M2L tiles take ranks up to 32");
  const bool k2 = f.lr_kp == 32;
  switch (f.n3p / 16) {
    case 2:
      if (k2) hipLaunchKernelGGL((k_m2l1t64_f64lr<2, 2>), grid, dim3(256), 0, s, n3, f.n_tiles, xrun, f.tile_tgt, f.tile_eoff, f.ent_oidx, f.ent_src, f.box_level, f.inv_half, f.lr_vt, f.lr_up, f.lr_rank, f.lr_kp, M, Lc, f.skip_box);
      else hipLaunchKernelGGL((k_m2l1t64_f64lr<2, 1>), grid, dim3(256), 0, s, n3, f.n_tiles, xrun, f.tile_tgt, f.tile_eoff, f.ent_oidx, f.ent_src, f.box_level, f.inv_half, f.lr_vt, f.lr_up, f.lr_rank, f.lr_kp, M, Lc, f.skip_box);
      break;
    case 4:
      if (k2) hipLaunchKernelGGL((k_m2l1t64_f64lr<4, 2>), grid, dim3(256), 0, s, n3, f.n_tiles, xrun, f.tile_tgt, f.tile_eoff, f.ent_oidx, f.ent_src, f.box_level, f.inv_half, f.lr_vt, f.lr_up, f.lr_rank, f.lr_kp, M, Lc, f.skip_box);
      else hipLaunchKernelGGL((k_m2l1t64_f64lr<4, 1>), grid, dim3(256), 0, s, n3, f.n_tiles, xrun, f.tile_tgt, f.tile_eoff, f.ent_oidx, f.ent_src, f.box_level, f.inv_half, f.lr_vt, f.lr_up, f.lr_rank, f.lr_kp, M, Lc, f.skip_box);
      break;
    default: throw Error(GP_ERR_INTERNAL, "no FP64 M2L tiles for this n3p");
  }
  GP_LAUNCH_CHECK();
}

void launch_m2l_tiles_f64lrs(const FmmDev<double> &f, int n3, const double *M, double *Lc,
                             int xrun, bool pb, hipStream_t s) {
  const dim3 grid(f.n_tiles);
  const int kr = f.lr_kmax, NP = f.n3p;
  const size_t lds = 2 * ((size_t)kr * (NP + 2) + (size_t)NP * (kr + 1)) * sizeof(double);
  if (kr < 1 || kr > 32 || kr > f.lr_kp || lds > 65536)
    throw Error(GP_ERR_INTERNAL, "staged factored FP64 M2L tiles take ranks up to 32 and 64 KB");
#define GP_LRS(MT_, KRB_)                                                                        \
  if (pb) hipLaunchKernelGGL((k_m2l1t64_f64lrs<MT_, KRB_, true>), grid, dim3(256), lds, s, n3,     \
                     f.n_tiles, xrun, f.tile_tgt, f.tile_eoff, f.ent_oidx, f.ent_src,             \
                     f.box_level, f.inv_half, f.lr_vt, f.lr_u, f.lr_rank, f.lr_kp, kr, M, Lc,     \
                     f.skip_box);                                                                 \
  else hipLaunchKernelGGL((k_m2l1t64_f64lrs<MT_, KRB_, false>), grid, dim3(256), lds, s, n3, f.n_tiles, xrun, \
                     f.tile_tgt, f.tile_eoff, f.ent_oidx, f.ent_src, f.box_level, f.inv_half,     \
                     f.lr_vt, f.lr_u, f.lr_rank, f.lr_kp, kr, M, Lc, f.skip_box)
  if (NP == 32) {
    if (kr <= 16) GP_LRS(2, 16);
    else if (kr <= 20) GP_LRS(2, 20);
    else GP_LRS(2, 32);
  } else if (NP == 64) {
    if (kr <= 16) GP_LRS(4, 16);
    else if (kr <= 20) GP_LRS(4, 20);
    else GP_LRS(4, 32);
  } else {
    throw Error(GP_ERR_INTERNAL, "no FP64 M2L tiles for this n3p");
  }
#undef GP_LRS
  GP_LAUNCH_CHECK();
}

void launch_m2l_tiles_f64(const FmmDev<double> &f, int n3, const double *M, double *Lc,
                          int xrun, hipStream_t s) {
  const dim3 grid(f.n_tiles);
  switch (f.n3p / 16) {
    case 2: hipLaunchKernelGGL((k_m2l1t64_f64<2>), grid, dim3(256), 0, s, n3, f.n_tiles, xrun, f.tile_tgt, f.tile_eoff, f.ent_oidx, f.ent_src, f.box_level, f.inv_half, f.G, M, Lc, f.skip_box); break;
    case 4: hipLaunchKernelGGL((k_m2l1t64_f64<4>), grid, dim3(256), 0, s, n3, f.n_tiles, xrun, f.tile_tgt, f.tile_eoff, f.ent_oidx, f.ent_src, f.box_level, f.inv_half, f.G, M, Lc, f.skip_box); break;
    default: throw Error(GP_ERR_INTERNAL, "no FP64 M2L tiles for this n3p");
  }
  GP_LAUNCH_CHECK();
}

template <int MT, int RTM>
void launch_m2l_bf16(const FmmDev<float> &f, int n3, int nr, int ldm, const __bf16 *Mb, float *Lc,
                     int xrun, hipStream_t s) {
  hipLaunchKernelGGL((k_m2l_bf16_pipe<MT, RTM>),
                     dim3(f.n_targets, ((nr + 15) / 16 + RTM - 1) / RTM), dim3(256), 0, s, n3, nr,
                     ldm, xrun,
                     f.targets, f.il_off, f.il_src, f.il_oidx, f.box_level, f.inv_half,
                     reinterpret_cast<const __bf16 *>(f.Gbf), Mb, Lc);
}
void launch_m2l_pairs(const FmmDev<float> &f, int n3, int nr, int ldm, const __bf16 *Mb,
                      float *Lc, int xrun, hipStream_t s) {
  constexpr int RTM = 4;
  const dim3 grid(f.n_pairs2, ((nr + 15) / 16 + RTM - 1) / RTM);
  const __bf16 *G = reinterpret_cast<const __bf16 *>(f.Gbf);
  switch (f.n3p / 16) {
    case 2: hipLaunchKernelGGL((k_m2l_bf16_pair<2, RTM>), grid, dim3(256), 0, s, n3, nr, ldm, xrun, f.pair_tgt, f.pair_off, f.pe_src, f.pe_o1, f.pe_o2, f.box_level, f.inv_half, G, Mb, Lc); break;
    case 4: hipLaunchKernelGGL((k_m2l_bf16_pair<4, RTM>), grid, dim3(256), 0, s, n3, nr, ldm, xrun, f.pair_tgt, f.pair_off, f.pe_src, f.pe_o1, f.pe_o2, f.box_level, f.inv_half, G, Mb, Lc); break;
    case 8: hipLaunchKernelGGL((k_m2l_bf16_pair<8, RTM>), grid, dim3(256), 0, s, n3, nr, ldm, xrun, f.pair_tgt, f.pair_off, f.pe_src, f.pe_o1, f.pe_o2, f.box_level, f.inv_half, G, Mb, Lc); break;
    default: throw Error(GP_ERR_INTERNAL, "no sibling-pair M2L for this n3p");
  }
  GP_LAUNCH_CHECK();
}

template <int MT, int RTM>
void launch_m2l_bf16x3(const FmmDev<float> &f, int n3, int nr, int ldm, const __bf16 *Mb,
                       const __bf16 *Mbl, float *Lc, int xrun, hipStream_t s) {
  hipLaunchKernelGGL((k_m2l_bf16x3<MT, RTM>),
                     dim3(f.n_targets, ((nr + 15) / 16 + RTM - 1) / RTM), dim3(256), 0, s, n3, nr,
                     ldm, xrun, f.targets, f.il_off, f.il_src, f.il_oidx, f.box_level, f.inv_half,
                     reinterpret_cast<const __bf16 *>(f.Gbf),
                     reinterpret_cast<const __bf16 *>(f.Gcbf_lo), Mb, Mbl, Lc);
}
template <int MT, int RTM>
void launch_m2l_bf16x3_d2(const FmmDev<float> &f, int n3, int nr, int ldm, const __bf16 *Mb,
                          const __bf16 *Mbl, float *Lc, int xrun, hipStream_t s) {
  hipLaunchKernelGGL((k_m2l_bf16x3<MT, RTM, 2>),
                     dim3(f.n_targets, ((nr + 15) / 16 + RTM - 1) / RTM), dim3(256), 0, s, n3, nr,
                     ldm, xrun, f.targets, f.il_off, f.il_src, f.il_oidx, f.box_level, f.inv_half,
                     reinterpret_cast<const __bf16 *>(f.Gbf),
                     reinterpret_cast<const __bf16 *>(f.Gcbf_lo), Mb, Mbl, Lc);
}
template <int MT, int RTM>
void launch_m2l_bf16x3_d4(const FmmDev<float> &f, int n3, int nr, int ldm, const __bf16 *Mb,
                          const __bf16 *Mbl, float *Lc, int xrun, hipStream_t s) {
  hipLaunchKernelGGL((k_m2l_bf16x3<MT, RTM, 4>),
                     dim3(f.n_targets, ((nr + 15) / 16 + RTM - 1) / RTM), dim3(256), 0, s, n3, nr,
                     ldm, xrun, f.targets, f.il_off, f.il_src, f.il_oidx, f.box_level, f.inv_half,
                     reinterpret_cast<const __bf16 *>(f.Gbf),
                     reinterpret_cast<const __bf16 *>(f.Gcbf_lo), Mb, Mbl, Lc);
}
template <int MT, int RTM>
void launch_m2l_bf16_d2(const FmmDev<float> &f, int n3, int nr, int ldm, const __bf16 *Mb,
                        float *Lc, int xrun, hipStream_t s) {
  hipLaunchKernelGGL((k_m2l_bf16_pipe<MT, RTM, 2>),
                     dim3(f.n_targets, ((nr + 15) / 16 + RTM - 1) / RTM), dim3(256), 0, s, n3, nr,
                     ldm, xrun, f.targets, f.il_off, f.il_src, f.il_oidx, f.box_level, f.inv_half,
                     reinterpret_cast<const __bf16 *>(f.Gbf), Mb, Lc);
}
template <int MT, int RTM>
void launch_m2l_f64(const FmmDev<double> &f, int n3, int nr, int ldm, const double *M, double *Lc,
                    int xrun, hipStream_t s) {
  hipLaunchKernelGGL((k_m2l_f64_pipe<MT, RTM>), dim3(f.n_targets), dim3(256), 0, s, n3, nr, ldm,
                     xrun,
                     f.targets, f.il_off, f.il_src, f.il_oidx, f.box_level, f.inv_half, f.G, M, Lc);
}
// Instantiation by (MT, RTM): MT = n3p / 16 in {2, 4, 8, 14} (orders 2-6),
// RTM in {2, 4, 8} the smallest bound on the RHS tiles.
#define GP_M2L_DISPATCH(FN, MT_, RT_, ...)                                          \
  do {                                                                             \
    const int rtm_ = (RT_) <= 2 ? 2 : (RT_) <= 4 ? 4 : 8;                          \
    switch ((MT_) * 16 + rtm_) {                                                   \
      case 2 * 16 + 2: FN<2, 2>(__VA_ARGS__); break;                               \
      case 2 * 16 + 4: FN<2, 4>(__VA_ARGS__); break;                               \
      case 2 * 16 + 8: FN<2, 8>(__VA_ARGS__); break;                               \
      case 4 * 16 + 2: FN<4, 2>(__VA_ARGS__); break;                               \
      case 4 * 16 + 4: FN<4, 4>(__VA_ARGS__); break;                               \
      case 4 * 16 + 8: FN<4, 8>(__VA_ARGS__); break;                               \
      case 8 * 16 + 2: FN<8, 2>(__VA_ARGS__); break;                               \
      case 8 * 16 + 4: FN<8, 4>(__VA_ARGS__); break;                               \
      case 8 * 16 + 8: FN<8, 8>(__VA_ARGS__); break;                               \
      case 14 * 16 + 2: FN<14, 2>(__VA_ARGS__); break;                             \
      case 14 * 16 + 4: FN<14, 4>(__VA_ARGS__); break;                             \
      case 14 * 16 + 8: FN<14, 8>(__VA_ARGS__); break;                             \
      default: throw Error(GP_ERR_INTERNAL, "no matrix-core M2L for this n3p");    \
    }                                                                              \
  } while (0)

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
__global__ __launch_bounds__(256) void k_l2p(int n, const int *__restrict__ leaf_lo, const int *__restrict__ leaf_box, int leaf_box0,
                                             const T *__restrict__ w1d, int nr,
                                             const T *__restrict__ Lc, int ldm,
                                             double *__restrict__ y, int ldy) {
  extern __shared__ unsigned char smem_raw[];
  T *s0 = reinterpret_cast<T *>(smem_raw);
  const int li = blockIdx.x, r0 = blockIdx.y * FM_RC;
  const int rc = min(FM_RC, nr - r0);
  const int n3 = n * n * n, nn = n * n;
  const T *ll = Lc + (size_t)(leaf_box ? leaf_box[li] : leaf_box0 + li) * n3 * ldm;
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

// ---------------------------------------------------------------- tiled kernels
// Used when the RHS block ldm is a multiple of 16 (<= 128).  Thread t of a
// 256-thread group: mg = t % 16 (node / point lane), rg = t / 16 (RHS group
// of RB = ldm / 16 consecutive columns).

constexpr int TP_KP = 16;  // points per LDS stage (P2M)
constexpr int TL_PT = 32;  // points per tile (L2P)
constexpr int TL_KM = 32;  // nodes per LDS stage (L2P)
constexpr int TM_RC = 32;  // RHS columns per work-group (M2M, L2L)

// Tensor weight S(p, m) = wx[a] wy[b] wz[c] of a point's 1-D weights.
template <typename T>
__device__ inline T tensor_weight(const T *__restrict__ w, int n, int m) {
  const int nn = n * n;
  return w[m / nn] * w[n + (m / n) % n] * w[2 * n + m % n];
}

// P2M as a per-leaf product M^T[m][r] = sum_p S(p, m) area_p z[p][r]; S and
// q = area * z are staged for TP_KP points at a time, each thread owns MB
// nodes (mg + 16 i) x RB columns (rg RB + j) in registers.
template <typename T, int MB>
__global__ __launch_bounds__(256) void k_p2m_tiled(int n, int n3p, const int *__restrict__ leaf_lo,
                                                   const int *__restrict__ leaf_box, int leaf_box0, const T *__restrict__ w1d,
                                                   const double *__restrict__ area,
                                                   const T *__restrict__ z, int ldz, int nr,
                                                   T *__restrict__ M, __bf16 *__restrict__ Mb,
                                                   int ldm) {
  __shared__ T sS[TP_KP][16 * MB];
  __shared__ T sQ[TP_KP][128];
  __shared__ T sW[TP_KP * 18];  // the stage's 1-D weights (3 n <= 18 per point)
  const int li = blockIdx.x, tid = threadIdx.x;
  const int mg = tid & 15, rg = tid >> 4;
  const int RB = ldm >> 4;
  const int n3 = n * n * n, NP = 16 * MB;
  T acc[MB][8];
#pragma unroll
  for (int i = 0; i < MB; ++i)
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[i][j] = 0;
  const int pb = leaf_lo[li], pe = leaf_lo[li + 1];
  for (int p0 = pb; p0 < pe; p0 += TP_KP) {
    const int np = min(TP_KP, pe - p0);
    for (int e = tid; e < np * 3 * n; e += 256) sW[e] = w1d[(size_t)p0 * 3 * n + e];
    __syncthreads();
    for (int e = tid; e < TP_KP * NP; e += 256) {
      int pp = e / NP, m = e - pp * NP;
      sS[pp][m] = (pp < np && m < n3) ? tensor_weight(sW + pp * 3 * n, n, m) : (T)0;
    }
    for (int e = tid; e < TP_KP * ldm; e += 256) {
      int pp = e / ldm, r = e - pp * ldm;
      sQ[pp][r] = (pp < np && r < nr) ? (T)area[p0 + pp] * z[(size_t)(p0 + pp) * ldz + r] : (T)0;
    }
    __syncthreads();
    for (int pp = 0; pp < np; ++pp) {
      T sv[MB];
#pragma unroll
      for (int i = 0; i < MB; ++i) sv[i] = sS[pp][mg + 16 * i];
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (j < RB) {
          const T q = sQ[pp][rg * RB + j];
#pragma unroll
          for (int i = 0; i < MB; ++i) acc[i][j] += sv[i] * q;
        }
    }
    __syncthreads();
  }
  const size_t base = (size_t)(leaf_box ? leaf_box[li] : leaf_box0 + li) * ldm * n3p;
#pragma unroll
  for (int i = 0; i < MB; ++i)
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int m = mg + 16 * i, r = rg * RB + j;
      if (j < RB && m < n3 && r < nr) store_mult(M, Mb, base + (size_t)r * n3p + m, acc[i][j]);
    }
}

// L2P as a per-leaf product y[p][r] += sum_m S(p, m) Lc[m][r]: node chunks of
// Lc and S staged in LDS; thread (pg, rc) owns points pg, pg + 16 and the RB
// columns rc + 16 j, so the 16 lanes of a point update 16 consecutive
// columns of y.
template <typename T>
__global__ __launch_bounds__(256) void k_l2p_tiled(int n, const int *__restrict__ leaf_lo,
                                                   const int *__restrict__ leaf_box, int leaf_box0, const T *__restrict__ w1d, int nr,
                                                   const T *__restrict__ Lc, int ldm,
                                                   double *__restrict__ y, int ldy) {
  __shared__ T sL[TL_KM][128];
  __shared__ T sS[TL_PT][TL_KM + 1];
  const int li = blockIdx.x, tid = threadIdx.x;
  const int pg = tid >> 4, rc = tid & 15;
  const int RB = ldm >> 4;
  const int n3 = n * n * n;
  const T *ll = Lc + (size_t)(leaf_box ? leaf_box[li] : leaf_box0 + li) * n3 * ldm;
  const int pb = leaf_lo[li], pe = leaf_lo[li + 1];
  for (int p0 = pb; p0 < pe; p0 += TL_PT) {
    const int np = min(TL_PT, pe - p0);
    T acc[2][8];
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int j = 0; j < 8; ++j) acc[q][j] = 0;
    for (int m0 = 0; m0 < n3; m0 += TL_KM) {
      for (int e = tid; e < TL_KM * ldm; e += 256) {
        int mm = e / ldm, r = e - mm * ldm;
        sL[mm][r] = (m0 + mm < n3) ? ll[(size_t)(m0 + mm) * ldm + r] : (T)0;
      }
      for (int e = tid; e < TL_PT * TL_KM; e += 256) {
        int pp = e / TL_KM, mm = e - pp * TL_KM;
        sS[pp][mm] = (pp < np && m0 + mm < n3)
                         ? tensor_weight(w1d + (size_t)(p0 + pp) * 3 * n, n, m0 + mm)
                         : (T)0;
      }
      __syncthreads();
#pragma unroll 4
      for (int mm = 0; mm < TL_KM; ++mm) {
        const T s0 = sS[pg][mm], s1 = sS[pg + 16][mm];
#pragma unroll
        for (int j = 0; j < 8; ++j)
          if (j < RB) {
            const T l = sL[mm][rc + 16 * j];
            acc[0][j] += s0 * l;
            acc[1][j] += s1 * l;
          }
      }
      __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int pp = pg + 16 * q, r = rc + 16 * j;
        if (j < RB && pp < np && r < nr) y[(size_t)(p0 + pp) * ldy + r] += (double)acc[q][j];
      }
  }
}

// L2P with one wavefront per leaf (FP32 locals, column groups of 4):
// y[p][r] += sum_m S(p, m) Lc[m][r].  The wavefront's lanes are qn = ld / 4
// column groups x pg = 64 / qn point groups: lane (ph, cq) owns columns
// 4 cq .. 4 cq + 3 and points ph, ph + pg, ... of a chunk of 16 (32 columns:
// 8 groups x 8 point groups, so no lane idles for narrow blocks); the leaf's
// locals stream from memory one 16-byte group per node, N node rows issued
// at a time, and the chunk's tensor weights sit in LDS (read as
// broadcasts).  The weights come from the points' 1-D weights, staged first.
template <int N, int PPL>  // PPL = ceil(16 / pg) points per lane and chunk
__global__ __launch_bounds__(64) void k_l2p_wave(const int *__restrict__ leaf_lo,
                                                 const int *__restrict__ leaf_box, int leaf_box0,
                                                 const float *__restrict__ w1d, int nr,
                                                 const float *__restrict__ Lc, int ldm,
                                                 double *__restrict__ y, int ldy) {
  constexpr int N3 = N * N * N, PC = 16, NW = PC * 3 * N;
  __shared__ float sW[NW];
  __shared__ float sS[PC][N3];
  const int li = blockIdx.x, lane = threadIdx.x;
  const int qn = ldm >> 2, pg = 64 / qn;  // caller: ldm % 4 == 0, 16 <= ldm <= 128
  const int cq = lane % qn, ph = lane / qn;
  const int r = 4 * cq;
  const bool rok = ph < pg;
  const float *ll = Lc + (size_t)(leaf_box ? leaf_box[li] : leaf_box0 + li) * N3 * ldm + r;
  const int pb = leaf_lo[li], pe = leaf_lo[li + 1];
  for (int p0 = pb; p0 < pe; p0 += PC) {
    const int np = min(PC, pe - p0);
    __syncthreads();  // the previous chunk's weights are no longer read
    for (int e = lane; e < NW; e += 64)
      sW[e] = e < np * 3 * N ? w1d[(size_t)p0 * 3 * N + e] : 0.f;
    __syncthreads();
    for (int e = lane; e < PC * N3; e += 64) {
      const int pp = e / N3, m = e - pp * N3;
      const float *w = sW + pp * 3 * N;
      sS[pp][m] = w[m / (N * N)] * w[N + (m / N) % N] * w[2 * N + m % N];
    }
    __syncthreads();
    f32x4 acc[PPL];
    const float *srow[PPL];  // the lane's weight rows (clamped; extra rows are never stored)
#pragma unroll
    for (int i = 0; i < PPL; ++i) {
      acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
      srow[i] = &sS[min(ph + pg * i, PC - 1)][0];
    }
    if (rok) {
      // two node groups in flight for narrow blocks (few points per lane)
      constexpr int UNR = PPL <= 4 ? 2 : 1;
#pragma unroll UNR
      for (int m = 0; m < N3; m += N) {
        f32x4 L[N];
#pragma unroll
        for (int c = 0; c < N; ++c) L[c] = *reinterpret_cast<const f32x4 *>(ll + (size_t)(m + c) * ldm);
#pragma unroll
        for (int c = 0; c < N; ++c)
#pragma unroll
          for (int i = 0; i < PPL; ++i) acc[i] += srow[i][m + c] * L[c];
      }
#pragma unroll
      for (int i = 0; i < PPL; ++i) {
        const int pp = ph + pg * i;
        if (pp >= np) continue;
        double *yr = y + (size_t)(p0 + pp) * ldy + r;
#pragma unroll
        for (int v = 0; v < 4; ++v)
          if (r + v < nr) yr[v] += (double)acc[i][v];
      }
    }
  }
}

// The finest L2L and L2P fused (FP32 blocks of RHS, every leaf at the
// finest level): one work-group per leaf, per 32-column chunk the transposed
// 1-D interpolations of the parent's locals added to the leaf's own (the
// sums of k_l2l_tiled) into LDS, then the tensor-product sum of each panel
// over the chunk's locals (the sums of k_l2p_wave) added to y.  The leaf's
// final locals never go back to memory: k_l2l_tiled wrote them and
// k_l2p_wave read them again.
template <int N>
__global__ __launch_bounds__(256) void k_l2l_l2p_leaf(int leaf0, const int *__restrict__ parent,
                                                      const int *__restrict__ octant,
                                                      const float *__restrict__ m2m1d,
                                                      const int *__restrict__ leaf_lo,
                                                      const float *__restrict__ w1d, int nr,
                                                      const float *__restrict__ Lc, int ldm,
                                                      double *__restrict__ y, int ldy) {
  constexpr int NN = N * N, N3 = N * N * N, PC = 16, NW = PC * 3 * N;
  constexpr int ITEMS = (TM_RC * NN + 255) / 256;
  __shared__ float t2[TM_RC * N * NN];
  __shared__ float sL[N3 * TM_RC];
  __shared__ float sW[NW];
  __shared__ float sS[PC][N3];
  const int li = blockIdx.x, ch = leaf0 + li, tid = threadIdx.x;
  const int p = parent[ch], o = octant[ch];
  const float *mx = m2m1d + (o * 3 + 0) * NN, *my = m2m1d + (o * 3 + 1) * NN,
              *mz = m2m1d + (o * 3 + 2) * NN;
  const float *lp = Lc + (size_t)p * N3 * ldm;
  const float *lc = Lc + (size_t)ch * N3 * ldm;
  const int pb = leaf_lo[li], pe = leaf_lo[li + 1];
  for (int r0 = 0; r0 < nr; r0 += TM_RC) {
    const int rc = min(TM_RC, nr - r0);
    __syncthreads();  // the previous chunk's t2 and sL are no longer read
    float cur[ITEMS][N];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      const int item = tid + 256 * k;
      const int r = item % rc, bc = item / rc;
#pragma unroll
      for (int ap = 0; ap < N; ++ap)
        cur[k][ap] = item < rc * NN ? lc[(size_t)(ap * NN + bc) * ldm + r0 + r] : 0.f;
    }
    if (tid < rc * N) {
      const int r = tid % rc, a = tid / rc;
      float in[NN], u[NN];
#pragma unroll
      for (int e = 0; e < NN; ++e) in[e] = lp[(size_t)(a * NN + e) * ldm + r0 + r];
#pragma unroll
      for (int bb = 0; bb < N; ++bb)
#pragma unroll
        for (int cp = 0; cp < N; ++cp) {
          float acc = 0;
#pragma unroll
          for (int c = 0; c < N; ++c) acc += mz[c * N + cp] * in[bb * N + c];
          u[bb * N + cp] = acc;
        }
      float *dst = t2 + ((size_t)r * N + a) * NN;
#pragma unroll
      for (int bp = 0; bp < N; ++bp)
#pragma unroll
        for (int cp = 0; cp < N; ++cp) {
          float acc = 0;
#pragma unroll
          for (int bb = 0; bb < N; ++bb) acc += my[bb * N + bp] * u[bb * N + cp];
          dst[bp * N + cp] = acc;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      const int item = tid + 256 * k;
      if (item < rc * NN) {
        const int r = item % rc, bc = item / rc;
        float tv[N];
#pragma unroll
        for (int a = 0; a < N; ++a) tv[a] = t2[((size_t)r * N + a) * NN + bc];
#pragma unroll
        for (int ap = 0; ap < N; ++ap) {
          float acc = 0;
#pragma unroll
          for (int a = 0; a < N; ++a) acc += mx[a * N + ap] * tv[a];
          sL[(ap * NN + bc) * TM_RC + r] = cur[k][ap] + acc;
        }
      }
    }
    // L2P of the chunk: 16 panels at a time, thread = (panel, column) pairs
    for (int p0 = pb; p0 < pe; p0 += PC) {
      const int np = min(PC, pe - p0);
      __syncthreads();  // sL complete; the previous panels' weights no longer read
      for (int e = tid; e < NW; e += 256) sW[e] = e < np * 3 * N ? w1d[(size_t)p0 * 3 * N + e] : 0.f;
      __syncthreads();
      for (int e = tid; e < PC * N3; e += 256) {
        const int pp = e / N3, m = e - pp * N3;
        const float *w = sW + pp * 3 * N;
        sS[pp][m] = w[m / (N * N)] * w[N + (m / N) % N] * w[2 * N + m % N];
      }
      __syncthreads();
      for (int e = tid; e < np * TM_RC; e += 256) {
        const int pp = e / TM_RC, r = e - pp * TM_RC;
        if (r >= rc) continue;
        float acc = 0.f;
#pragma unroll 5
        for (int m = 0; m < N3; ++m) acc += sS[pp][m] * sL[m * TM_RC + r];
        y[(size_t)(p0 + pp) * ldy + r0 + r] += (double)acc;
      }
    }
  }
}

// M2M with the order N fixed at compile time: per parent box and 32-column
// RHS chunk, for each child the z and y contractions of every (r, a) slice
// run in registers, one LDS exchange, then the x contraction accumulates the
// parent's (r, bc) outputs in registers.
template <typename T, int N>
__global__ __launch_bounds__(256) void k_m2m_tiled(int n3p, int box0,
                                                   const int *__restrict__ child_off,
                                                   const int *__restrict__ child_idx,
                                                   const int *__restrict__ octant,
                                                   const T *__restrict__ m2m1d, int nr,
                                                   T *__restrict__ M, __bf16 *__restrict__ Mb,
                                                   int ldm) {
  constexpr int NN = N * N;
  constexpr int ITEMS = (TM_RC * NN + 255) / 256;
  __shared__ T t2[TM_RC * N * NN];
  const int b = box0 + blockIdx.x, r0 = blockIdx.y * TM_RC;
  const int rc = min(TM_RC, nr - r0);
  const int tid = threadIdx.x;
  T acc[ITEMS][N];
#pragma unroll
  for (int k = 0; k < ITEMS; ++k)
#pragma unroll
    for (int a = 0; a < N; ++a) acc[k][a] = 0;
  for (int q = child_off[b]; q < child_off[b + 1]; ++q) {
    const int ch = child_idx[q], o = octant[ch];
    const T *mx = m2m1d + (o * 3 + 0) * NN, *my = m2m1d + (o * 3 + 1) * NN,
            *mz = m2m1d + (o * 3 + 2) * NN;
    if (tid < rc * N) {
      const int r = tid / N, a = tid - r * N;
      const T *src = M + (size_t)ch * ldm * n3p + (size_t)(r0 + r) * n3p + a * NN;
      T in[NN], u[NN];
#pragma unroll
      for (int e = 0; e < NN; ++e) in[e] = src[e];
      // z: u[b][cp] = sum_c mz[cp][c] in[b][c]
#pragma unroll
      for (int bb = 0; bb < N; ++bb)
#pragma unroll
        for (int cp = 0; cp < N; ++cp) {
          T s = 0;
#pragma unroll
          for (int c = 0; c < N; ++c) s += mz[cp * N + c] * in[bb * N + c];
          u[bb * N + cp] = s;
        }
      // y: t2[r][a][bp][cp] = sum_b my[bp][b] u[b][cp]
      T *dst = t2 + ((size_t)r * N + a) * NN;
#pragma unroll
      for (int bp = 0; bp < N; ++bp)
#pragma unroll
        for (int cp = 0; cp < N; ++cp) {
          T s = 0;
#pragma unroll
          for (int bb = 0; bb < N; ++bb) s += my[bp * N + bb] * u[bb * N + cp];
          dst[bp * N + cp] = s;
        }
    }
    __syncthreads();
    // x: acc[(r, bc)][ap] += sum_a mx[ap][a] t2[r][a][bc]
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      const int item = tid + 256 * k;
      if (item < rc * NN) {
        const int r = item / NN, bc = item - r * NN;
        T tv[N];
#pragma unroll
        for (int a = 0; a < N; ++a) tv[a] = t2[((size_t)r * N + a) * NN + bc];
#pragma unroll
        for (int ap = 0; ap < N; ++ap) {
          T s = 0;
#pragma unroll
          for (int a = 0; a < N; ++a) s += mx[ap * N + a] * tv[a];
          acc[k][ap] += s;
        }
      }
    }
    __syncthreads();
  }
  const size_t base = (size_t)b * ldm * n3p;
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const int item = tid + 256 * k;
    if (item < rc * NN) {
      const int r = item / NN, bc = item - r * NN;
#pragma unroll
      for (int ap = 0; ap < N; ++ap)
        store_mult(M, Mb, base + (size_t)(r0 + r) * n3p + ap * NN + bc, acc[k][ap]);
    }
  }
}

// L2L with N fixed: per child box and 32-column chunk, the transposed 1-D
// interpolations of the parent's local expansion: z^T and y^T in registers
// per (r, a) slice, one LDS exchange, x^T per (r, bc), added to the child.
// (The reference applies x^T first; the three 1-D maps commute.)
template <typename T, int N>
__global__ __launch_bounds__(256) void k_l2l_tiled(int box0, const int *__restrict__ parent,
                                                   const int *__restrict__ octant,
                                                   const T *__restrict__ m2m1d, int nr,
                                                   T *__restrict__ Lc, int ldm) {
  constexpr int NN = N * N, N3 = N * N * N;
  constexpr int ITEMS = (TM_RC * NN + 255) / 256;
  __shared__ T t2[TM_RC * N * NN];
  const int ch = box0 + blockIdx.x, r0 = blockIdx.y * TM_RC;
  const int rc = min(TM_RC, nr - r0);
  const int tid = threadIdx.x;
  const int p = parent[ch], o = octant[ch];
  const T *mx = m2m1d + (o * 3 + 0) * NN, *my = m2m1d + (o * 3 + 1) * NN,
          *mz = m2m1d + (o * 3 + 2) * NN;
  const T *lp = Lc + (size_t)p * N3 * ldm;
  T *lc = Lc + (size_t)ch * N3 * ldm;
  // the child's current locals, read before the parent's passes so the
  // loads overlap them (phase 2 adds to these)
  T cur[ITEMS][N];
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const int item = tid + 256 * k;
    const int r = item % rc, bc = item / rc;
#pragma unroll
    for (int ap = 0; ap < N; ++ap)
      cur[k][ap] = item < rc * NN ? lc[(size_t)(ap * NN + bc) * ldm + r0 + r] : (T)0;
  }
  if (tid < rc * N) {
    const int r = tid % rc, a = tid / rc;
    T in[NN], u[NN];
#pragma unroll
    for (int e = 0; e < NN; ++e) in[e] = lp[(size_t)(a * NN + e) * ldm + r0 + r];
    // z^T: u[b][cp] = sum_c mz[c][cp] in[b][c]
#pragma unroll
    for (int bb = 0; bb < N; ++bb)
#pragma unroll
      for (int cp = 0; cp < N; ++cp) {
        T s = 0;
#pragma unroll
        for (int c = 0; c < N; ++c) s += mz[c * N + cp] * in[bb * N + c];
        u[bb * N + cp] = s;
      }
    // y^T: t2[r][a][bp][cp] = sum_b my[b][bp] u[b][cp]
    T *dst = t2 + ((size_t)r * N + a) * NN;
#pragma unroll
    for (int bp = 0; bp < N; ++bp)
#pragma unroll
      for (int cp = 0; cp < N; ++cp) {
        T s = 0;
#pragma unroll
        for (int bb = 0; bb < N; ++bb) s += my[bb * N + bp] * u[bb * N + cp];
        dst[bp * N + cp] = s;
      }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const int item = tid + 256 * k;
    if (item < rc * NN) {
      const int r = item % rc, bc = item / rc;
      T tv[N];
#pragma unroll
      for (int a = 0; a < N; ++a) tv[a] = t2[((size_t)r * N + a) * NN + bc];
#pragma unroll
      for (int ap = 0; ap < N; ++ap) {
        T s = 0;
#pragma unroll
        for (int a = 0; a < N; ++a) s += mx[a * N + ap] * tv[a];
        lc[(size_t)(ap * NN + bc) * ldm + r0 + r] = cur[k][ap] + s;
      }
    }
  }
}

template <typename T, int N>
void launch_m2m_tiled(int n3p, int box0, int nb, int nrc32, const FmmDev<T> &f, int nr, T *M,
                      __bf16 *Mb, int ldm, hipStream_t s) {
  hipLaunchKernelGGL((k_m2m_tiled<T, N>), dim3(nb, nrc32), dim3(256), 0, s, n3p, box0, f.child_off,
                     f.child_idx, f.octant, f.m2m1d, nr, M, Mb, ldm);
}
template <typename T, int N>
void launch_l2l_tiled(int box0, int nb, int nrc32, const FmmDev<T> &f, int nr, T *Lc, int ldm,
                      hipStream_t s) {
  hipLaunchKernelGGL((k_l2l_tiled<T, N>), dim3(nb, nrc32), dim3(256), 0, s, box0, f.parent,
                     f.octant, f.m2m1d, nr, Lc, ldm);
}

// In-place rounding to BF16 (round to nearest even) and back.
__global__ __launch_bounds__(256) void k_round_bf16(float *__restrict__ p, size_t n) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x)
    p[e] = (float)(__bf16)p[e];
}

// Zero fill (instead of a memset node, so the pass can be graph-captured).
template <typename T>
__global__ __launch_bounds__(256) void k_fill_zero(T *__restrict__ p, size_t n) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x)
    p[e] = (T)0;
}

// Low-rank M2L, projection: Mt[q] = U^T M[q] (U: n3 x KR, row-major) for the
// multipole rows q -> (box0 + q / nr, q % nr): a register-tiled GEMM, 64 rows
// x KR columns per work-group, the contraction over n3 in LDS stages of 32;
// thread (ty, tx) owns rows 4 ty .. 4 ty + 3 and columns tx + 16 c.  A BF16
// copy goes to Mtb when set.
template <int KR>
__global__ __launch_bounds__(256) void k_m2l_compress(int n3, int n3p, int box0, int nrow, int nr,
                                                      int ldm, const float *__restrict__ U,
                                                      const float *__restrict__ M,
                                                      float *__restrict__ Mt,
                                                      __bf16 *__restrict__ Mtb,
                                                      __bf16 *__restrict__ Mtbl) {
  constexpr int TC = KR / 16;
  __shared__ __align__(16) float sM[32][64 + 4];  // [k][row]
  __shared__ float sU[32][KR];                    // [k][col]
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int q0 = blockIdx.x * 64;
  float acc[4][TC];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int c = 0; c < TC; ++c) acc[i][c] = 0.f;
  for (int k0 = 0; k0 < n3; k0 += 32) {
    for (int e = threadIdx.x; e < 64 * 32; e += 256) {
      const int i = e >> 5, k = e & 31, q = q0 + i, m = k0 + k;
      sM[k][i] = (q < nrow && m < n3) ? M[((size_t)(box0 + q / nr) * ldm + q % nr) * n3p + m] : 0.f;
    }
    for (int e = threadIdx.x; e < 32 * KR; e += 256) {
      const int k = e / KR, j = e % KR, m = k0 + k;
      sU[k][j] = m < n3 ? U[(size_t)m * KR + j] : 0.f;
    }
    __syncthreads();
#pragma unroll 4
    for (int k = 0; k < 32; ++k) {
      const float4 a = *reinterpret_cast<const float4 *>(&sM[k][ty * 4]);
      float b[TC];
#pragma unroll
      for (int c = 0; c < TC; ++c) b[c] = sU[k][tx + 16 * c];
#pragma unroll
      for (int c = 0; c < TC; ++c) {
        acc[0][c] += a.x * b[c];
        acc[1][c] += a.y * b[c];
        acc[2][c] += a.z * b[c];
        acc[3][c] += a.w * b[c];
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int q = q0 + ty * 4 + i;
    if (q >= nrow) continue;
    const size_t rb = ((size_t)(box0 + q / nr) * ldm + q % nr) * KR;
#pragma unroll
    for (int c = 0; c < TC; ++c) {
      const float v = acc[i][c];
      Mt[rb + tx + 16 * c] = v;
      if (Mtb) {
        const __bf16 h = (__bf16)v;
        Mtb[rb + tx + 16 * c] = h;
        if (Mtbl) Mtbl[rb + tx + 16 * c] = (__bf16)(v - (float)h);
      }
    }
  }
}

// Low-rank M2L, expansion: Lc[b][m][r] = sum_j U[m][j] Lt[b][j][r] for the
// columns c -> (b, r) = (targets[c / nr], c % nr): a register-tiled GEMM,
// 64 nodes x 64 columns per work-group, the contraction over kr in LDS stages
// of 32; thread (ty, tx) owns nodes 4 ty .. 4 ty + 3 and columns tx + 16 c.
__global__ __launch_bounds__(256) void k_m2l_expand(int n3, int kr, int ncol, int nr, int ldm,
                                                    const int *__restrict__ targets,
                                                    const float *__restrict__ U,
                                                    const float *__restrict__ Lt,
                                                    float *__restrict__ Lc) {
  __shared__ __align__(16) float sU[32][64 + 4];  // [k][node]
  __shared__ float sB[32][64];                    // [k][column]
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int c0 = blockIdx.x * 64, m0 = blockIdx.y * 64;
  float acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[i][c] = 0.f;
  for (int k0 = 0; k0 < kr; k0 += 32) {
    for (int e = threadIdx.x; e < 64 * 32; e += 256) {
      const int i = e >> 5, k = e & 31, m = m0 + i;
      sU[k][i] = m < n3 ? U[(size_t)m * kr + k0 + k] : 0.f;
    }
    for (int e = threadIdx.x; e < 32 * 64; e += 256) {
      const int k = e >> 6, cc = e & 63, c = c0 + cc;
      sB[k][cc] = c < ncol ? Lt[((size_t)targets[c / nr] * kr + k0 + k) * ldm + c % nr] : 0.f;
    }
    __syncthreads();
#pragma unroll 4
    for (int k = 0; k < 32; ++k) {
      const float4 a = *reinterpret_cast<const float4 *>(&sU[k][ty * 4]);
      float b[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) b[c] = sB[k][tx + 16 * c];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        acc[0][c] += a.x * b[c];
        acc[1][c] += a.y * b[c];
        acc[2][c] += a.z * b[c];
        acc[3][c] += a.w * b[c];
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int cg = c0 + tx + 16 * c;
    if (cg >= ncol) continue;
    const int b = targets[cg / nr], r = cg % nr;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int m = m0 + ty * 4 + i;
      if (m < n3) Lc[((size_t)b * n3 + m) * ldm + r] = acc[i][c];
    }
  }
}

// The same projection and expansion on the FP32 matrix cores
// (v_mfma_f32_16x16x4f32: lane holds A[i = lane % 16][k = lane / 16],
// B[k = lane / 16][j = lane % 16], D[i = 4 (lane / 16) + v][j = lane % 16]).
// Both are HBM streams (FP32 expansions in, projected ones out, or back), so
// one wavefront takes one block of rows at a time with all its loads issued
// ahead, and U sits in LDS for the whole work-group.

// Projection: per wavefront 16 multipole rows q (column j of the product),
// D_ct = U_ct^T M^T over the n3 nodes in NS steps of 16.  Tile ct's output row
// i = 4 qq + v is projected column (KR / 4) qq + 4 ct + v, so lane (j, qq) ends
// with KR / 4 consecutive columns of row j (16-byte stores).  The k order
// inside a step is 4 qq + t, matching one 16-byte load of M per lane.
template <int KR, int NS>
__global__ __launch_bounds__(256) void k_m2l_compress_mfma(int n3, int box0, int nrow, int nr,
                                                           int ldm, const float *__restrict__ U,
                                                           const float *__restrict__ M,
                                                           float *__restrict__ Mt,
                                                           __bf16 *__restrict__ Mtb,
                                                           __bf16 *__restrict__ Mtbl) {
  constexpr int CT = KR / 16, QW = KR / 4, N3P = NS * 16, LDA = N3P + 4;
  extern __shared__ __align__(16) float sA[];  // [CT][16][LDA]: U_ct^T
  for (int e = threadIdx.x; e < CT * 16 * N3P; e += 256) {
    const int k = e % N3P, ci = e / N3P, ct = ci >> 4, i = ci & 15;
    const int col = QW * (i >> 2) + 4 * ct + (i & 3);
    sA[ci * LDA + k] = k < n3 ? U[(size_t)k * KR + col] : 0.f;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, j = lane & 15, qq = lane >> 4;
  const int ngroups = (nrow + 15) >> 4;
  const int stride = gridDim.x * 4;
  int g = blockIdx.x * 4 + (threadIdx.x >> 6);
  // row base of multipole row q (the caller's M rows are n3p = N3P wide)
  auto row_ptr = [&](int q) -> const float * {
    return M + ((size_t)(box0 + q / nr) * ldm + q % nr) * N3P;
  };
  f32x4 bn[NS];
  auto load = [&](int gg) {
    const int q = gg * 16 + j;
    const bool ok = gg < ngroups && q < nrow;
    const float *p = ok ? row_ptr(q) : M;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int k0 = 16 * s + 4 * qq;
      f32x4 v = {0.f, 0.f, 0.f, 0.f};
      if (ok && k0 < n3) {
        v = *reinterpret_cast<const f32x4 *>(p + k0);
#pragma unroll
        for (int t = 0; t < 4; ++t)
          if (k0 + t >= n3) v[t] = 0.f;  // node padding may hold anything
      }
      bn[s] = v;
    }
  };
  load(g);
  for (; g < ngroups; g += stride) {
    f32x4 b[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) b[s] = bn[s];
    load(g + stride);  // the next block's rows stream in under this one's products
    f32x4 acc[CT];
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) acc[ct] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < NS; ++s) {
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) {
        const f32x4 a = *reinterpret_cast<const f32x4 *>(&sA[(ct * 16 + j) * LDA + 16 * s + 4 * qq]);
#pragma unroll
        for (int t = 0; t < 4; ++t)
          acc[ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[t], b[s][t], acc[ct], 0, 0, 0);
      }
    }
    // lane (qq, v) of tile ct holds row j, projected column QW qq + 4 ct + v
    const int q = g * 16 + j;
    if (q < nrow) {
      const size_t rb = ((size_t)(box0 + q / nr) * ldm + q % nr) * KR + QW * qq;
      if (Mtb) {
        __bf16 h[QW], l[QW];
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const float x = acc[ct][v];
            const __bf16 hi = (__bf16)x;
            h[4 * ct + v] = hi;
            l[4 * ct + v] = (__bf16)(x - (float)hi);
          }
#pragma unroll
        for (int c = 0; c < QW; c += 8) {
          bf16x8 vh, vl;
#pragma unroll
          for (int e = 0; e < 8; ++e) vh[e] = h[c + e], vl[e] = l[c + e];
          *reinterpret_cast<bf16x8 *>(Mtb + rb + c) = vh;
          if (Mtbl) *reinterpret_cast<bf16x8 *>(Mtbl + rb + c) = vl;
        }
      } else {
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) *reinterpret_cast<f32x4 *>(Mt + rb + 4 * ct) = acc[ct];
      }
    }
  }
}

// Expansion: per wavefront one box b and 64 columns r0 .. r0 + 63 of it,
// Lc[b][m][r] = sum_k U[m][k] Lt[b][k][r] for m < n3 in chunks of 64 nodes
// (four 16-node tiles; U in LDS for the nodes rounded up to 16), every
// column r < ldm written (zero for r >= nr).
// Tile v holds columns r0 + 4 j + v, so lane (j, qq) stores 16 bytes per
// node row; the k order inside a step of 16 is 4 qq + t, matching the lane's
// four rows of Lt.  Boxes without an interaction list have Lt = 0 and get
// Lc = 0, so the caller needs no zero fill of Lc.
template <int KR>
__global__ __launch_bounds__(256) void k_m2l_expand_mfma(int n3, int nmc, int n_boxes, int ngr,
                                                         int nr, int ldm,
                                                         const float *__restrict__ U,
                                                         const float *__restrict__ Lt,
                                                         float *__restrict__ Lc) {
  constexpr int KS = KR / 16, LDU = KR + 4;
  const int m16 = (n3 + 15) & ~15;  // node rows held in LDS
  extern __shared__ __align__(16) float sU[];  // [m16][LDU]
  for (int e = threadIdx.x; e < m16 * KR; e += 256) {
    const int m = e / KR, k = e % KR;
    sU[m * LDU + k] = m < n3 ? U[(size_t)m * KR + k] : 0.f;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, j = lane & 15, qq = lane >> 4;
  const int nitems = n_boxes * ngr;
  for (int w = blockIdx.x * 4 + (threadIdx.x >> 6); w < nitems; w += gridDim.x * 4) {
    const int b = w / ngr, r = (w % ngr) * 64 + 4 * j;
    const bool rok = r < ldm;
    f32x4 bv[KS][4];
#pragma unroll
    for (int s = 0; s < KS; ++s)
#pragma unroll
      for (int t = 0; t < 4; ++t)
        bv[s][t] = rok ? *reinterpret_cast<const f32x4 *>(
                             Lt + ((size_t)b * KR + 16 * s + 4 * qq + t) * ldm + r)
                       : f32x4{0.f, 0.f, 0.f, 0.f};
    for (int mc = 0; mc < nmc; ++mc) {
      f32x4 acc[4][4];  // [node tile][column tile v]
#pragma unroll
      for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[mt][v] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        f32x4 a[4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
          a[mt] = mc * 64 + mt * 16 < m16
                      ? *reinterpret_cast<const f32x4 *>(&sU[(mc * 64 + mt * 16 + j) * LDU + 16 * s + 4 * qq])
                      : f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
          for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int v = 0; v < 4; ++v)
              acc[mt][v] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[mt][t], bv[s][t][v], acc[mt][v], 0, 0, 0);
      }
      if (rok) {
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
          for (int vv = 0; vv < 4; ++vv) {
            const int m = mc * 64 + mt * 16 + 4 * qq + vv;
            if (m >= n3) continue;
            f32x4 o;
#pragma unroll
            for (int v = 0; v < 4; ++v) o[v] = r + v < nr ? acc[mt][v][vv] : 0.f;
            *reinterpret_cast<f32x4 *>(Lc + ((size_t)b * n3 + m) * ldm + r) = o;
          }
      }
    }
  }
}

// One thread per (panel, axis): x = the centroid's coordinate in the leaf
// box's [-1, 1] frame, w[m] = (2/n) (1/2 + sum_{k=1}^{n-1} T_k(x_m) T_k(x))
// (chebyshev.rs:49-84), the Chebyshev polynomials by their recurrence.
__global__ __launch_bounds__(256) void k_interp_w1d(int N, int n,
                                                    const int *__restrict__ panel_leaf,
                                                    const double *__restrict__ geo,
                                                    const double *__restrict__ cen,
                                                    const double *__restrict__ nodes,
                                                    double *__restrict__ w64,
                                                    float *__restrict__ w32) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= N * 3) return;
  const int i = e / 3, d = e % 3;
  const double *g = geo + (size_t)panel_leaf[i] * 4;
  const double x = (cen[(size_t)i * 3 + d] - g[d]) / g[3];
  double tx[8];
  tx[0] = 1.0;
  tx[1] = x;
  for (int k = 2; k < n; ++k) tx[k] = 2.0 * x * tx[k - 1] - tx[k - 2];
  const double inv_n = 1.0 / n;
  for (int m = 0; m < n; ++m) {
    const double xm = nodes[m];
    double sum = 0.5, tkm1 = 1.0, tk = xm;
    for (int k = 1; k < n; ++k) {
      if (k > 1) {
        const double t = 2.0 * xm * tk - tkm1;
        tkm1 = tk;
        tk = t;
      }
      sum += tk * tx[k];
    }
    const double w = 2.0 * inv_n * sum;
    const size_t o = (size_t)e * n + m;
    if (w64) w64[o] = w;
    if (w32) w32[o] = (float)w;
  }
}

void launch_interp_w1d(int N, int n, const int *panel_leaf, const double *geo, const double *cen,
                       const double *nodes, double *w64, float *w32, hipStream_t s) {
  if (N == 0) return;
  if (n < 2 || n > 8) throw Error(GP_ERR_INTERNAL, "interpolation order out of range");
  hipLaunchKernelGGL(k_interp_w1d, dim3((N * 3 + 255) / 256), dim3(256), 0, s, N, n, panel_leaf,
                     geo, cen, nodes, w64, w32);
  GP_LAUNCH_CHECK();
}

// d/dx of each panel's 1-D interpolation weights (the sums of k_interp_w1d
// differentiated: T_k' = k U_{k-1}), divided by its leaf's half-width: the
// gradient weights of the local expansion in the unit frame.
__global__ __launch_bounds__(256) void k_interp_dw1d(int N, int n,
                                                     const int *__restrict__ panel_leaf,
                                                     const double *__restrict__ geo,
                                                     const double *__restrict__ cen,
                                                     const double *__restrict__ nodes,
                                                     double *__restrict__ dw) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= N * 3) return;
  const int i = e / 3, d = e % 3;
  const double *g = geo + (size_t)panel_leaf[i] * 4;
  const double x = (cen[(size_t)i * 3 + d] - g[d]) / g[3];
  double u[8], dt[8];  // U_k(x), T_k'(x)
  u[0] = 1.0;
  u[1] = 2.0 * x;
  for (int k = 2; k < n; ++k) u[k] = 2.0 * x * u[k - 1] - u[k - 2];
  dt[0] = 0.0;
  for (int k = 1; k < n; ++k) dt[k] = k * u[k - 1];
  for (int m = 0; m < n; ++m) {
    const double xm = nodes[m];
    double s = 0.0, tkm1 = 1.0, tk = xm;
    for (int k = 1; k < n; ++k) {
      if (k > 1) {
        const double t = 2.0 * xm * tk - tkm1;
        tkm1 = tk;
        tk = t;
      }
      s += tk * dt[k];
    }
    dw[(size_t)e * n + m] = 2.0 / n * s / g[3];
  }
}

void launch_interp_dw1d(int N, int n, const int *panel_leaf, const double *geo, const double *cen,
                        const double *nodes, double *dw, hipStream_t s) {
  if (N == 0) return;
  if (n < 2 || n > 8) throw Error(GP_ERR_INTERNAL, "interpolation order out of range");
  hipLaunchKernelGGL(k_interp_dw1d, dim3((N * 3 + 255) / 256), dim3(256), 0, s, N, n, panel_leaf,
                     geo, cen, nodes, dw);
  GP_LAUNCH_CHECK();
}

// The far field of the dielectric formulation, one thread per (panel, RHS):
// a conductor panel takes the potential yfar (from run_far_field), an
// interface panel minus the normal derivative of its leaf's local expansion
// at its centroid, sum over nodes of (n . grad S) L (Lc: [box][node][RHS],
// row stride ldm).
__global__ __launch_bounds__(256) void k_far_dielectric(
    int N, int n, int nr, int ldm, int ldy, const int *__restrict__ panel_leaf,
    const int *__restrict__ leaf_box, int leaf0, const double *__restrict__ w1d,
    const double *__restrict__ dw1d, const double *__restrict__ nrm,
    const unsigned char *__restrict__ iface, const double *__restrict__ Lc,
    const double *__restrict__ yfar, double *__restrict__ y) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)N * nr) return;
  const int p = (int)(e / nr), r = (int)(e % nr);
  const size_t o = (size_t)p * ldy + r;
  if (!iface[p]) {
    y[o] += yfar[o];
    return;
  }
  const int li = panel_leaf[p], n3 = n * n * n;
  const double *ll = Lc + (size_t)(leaf_box ? leaf_box[li] : leaf0 + li) * n3 * ldm;
  const double *w = w1d + (size_t)p * 3 * n, *dw = dw1d + (size_t)p * 3 * n;
  double gx = 0.0, gy = 0.0, gz = 0.0;
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < n; ++b)
      for (int c = 0; c < n; ++c) {
        const double l = ll[(size_t)((a * n + b) * n + c) * ldm + r];
        gx += dw[a] * w[n + b] * w[2 * n + c] * l;
        gy += w[a] * dw[n + b] * w[2 * n + c] * l;
        gz += w[a] * w[n + b] * dw[2 * n + c] * l;
      }
  y[o] -= nrm[3 * p] * gx + nrm[3 * p + 1] * gy + nrm[3 * p + 2] * gz;
}

void launch_far_dielectric(const FmmDev<double> &f, int nr, int ldm, int ldy, const double *dw1d,
                           const double *nrm, const unsigned char *iface, const double *Lc,
                           const double *yfar, double *y, hipStream_t s) {
  const size_t total = (size_t)f.n_panels * nr;
  if (total == 0) return;
  const int leaf0 = f.lvl_off[f.L];
  ProfScope ps("far_dielectric", s, 0.0, 0.0);
  hipLaunchKernelGGL(k_far_dielectric, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                     f.n_panels, f.n, nr, ldm, ldy, f.panel_leaf, f.leaf_box, leaf0, f.w1d, dw1d,
                     nrm, iface, Lc, yfar, y);
  GP_LAUNCH_CHECK();
}

template <typename T>
void run_far_field_t(const FmmDev<T> &f, const T *z, int ldz, int nr, T *M, __bf16 *Mb, T *Lc,
                     int ldm, int m2l, double *y, int ldy, hipStream_t s, T *Mt, __bf16 *Mtb,
                     T *Lt, __bf16 *Mtbl) {
  const int n = f.n, n3 = n * n * n, n3p = f.n3p;
  const int nrc = (nr + FM_RC - 1) / FM_RC;
  const size_t lds1 = (size_t)n3 * FM_RC * sizeof(T);
  const int L = f.L;
  const int leaf0 = f.lvl_off[L];
  const int nleaf = f.leaf_box ? f.n_leaves : f.lvl_off[L + 1] - leaf0;
  const bool f32 = sizeof(T) == 4;
  const double exp_bytes = (double)n3 * nr * sizeof(T);
  const bool want_bf16 = (m2l == GP_M2L_BF16);
  // single right-hand side kernels: target tiles, or pairs grouped by offset
  const bool by_offset = ldm == 1 && nr == 1 && (f.po_off != nullptr || f.tile_tgt != nullptr);
  const bool comp = f.kr > 0;
  if (comp && (!f32 || !Mt || !Lt || !f.U || (want_bf16 && !Mtb) || f.kr % 32 || f.kr > 128))
    throw Error(GP_ERR_INTERNAL, "low-rank M2L needs the FP32 operator, kr in {32, 64, 128} "
                                 "and its work buffers");
  if (want_bf16 && ((comp ? !f.Gcbf : (!Mb || !f.Gbf || n3p % 32)) ||
                    (!by_offset && (ldm % 16 || ldm > 128))))
    throw Error(GP_ERR_INTERNAL, "BF16 M2L needs 16 <= ld <= 128 (multiple of 16) and its tables");
  // BF16 multipoles from the upward pass (the low-rank M2L takes its BF16
  // operands from the projection instead)
  __bf16 *mb = (want_bf16 && !comp) ? Mb : nullptr;
  // Register-tiled kernels for RHS blocks that are a multiple of 16.
  const bool tiled = (ldm % 16 == 0) && ldm <= 128 && n >= 2 && n <= 6;
  const int nrc32 = (nr + TM_RC - 1) / TM_RC;
  {
    ProfScope ps(f32 ? "p2m_f32" : "p2m_f64", s, 2.0 * f.n_panels * n3 * nr * f.frac_panels,
                 ((double)f.n_panels * nr * sizeof(T) + nleaf * exp_bytes) * f.frac_panels);
    if (ldm == 1 && nr == 1 && ldz == 1) {
      // (orders <= 3: half-wavefront slots, two leaves to a wavefront)
      const int tpl = round_up(n3, n3 <= 32 ? 32 : 64), lpb = 256 / tpl;
      // one thread per (leaf, node), or (variant 1) one work-group per leaf
      // with its panels staged in LDS
      // 3, 4: k_p2m1x, k_p2m1mx (FP32, weights from the box coordinates)
      const bool wx = sizeof(T) == 4 && f.xl && f.wc && n >= 3 && n <= 6;
      tuned_launch(f32 ? "p2m1_f32" : "p2m1_f64", nleaf, n, 1,
                   n >= 3 && n <= 6 ? (wx ? 5 : 3) : 1, s, [&](int v) {
        if constexpr (sizeof(T) == 4) {
          if (wx && v >= 3) {
            const float4 *xl = reinterpret_cast<const float4 *>(f.xl);
#define GP_P2M1X(NF)                                                                             \
  if (v == 4)                                                                                    \
    hipLaunchKernelGGL((k_p2m1mx<NF>), dim3((nleaf + 256 / tpl - 1) / (256 / tpl)), dim3(256), 0, \
                       s, n3p, nleaf, f.leaf_lo, f.leaf_box, leaf0, xl, f.wc, z, M, mb,           \
                       f.skip_leaf);                                                              \
  else                                                                                           \
    hipLaunchKernelGGL((k_p2m1x<NF>), dim3((nleaf + lpb - 1) / lpb), dim3(256), 0, s, n3p, tpl,   \
                       nleaf, f.leaf_lo, f.leaf_box, leaf0, xl, f.wc, z, M, mb, f.skip_leaf)
            switch (n) {
              case 3: GP_P2M1X(3); break;
              case 4: GP_P2M1X(4); break;
              case 5: GP_P2M1X(5); break;
              default: GP_P2M1X(6); break;
            }
#undef GP_P2M1X
            return;
          }
        }
        switch (n) {
#define GP_P2M1(NF)                                                                              \
  if (v == 2 && NF > 0)                                                                          \
    hipLaunchKernelGGL((k_p2m1m<T, (NF > 0 ? NF : 3)>), dim3((nleaf + 256 / tpl - 1) / (256 / tpl)), \
                       dim3(256), 0, s, n3p, nleaf, f.leaf_lo, f.leaf_box, leaf0, f.w1d, f.area, z, \
                       M, mb, f.skip_leaf);                                                       \
  else if (v == 1 && NF > 0)                                                                     \
    hipLaunchKernelGGL((k_p2m1s<T, (NF > 0 ? NF : 3)>), dim3(nleaf), dim3(tpl), 0, s, n3p, nleaf, \
                       f.leaf_lo, f.leaf_box, leaf0, f.w1d, f.area, z, M, mb, f.skip_leaf);       \
  else                                                                                           \
    hipLaunchKernelGGL((k_p2m1<T, NF>), dim3((nleaf + lpb - 1) / lpb), dim3(256), 0, s, n, n3p,  \
                       tpl, nleaf, f.leaf_lo, f.leaf_box, leaf0, f.w1d, f.area, z, M, mb,         \
                       f.skip_leaf)
          case 3: GP_P2M1(3); break;
          case 4: GP_P2M1(4); break;
          case 5: GP_P2M1(5); break;
          case 6: GP_P2M1(6); break;
          default: GP_P2M1(0); break;
#undef GP_P2M1
        }
      });
    } else if (tiled) {
      const int MB = n3p / 16;
      if (MB <= 2)
        hipLaunchKernelGGL((k_p2m_tiled<T, 2>), dim3(nleaf), dim3(256), 0, s, n, n3p, f.leaf_lo,
                           f.leaf_box, leaf0, f.w1d, f.area, z, ldz, nr, M, mb, ldm);
      else if (MB <= 4)
        hipLaunchKernelGGL((k_p2m_tiled<T, 4>), dim3(nleaf), dim3(256), 0, s, n, n3p, f.leaf_lo,
                           f.leaf_box, leaf0, f.w1d, f.area, z, ldz, nr, M, mb, ldm);
      else if (MB <= 8)
        hipLaunchKernelGGL((k_p2m_tiled<T, 8>), dim3(nleaf), dim3(256), 0, s, n, n3p, f.leaf_lo,
                           f.leaf_box, leaf0, f.w1d, f.area, z, ldz, nr, M, mb, ldm);
      else
        hipLaunchKernelGGL((k_p2m_tiled<T, 14>), dim3(nleaf), dim3(256), 0, s, n, n3p, f.leaf_lo,
                           f.leaf_box, leaf0, f.w1d, f.area, z, ldz, nr, M, mb, ldm);
    } else {
      hipLaunchKernelGGL((k_p2m<T>), dim3(nleaf, nrc), dim3(256), 0, s, n, n3p, f.leaf_lo, f.leaf_box, leaf0,
                         f.w1d, f.area, z, ldz, nr, M, mb, ldm);
    }
    GP_LAUNCH_CHECK();
  }
  // M2M, finest to level 2 (multipoles above level 2 are never used), over
  // the boxes of each level that have children (listed first in the level);
  // one RHS with subtrees built: one launch, a work-group per subtree.
  const bool by_sub = ldm == 1 && nr == 1 && n3 <= 256 && f.n_sub > 0 && f.sub_root;
  if (by_sub) {
    int npar_all = 0, nch_all = 0;
    for (int l = 2; l < L; ++l) {
      npar_all += f.npar ? f.npar[l] : f.lvl_off[l + 1] - f.lvl_off[l];
      nch_all += f.lvl_off[l + 2] - f.lvl_off[l + 1];
    }
    ProfScope pm(f32 ? "m2m_f32" : "m2m_f64", s, 6.0 * n * n3 * nch_all * f.frac_boxes,
                 (npar_all + nch_all) * exp_bytes * f.frac_boxes);
    const int tpb = round_up(n3, n3 <= 32 ? 32 : 64);
#define GP_M2M1S(NF)                                                                            \
  hipLaunchKernelGGL((k_m2m1_sub<T, NF>), dim3(f.n_sub), dim3(256), 4096 * sizeof(T), s, n, n3p, \
                     tpb, L, f.sub_root, f.sub_leaf, f.sub_rng, f.child_off, f.child_idx,        \
                     f.octant, f.m2m1d, M, mb, f.skip_box)
    switch (n) {
      case 3: GP_M2M1S(3); break;
      case 4: GP_M2M1S(4); break;
      case 5: GP_M2M1S(5); break;
      case 6: GP_M2M1S(6); break;
      default: GP_M2M1S(0); break;
    }
#undef GP_M2M1S
    GP_LAUNCH_CHECK();
  }
  for (int l = L - 1; l >= 2 && !by_sub; --l) {
    int nb = f.npar ? f.npar[l] : f.lvl_off[l + 1] - f.lvl_off[l];
    if (nb == 0) continue;
    int nch = f.lvl_off[l + 2] - f.lvl_off[l + 1];
    ProfScope pm(f32 ? "m2m_f32" : "m2m_f64", s, 6.0 * n * n3 * nr * nch * f.frac_boxes,
                 (nb + nch) * exp_bytes * f.frac_boxes);
    if (ldm == 1 && nr == 1 && n3 <= 256) {  // orders <= 6: one slot fits a work-group
      // (orders <= 3: half-wavefront slots, two boxes to a wavefront)
      const int tpb = round_up(n3, n3 <= 32 ? 32 : 64), bpb = 256 / tpb;
#define GP_M2M1(NF)                                                                             \
  hipLaunchKernelGGL((k_m2m1<T, NF>), dim3((nb + bpb - 1) / bpb), dim3(256), 4096 * sizeof(T), s, \
                     n, n3p, tpb, f.lvl_off[l], nb, f.child_off, f.child_idx, f.octant, f.m2m1d,   \
                     M, mb, f.skip_box)
      switch (n) {
        case 3: GP_M2M1(3); break;
        case 4: GP_M2M1(4); break;
        case 5: GP_M2M1(5); break;
        case 6: GP_M2M1(6); break;
        default: GP_M2M1(0); break;
      }
#undef GP_M2M1
    } else if (tiled) {
      const int b0 = f.lvl_off[l];
      switch (n) {
        case 2: launch_m2m_tiled<T, 2>(n3p, b0, nb, nrc32, f, nr, M, mb, ldm, s); break;
        case 3: launch_m2m_tiled<T, 3>(n3p, b0, nb, nrc32, f, nr, M, mb, ldm, s); break;
        case 4: launch_m2m_tiled<T, 4>(n3p, b0, nb, nrc32, f, nr, M, mb, ldm, s); break;
        case 5: launch_m2m_tiled<T, 5>(n3p, b0, nb, nrc32, f, nr, M, mb, ldm, s); break;
        default: launch_m2m_tiled<T, 6>(n3p, b0, nb, nrc32, f, nr, M, mb, ldm, s); break;
      }
    } else {
      hipLaunchKernelGGL((k_m2m<T>), dim3(nb, nrc), dim3(256), 2 * lds1, s, n, n3p, f.lvl_off[l],
                         f.child_off, f.child_idx, f.octant, f.m2m1d, nr, M, mb, ldm);
    }
    GP_LAUNCH_CHECK();
  }
  // M2L over every box of level >= 2 with a non-empty list; boxes without
  // one keep a zero local expansion.  A low-rank operator (f.kr > 0) runs the
  // M2L on the projected multipoles Mt and locals Lt (kr values per box and
  // RHS column) with the kr x kr tables, then expands Lt into Lc.
  FmmDev<T> fm = f;
  int m3 = n3;
  T *Mo = M, *Lo = Lc;
  __bf16 *mbo = mb;
  // split-BF16 products for a block of RHS on the projected operator
  const bool split = comp && want_bf16 && !by_offset && Mtbl && f.Gcbf_lo;
  // projection and expansion on the FP32 matrix cores: rank 32 or 64, node
  // counts padded to 32, 64, 128 or 224, 16-byte column groups
  const int pns = n3p / 16;
  const bool proj_mfma = comp && sizeof(T) == 4 && (f.kr == 32 || f.kr == 64) &&
                         n3p % 16 == 0 && (pns == 2 || pns == 4 || pns == 8 || pns == 14) &&
                         ldm % 4 == 0;
  __bf16 *mbl = split ? Mtbl : nullptr;
  if (comp) {
    fm.n3p = f.kr;
    fm.G = f.Gc;
    fm.Gbf = f.Gcbf;
    fm.Gr = f.Gcr;
    m3 = f.kr;
    Mo = Mt;
    Lo = Lt;
    mbo = want_bf16 ? Mtb : nullptr;
    const int b0 = f.lvl_off[2], nrow = (f.n_boxes - b0) * nr;
    if (nrow > 0) {
      ProfScope pc("m2l_compress_f32", s, 2.0 * nrow * n3 * f.kr,
                   (double)nrow * (n3 + f.kr) * sizeof(T));
      if constexpr (sizeof(T) == 4) {
        const dim3 grid((nrow + 63) / 64);
        if (proj_mfma) {
          const int ng = (nrow + 15) / 16;
          const dim3 gm((unsigned)std::min(ng / 4 + 1, 2048));
          const size_t lds = (size_t)f.kr * (n3p + 4) * sizeof(float);
#define GP_COMPRESS_MFMA(KR, NS)                                                              \
  hipLaunchKernelGGL((k_m2l_compress_mfma<KR, NS>), gm, dim3(256), lds, s, n3, b0, nrow, nr, \
                     ldm, f.U, M, Mt, mbo, mbl)
          const int ns = n3p / 16;
          if (f.kr == 32) {
            if (ns == 2) GP_COMPRESS_MFMA(32, 2); else if (ns == 4) GP_COMPRESS_MFMA(32, 4);
            else if (ns == 8) GP_COMPRESS_MFMA(32, 8); else GP_COMPRESS_MFMA(32, 14);
          } else {
            if (ns == 2) GP_COMPRESS_MFMA(64, 2); else if (ns == 4) GP_COMPRESS_MFMA(64, 4);
            else if (ns == 8) GP_COMPRESS_MFMA(64, 8); else GP_COMPRESS_MFMA(64, 14);
          }
#undef GP_COMPRESS_MFMA
        } else
        if (f.kr == 32)
          hipLaunchKernelGGL((k_m2l_compress<32>), grid, dim3(256), 0, s, n3, n3p, b0, nrow, nr,
                             ldm, f.U, M, Mt, mbo, mbl);
        else if (f.kr == 64)
          hipLaunchKernelGGL((k_m2l_compress<64>), grid, dim3(256), 0, s, n3, n3p, b0, nrow, nr,
                             ldm, f.U, M, Mt, mbo, mbl);
        else
          hipLaunchKernelGGL((k_m2l_compress<128>), grid, dim3(256), 0, s, n3, n3p, b0, nrow, nr,
                             ldm, f.U, M, Mt, mbo, mbl);
        GP_LAUNCH_CHECK();
      }
    }
  }
  const int m3p = fm.n3p;
  {
    size_t cnt = (size_t)f.n_boxes * n3 * ldm;
    int blocks = (int)std::min<size_t>((cnt + 255) / 256, 8192);
    if (!proj_mfma) {  // the matrix-core expansion writes every box's locals
      hipLaunchKernelGGL((k_fill_zero<T>), dim3(blocks), dim3(256), 0, s, Lc, cnt);
      GP_LAUNCH_CHECK();
    }
    if (comp) {
      cnt = (size_t)f.n_boxes * m3 * ldm;
      blocks = (int)std::min<size_t>((cnt + 255) / 256, 8192);
      hipLaunchKernelGGL((k_fill_zero<T>), dim3(blocks), dim3(256), 0, s, Lt, cnt);
      GP_LAUNCH_CHECK();
    }
  }
  const bool bf16_valu = (m2l == GP_M2L_BF16_VALU);
  if (bf16_valu) {
    if constexpr (sizeof(T) == 4) {
      if (!fm.Gr) throw Error(GP_ERR_INTERNAL, "BF16-rounded M2L table missing");
      size_t cnt = (size_t)f.n_boxes * ldm * m3p;
      int blocks = (int)std::min<size_t>((cnt + 255) / 256, 8192);
      hipLaunchKernelGGL(k_round_bf16, dim3(blocks), dim3(256), 0, s, Mo, cnt);
      GP_LAUNCH_CHECK();
    } else {
      throw Error(GP_ERR_INTERNAL, "BF16 M2L runs on the FP32 operator only");
    }
  }
  if (f.n_targets > 0) {
    const double fl = 2.0 * (double)f.n_pairs * m3 * m3 * nr;
    const double mbytes = (double)m3 * nr * sizeof(T);
    const double by = (double)(f.n_boxes + f.n_targets) * mbytes;
    const int MT = m3p / 16, RT = (nr + 15) / 16;
    if (want_bf16 && by_offset && f.tile_tgt) {
      if constexpr (sizeof(T) == 4) {
        ProfScope pl("m2l1t_bf16_mfma", s, fl * f.frac_pairs,
                     ((double)f.n_pairs * m3p * 2 + f.n_targets * mbytes) * f.frac_pairs);
        tuned_launch("m2l1t_bf16", f.n_tiles, m3p, f.tile_size, 2, s,
                     [&](int v) { launch_m2l_tiles(fm, m3, mbo, Lo, v ? 16 : 0, s); });
      }
    } else if (want_bf16 && by_offset) {
      throw Error(GP_ERR_INTERNAL, "single-RHS BF16 M2L needs its target tiles");
    } else if (m2l == GP_M2L_FP64 && sizeof(T) == 8 && by_offset && f.tile_tgt &&
               f.tile_size == 64 && m3p <= 64) {
      if constexpr (sizeof(T) == 8) {
        ProfScope pl("m2l1t_f64_mfma", s, fl * f.frac_pairs,
                     ((double)f.n_pairs * m3p * 8 + f.n_targets * mbytes) * f.frac_pairs);
        // 2, 3: the truncated blocks applied in factored form, by independent
        // wavefronts; 4-7: the same with the factors staged in LDS (4, 5
        // with the sources one entry ahead in registers, 6, 7 without), where
        // the ranks and the staging buffers fit
        const bool lr_ok = f.lr_vt && (f.lr_kp == 16 || f.lr_kp == 32);
        const size_t lrs_lds =
            2 * ((size_t)f.lr_kmax * (m3p + 2) + (size_t)m3p * (f.lr_kmax + 1)) * sizeof(double);
        const bool lrs_ok = lr_ok && f.lr_kmax >= 1 && f.lr_kmax <= 32 && lrs_lds <= 65536;
        tuned_launch("m2l1t_f64", f.n_tiles, m3p, f.tile_size, lrs_ok ? 8 : lr_ok ? 4 : 2, s,
                     [&](int v) {
          if (v >= 4) launch_m2l_tiles_f64lrs(fm, m3, Mo, Lo, v % 2 ? 16 : 0, v < 6, s);
          else if (v >= 2) launch_m2l_tiles_f64lr(fm, m3, Mo, Lo, v == 3 ? 16 : 0, s);
          else launch_m2l_tiles_f64(fm, m3, Mo, Lo, v ? 16 : 0, s);
        });
      }
    } else if (m2l == GP_M2L_FP64 && sizeof(T) == 8 && by_offset) {
      if (!f.po_off) throw Error(GP_ERR_INTERNAL, "single-RHS M2L pairs by offset not built");
      if constexpr (sizeof(T) == 8) {
        ProfScope pl("m2l1_f64_mfma", s, fl * f.frac_pairs,
                     (double)f.n_pairs * (m3p * 8 + m3 * 16) * f.frac_pairs);
        launch_m2l_by_offset(fm, m3, Mo, Lo, s);
      }
    } else if (split) {
      if constexpr (sizeof(T) == 4) {
        // the algorithm's matrix-core work: three BF16 products per projected pair
        ProfScope pl("m2l_bf16x3_mfma", s, 3.0 * fl,
                     (double)f.n_boxes * m3 * nr * 4 + f.n_targets * mbytes);
        // one, two or four pairs of prefetch (four: rank 32 and 64 only, the
        // register sets of larger tables do not fit)
        tuned_launch("m2l_bf16x3", f.n_targets, MT * 100 + RT, ldm, MT <= 4 ? 3 : 2, s, [&](int v) {
          if (v == 0)
            GP_M2L_DISPATCH(launch_m2l_bf16x3, MT, RT, fm, m3, nr, ldm, mbo, mbl, Lo, 16, s);
          else if (v == 1)
            GP_M2L_DISPATCH(launch_m2l_bf16x3_d2, MT, RT, fm, m3, nr, ldm, mbo, mbl, Lo, 16, s);
          else
            GP_M2L_DISPATCH(launch_m2l_bf16x3_d4, MT, RT, fm, m3, nr, ldm, mbo, mbl, Lo, 16, s);
        });
      }
    } else if (want_bf16) {
      if constexpr (sizeof(T) == 4) {
        ProfScope pl("m2l_bf16_mfma", s, fl, (double)f.n_boxes * m3 * nr * 2 + f.n_targets * mbytes);
        // one target a work-group (XCD runs of 16 or 64, or the hardware
        // order), the RHS tiles split over two work-groups per target (half
        // the registers and LDS, twice the G_off loads), or sibling pairs
        const bool pairs = f.pair_tgt && m3p <= 128;
        const bool split = RT > 4;
        std::vector<int> var = {0, 1, 2, 6};
        if (split) var.push_back(3);
        if (pairs) var.push_back(4), var.push_back(5);
        tuned_launch("m2l_bf16_pipe", f.n_targets, MT * 100 + RT, ldm, (int)var.size(), s,
                     [&](int vi) {
          const int v = var[vi];
          if (v == 6) {  // source rows two pairs ahead
            GP_M2L_DISPATCH(launch_m2l_bf16_d2, MT, RT, fm, m3, nr, ldm, mbo, Lo, 16, s);
          } else if (v >= 4) {
            launch_m2l_pairs(fm, m3, nr, ldm, mbo, Lo, v == 4 ? 16 : 0, s);
          } else if (v == 3) {
            GP_M2L_DISPATCH(launch_m2l_bf16, MT, 4, fm, m3, nr, ldm, mbo, Lo, 16, s);
          } else {
            const int xr = v == 0 ? 16 : v == 1 ? 0 : 64;
            GP_M2L_DISPATCH(launch_m2l_bf16, MT, RT, fm, m3, nr, ldm, mbo, Lo, xr, s);
          }
        });
      } else {
        throw Error(GP_ERR_INTERNAL, "BF16 M2L runs on the FP32 operator only");
      }
    } else if (m2l == GP_M2L_FP64 && sizeof(T) == 8 && ldm % 16 == 0 && ldm <= 128) {
      if constexpr (sizeof(T) == 8) {
        ProfScope pl("m2l_f64_mfma", s, fl, by);
        tuned_launch("m2l_f64_pipe", f.n_targets, MT * 100 + RT, ldm, 3, s, [&](int v) {
          const int xr = v == 0 ? 16 : v == 1 ? 0 : 64;
          GP_M2L_DISPATCH(launch_m2l_f64, MT, RT, fm, m3, nr, ldm, Mo, Lo, xr, s);
        });
      }
    } else {
      ProfScope pl(bf16_valu ? "m2l_bf16_valu" : f32 ? "m2l_f32" : "m2l_f64", s, fl, by);
      hipLaunchKernelGGL((k_m2l<T>), dim3(f.n_targets, nrc), dim3(256),
                         (size_t)m3 * FM_RC * sizeof(T), s, m3, m3p, f.targets, f.il_off, f.il_src,
                         f.il_oidx, f.box_level, f.inv_half, bf16_valu ? fm.Gr : fm.G, nr, Mo, Lo,
                         ldm);
    }
    GP_LAUNCH_CHECK();
    if (comp) {
      const int ncol = f.n_targets * nr;
      ProfScope pe("m2l_expand_f32", s, 2.0 * ncol * n3 * f.kr,
                   (double)ncol * (n3 + f.kr) * sizeof(T));
      if constexpr (sizeof(T) == 4) {
        if (proj_mfma) {
          const int ngr = (ldm + 63) / 64, nmc = (n3 + 63) / 64;
          const int items = f.n_boxes * ngr;
          const dim3 gm((unsigned)std::min(items / 4 + 1, 2048));
          const size_t lds = (size_t)((n3 + 15) & ~15) * (f.kr + 4) * sizeof(float);
          if (f.kr == 32)
            hipLaunchKernelGGL((k_m2l_expand_mfma<32>), gm, dim3(256), lds, s, n3, nmc,
                               f.n_boxes, ngr, nr, ldm, f.U, Lt, Lc);
          else
            hipLaunchKernelGGL((k_m2l_expand_mfma<64>), gm, dim3(256), lds, s, n3, nmc,
                               f.n_boxes, ngr, nr, ldm, f.U, Lt, Lc);
        } else {
          hipLaunchKernelGGL(k_m2l_expand, dim3((ncol + 63) / 64, (n3 + 63) / 64), dim3(256), 0,
                             s, n3, f.kr, ncol, nr, ldm, f.targets, f.U, Lt, Lc);
        }
        GP_LAUNCH_CHECK();
      }
    }
  }
  // The finest L2L fused with L2P (FP32 blocks of RHS, orders up to 5, whose
  // work-group needs 41 KB of LDS).  One tree (no batch leaf map) has every
  // leaf at the finest level, the layout the fused kernel walks.
  const bool fuse_leaf = sizeof(T) == 4 && tiled && !f.leaf_box && f.panel_leaf && ldm % 4 == 0 &&
                         ldm >= 16 && n >= 3 && n <= 5 && L >= 3 && nr > 1 && f.fuse_leaf;
  // L2L, level 2 to the leaves (one RHS with subtrees built: one launch, a
  // work-group per subtree).
  if (by_sub) {
    int nch_all = 0;
    for (int l = 2; l < L; ++l) nch_all += f.lvl_off[l + 2] - f.lvl_off[l + 1];
    ProfScope pm(f32 ? "l2l_f32" : "l2l_f64", s, 6.0 * n * n3 * nch_all * f.frac_boxes,
                 2.0 * nch_all * exp_bytes * f.frac_boxes);
    const int tpb = round_up(n3, n3 <= 32 ? 32 : 64);
#define GP_L2L1S(NF)                                                                            \
  hipLaunchKernelGGL((k_l2l1_sub<T, NF>), dim3(f.n_sub), dim3(256), 512 * sizeof(T), s, n, tpb, \
                     L, f.sub_root, f.sub_leaf, f.sub_rng, f.parent, f.octant, f.m2m1d, Lc,      \
                     f.skip_box)
    switch (n) {
      case 3: GP_L2L1S(3); break;
      case 4: GP_L2L1S(4); break;
      case 5: GP_L2L1S(5); break;
      case 6: GP_L2L1S(6); break;
      default: GP_L2L1S(0); break;
    }
#undef GP_L2L1S
    GP_LAUNCH_CHECK();
  }
  for (int l = 2; l < L && !by_sub; ++l) {
    int nb = f.lvl_off[l + 2] - f.lvl_off[l + 1];
    if (nb == 0) continue;
    if (fuse_leaf && l == L - 1) continue;  // in k_l2l_l2p_leaf
    ProfScope pm(f32 ? "l2l_f32" : "l2l_f64", s, 6.0 * n * n3 * nr * nb * f.frac_boxes,
                 2.0 * nb * exp_bytes * f.frac_boxes);
    if (ldm == 1 && nr == 1 && n3 <= 256) {  // orders <= 6: one slot fits a work-group
      // (orders <= 3: half-wavefront slots, two boxes to a wavefront)
      const int tpb = round_up(n3, n3 <= 32 ? 32 : 64), bpb = 256 / tpb;
#define GP_L2L1(NF)                                                                            \
  hipLaunchKernelGGL((k_l2l1<T, NF>), dim3((nb + bpb - 1) / bpb), dim3(256), 512 * sizeof(T), s, \
                     n, tpb, f.lvl_off[l + 1], nb, f.parent, f.octant, f.m2m1d, Lc, f.skip_box)
      switch (n) {
        case 3: GP_L2L1(3); break;
        case 4: GP_L2L1(4); break;
        case 5: GP_L2L1(5); break;
        case 6: GP_L2L1(6); break;
        default: GP_L2L1(0); break;
      }
#undef GP_L2L1
    } else if (tiled) {
      const int b0 = f.lvl_off[l + 1];
      switch (n) {
        case 2: launch_l2l_tiled<T, 2>(b0, nb, nrc32, f, nr, Lc, ldm, s); break;
        case 3: launch_l2l_tiled<T, 3>(b0, nb, nrc32, f, nr, Lc, ldm, s); break;
        case 4: launch_l2l_tiled<T, 4>(b0, nb, nrc32, f, nr, Lc, ldm, s); break;
        case 5: launch_l2l_tiled<T, 5>(b0, nb, nrc32, f, nr, Lc, ldm, s); break;
        default: launch_l2l_tiled<T, 6>(b0, nb, nrc32, f, nr, Lc, ldm, s); break;
      }
    } else {
      hipLaunchKernelGGL((k_l2l<T>), dim3(nb, nrc), dim3(256), 2 * lds1, s, n, f.lvl_off[l + 1],
                         f.parent, f.octant, f.m2m1d, nr, Lc, ldm);
    }
    GP_LAUNCH_CHECK();
  }
  if (fuse_leaf) {
    if constexpr (sizeof(T) == 4) {
      ProfScope pf("l2l_l2p_f32", s, 2.0 * f.n_panels * n3 * nr + 6.0 * n * n3 * nr * nleaf,
                   (double)f.n_panels * nr * 8.0 + 2.0 * nleaf * exp_bytes);
      switch (n) {
#define GP_L2LP(NF)                                                                               \
  hipLaunchKernelGGL((k_l2l_l2p_leaf<NF>), dim3(nleaf), dim3(256), 0, s, leaf0, f.parent,        \
                     f.octant, f.m2m1d, f.leaf_lo, f.w1d, nr, Lc, ldm, y, ldy)
        case 3: GP_L2LP(3); break;
        case 4: GP_L2LP(4); break;
        default: GP_L2LP(5); break;
#undef GP_L2LP
      }
      GP_LAUNCH_CHECK();
    }
    return;
  }
  ProfScope pp(f32 ? "l2p_f32" : "l2p_f64", s, 2.0 * f.n_panels * n3 * nr * f.frac_panels,
               ((double)f.n_panels * nr * 16.0 + nleaf * exp_bytes) * f.frac_panels);
  // (L2P accumulates into y, so it is not a tuner candidate: the tuner
  // launches every variant on the same outputs)
  if (ldm == 1 && nr == 1 && ldy == 1 && n == 4 && sizeof(T) == 4 && f.panel_leaf) {
    if constexpr (sizeof(T) == 4)
      hipLaunchKernelGGL(k_l2p1v4, dim3((f.n_panels + 255) / 256), dim3(256), 0, s, f.n_panels,
                         f.panel_leaf, f.leaf_box, leaf0, f.w1d, Lc, y, f.skip_leaf);
  } else if (ldm == 1 && nr == 1 && ldy == 1 && f.panel_leaf && n >= 2 && n <= 6) {
    const dim3 grid((f.n_panels + 255) / 256);
    // (FP32 with box coordinates: the weights formed on the fly)
    bool done = false;
    if constexpr (sizeof(T) == 4) {
      if (f.xl && f.wc && n >= 3) {
        const float4 *xl = reinterpret_cast<const float4 *>(f.xl);
#define GP_L2P1RX(NF)                                                                             \
  hipLaunchKernelGGL((k_l2p1rx<NF>), grid, dim3(256), 0, s, f.n_panels, f.panel_leaf, f.leaf_box, \
                     leaf0, xl, f.wc, Lc, y, f.skip_leaf)
        switch (n) {
          case 3: GP_L2P1RX(3); break;
          case 4: GP_L2P1RX(4); break;
          case 5: GP_L2P1RX(5); break;
          default: GP_L2P1RX(6); break;
        }
#undef GP_L2P1RX
        done = true;
      }
    }
#define GP_L2P1R(NF)                                                                              \
  hipLaunchKernelGGL((k_l2p1r<T, NF>), grid, dim3(256), 0, s, f.n_panels, f.panel_leaf,          \
                     f.leaf_box, leaf0, f.w1d, Lc, y, f.skip_leaf)
    if (!done) switch (n) {
      case 2: GP_L2P1R(2); break;
      case 3: GP_L2P1R(3); break;
      case 4: GP_L2P1R(4); break;
      case 5: GP_L2P1R(5); break;
      default: GP_L2P1R(6); break;
    }
#undef GP_L2P1R
  } else if (ldm == 1 && nr == 1 && ldy == 1 && n3 <= 64)
    hipLaunchKernelGGL((k_l2p1w<T>), dim3((nleaf + 3) / 4), dim3(256), 0, s, n, nleaf, f.leaf_lo,
                       f.leaf_box, leaf0, f.w1d, Lc, y, f.skip_leaf);
  else if (ldm == 1 && nr == 1 && ldy == 1 && f.panel_leaf)
    hipLaunchKernelGGL((k_l2p1<T>), dim3((f.n_panels + 255) / 256), dim3(256), 0, s, n, f.n_panels,
                       f.panel_leaf, f.leaf_box, leaf0, f.w1d, Lc, y, f.skip_leaf);
  else if (tiled && sizeof(T) == 4 && n >= 3 && n <= 6 && ldm % 4 == 0 && ldm >= 16) {
    if constexpr (sizeof(T) == 4) {
      // points per lane: 16 / (64 / (ldm / 4)) rounded up
      const int pg = 64 / (ldm / 4), ppl = (16 + pg - 1) / pg;
#define GP_L2P_WAVE(N, P)                                                                         \
  hipLaunchKernelGGL((k_l2p_wave<N, P>), dim3(nleaf), dim3(64), 0, s, f.leaf_lo, f.leaf_box, leaf0, \
                     f.w1d, nr, Lc, ldm, y, ldy)
#define GP_L2P_WAVE_N(N)                      \
  do {                                        \
    if (ppl <= 1) GP_L2P_WAVE(N, 1);          \
    else if (ppl <= 2) GP_L2P_WAVE(N, 2);     \
    else if (ppl <= 4) GP_L2P_WAVE(N, 4);     \
    else GP_L2P_WAVE(N, 8);                   \
  } while (0)
      switch (n) {
        case 3: GP_L2P_WAVE_N(3); break;
        case 4: GP_L2P_WAVE_N(4); break;
        case 5: GP_L2P_WAVE_N(5); break;
        default: GP_L2P_WAVE_N(6); break;
      }
#undef GP_L2P_WAVE_N
#undef GP_L2P_WAVE
    }
  } else if (tiled)
    hipLaunchKernelGGL((k_l2p_tiled<T>), dim3(nleaf), dim3(256), 0, s, n, f.leaf_lo, f.leaf_box, leaf0, f.w1d,
                       nr, Lc, ldm, y, ldy);
  else
    hipLaunchKernelGGL((k_l2p<T>), dim3(nleaf, nrc), dim3(256), lds1, s, n, f.leaf_lo, f.leaf_box, leaf0,
                       f.w1d, nr, Lc, ldm, y, ldy);
  GP_LAUNCH_CHECK();
}

void run_far_field(const FmmDev<double> &f, const double *z, int ldz, int nr, double *M,
                   double *Lc, int ldm, int m2l, double *y, int ldy, hipStream_t s) {
  run_far_field_t<double>(f, z, ldz, nr, M, nullptr, Lc, ldm, m2l, y, ldy, s, nullptr, nullptr,
                          nullptr, nullptr);
}
void run_far_field(const FmmDev<float> &f, const float *z, int ldz, int nr, float *M, void *Mb,
                   float *Lc, int ldm, int m2l, double *y, int ldy, hipStream_t s, float *Mt,
                   void *Mtb, float *Lt, void *Mtbl) {
  run_far_field_t<float>(f, z, ldz, nr, M, reinterpret_cast<__bf16 *>(Mb), Lc, ldm, m2l, y, ldy,
                         s, Mt, reinterpret_cast<__bf16 *>(Mtb), Lt,
                         reinterpret_cast<__bf16 *>(Mtbl));
}

}  // namespace gp
