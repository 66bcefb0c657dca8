// Near-field operator apply: the exact-Wilton leaf tiles of the precorrected
// FMM operator (cap/fmm_solver.rs:116-128) times a block of right-hand sides.
//
// Layout: tile of target leaf T is row-major nt x S at nf_off[T]; its S
// columns are panels cols[col_off[T] ..] (the leaf and its neighbours, runs of
// consecutive sorted panels).  Vectors are N x ld row-major with the RHS index
// fastest.  The output y (FP64) is written, not accumulated: the far field is
// added afterwards by the L2P kernel.
//
// Multi-RHS (k_near_spmm): one work-group per (target leaf, 16-column RHS
// chunk); 64-row target tiles x 32-column source chunks staged in LDS, each
// thread owns one target row and 4 RHS columns.  Single RHS
// (k_near_spmv_leaf): one work-group per target leaf gathers the leaf's S
// source values into LDS once, then each wavefront streams whole tile rows
// with 16-byte loads; k_near_spmv (one wavefront per row, x gathered per row)
// covers tiles wider than the LDS budget.

#include <type_traits>

#include "gp_common.h"
#include "gp_kernels.h"

namespace gp {

template <typename T>
struct Acc;
template <>
struct Acc<float> {
  using type = float;
};
template <>
struct Acc<double> {
  using type = double;
};

constexpr int NF_TM = 64;  // target rows per tile
constexpr int NF_KS = 32;  // source columns per LDS stage
constexpr int NF_RC = 16;  // RHS columns per work-group
constexpr int NF_CPT = 4;  // RHS columns per thread
constexpr int XCD_RUN_NEAR = 16;  // consecutive (leaf, RHS chunk) blocks per XCD run

template <typename TV, typename TX>
__global__ __launch_bounds__(256) void k_near_spmm(
    const int *__restrict__ leaf_lo, const int64_t *__restrict__ nf_off,
    const int *__restrict__ col_off, const int *__restrict__ cols, const TV *__restrict__ vals,
    const TX *__restrict__ x, int ldx, double *__restrict__ y, int ldy, int nr) {
  using A = typename Acc<TX>::type;
  __shared__ TV sA[NF_TM][NF_KS + 1];
  __shared__ TX sB[NF_KS][NF_RC];
  // logical block = (leaf, RHS chunk), chunks fastest: the chunks of one
  // leaf run together on one XCD and read its tile from that L2
  const int nch = (nr + NF_RC - 1) / NF_RC;
  const int lb = xcd_block(blockIdx.x, gridDim.x, XCD_RUN_NEAR);
  const int T = lb / nch;
  const int r0 = (lb % nch) * NF_RC;
  const int t0 = leaf_lo[T], nt = leaf_lo[T + 1] - t0;
  const int c0 = col_off[T], S = col_off[T + 1] - c0;
  const int64_t base = nf_off[T];
  const int tid = threadIdx.x;
  const int row = tid >> 2, cg = (tid & 3) * NF_CPT;
  for (int tm = 0; tm < nt; tm += NF_TM) {
    A acc[NF_CPT];
#pragma unroll
    for (int c = 0; c < NF_CPT; ++c) acc[c] = 0;
    for (int ks = 0; ks < S; ks += NF_KS) {
      // stage A: NF_TM x NF_KS
      for (int e = tid; e < NF_TM * NF_KS; e += 256) {
        int rr = e / NF_KS, ss = e % NF_KS;
        int t = tm + rr, s = ks + ss;
        sA[rr][ss] = (t < nt && s < S) ? vals[base + (int64_t)t * S + s] : (TV)0;
      }
      // stage B: NF_KS x NF_RC
      for (int e = tid; e < NF_KS * NF_RC; e += 256) {
        int ss = e / NF_RC, cc = e % NF_RC;
        int s = ks + ss, r = r0 + cc;
        sB[ss][cc] = (s < S && r < nr) ? x[(int64_t)cols[c0 + s] * ldx + r] : (TX)0;
      }
      __syncthreads();
      int kmax = min(NF_KS, S - ks);
      for (int s = 0; s < kmax; ++s) {
        A a = (A)sA[row][s];
#pragma unroll
        for (int c = 0; c < NF_CPT; ++c) acc[c] += a * (A)sB[s][cg + c];
      }
      __syncthreads();
    }
    int t = tm + row;
    if (t < nt) {
#pragma unroll
      for (int c = 0; c < NF_CPT; ++c) {
        int r = r0 + cg + c;
        if (r < nr) y[(int64_t)(t0 + t) * ldy + r] = (double)acc[c];
      }
    }
  }
}

// One wavefront per target row (sorted panel i), single RHS.
template <typename TV, typename TX>
__global__ __launch_bounds__(256) void k_near_spmv(
    int n, const int *__restrict__ panel_leaf, const int *__restrict__ leaf_lo,
    const int64_t *__restrict__ nf_off, const int *__restrict__ col_off,
    const int *__restrict__ cols, const TV *__restrict__ vals, const TX *__restrict__ x,
    double *__restrict__ y, const unsigned char *__restrict__ skip,
    const float *__restrict__ rscale) {
  using A = typename Acc<TX>::type;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= n) return;
  const int T = panel_leaf[i];
  if (skip && skip[T]) return;  // wavefront-uniform
  const int t = i - leaf_lo[T];
  const int c0 = col_off[T], S = col_off[T + 1] - c0;
  const TV *row = vals + nf_off[T] + (int64_t)t * S;
  A acc = 0;
  for (int s = lane; s < S; s += 64) acc += (A)row[s] * (A)x[cols[c0 + s]];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off);
  if (lane == 0) y[i] = rscale ? (double)acc * (double)rscale[i] : (double)acc;
}

// Multi-RHS, register-tiled: one work-group per (target leaf, RHS chunk of
// 16 CPT columns); 64-row x 32-source tiles of the leaf's values (stored
// source-major in LDS) and 32 x 16 CPT slabs of x; thread (ty, tx) owns rows
// 4 ty .. 4 ty + 3 and columns CPT tx .. CPT tx + CPT - 1, read as vectors.
template <typename T, int W>
struct VecN {
  typedef T type __attribute__((ext_vector_type(W)));
};
template <typename TV, typename TX, int CPT>
__global__ __launch_bounds__(256) void k_near_spmm_rt(
    const int *__restrict__ leaf_lo, const int64_t *__restrict__ nf_off,
    const int *__restrict__ col_off, const int *__restrict__ cols, const TV *__restrict__ vals,
    const TX *__restrict__ x, int ldx, double *__restrict__ y, int ldy, int nr) {
  using A = typename Acc<TX>::type;
  using V4 = typename VecN<A, 4>::type;
  using VC = typename VecN<A, CPT>::type;
  constexpr int RC = 16 * CPT;
  __shared__ __attribute__((aligned(32))) A sA[32][64 + 4];  // [source][row]
  __shared__ __attribute__((aligned(32))) A sB[32][RC];      // [source][column]
  const int nch = (nr + RC - 1) / RC;
  const int lb = xcd_block(blockIdx.x, gridDim.x, XCD_RUN_NEAR);
  const int T = lb / nch, r0 = (lb % nch) * RC;
  const int t0 = leaf_lo[T], nt = leaf_lo[T + 1] - t0;
  const int c0 = col_off[T], S = col_off[T + 1] - c0;
  const int64_t base = nf_off[T];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  for (int tm = 0; tm < nt; tm += 64) {
    A acc[4][CPT];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int c = 0; c < CPT; ++c) acc[i][c] = 0;
    for (int ks = 0; ks < S; ks += 32) {
      for (int e = threadIdx.x; e < 64 * 32; e += 256) {
        const int rr = e >> 5, ss = e & 31, t = tm + rr, s = ks + ss;
        sA[ss][rr] = (t < nt && s < S) ? (A)vals[base + (int64_t)t * S + s] : (A)0;
      }
      for (int e = threadIdx.x; e < 32 * RC; e += 256) {
        const int ss = e / RC, cc = e % RC, s = ks + ss, r = r0 + cc;
        sB[ss][cc] = (s < S && r < nr) ? (A)x[(int64_t)cols[c0 + s] * ldx + r] : (A)0;
      }
      __syncthreads();
#pragma unroll 8
      for (int s = 0; s < 32; ++s) {
        const V4 a = *reinterpret_cast<const V4 *>(&sA[s][4 * ty]);
        const VC b = *reinterpret_cast<const VC *>(&sB[s][CPT * tx]);
#pragma unroll
        for (int c = 0; c < CPT; ++c) {
          acc[0][c] += a[0] * b[c];
          acc[1][c] += a[1] * b[c];
          acc[2][c] += a[2] * b[c];
          acc[3][c] += a[3] * b[c];
        }
      }
      __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int t = tm + 4 * ty + i;
      if (t >= nt) continue;
#pragma unroll
      for (int c = 0; c < CPT; ++c) {
        const int r = r0 + CPT * tx + c;
        if (r < nr) y[(int64_t)(t0 + t) * ldy + r] = (double)acc[i][c];
      }
    }
  }
}

// Multi-RHS on the FP32 matrix cores (v_mfma_f32_16x16x4f32: lane holds
// A[i = lane % 16][k = lane / 16], B[k = lane / 16][j = lane % 16],
// D[4 (lane / 16) + v][lane % 16]).  One wavefront per (target leaf, 64-column
// RHS chunk), four per work-group, no LDS and no barriers: D (rows x columns)
// = tile (rows x sources) * x[cols] (sources x columns) in 16-row tiles, RTM
// per pass and only as many as the leaf has rows (leaves average a few panels, so fixed 64-row
// tiles would mostly multiply padding).  Per step of 16 sources, k-slot q of
// MFMA t is source 16 step + 4 q + t; tile v holds columns r0 + 4 j + v, so
// one 16-byte load of an x row feeds four tiles.  The source indices are
// loaded two steps ahead (one per lane, spread by shuffles) and the x rows
// and values one step ahead of the products.
typedef float nf_f32x4 __attribute__((ext_vector_type(4)));
template <int RTM, int DEPTH = 1>  // row tiles per pass (16 RTM rows), steps of prefetch
__global__ __launch_bounds__(256) void k_near_spmm_mfma(
    int nitems, const int *__restrict__ leaf_lo, const int64_t *__restrict__ nf_off,
    const int *__restrict__ col_off, const int *__restrict__ cols, const float *__restrict__ vals,
    const float *__restrict__ x, int ldx, double *__restrict__ y, int ldy, int nr) {
  const int lane = threadIdx.x & 63, j = lane & 15, kq = lane >> 4;
  const int nch = (nr + 63) >> 6;
  const int w = xcd_block(blockIdx.x, gridDim.x, XCD_RUN_NEAR / 4) * 4 + (threadIdx.x >> 6);
  if (w >= nitems) return;  // whole wavefront; the kernel has no barriers
  const int T = w / nch, r0 = (w % nch) * 64;
  const int t0 = leaf_lo[T], nt = leaf_lo[T + 1] - t0;
  const int c0 = col_off[T], S = col_off[T + 1] - c0;
  const float *tile = vals + nf_off[T];
  const int rc = r0 + 4 * j;
  const bool cok = rc < ldx;
  const int nsteps = (S + 15) >> 4;
  for (int tm = 0; tm < nt; tm += 16 * RTM) {
    const int rtn = min(RTM, (nt - tm + 15) >> 4);
    nf_f32x4 acc[RTM][4];
#pragma unroll
    for (int rt = 0; rt < RTM; ++rt)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[rt][v] = nf_f32x4{0.f, 0.f, 0.f, 0.f};
    auto load_col = [&](int st) {
      const int s = 16 * st + (lane & 15);
      return (st < nsteps && s < S) ? cols[c0 + s] : -1;
    };
    // x rows and values of step st; src holds that step's source indices
    // (one per lane, lane l: source 16 st + l % 16)
    auto load_step = [&](int st, int src, nf_f32x4 (&xb)[4], float (&av)[RTM][4]) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int sidx = __shfl(src, 4 * kq + u);
        xb[u] = (cok && sidx >= 0) ? *reinterpret_cast<const nf_f32x4 *>(x + (int64_t)sidx * ldx + rc)
                                   : nf_f32x4{0.f, 0.f, 0.f, 0.f};
      }
      const int s = 16 * st + 4 * kq;
#pragma unroll
      for (int rt = 0; rt < RTM; ++rt) {
        const int row = tm + 16 * rt + j;
        const bool rok = rt < rtn && row < nt;
        const float *vr = tile + (int64_t)row * S + s;
#pragma unroll
        for (int u = 0; u < 4; ++u) av[rt][u] = (rok && s + u < S) ? vr[u] : 0.f;
      }
    };
    auto mma = [&](const nf_f32x4 (&xb)[4], const float (&av)[RTM][4]) {
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int rt = 0; rt < RTM; ++rt)
          if (rt < rtn)
#pragma unroll
            for (int v = 0; v < 4; ++v)
              acc[rt][v] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[rt][u], xb[u][v], acc[rt][v], 0, 0, 0);
    };
    if constexpr (DEPTH == 1) {
      nf_f32x4 xb[4], xn[4];
      float av[RTM][4], an[RTM][4];
      int cnext = load_col(1);
      load_step(0, load_col(0), xb, av);
      for (int st = 0; st < nsteps; ++st) {
        const int cnn = load_col(st + 2);
        if (st + 1 < nsteps) load_step(st + 1, cnext, xn, an);
        mma(xb, av);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          xb[u] = xn[u];
#pragma unroll
          for (int rt = 0; rt < RTM; ++rt) av[rt][u] = an[rt][u];
        }
        cnext = cnn;
      }
    } else {
      // two register sets: the x rows and values of step st + 2 load into
      // the set step st has just used; source indices three steps ahead
      nf_f32x4 x0[4], x1[4];
      float a0[RTM][4], a1[RTM][4];
      int c2 = load_col(2);
      load_step(0, load_col(0), x0, a0);
      if (1 < nsteps) load_step(1, load_col(1), x1, a1);
      int st = 0;
      for (; st + 1 < nsteps; st += 2) {
        const int c3 = load_col(st + 3);
        mma(x0, a0);
        if (st + 2 < nsteps) load_step(st + 2, c2, x0, a0);
        const int c4 = load_col(st + 4);
        mma(x1, a1);
        if (st + 3 < nsteps) load_step(st + 3, c3, x1, a1);
        c2 = c4;
      }
      if (st < nsteps) mma(x0, a0);
    }
    // lane (j, kq) of tile (rt, v): row tm + 16 rt + 4 kq + vv, column rc + v
#pragma unroll
    for (int rt = 0; rt < RTM; ++rt)
#pragma unroll
      for (int vv = 0; vv < 4; ++vv) {
        const int t = tm + 16 * rt + 4 * kq + vv;
        if (rt >= rtn || t >= nt) continue;
        double *yr = y + (int64_t)(t0 + t) * ldy + rc;
#pragma unroll
        for (int v = 0; v < 4; ++v)
          if (rc + v < nr) yr[v] = (double)acc[rt][v][vv];
      }
  }
}

// One lane's vector load of tile values: 16 bytes of FP32 or FP64, 8 bytes
// of FP16 (as many values per lane as FP32, so a row keeps as many lanes
// busy, at half the bytes)
template <typename TV>
struct Vec16;
template <>
struct Vec16<float> {
  typedef float type __attribute__((ext_vector_type(4)));
  static constexpr int n = 4;
};
template <>
struct Vec16<double> {
  typedef double type __attribute__((ext_vector_type(2)));
  static constexpr int n = 2;
};
template <>
struct Vec16<_Float16> {
  typedef _Float16 type __attribute__((ext_vector_type(4)));
  static constexpr int n = 4;
};

// One work-group (WPB wavefronts) per target leaf, single RHS: xs = x[cols]
// in LDS, then wavefront w takes rows w, w + WPB, ...; a row is an unaligned
// head (< VN values), 16-byte aligned vectors, and a tail.
template <typename TV, typename TX, int WPB>
__global__ __launch_bounds__(64 * WPB) void k_near_spmv_leaf(
    const int *__restrict__ leaf_lo, const int64_t *__restrict__ nf_off,
    const int *__restrict__ col_off, const int *__restrict__ cols, const TV *__restrict__ vals,
    const TX *__restrict__ x, double *__restrict__ y, const unsigned char *__restrict__ skip,
    const float *__restrict__ rscale) {
  using A = typename Acc<TX>::type;
  using V = typename Vec16<TV>::type;
  constexpr int VN = Vec16<TV>::n;
  extern __shared__ unsigned char smem_raw[];
  A *xs = reinterpret_cast<A *>(smem_raw);
  const int T = blockIdx.x;
  if (skip && skip[T]) return;  // work-group-uniform, before the barrier
  const int t0 = leaf_lo[T], nt = leaf_lo[T + 1] - t0;
  const int c0 = col_off[T], S = col_off[T + 1] - c0;
  const int64_t base = nf_off[T];
  const int lane = threadIdx.x & 63;
  for (int s = threadIdx.x; s < S; s += 64 * WPB) xs[s] = (A)x[cols[c0 + s]];
  __syncthreads();
  for (int t = threadIdx.x >> 6; t < nt; t += WPB) {
    const int64_t st = base + (int64_t)t * S;
    const TV *row = vals + st;
    const int head = min((int)((VN - (st & (VN - 1))) & (VN - 1)), S);
    const int nv = (S - head) / VN;
    A acc = 0;
    if (lane < head) acc += (A)row[lane] * xs[lane];
    const V *rv = reinterpret_cast<const V *>(row + head);
#pragma unroll 4
    for (int q = lane; q < nv; q += 64) {
      const V v = rv[q];
      const A *xq = xs + head + q * VN;
#pragma unroll
      for (int k = 0; k < VN; ++k) acc += (A)v[k] * xq[k];
    }
    const int tl = head + nv * VN + lane;
    if (tl < S) acc += (A)row[tl] * xs[tl];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off);
    if (lane == 0) y[t0 + t] = rscale ? (double)acc * (double)rscale[t0 + t] : (double)acc;
  }
}

// k_near_spmv_leaf with the values issued early: each wavefront loads the
// first 16-byte vector (and the unaligned head and tail) of its first R rows
// before the x gather, so the value loads overlap the column -> x -> LDS
// chain instead of following it.  Same sums in the same order.
template <typename TV, typename TX, int WPB, int R>
__global__ __launch_bounds__(64 * WPB) void k_near_spmv_leaf_pf(
    const int *__restrict__ leaf_lo, const int64_t *__restrict__ nf_off,
    const int *__restrict__ col_off, const int *__restrict__ cols, const TV *__restrict__ vals,
    const TX *__restrict__ x, double *__restrict__ y, const unsigned char *__restrict__ skip,
    const float *__restrict__ rscale) {
  using A = typename Acc<TX>::type;
  using V = typename Vec16<TV>::type;
  constexpr int VN = Vec16<TV>::n;
  extern __shared__ unsigned char smem_raw[];
  A *xs = reinterpret_cast<A *>(smem_raw);
  const int T = blockIdx.x;
  if (skip && skip[T]) return;  // work-group-uniform, before the barrier
  const int t0 = leaf_lo[T], nt = leaf_lo[T + 1] - t0;
  const int c0 = col_off[T], S = col_off[T + 1] - c0;
  const int64_t base = nf_off[T];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  V pv[R];
  TV ph[R], pt[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int t = w + r * WPB;
#pragma unroll
    for (int k = 0; k < VN; ++k) pv[r][k] = 0;
    ph[r] = 0;
    pt[r] = 0;
    if (t < nt) {
      const int64_t st = base + (int64_t)t * S;
      const TV *row = vals + st;
      const int head = min((int)((VN - (st & (VN - 1))) & (VN - 1)), S);
      const int nv = (S - head) / VN;
      if (lane < head) ph[r] = row[lane];
      if (lane < nv) pv[r] = reinterpret_cast<const V *>(row + head)[lane];
      const int tl = head + nv * VN + lane;
      if (tl < S) pt[r] = row[tl];
    }
  }
  for (int s = threadIdx.x; s < S; s += 64 * WPB) xs[s] = (A)x[cols[c0 + s]];
  __syncthreads();
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int t = w + r * WPB;
    if (t < nt) {
      const int64_t st = base + (int64_t)t * S;
      const TV *row = vals + st;
      const int head = min((int)((VN - (st & (VN - 1))) & (VN - 1)), S);
      const int nv = (S - head) / VN;
      A acc = 0;
      if (lane < head) acc += (A)ph[r] * xs[lane];
      if (lane < nv) {
        const A *xq = xs + head + lane * VN;
#pragma unroll
        for (int k = 0; k < VN; ++k) acc += (A)pv[r][k] * xq[k];
      }
      const V *rv = reinterpret_cast<const V *>(row + head);
      for (int q = lane + 64; q < nv; q += 64) {
        const V v = rv[q];
        const A *xq = xs + head + q * VN;
#pragma unroll
        for (int k = 0; k < VN; ++k) acc += (A)v[k] * xq[k];
      }
      const int tl = head + nv * VN + lane;
      if (tl < S) acc += (A)pt[r] * xs[tl];
      for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off);
      if (lane == 0) y[t0 + t] = rscale ? (double)acc * (double)rscale[t0 + t] : (double)acc;
    }
  }
  for (int t = w + R * WPB; t < nt; t += WPB) {
    const int64_t st = base + (int64_t)t * S;
    const TV *row = vals + st;
    const int head = min((int)((VN - (st & (VN - 1))) & (VN - 1)), S);
    const int nv = (S - head) / VN;
    A acc = 0;
    if (lane < head) acc += (A)row[lane] * xs[lane];
    const V *rv = reinterpret_cast<const V *>(row + head);
#pragma unroll 4
    for (int q = lane; q < nv; q += 64) {
      const V v = rv[q];
      const A *xq = xs + head + q * VN;
#pragma unroll
      for (int k = 0; k < VN; ++k) acc += (A)v[k] * xq[k];
    }
    const int tl = head + nv * VN + lane;
    if (tl < S) acc += (A)row[tl] * xs[tl];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off);
    if (lane == 0) y[t0 + t] = rscale ? (double)acc * (double)rscale[t0 + t] : (double)acc;
  }
}

// Single RHS over packs of consecutive leaves of one problem (pack[P] ..
// pack[P + 1]: one leaf with a wide tile, or up to 32 small leaves whose
// columns add up to at most 1024): one gather of the pack's x[cols] into LDS
// (the pack's column lists are contiguous), then 16 lanes per row, four rows
// of the pack at a time per wavefront, each row against its own leaf's
// columns.  Most leaves of a one-RHS batch hold one to a few panels, so a
// work-group per leaf with a wavefront per row spent its time on load chains
// (offsets -> columns -> x -> LDS, then offsets -> values) rather than on the
// tile; here the chains of a whole pack overlap.  Per row the same products
// as the leaf kernels, summed in another order.
template <typename TV, typename TX, int WPB>
__global__ __launch_bounds__(64 * WPB) void k_near_spmv_pack(
    const int *__restrict__ pack, const int *__restrict__ leaf_lo,
    const int64_t *__restrict__ nf_off, const int *__restrict__ col_off,
    const int *__restrict__ cols, const int *__restrict__ panel_leaf,
    const TV *__restrict__ vals, const TX *__restrict__ x, double *__restrict__ y,
    const unsigned char *__restrict__ skip, const float *__restrict__ rscale) {
  using A = typename Acc<TX>::type;
  using V = typename Vec16<TV>::type;
  constexpr int VN = Vec16<TV>::n;
  extern __shared__ unsigned char smem_raw[];
  A *xs = reinterpret_cast<A *>(smem_raw);
  const int P = blockIdx.x;
  const int T0 = pack[P], T1 = pack[P + 1];
  if (skip && skip[T0]) return;  // one problem per pack: work-group-uniform
  const int c0 = col_off[T0], SC = col_off[T1] - c0;
  const int r0 = leaf_lo[T0], r1 = leaf_lo[T1];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int gl = lane & 15, slot = lane >> 4;
  for (int c = threadIdx.x; c < SC; c += 64 * WPB) xs[c] = (A)x[cols[c0 + c]];
  __syncthreads();
  for (int rb = r0 + 4 * w; rb < r1; rb += 4 * WPB) {
    const int r = rb + slot;
    A acc = 0;
    if (r < r1) {
      const int T = panel_leaf[r];
      const int cT = col_off[T], S = col_off[T + 1] - cT;
      const int64_t st = nf_off[T] + (int64_t)(r - leaf_lo[T]) * S;
      const TV *row = vals + st;
      const A *xr = xs + (cT - c0);
      const int head = min((int)((VN - (st & (VN - 1))) & (VN - 1)), S);
      const int nv = (S - head) / VN;
      if (gl < head) acc += (A)row[gl] * xr[gl];
      const V *rv = reinterpret_cast<const V *>(row + head);
#pragma unroll 4
      for (int v = gl; v < nv; v += 16) {
        const V vv = rv[v];
        const A *xq = xr + head + v * VN;
#pragma unroll
        for (int k = 0; k < VN; ++k) acc += (A)vv[k] * xq[k];
      }
      const int tl = head + nv * VN + gl;
      if (tl < S) acc += (A)row[tl] * xr[tl];
    }
    for (int off = 8; off > 0; off >>= 1) acc += __shfl_down(acc, off, 16);
    if (gl == 0 && r < r1) y[r] = rscale ? (double)acc * (double)rscale[r] : (double)acc;
  }
}

// Packs of small leaves with the pack's leaf table in LDS: before the
// barrier, besides gathering x[cols], the work-group stores for each leaf of
// the pack (at most 32) its first row, column window and value offset; a row
// then finds its leaf by a binary search of that table, so its value loads
// issue right after the barrier with no dependent global loads
// (k_near_spmv_pack reads panel_leaf, col_off, nf_off and leaf_lo per row).
// The same sums in the same order as k_near_spmv_pack.
template <typename TV, typename TX, int WPB>
__global__ __launch_bounds__(64 * WPB) void k_near_spmv_pack_lt(
    const int *__restrict__ pack, const int *__restrict__ leaf_lo,
    const int64_t *__restrict__ nf_off, const int *__restrict__ col_off,
    const int *__restrict__ cols, const TV *__restrict__ vals, const TX *__restrict__ x,
    double *__restrict__ y, const unsigned char *__restrict__ skip,
    const float *__restrict__ rscale) {
  using A = typename Acc<TX>::type;
  using V = typename Vec16<TV>::type;
  constexpr int VN = Vec16<TV>::n;
  extern __shared__ unsigned char smem_raw[];
  A *xs = reinterpret_cast<A *>(smem_raw);
  __shared__ int64_t lt_off[33];
  __shared__ int lt_row[33], lt_col[33];
  const int P = blockIdx.x;
  const int T0 = pack[P], T1 = pack[P + 1];
  if (skip && skip[T0]) return;  // one problem per pack: work-group-uniform
  const int nl = T1 - T0;         // 1 .. 32
  const int c0 = col_off[T0], SC = col_off[T1] - c0;
  const int r0 = leaf_lo[T0], R = leaf_lo[T1] - r0;
  if ((int)threadIdx.x <= nl) {
    const int T = T0 + threadIdx.x;
    lt_row[threadIdx.x] = leaf_lo[T] - r0;
    lt_col[threadIdx.x] = col_off[T] - c0;
    lt_off[threadIdx.x] = nf_off[T];
  }
  for (int c = threadIdx.x; c < SC; c += 64 * WPB) xs[c] = (A)x[cols[c0 + c]];
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int gl = lane & 15, slot = lane >> 4;
  for (int rb = 4 * w; rb < R; rb += 4 * WPB) {
    const int r = rb + slot;
    A acc = 0;
    if (r < R) {
      int lo = 0, hi = nl - 1;  // the last leaf whose first row is <= r
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (lt_row[mid] <= r) lo = mid;
        else hi = mid - 1;
      }
      const int S = lt_col[lo + 1] - lt_col[lo];
      const int64_t st = lt_off[lo] + (int64_t)(r - lt_row[lo]) * S;
      const TV *row = vals + st;
      const A *xr = xs + lt_col[lo];
      const int head = min((int)((VN - (st & (VN - 1))) & (VN - 1)), S);
      const int nv = (S - head) / VN;
      if (gl < head) acc += (A)row[gl] * xr[gl];
      const V *rv = reinterpret_cast<const V *>(row + head);
#pragma unroll 4
      for (int v = gl; v < nv; v += 16) {
        const V vv = rv[v];
        const A *xq = xr + head + v * VN;
#pragma unroll
        for (int k = 0; k < VN; ++k) acc += (A)vv[k] * xq[k];
      }
      const int tl = head + nv * VN + gl;
      if (tl < S) acc += (A)row[tl] * xr[tl];
    }
    for (int off = 8; off > 0; off >>= 1) acc += __shfl_down(acc, off, 16);
    if (gl == 0 && r < R)
      y[r0 + r] = rscale ? (double)acc * (double)rscale[r0 + r] : (double)acc;
  }
}

// Single RHS: the leaf-staged kernels with 4, 8, 2 or 1 wavefronts a leaf
// (fewer: more leaves in flight per CU for leaves of a few panels), or one
// wavefront per row; 5-7: leaf-staged with the first rows' values issued
// before the gather.  rscale (FP16 values) multiplies each row's sum.
template <typename TV, typename TX>
void launch_spmv1(const char *tname, const NearTiles &nf, const TV *vals, const float *rscale,
                  const TX *x, double *y, hipStream_t s) {
  using A = typename Acc<TX>::type;
  const size_t xs_bytes = (size_t)nf.max_cols * sizeof(A);
  const bool leaf_ok = xs_bytes <= 48 * 1024;
  const int nvar = !leaf_ok ? 1 : nf.n_packs > 0 ? 12 : 8;
  tuned_launch(tname, nf.n_leaves, nf.max_cols, 1, nvar, s, [&](int v) {
    if (v == 10 || v == 11) {  // packs with their leaf table in LDS, 4 or 2 wavefronts
      if (v == 10)
        hipLaunchKernelGGL((k_near_spmv_pack_lt<TV, TX, 4>), dim3(nf.n_packs), dim3(256), xs_bytes,
                           s, nf.pack, nf.leaf_lo, nf.nf_off, nf.col_off, nf.cols, vals, x, y,
                           nf.skip_leaf, rscale);
      else
        hipLaunchKernelGGL((k_near_spmv_pack_lt<TV, TX, 2>), dim3(nf.n_packs), dim3(128), xs_bytes,
                           s, nf.pack, nf.leaf_lo, nf.nf_off, nf.col_off, nf.cols, vals, x, y,
                           nf.skip_leaf, rscale);
    } else if (v == 8 || v == 9) {  // packs of small leaves, 2 or 4 wavefronts, 16 lanes a row
      if (v == 8)
        hipLaunchKernelGGL((k_near_spmv_pack<TV, TX, 2>), dim3(nf.n_packs), dim3(128), xs_bytes, s,
                           nf.pack, nf.leaf_lo, nf.nf_off, nf.col_off, nf.cols, nf.panel_leaf, vals,
                           x, y, nf.skip_leaf, rscale);
      else
        hipLaunchKernelGGL((k_near_spmv_pack<TV, TX, 4>), dim3(nf.n_packs), dim3(256), xs_bytes, s,
                           nf.pack, nf.leaf_lo, nf.nf_off, nf.col_off, nf.cols, nf.panel_leaf, vals,
                           x, y, nf.skip_leaf, rscale);
    } else if (!leaf_ok || v == 2)
      hipLaunchKernelGGL((k_near_spmv<TV, TX>), dim3((nf.n + 3) / 4), dim3(256), 0, s, nf.n,
                         nf.panel_leaf, nf.leaf_lo, nf.nf_off, nf.col_off, nf.cols, vals, x, y,
                         nf.skip_leaf, rscale);
    else if (v == 0)
      hipLaunchKernelGGL((k_near_spmv_leaf<TV, TX, 4>), dim3(nf.n_leaves), dim3(256), xs_bytes, s,
                         nf.leaf_lo, nf.nf_off, nf.col_off, nf.cols, vals, x, y, nf.skip_leaf, rscale);
    else if (v == 1)
      hipLaunchKernelGGL((k_near_spmv_leaf<TV, TX, 8>), dim3(nf.n_leaves), dim3(512), xs_bytes, s,
                         nf.leaf_lo, nf.nf_off, nf.col_off, nf.cols, vals, x, y, nf.skip_leaf, rscale);
    else if (v == 3)
      hipLaunchKernelGGL((k_near_spmv_leaf<TV, TX, 2>), dim3(nf.n_leaves), dim3(128), xs_bytes, s,
                         nf.leaf_lo, nf.nf_off, nf.col_off, nf.cols, vals, x, y, nf.skip_leaf, rscale);
    else if (v == 4)
      hipLaunchKernelGGL((k_near_spmv_leaf<TV, TX, 1>), dim3(nf.n_leaves), dim3(64), xs_bytes, s,
                         nf.leaf_lo, nf.nf_off, nf.col_off, nf.cols, vals, x, y, nf.skip_leaf, rscale);
    else if (v == 5)
      hipLaunchKernelGGL((k_near_spmv_leaf_pf<TV, TX, 2, 2>), dim3(nf.n_leaves), dim3(128),
                         xs_bytes, s, nf.leaf_lo, nf.nf_off, nf.col_off, nf.cols, vals, x, y,
                         nf.skip_leaf, rscale);
    else if (v == 6)
      hipLaunchKernelGGL((k_near_spmv_leaf_pf<TV, TX, 2, 3>), dim3(nf.n_leaves), dim3(128),
                         xs_bytes, s, nf.leaf_lo, nf.nf_off, nf.col_off, nf.cols, vals, x, y,
                         nf.skip_leaf, rscale);
    else
      hipLaunchKernelGGL((k_near_spmv_leaf_pf<TV, TX, 4, 2>), dim3(nf.n_leaves), dim3(256),
                         xs_bytes, s, nf.leaf_lo, nf.nf_off, nf.col_off, nf.cols, vals, x, y,
                         nf.skip_leaf, rscale);
  });
}

// FP16 copy of the near tiles for the single-RHS inner operator: per row
// the scale is the largest magnitude (1 for a zero row) and each value is
// stored as fp16(v / scale), one rounding from FP64; a work-group per leaf,
// a wavefront per row.
__global__ __launch_bounds__(256) void k_near_to_f16(const int *__restrict__ leaf_lo,
                                                      const int64_t *__restrict__ nf_off,
                                                      const int *__restrict__ col_off,
                                                      const double *__restrict__ v64,
                                                      _Float16 *__restrict__ v16,
                                                      float *__restrict__ rscale) {
  const int T = blockIdx.x;
  const int t0 = leaf_lo[T], nt = leaf_lo[T + 1] - t0;
  const int S = col_off[T + 1] - col_off[T];
  const int64_t base = nf_off[T];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int t = w; t < nt; t += 4) {
    const double *row = v64 + base + (int64_t)t * S;
    double m = 0.0;
    for (int c = lane; c < S; c += 64) m = fmax(m, fabs(row[c]));
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    const float sc = m > 0.0 ? (float)m : 1.0f;
    const double inv = 1.0 / (double)sc;
    _Float16 *h = v16 + base + (int64_t)t * S;
    for (int c = lane; c < S; c += 64) h[c] = (_Float16)(row[c] * inv);
    if (lane == 0) rscale[t0 + t] = sc;
  }
}

void launch_near_to_f16(const NearTiles &nf, const double *v64, _Float16 *v16, float *rscale,
                        hipStream_t s) {
  if (nf.n_leaves == 0) return;
  hipLaunchKernelGGL(k_near_to_f16, dim3(nf.n_leaves), dim3(256), 0, s, nf.leaf_lo, nf.nf_off,
                     nf.col_off, v64, v16, rscale);
  GP_LAUNCH_CHECK();
}

void launch_near_apply_f16(const NearTiles &nf, const _Float16 *vals, const float *rscale,
                           const float *x, double *y, hipStream_t s) {
  if (nf.n_leaves == 0) return;
  // the values, and per row its scale, x and y (the rows not skipped)
  const double flops = 2.0 * nf.nnz * nf.frac_nnz;
  const double bytes = (double)nf.nnz * sizeof(_Float16) * nf.frac_nnz +
                       (double)nf.n * (sizeof(float) + sizeof(float) + sizeof(double)) * nf.frac_rows;
  ProfScope ps("near_spmv_f16", s, flops, bytes);
  launch_spmv1<_Float16, float>("near_spmv_f16", nf, vals, rscale, x, y, s);
  GP_LAUNCH_CHECK();
}

template <typename TV, typename TX>
void launch_near_apply_t(const NearTiles &nf, const TV *vals, const TX *x, int ldx, double *y,
                         int ldy, int nr, hipStream_t s) {
  if (nf.n_leaves == 0) return;
  // the work of the leaves not skipped (frac_nnz, frac_rows: their share)
  const double flops = 2.0 * nf.nnz * nr * nf.frac_nnz;
  const double bytes = (double)nf.nnz * sizeof(TV) * nf.frac_nnz +
                       (double)nf.n * nr * (sizeof(TX) + sizeof(double)) * nf.frac_rows;
  ProfScope ps(sizeof(TV) == 4 ? (nr == 1 ? "near_spmv_f32" : "near_spmm_f32")
                               : (nr == 1 ? "near_spmv_f64" : "near_spmm_f64"),
               s, flops, bytes);
  // kernel variants, fastest per shape picked by the context's autotuner
  auto spmm16 = [&] {
    const dim3 grid(nf.n_leaves * ((nr + NF_RC - 1) / NF_RC));
    hipLaunchKernelGGL((k_near_spmm<TV, TX>), grid, dim3(256), 0, s, nf.leaf_lo, nf.nf_off,
                       nf.col_off, nf.cols, vals, x, ldx, y, ldy, nr);
  };
  auto spmm_rt = [&](int cpt) {
    if (cpt == 4)
      hipLaunchKernelGGL((k_near_spmm_rt<TV, TX, 4>), dim3(nf.n_leaves * ((nr + 63) / 64)),
                         dim3(256), 0, s, nf.leaf_lo, nf.nf_off, nf.col_off, nf.cols, vals, x, ldx,
                         y, ldy, nr);
    else
      hipLaunchKernelGGL((k_near_spmm_rt<TV, TX, 2>), dim3(nf.n_leaves * ((nr + 31) / 32)),
                         dim3(256), 0, s, nf.leaf_lo, nf.nf_off, nf.col_off, nf.cols, vals, x, ldx,
                         y, ldy, nr);
  };
  const char *tname = sizeof(TV) == 4 ? (nr == 1 ? "near_spmv_f32" : "near_spmm_f32")
                                      : (nr == 1 ? "near_spmv_f64" : "near_spmm_f64");
  if (nr == 1 && ldx == 1 && ldy == 1) {
    launch_spmv1<TV, TX>(tname, nf, vals, nullptr, x, y, s);
  } else if (nr > 16) {
    // FP32 matrix cores (64-column chunks, 16-row tiles), register tiles of
    // 64 or 32 RHS columns, or 16-column work-groups
    const int c0 = nr > 32 ? 4 : 2;
    const bool mfma = std::is_same<TV, float>::value && std::is_same<TX, float>::value &&
                      ldx % 4 == 0;
    tuned_launch(tname, nf.n_leaves, nr, ldx, mfma ? 6 : 3, s, [&](int v) {
      if (v >= 3) {  // 32 or 64 rows a pass; 32 rows with two steps of prefetch
        if constexpr (std::is_same<TV, float>::value && std::is_same<TX, float>::value) {
          const int items = nf.n_leaves * ((nr + 63) / 64);
          if (v == 3)
            hipLaunchKernelGGL((k_near_spmm_mfma<2>), dim3((items + 3) / 4), dim3(256), 0, s, items,
                               nf.leaf_lo, nf.nf_off, nf.col_off, nf.cols, vals, x, ldx, y, ldy, nr);
          else if (v == 4)
            hipLaunchKernelGGL((k_near_spmm_mfma<4>), dim3((items + 3) / 4), dim3(256), 0, s, items,
                               nf.leaf_lo, nf.nf_off, nf.col_off, nf.cols, vals, x, ldx, y, ldy, nr);
          else
            hipLaunchKernelGGL((k_near_spmm_mfma<2, 2>), dim3((items + 3) / 4), dim3(256), 0, s,
                               items, nf.leaf_lo, nf.nf_off, nf.col_off, nf.cols, vals, x, ldx, y,
                               ldy, nr);
        }
      } else if (v == 0)
        spmm_rt(c0);
      else if (v == 1)
        spmm_rt(6 - c0);
      else
        spmm16();
    });
  } else {
    spmm16();
  }
  GP_LAUNCH_CHECK();
}

// Transposed near field y = P_near^T x (the adjoint solve).  The near
// relation is symmetric, so the rows of leaf S in P_near^T are the columns
// of S in the tiles of S's near leaves T: block (T, S) of T's row-major tile
// starts at tr_base and has row stride tr_stride (T's column count).  One
// work-group per (leaf S, 16-column chunk); consecutive threads take
// consecutive panels of S, so each tile row is read contiguously.
template <typename TV, typename TX>
__global__ __launch_bounds__(256) void k_near_spmm_tr(
    const int *__restrict__ leaf_lo, const int *__restrict__ tr_off, const int *__restrict__ tr_leaf,
    const int64_t *__restrict__ tr_base, const int *__restrict__ tr_stride,
    const TV *__restrict__ vals, const TX *__restrict__ x, int ldx, double *__restrict__ y,
    int ldy, int nr) {
  using A = typename Acc<TX>::type;
  const int S = blockIdx.x, r0 = blockIdx.y * NF_RC;
  const int rc = min(NF_RC, nr - r0);
  const int s0 = leaf_lo[S], ns = leaf_lo[S + 1] - s0;
  for (int e = threadIdx.x; e < ns * rc; e += blockDim.x) {
    const int s = e % ns, r = r0 + e / ns;
    A acc = 0;
    for (int q = tr_off[S]; q < tr_off[S + 1]; ++q) {
      const int T = tr_leaf[q], t0 = leaf_lo[T], nt = leaf_lo[T + 1] - t0;
      const TV *blk = vals + tr_base[q] + s;
      const int st = tr_stride[q];
      for (int t = 0; t < nt; ++t) acc += (A)blk[(int64_t)t * st] * (A)x[(int64_t)(t0 + t) * ldx + r];
    }
    y[(int64_t)(s0 + s) * ldy + r] = (double)acc;
  }
}

// Transposed near field on the FP32 matrix cores: one wavefront per (leaf
// S, 64-column RHS chunk), D (rows of S x columns) = sum over S's near leaves
// T of block(T, S)^T (rows of S x rows of T) * x[rows of T], in 16-row tiles
// of S (RTM per pass, only as many as S has).  Per step of 16 rows of T,
// k-slot q of MFMA u is T's row 16 step + 4 q + u; tile v holds columns
// r0 + 4 j + v (one 16-byte x load feeds four tiles).  A block row (fixed t)
// is contiguous along S's panels, so the 16 lanes of a k-slot read 16
// consecutive values.
template <int RTM>
__global__ __launch_bounds__(256) void k_near_spmm_tr_mfma(
    int nitems, const int *__restrict__ leaf_lo, const int *__restrict__ tr_off,
    const int *__restrict__ tr_leaf, const int64_t *__restrict__ tr_base,
    const int *__restrict__ tr_stride, const float *__restrict__ vals,
    const float *__restrict__ x, int ldx, double *__restrict__ y, int ldy, int nr) {
  const int lane = threadIdx.x & 63, j = lane & 15, kq = lane >> 4;
  const int nch = (nr + 63) >> 6;
  const int w = xcd_block(blockIdx.x, gridDim.x, XCD_RUN_NEAR / 4) * 4 + (threadIdx.x >> 6);
  if (w >= nitems) return;  // whole wavefront; the kernel has no barriers
  const int S = w / nch, r0 = (w % nch) * 64;
  const int s0 = leaf_lo[S], ns = leaf_lo[S + 1] - s0;
  const int rc = r0 + 4 * j;
  const bool cok = rc < ldx;
  const int q0 = tr_off[S], q1 = tr_off[S + 1];
  for (int tm = 0; tm < ns; tm += 16 * RTM) {
    const int rtn = min(RTM, (ns - tm + 15) >> 4);
    nf_f32x4 acc[RTM][4];
#pragma unroll
    for (int rt = 0; rt < RTM; ++rt)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[rt][v] = nf_f32x4{0.f, 0.f, 0.f, 0.f};
    for (int q = q0; q < q1; ++q) {
      const int T = tr_leaf[q], t0 = leaf_lo[T], nt = leaf_lo[T + 1] - t0;
      const float *blk = vals + tr_base[q];
      const int st = tr_stride[q];
      for (int kb = 0; kb < nt; kb += 16) {
        nf_f32x4 xb[4];
        float av[RTM][4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int tt = kb + 4 * kq + u;
          xb[u] = (cok && tt < nt) ? *reinterpret_cast<const nf_f32x4 *>(x + (int64_t)(t0 + tt) * ldx + rc)
                                   : nf_f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int rt = 0; rt < RTM; ++rt) {
            const int sr = tm + 16 * rt + j;
            av[rt][u] = (rt < rtn && sr < ns && tt < nt) ? blk[(int64_t)tt * st + sr] : 0.f;
          }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int rt = 0; rt < RTM; ++rt)
            if (rt < rtn)
#pragma unroll
              for (int v = 0; v < 4; ++v)
                acc[rt][v] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[rt][u], xb[u][v], acc[rt][v], 0, 0, 0);
      }
    }
#pragma unroll
    for (int rt = 0; rt < RTM; ++rt)
#pragma unroll
      for (int vv = 0; vv < 4; ++vv) {
        const int sr = tm + 16 * rt + 4 * kq + vv;
        if (rt >= rtn || sr >= ns) continue;
        double *yr = y + (int64_t)(s0 + sr) * ldy + rc;
#pragma unroll
        for (int v = 0; v < 4; ++v)
          if (rc + v < nr) yr[v] = (double)acc[rt][v][vv];
      }
  }
}

template <typename TV, typename TX>
void launch_near_apply_tr_t(const NearTiles &nf, const NearTranspose &tr, const TV *vals,
                            const TX *x, int ldx, double *y, int ldy, int nr, hipStream_t s) {
  if (nf.n_leaves == 0) return;
  ProfScope ps(sizeof(TV) == 4 ? "near_tr_f32" : "near_tr_f64", s, 2.0 * nf.nnz * nr,
               (double)nf.nnz * sizeof(TV) + (double)nf.n * nr * (sizeof(TX) + sizeof(double)));
  auto valu = [&] {
    dim3 grid(nf.n_leaves, (nr + NF_RC - 1) / NF_RC);
    hipLaunchKernelGGL((k_near_spmm_tr<TV, TX>), grid, dim3(256), 0, s, nf.leaf_lo, tr.off,
                       tr.leaf, tr.base, tr.stride, vals, x, ldx, y, ldy, nr);
  };
  if constexpr (std::is_same<TV, float>::value && std::is_same<TX, float>::value) {
    if (nr > 16 && ldx % 4 == 0) {
      // FP32 matrix cores (32 rows of S a pass), or the VALU kernel
      tuned_launch("near_tr_f32", nf.n_leaves, nr, ldx, 2, s, [&](int v) {
        if (v == 0) {
          const int items = nf.n_leaves * ((nr + 63) / 64);
          hipLaunchKernelGGL((k_near_spmm_tr_mfma<2>), dim3((items + 3) / 4), dim3(256), 0, s,
                             items, nf.leaf_lo, tr.off, tr.leaf, tr.base, tr.stride, vals, x, ldx,
                             y, ldy, nr);
        } else {
          valu();
        }
      });
      GP_LAUNCH_CHECK();
      return;
    }
  }
  valu();
  GP_LAUNCH_CHECK();
}

void launch_near_apply_tr(const NearTiles &nf, const NearTranspose &tr, const double *vals,
                          const double *x, int ldx, double *y, int ldy, int nr, hipStream_t s) {
  launch_near_apply_tr_t<double, double>(nf, tr, vals, x, ldx, y, ldy, nr, s);
}
void launch_near_apply_tr(const NearTiles &nf, const NearTranspose &tr, const float *vals,
                          const float *x, int ldx, double *y, int ldy, int nr, hipStream_t s) {
  launch_near_apply_tr_t<float, float>(nf, tr, vals, x, ldx, y, ldy, nr, s);
}

void launch_near_apply(const NearTiles &nf, const double *vals, const double *x, int ldx,
                       double *y, int ldy, int nr, hipStream_t s) {
  launch_near_apply_t<double, double>(nf, vals, x, ldx, y, ldy, nr, s);
}
void launch_near_apply(const NearTiles &nf, const float *vals, const float *x, int ldx,
                       double *y, int ldy, int nr, hipStream_t s) {
  launch_near_apply_t<float, float>(nf, vals, x, ldx, y, ldy, nr, s);
}

}  // namespace gp
