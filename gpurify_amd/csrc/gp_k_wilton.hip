// Collocation assembly on the device: the Wilton closed-form panel integral
// in FP64, the dense potential-coefficient matrix, the near-field leaf tiles
// of the precorrected FMM operator and the Jacobi diagonal.
//
// All geometry here is in scaled units (metres divided by the root half-width
// L0, centred on the root cube), so P' = P / (k_e L0) with k_e = 1/(4 pi eps0).

#include "gp_common.h"
#include "gp_kernels.h"

namespace gp {

// Closed-form I(p) = integral over a flat quadrilateral of 1/|p - r'| da'
// (Wilton et al. 1984), following crates/pex/src/quasistatic/integrals/
// panel.rs:147-208 term by term: Newell normal (panel.rs:256-269), per edge
// the in-plane distance, the log term with its asinh fallback
// (panel.rs:271-285) and the solid-angle atan term.  v: 12 doubles.
__device__ double wilton(const double *__restrict__ v, double px, double py, double pz) {
  double vx[4], vy[4], vz[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    vx[k] = v[3 * k];
    vy[k] = v[3 * k + 1];
    vz[k] = v[3 * k + 2];
  }
  double nx = 0.0, ny = 0.0, nz = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    int j = (i + 1) & 3;
    nx += (vy[i] - vy[j]) * (vz[i] + vz[j]);
    ny += (vz[i] - vz[j]) * (vx[i] + vx[j]);
    nz += (vx[i] - vx[j]) * (vy[i] + vy[j]);
  }
  double nn = sqrt(nx * nx + ny * ny + nz * nz);
  if (nn == 0.0) {
    nx = ny = nz = 0.0;
  } else {
    nx /= nn;
    ny /= nn;
    nz /= nn;
  }
  double d = (px - vx[0]) * nx + (py - vy[0]) * ny + (pz - vz[0]) * nz;
  double p0x = px - nx * d, p0y = py - ny * d, p0z = pz - nz * d;
  double absd = fabs(d);
  double total = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    int j = (i + 1) & 3;
    double ex = vx[j] - vx[i], ey = vy[j] - vy[i], ez = vz[j] - vz[i];
    double el = sqrt(ex * ex + ey * ey + ez * ez);
    if (el == 0.0) continue;
    double lx = ex / el, ly = ey / el, lz = ez / el;
    double ux = ly * nz - lz * ny, uy = lz * nx - lx * nz, uz = lx * ny - ly * nx;
    double ax = vx[i] - p0x, ay = vy[i] - p0y, az = vz[i] - p0z;
    double bx = vx[j] - p0x, by = vy[j] - p0y, bz = vz[j] - p0z;
    double perp = ax * ux + ay * uy + az * uz;
    double sm = ax * lx + ay * ly + az * lz;
    double sp = bx * lx + by * ly + bz * lz;
    double r0sq = perp * perp + d * d;
    double rp = sqrt(sp * sp + r0sq);
    double rm = sqrt(sm * sm + r0sq);
    double num = rp + sp, den = rm + sm;
    double logt;
    if (num > 0.0 && den > 0.0) {
      logt = perp * log(num / den);
    } else if (fabs(den) > 0.0) {
      double r0 = sqrt(r0sq);
      logt = r0 == 0.0 ? 0.0 : perp * (asinh(sp / r0) - asinh(sm / r0));
    } else {
      logt = 0.0;
    }
    double att = 0.0;
    if (absd > 0.0 && perp != 0.0) {
      double tp = (perp * sp) / (r0sq + absd * rp);
      double tm = (perp * sm) / (r0sq + absd * rm);
      att = absd * (atan(tp) - atan(tm));
    }
    total += logt - att;
  }
  return total;
}

// P'[k][l] = I_l(c_k), cap/solver.rs:81-97 (row k = collocation at centroid k).
__global__ __launch_bounds__(256) void k_dense_assemble(int n, const double *__restrict__ verts,
                                                        const double *__restrict__ cen,
                                                        double *__restrict__ P) {
  long total = (long)n * n;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    int k = (int)(e / n), l = (int)(e % n);
    P[e] = wilton(verts + 12L * l, cen[3 * k], cen[3 * k + 1], cen[3 * k + 2]);
  }
}

// Near-field tile of target leaf T: rows = T's panels, columns = the panels of
// T and its neighbour leaves (cap/fmm_solver.rs:52-91), row-major at nf_off[T].
// Writes the FP64 values and, when v32 != nullptr, their FP32 rounding.
__global__ __launch_bounds__(256) void k_near_assemble(
    const int *__restrict__ leaf_lo, const int64_t *__restrict__ nf_off,
    const int *__restrict__ col_off, const int *__restrict__ cols,
    const double *__restrict__ verts, const double *__restrict__ cen, double *__restrict__ v64,
    float *__restrict__ v32) {
  int T = blockIdx.x;
  int t0 = leaf_lo[T], nt = leaf_lo[T + 1] - t0;
  int c0 = col_off[T], S = col_off[T + 1] - c0;
  long cnt = (long)nt * S;
  int64_t base = nf_off[T];
  for (long e = threadIdx.x; e < cnt; e += blockDim.x) {
    int t = (int)(e / S), s = (int)(e % S);
    int j = cols[c0 + s];
    int i = t0 + t;
    double w = wilton(verts + 12L * j, cen[3 * i], cen[3 * i + 1], cen[3 * i + 2]);
    v64[base + e] = w;
    if (v32) v32[base + e] = (float)w;
  }
}

// Self term P'_ii for the Jacobi preconditioner (cap/fmm_solver.rs:150-162).
// inv = 1/d unless |k_e L0 d| <= 1e-30, the reference's threshold in SI units.
__global__ void k_self_inv_diag(int n, const double *__restrict__ verts,
                                const double *__restrict__ cen, double si_scale,
                                double *__restrict__ inv) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double d = wilton(verts + 12L * i, cen[3 * i], cen[3 * i + 1], cen[3 * i + 2]);
  inv[i] = fabs(d * si_scale) > 1e-30 ? 1.0 / d : 1.0;
}

void launch_dense_assemble(int n, const double *verts, const double *cen, double *P,
                           hipStream_t s) {
  long total = (long)n * n;
  int blocks = (int)std::min<long>((total + 255) / 256, 65536);
  ProfScope ps("dense_assemble", s, 0.0, 8.0 * total);
  hipLaunchKernelGGL(k_dense_assemble, dim3(blocks), dim3(256), 0, s, n, verts, cen, P);
  GP_LAUNCH_CHECK();
}

void launch_near_assemble(int n_leaves, const int *leaf_lo, const int64_t *nf_off,
                          const int *col_off, const int *cols, const double *verts,
                          const double *cen, double *v64, float *v32, hipStream_t s) {
  if (n_leaves == 0) return;
  ProfScope ps("near_assemble", s, 0.0, 0.0);
  hipLaunchKernelGGL(k_near_assemble, dim3(n_leaves), dim3(256), 0, s, leaf_lo, nf_off, col_off,
                     cols, verts, cen, v64, v32);
  GP_LAUNCH_CHECK();
}

void launch_self_inv_diag(int n, const double *verts, const double *cen, double si_scale,
                          double *inv, hipStream_t s) {
  hipLaunchKernelGGL(k_self_inv_diag, dim3((n + 255) / 256), dim3(256), 0, s, n, verts, cen,
                     si_scale, inv);
  GP_LAUNCH_CHECK();
}

}  // namespace gp
