// Launch wrappers of the device kernels (defined in gp_k_*.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace gp {

// ---------------------------------------------------------------- assembly
// Dielectric-interface rows drow[0..nd) of a dense system (oriented normals
// nrm, 3 per row; diagonal values diag), over the rows launch_dense_assemble
// wrote.
void launch_dielectric_rows(int n, int nd, const int *drow, const double *nrm, const double *diag,
                            const double *verts, const double *cen, double *P, hipStream_t s);
// P'[k][l] = I_l(c_k) over n panels; image: minus the integral over panel
// l's mirror image in the plane z = zm2 / 2 (a grounded plane by images).
void launch_dense_assemble(int n, const double *verts, const double *cen, double *P,
                           hipStream_t s, bool image = false, double zm2 = 0.0);
void launch_near_assemble(int n_leaves, const int *leaf_lo, const int64_t *nf_off,
                          const int *col_off, const int *cols, const double *verts,
                          const double *cen, double *v64, float *v32, hipStream_t s);
// Dielectric-interface rows of the near tiles (iface, nrm, diag per panel in
// device order) and 1/diag as their Jacobi preconditioner in inv_diag.
void launch_near_dielectric_rows(int n_leaves, int n, const int *leaf_lo, const int64_t *nf_off,
                                 const int *col_off, const int *cols, const unsigned char *iface,
                                 const double *nrm, const double *diag, const double *verts,
                                 const double *cen, double *v64, double *inv_diag, hipStream_t s);
// Jacobi entries of every problem of a batch in one launch: off (np + 1
// panel offsets) and si_scale (np threshold scales) are device arrays.
void launch_self_inv_diag(int n, int np, const int *off, const double *si_scale,
                          const double *verts, const double *cen, double *inv, hipStream_t s);

// Geometric sensitivities of C' for a group of nd (1 or 2) parameters that
// move the same panels (device arrays; scaled units).  dv: direction t at
// dv + t * dstride.  sigma: N x lds forward solutions (nc columns), lambda:
// N x ldl adjoint solutions (P'^T lambda_i = scaled areas of conductor
// cond0 + i).  mv_of (N ints) and da (N doubles) must hold -1 / 0 on entry
// and are restored.  scratch: sensitivity_scratch(N, nm, nc, nd) doubles;
// R: nd x N x nc; dcen: 3 nd nm; darea: nd nm; dgq: 16 nd nm.  J: nd x nc x nc.
// gq: launch_gauss_table of the panels; entries whose source centroid lies
// beyond sqrt(far2) source diagonals use its 2 x 2 Gauss rule, beyond
// sqrt(far1) diagonals the centroid rule A / |p - c| (0: not used; both 0:
// every entry by the closed form).
struct SensArgs {
  int N = 0, nm = 0, nc = 0, nd = 1, lds = 0, ldl = 0, cond0 = 0;
  long dstride = 0;
  double far2 = 0.0, far1 = 0.0;
  const double *gq = nullptr;
  double *dgq = nullptr;
  const int *moved = nullptr;
  const double *dv = nullptr, *verts = nullptr, *cen = nullptr, *sigma = nullptr,
               *lambda = nullptr;
  const uint32_t *cond = nullptr;
  int *mv_of = nullptr;
  double *da = nullptr, *dcen = nullptr, *darea = nullptr, *R = nullptr, *scratch = nullptr,
         *J = nullptr;
};
void launch_sensitivity(const SensArgs &a, hipStream_t s);
size_t sensitivity_scratch(int N, int nm_max, int nc, int nd);
void launch_gauss_table(int n, const double *verts, double *gq, hipStream_t s);

// ---------------------------------------------------------------- dense LU
void launch_lu_solve(int n, int nrhs, double *A, int *piv, double *B, double *work, int *info,
                     hipStream_t s);
void launch_transpose(int n, const double *A, double *AT, hipStream_t s);
// A batch of independent dense problems (device arrays): problem p has n[p]
// panels at verts + 12 n_off[p] / cen + 3 n_off[p], its matrix at
// A + a_off[p], its n[p] x nrhs[p] right-hand sides at B + b_off[p] (work
// alike), pivots at piv + n_off[p], status info[p].  Assembles every P'_p and
// solves P'_p X = B_p in place, one work-group per problem.
struct DenseBatch {
  int count = 0;
  const int *n = nullptr, *nrhs = nullptr, *n_off = nullptr;
  const long *a_off = nullptr, *b_off = nullptr;
  const double *verts = nullptr, *cen = nullptr;
  double *A = nullptr, *B = nullptr, *work = nullptr;
  int *piv = nullptr, *info = nullptr;
  const double *zm2 = nullptr;  // per problem 2 z of its ground plane (unit frame); null: none
};
void launch_dense_assemble_batched(const DenseBatch &b, hipStream_t s);
void launch_dense_batch(const DenseBatch &b, hipStream_t s);
void launch_cap_accumulate(int ncond, int nrhs, int ldx, const int *coff, const int *cidx,
                           const double *w, const double *X, double *C, hipStream_t s);

// ---------------------------------------------------------------- near field
struct NearTiles {
  int n = 0;         // panels
  long nnz = 0;      // near-field values
  int n_leaves = 0;  // target leaves
  const int *leaf_lo = nullptr;     // n_leaves + 1
  const int64_t *nf_off = nullptr;  // n_leaves + 1 (value offsets)
  const int *col_off = nullptr;     // n_leaves + 1
  const int *cols = nullptr;        // source panel per column
  const int *panel_leaf = nullptr;  // n
  int max_cols = 0;                  // widest tile (source columns of one leaf)
  // Single RHS: packs of consecutive leaves of one problem (n_packs + 1
  // first leaves; each pack's columns fit max_cols), or none.
  int n_packs = 0;
  const int *pack = nullptr;
  // Single right-hand side: leaves flagged 1 are skipped (null = none); the
  // profiler counts frac_nnz of the values and frac_rows of the rows.
  const unsigned char *skip_leaf = nullptr;
  double frac_nnz = 1.0, frac_rows = 1.0;
};
void launch_near_apply(const NearTiles &nf, const double *vals, const double *x, int ldx,
                       double *y, int ldy, int nr, hipStream_t s);
void launch_near_apply(const NearTiles &nf, const float *vals, const float *x, int ldx,
                       double *y, int ldy, int nr, hipStream_t s);
// Single RHS on FP16 tiles scaled per row (rscale: one FP32 scale per panel).
void launch_near_apply_f16(const NearTiles &nf, const _Float16 *vals, const float *rscale,
                           const float *x, double *y, hipStream_t s);
// The FP16 tiles and row scales from the FP64 values.
void launch_near_to_f16(const NearTiles &nf, const double *v64, _Float16 *v16, float *rscale,
                        hipStream_t s);
// Transposed tiles: per leaf S (CSR tr.off), its near leaves T with the
// value offset and row stride of block (T, S) inside T's tile.
struct NearTranspose {
  const int *off = nullptr, *leaf = nullptr, *stride = nullptr;
  const int64_t *base = nullptr;
};
void launch_near_apply_tr(const NearTiles &nf, const NearTranspose &tr, const double *vals,
                          const double *x, int ldx, double *y, int ldy, int nr, hipStream_t s);
void launch_near_apply_tr(const NearTiles &nf, const NearTranspose &tr, const float *vals,
                          const float *x, int ldx, double *y, int ldy, int nr, hipStream_t s);

// ---------------------------------------------------------------- far field
template <typename T>
struct FmmDev {
  int n = 0;          // Chebyshev order
  int L = 0;          // leaf level
  int n_boxes = 0;
  int n_targets = 0;  // boxes with a non-empty interaction list
  long n_pairs = 0;   // M2L (target, source) pairs
  int n_panels = 0;
  const int *lvl_off = nullptr;  // HOST array, L + 2 entries
  // Several independent trees (a batch) share one box numbering: each level
  // lists its boxes with children first (npar, HOST, L + 1 entries) and the
  // leaves of every depth are listed in leaf_box (device, n_leaves).  Both
  // null for one tree whose leaves are the boxes of level L.
  const int *npar = nullptr;
  const int *leaf_box = nullptr;
  int n_leaves = 0;
  const int *leaf_lo = nullptr;  // device, n_leaves + 1
  const int *panel_leaf = nullptr;  // device, per sorted panel: its leaf index
  const int *parent = nullptr, *octant = nullptr, *child_off = nullptr, *child_idx = nullptr;
  const int *targets = nullptr, *il_off = nullptr, *il_src = nullptr, *il_oidx = nullptr;
  const int *box_level = nullptr;
  const double *inv_half = nullptr;  // per level, scaled units
  const double *area = nullptr;      // per sorted panel, scaled units
  int n3p = 0;                       // n^3 rounded up to a multiple of 32
  int fuse_leaf = 1;                 // FP32 blocks of RHS: finest L2L + L2P as one kernel
  // Single right-hand side: work of boxes / leaves flagged 1 is skipped
  // (skip_box per global box, skip_leaf per leaf index; null = none), for
  // the problems of a batch whose GMRES has converged.  Problems never share
  // a box or a near tile, so the others' results are unchanged.
  const unsigned char *skip_box = nullptr, *skip_leaf = nullptr;
  // the profiler's share of the work done while boxes are skipped
  double frac_panels = 1.0, frac_boxes = 1.0, frac_pairs = 1.0;
  const T *w1d = nullptr;            // per sorted panel 3 * n
  // FP32 single-RHS operator: per panel (x, y, z, area) in its leaf box's
  // frame and the n x n weight coefficients, for P2M / L2P forming the
  // weights on the fly (null: they read w1d)
  const float *xl = nullptr, *wc = nullptr;
  const T *m2m1d = nullptr;          // 24 * n * n
  const T *G = nullptr;              // 343 * n3p * n3p (zero padded)
  const void *Gbf = nullptr;         // same, BF16 (FP32 operator only)
  const T *Gr = nullptr;             // same, BF16-rounded, widened to T (GP_M2L_BF16_VALU)
  // Low-rank M2L (FP32 operator only): rank kr (0 = off), U: n3 x kr
  // (row-major), tables Gc / Gcbf / Gcr as G / Gbf / Gr with kr x kr blocks.
  int kr = 0;
  const T *U = nullptr, *Gc = nullptr, *Gcr = nullptr;
  const void *Gcbf = nullptr, *Gcbf_lo = nullptr;  // BF16 pair hi + lo of Gc
  // Single right-hand side: the M2L pairs grouped by offset index o, pairs
  // [po_off[o], po_off[o+1]) (HOST, 344 entries) of device lists po_tgt /
  // po_src; null when not built.
  const int *po_off = nullptr;
  const int *po_tgt = nullptr, *po_src = nullptr;
  // Single right-hand side, target tiles (BF16 M2L): tile T holds 16 targets
  // of one octant (tile_tgt, -1 padded) and the entries [tile_eoff[T],
  // tile_eoff[T + 1]): the union of their M2L offsets ascending (ent_oidx),
  // each with the source box of every target (ent_src, 16 per entry, -1 for
  // none).  Null when not built.
  // Multi-RHS sibling pairs (BF16 M2L): pair g = targets pair_tgt[2g], [2g+1]
  // (-1: none), entries [pair_off[g], pair_off[g + 1]) by source box with the
  // offsets pe_o1 / pe_o2 of the source for each target (-1: not listed).
  int n_pairs2 = 0;
  const int *pair_tgt = nullptr, *pair_off = nullptr, *pe_src = nullptr, *pe_o1 = nullptr,
            *pe_o2 = nullptr;
  // FP64 operator with truncated M2L blocks: their factors (M2lFactors of
  // gp_m2l_cpqr.h: lr_vt, lr_up, lr_u, per-offset lr_rank, padded and
  // largest rank lr_kp, lr_kmax), for the factored single-RHS tile kernels;
  // null when not built.
  const double *lr_vt = nullptr, *lr_up = nullptr, *lr_u = nullptr;
  const int *lr_rank = nullptr;
  int lr_kp = 0, lr_kmax = 0;
  // Single-RHS tree passes by subtree (k_m2m1_sub, k_l2l1_sub): the level-2
  // boxes sub_root, the level of their leaves sub_leaf, and per level l <= L
  // the global range of each one's descendants, sub_rng[(r (L + 1) + l) 2 ..
  // + 1]; n_sub = 0: a launch per level.
  int n_sub = 0;
  const int *sub_root = nullptr, *sub_leaf = nullptr, *sub_rng = nullptr;
  int n_tiles = 0, tile_size = 16;  // targets per tile: 16, or 64 (n3p <= 128)
  const int *tile_tgt = nullptr, *tile_eoff = nullptr, *ent_oidx = nullptr, *ent_src = nullptr;
};
// Chebyshev interpolation weights of every sorted panel's centroid in its
// leaf box: w1d[i][d][m] for axis d and node m (3 n per panel), FP64 and, when
// w32 is set, FP32.  geo: per leaf the box centre (3) and half-width in the
// problem's unit frame; cen: scaled centroids; nodes: the n Chebyshev nodes.
// xl[4 p ..] = panel p's centroid in its leaf box's frame and its area (FP32).
void launch_local_coords(int N, const int *panel_leaf, const double *geo, const double *cen,
                         const double *area, float *xl, hipStream_t s);
void launch_interp_w1d(int N, int n, const int *panel_leaf, const double *geo, const double *cen,
                       const double *nodes, double *w64, float *w32, hipStream_t s);
// d/dx of the 1-D interpolation weights over the leaf half-width (3n per panel).
void launch_interp_dw1d(int N, int n, const int *panel_leaf, const double *geo, const double *cen,
                        const double *nodes, double *dw, hipStream_t s);
// Far field y += P_far z.  Multipoles M: n_boxes x ldm x n3p; locals Lc:
// n_boxes x n3 x ldm.  m2l selects the M2L kernel (GP_M2L_*): BF16 needs Mb
// (BF16 multipoles, same layout), FP64 uses matrix cores when ldm allows,
// BF16_VALU rounds M to BF16 in place after the upward pass and runs the
// FP32 VALU M2L on the table Gr.  A low-rank operator (f.kr > 0) needs the
// projected multipoles Mt (n_boxes x ldm x kr, BF16 copy Mtb for the BF16
// M2L, with Mtbl its rounding residual for the split-BF16 products of a
// block of RHS) and locals Lt (n_boxes x kr x ldm); Mb is then unused.
// y += yfar on conductor panels, y -= n . grad(local expansion) on
// interface panels (the dielectric formulation's far field; FP64).
void launch_far_dielectric(const FmmDev<double> &f, int nr, int ldm, int ldy, const double *dw1d,
                           const double *nrm, const unsigned char *iface, const double *Lc,
                           const double *yfar, double *y, hipStream_t s);
void run_far_field(const FmmDev<double> &f, const double *z, int ldz, int nr, double *M,
                   double *Lc, int ldm, int m2l, double *y, int ldy, hipStream_t s);
void run_far_field(const FmmDev<float> &f, const float *z, int ldz, int nr, float *M, void *Mb,
                   float *Lc, int ldm, int m2l, double *y, int ldy, hipStream_t s,
                   float *Mt = nullptr, void *Mtb = nullptr, float *Lt = nullptr,
                   void *Mtbl = nullptr);

// ---------------------------------------------------------------- Krylov
// Basis types TV / TX / TO: double or float (explicitly instantiated).
size_t kr_partial_size(int nvec_max, int ld);
template <typename TX>
void kr_dots(const TX *X, size_t strideX, int nvec, const double *Y, int N, int ld, int nr,
             double *partial, double *out, bool take_sqrt, hipStream_t s);
template <typename TV>
void kr_axpy_multi(double *w, const TV *V, size_t strideV, const double *coef, int nvec,
                   int N, int ld, int nr, double alpha, hipStream_t s);
template <typename TO>
void kr_scale_cols(const double *in, TO *out, const double *scale, const int *mask, int N,
                   int ld, int nr, hipStream_t s);
void kr_axpy(double *x, const double *y, double alpha, int N, int ld, int nr, hipStream_t s);
// b = unit right-hand sides of conductors j0 .. j0 + nr - 1 (sign[row] in
// place of 1 when given: a signed panel set).
void kr_unit_rhs(const uint32_t *cond, int j0, int N, int ld, int nr, double *b, hipStream_t s,
                 const double *sign = nullptr);
// skip_box[b] = !active[box_prob[b]] over nb boxes, skip_leaf[T] likewise
// for the nl leaves (leaf T is box leaf_box[T]).
void kr_skip_flags(int nb, int nl, const int *box_prob, const int *leaf_box, const int *active,
                   unsigned char *skip_box, unsigned char *skip_leaf, hipStream_t s);
void kr_weight_rhs(const uint32_t *cond, const double *w, int j0, int N, int ld, int nr, double *b,
                   hipStream_t s);
void kr_rows_axpy_clear(double *y, double *t, const double *w, int N, int ld, int nr,
                        hipStream_t s);
void kr_sub(const double *a, const double *b, double *out, int N, int ld, int nr, hipStream_t s);
// TI -> T: (double, double), (double, float), (float, float)
template <typename TI, typename T>
void kr_precond_cast(const TI *v, const double *dinv, T *z, int N, int ld, int nr, hipStream_t s);
void kr_gmres_start(int m, int nr, int ld, const double *beta, const double *bnorm, double tol,
                    const int *live, int *active, int *kused, double *g, double *inv_beta,
                    double *est, hipStream_t s);
void kr_gmres_givens(int k, int m, int nr, int ld, const double *h1, const double *h2,
                     const double *hn, double *H, double *cs, double *sn, double *g,
                     const double *bnorm, double tol, int *active, int *kused, double *est,
                     double *inv_h, hipStream_t s);
void kr_gmres_backsolve(int m, int nr, int ld, const double *H, const double *g, const int *kused,
                        double *Y, hipStream_t s);
// Segmented vectors: a batch of independent systems stacked in one column of
// N rows (device arrays; seg per row, a fixed block map for reductions).
struct SegMap {
  int N = 0, nsys = 0, nblk = 0;
  const int *seg = nullptr;      // N: system of each row
  const int *blk_lo = nullptr;   // nblk: first row of each reduction block
  const int *blk_hi = nullptr;   // nblk: end row of each reduction block
  const int *sys_blk = nullptr;  // nsys + 1: blocks of each system
};
size_t seg_partial_size(const SegMap &sm, int nvec_max);
// active (optional, per system): systems flagged 0 are skipped (dots 0,
// rows untouched)
template <typename TX>
void seg_dots(const SegMap &sm, const TX *X, size_t strideX, int nvec, const double *Y,
              double *partial, double *out, bool take_sqrt, hipStream_t s,
              const int *active = nullptr);
template <typename TV>
void seg_axpy_multi(const SegMap &sm, double *w, const TV *V, size_t strideV, const double *coef,
                    int nvec, double alpha, hipStream_t s, const int *active = nullptr);
template <typename TO>
void seg_scale(const SegMap &sm, const double *in, TO *out, const double *scale, const int *mask,
               hipStream_t s);
template <typename TV>
void seg_update_x(const SegMap &sm, double *x, const TV *V, size_t strideV, const double *Y,
                  const int *kused, int kmax, const double *dinv, hipStream_t s);
void seg_unit_rhs(const SegMap &sm, const uint32_t *cond, const int *rhs_cond, double *b,
                  hipStream_t s, const double *sign = nullptr);

template <typename TV>
void kr_gmres_update_x(double *x, const TV *V, size_t strideV, const double *Y,
                       const int *kused, int kmax, const double *dinv, int N, int ld, int nr,
                       hipStream_t s);

}  // namespace gp
