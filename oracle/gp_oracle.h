/*
 * gp_oracle.h - CPU FP64 restatement of the reference quasi-static
 * capacitance path (UW-ASIC/GPurify, crates/pex).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * by nothing else.  The product (libgpurify) never links or calls it.
 *
 * Every function states the reference file:line it restates.  Paths are
 * relative to the reference repository root.  Geometry is in metres,
 * capacitances in farads (vacuum unless a dielectric factor is applied),
 * matrices row-major.
 */
#ifndef GP_ORACLE_H
#define GP_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* crates/pex/src/quasistatic/integrals/panel.rs:147-208 (Wilton closed form).
 * verts: nv*3 doubles. */
double or_potential_integral(const double *verts, int nv, const double *p);

/* panel.rs:232-253 (7-point normal-field quadrature). */
double or_field_integral(const double *verts, int nv, const double *p, const double *n_obs);

/* panel.rs:213-221: the 7-point triangle rule as 7 rows (l0, l1, l2, weight). */
void or_tri7(double *tab);

/* linalg.rs:277-381 (LowRankApprox::compress): column-pivoted MGS QR of the
 * m x n row-major a, truncated at squared column norm < eps * max.
 * u: m*k row-major, vt: k*n row-major (room for k = min(m, n)); returns k. */
int or_lowrank_compress(const double *a, int m, int n, double eps, double *u, double *vt);

/* bbfmm.rs:150 and 156-166: the truncation eps of order n and the dense M2L
 * matrix of offset (o0, o1, o2) on the reference box (n^3 x n^3). */
double or_m2l_eps(int order);
void or_m2l_matrix(int order, int o0, int o1, int o2, double *G);

/* panel.rs:29-67 + 95-119: centroid (fan, area weighted), area, Newell normal.
 * verts: n*4*3 doubles (quads; a triangle repeats its last vertex). */
void or_panel_props(int n, const double *verts4, double *cx, double *cy, double *cz,
                    double *area, double *nrm /* n*3, may be NULL */);

/* crates/pex/src/quasistatic/cap/solver.rs:81-97.  P (n*n, row-major, F^-1 m^-2 * m^2 units of V/(C/m^2)) */
void or_assemble_p(int n, const double *verts4, double *P);

/* crates/pex/src/quasistatic/linalg.rs:172-252: partial-pivot LU solve of
 * A X = B for nrhs columns stored as B[col*n + row].  Returns 0, or -1 when singular. */
int or_lu_solve(int n, const double *A, int nrhs, double *B);

/* cap/solver.rs:100-134.  method 0 = Direct (LU), 1 = Iterative (GMRES tol 1e-8,
 * restart 100, no preconditioner).  C: ncond*ncond row-major [F, vacuum]. */
int or_cap_solve_dense(int n, const double *verts4, const uint32_t *cond, int ncond,
                       int method, double *C, double *P_out /* optional n*n */);

/* cap/fmm_solver.rs:147-172 with LaplaceFmm (fmm/bbfmm.rs) and gmres
 * (krylov.rs:50-191).  m2l_mode 0 = reference CPQR-compressed M2L
 * (bbfmm.rs:132-174, linalg.rs:277-400); 1 = uncompressed dense M2L (same
 * interpolant, no truncation).  iters_out: per-conductor GMRES iteration
 * counts (ncond ints, may be NULL); resid_out likewise (doubles). */
int or_cap_solve_fmm(int n, const double *verts4, const uint32_t *cond, int ncond,
                     int order, int max_level, double tol, int m2l_mode,
                     double *C, int *iters_out, double *resid_out);

/* Ground plane z = zg by images (no reference counterpart): dense LU
 * (or_cap_solve_dense_image) and the FMM form over the panels and their
 * mirrors with an antisymmetric right-hand side (or_cap_solve_fmm_image).
 * C: ncond^2, farads, the conductors' Maxwell matrix with the plane as the
 * grounded reference.  -7 when a vertex is not above the plane. */
int or_cap_solve_dense_image(int n, const double *verts4, const uint32_t *cond, int ncond,
                             double zg, double *C);
int or_cap_solve_fmm_image(int n, const double *verts4, const uint32_t *cond, int ncond, double zg,
                           int order, int max_level, double tol, int m2l_mode, double *C,
                           int *iters_out, double *resid_out);
/* As or_cap_solve_fmm for the conductor columns [j0, j0 + nj) only (the
 * bounded CPU sample).  seconds_out[0] = operator setup, [1] = GMRES. */
int or_cap_solve_fmm_cols(int n, const double *verts4, const uint32_t *cond, int ncond,
                          int order, int max_level, double tol, int m2l_mode, int j0, int nj,
                          double *C, int *iters_out, double *seconds_out);

/* The precorrected operator y = P sigma of cap/fmm_solver.rs:105-130 for a
 * given sigma (one column). */
int or_fmm_cap_apply(int n, const double *verts4, int order, int max_level, int m2l_mode,
                     const double *sigma, double *y);

/* fmm/bbfmm.rs:176-187 + 209-448: point-to-point Laplace FMM matvec
 * y_i = sum_{j != i} q_j / |x_i - x_j| (with_p2p=1) or its far-field part (0). */
int or_fmm_matvec(int npts, const double *pts /* npts*3 */, const double *q, int order,
                  int max_level, int with_p2p, int m2l_mode, double *y);

/* Direct O(N^2) sum y_i = sum_{j != i} q_j/|x_i-x_j| (crates/pex/tests/fmm.rs:6-19). */
void or_dense_p2p(int npts, const double *pts, const double *q, double *y);

/* crates/pex/tests/fmm.rs:21-29 (the LCG point set). */
void or_lcg_points(int n, uint64_t seed, double *pts);

/* fmm/chebyshev.rs:11-17, 40-74. */
void or_cheb_nodes(int n, double *nodes);
void or_interp_weights(double x, int n, double *w);

/* fmm/tree.rs:41-210: octree statistics used by tests.
 * out[0]=#nodes out[1]=#leaves out[2]=sum leaf points out[3]=#interaction pairs
 * out[4]=min interaction centre distance / box width (over leaf level) out[5]=#near pairs */
int or_octree_stats(int npts, const double *pts, int max_level, double *out);

/* Near-field CSR of the precorrected operator, fmm_solver.rs:52-91 (for
 * comparing tree/near lists).  Returns nnz.  If off/idx are non-NULL they are
 * filled (off: n+1, idx: nnz). */
long or_fmm_near_pattern(int n, const double *verts4, int order, int max_level,
                         int64_t *off, uint32_t *idx);

/* crates/pex/src/quasistatic/cap/mesh.rs:12-43 (plate) and 50-116 (icosphere).
 * Returns the panel count; verts4 may be NULL to query. Triangles repeat v2 as v3. */
int or_mesh_plate(double cx, double cy, double z0, double lx, double ly, int nx, int ny,
                  double *verts4);
int or_mesh_icosphere(double cx, double cy, double cz, double r, int subdiv, double *verts4);

/* crates/pex/src/analytical/process_stack.rs:31-76: resolve the heuristic stack.
 * Inputs are sorted by LayerId.  Outputs thickness/height/k per layer. */
void or_process_stack(int nlayers, const double *thick_in, const double *height_in,
                      const double *k_in, double *thick, double *height, double *k);

/* crates/pex/src/bridge.rs:307-443 + 445-652 (solve_capacitance with the
 * union-grid mesher, ground plate, dense/FMM rule, eps scaling).  Groups are
 * supplied in (net, layer) order.  rect_off: n_groups+1 offsets into rects
 * (4 int32 per rect: x0,y0,x1,y1 DBU).  refine >= 1 splits every reference
 * panel refine x refine (refine = 1 is the reference mesh).  max_panels 0
 * means the reference cap (20000).
 * mode: 0 = reference rule (dense < 500 else FMM(4,4,1e-8)),
 *       1 = force dense LU, 2 = force FMM with (order, max_level, tol).
 * Outputs: net ids (sorted, n_nets) and cap_af per net; C (n_nets+1)^2 [F, eps applied].
 * Returns 0, or a negative error code (-2 grid cap, -3 panel cap, -4 non-finite,
 * -5 inhomogeneous dielectric >= 500 panels, -6 empty). */
int or_mesh_groups(int n_groups, const uint32_t *net, const double *height_nm,
                   const double *thickness_nm, const double *k, const int64_t *rect_off,
                   const int32_t *rects, double xy_scale, int refine, int max_panels,
                   int *n_panels_out, double *verts4, uint32_t *cond, double *eps);
int or_solve_capacitance(int n_groups, const uint32_t *net, const double *height_nm,
                         const double *thickness_nm, const double *k, const int64_t *rect_off,
                         const int32_t *rects, double xy_scale, int refine, int max_panels,
                         int mode, int order, int max_level, double tol,
                         int *n_nets_out, uint32_t *nets_out, double *cap_af_out, double *C_out,
                         int *n_panels_out);

/* Number of OpenMP threads the oracle uses (0 = library default). */
void or_set_threads(int n);
int or_get_threads(void);

/* No reference counterpart: the dielectric formulation of
 * cap/dielectric.rs:76-151 on the FMM operator (near rows as the dense
 * rows over the leaf and its neighbours; far field: the centroid-charge
 * potential for conductor rows, -k_e n . its gradient for interface rows),
 * Jacobi-preconditioned GMRES per conductor.  Arguments as
 * or_solve_dielectric plus the FMM order, depth, tolerance and m2l_mode. */
int or_solve_dielectric_fmm(int n, const double *verts4, const int *role, const uint32_t *cond,
                            const double *eps, const double *eps_out, const double *eps_in,
                            const double *ref, int ncond, int order, int max_level, double tol,
                            int m2l_mode, double *C);

/* d/dx of the 1-D Chebyshev interpolation weights at x (n values). */
void or_interp_dweights(double x, int n, double *dw);

#ifdef __cplusplus
}
#endif
#endif
