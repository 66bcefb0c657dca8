/*
 * gpurify.h - C ABI of libgpurify, the MI355X (gfx950) quasi-static
 * capacitance engine behind the field-solver PEX path.
 *
 * The library replaces the Maxwell BEM solve of the reference
 * (UW-ASIC/GPurify, Rust crate crates/pex).  Every entry point below names
 * the reference interface it stands in for.  All pointers are plain host
 * pointers owned by the caller and borrowed for the duration of the call
 * only; the library owns every device allocation, scoped to a context.
 * No C++ type or exception crosses this boundary: every function returns a
 * status code (GP_OK = 0) and gp_last_error() holds the message.
 *
 * Units: geometry in metres (panels) or database units scaled by xy_scale
 * metres/DBU (volume groups), heights and thicknesses in nm, capacitance in
 * farads (vacuum permittivity unless stated), cap_af in attofarads.
 * Matrices are row-major.
 */
#ifndef GPURIFY_H
#define GPURIFY_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status codes
 * Mirror the failure classes of BridgeError (crates/pex/src/bridge.rs:34-56)
 * plus the device classes the reference has no counterpart for. */
#define GP_OK 0
#define GP_ERR_INVALID_INPUT 1   /* bad panels, sizes, non-finite geometry or options */
#define GP_ERR_NO_DEVICE 2       /* no usable gfx950 device (or GPURIFY_FORCE_NO_DEVICE=1) */
#define GP_ERR_OOM 3             /* device allocation failed or exceeded GPURIFY_MAX_DEVICE_BYTES */
#define GP_ERR_NOT_CONVERGED 4   /* GMRES did not reach tol and opts.strict != 0 */
#define GP_ERR_NON_FINITE 5      /* BridgeError::NonFiniteResult (bridge.rs:54-55, 432-439) */
#define GP_ERR_INTERNAL 6        /* HIP runtime failure or an unexpected internal state */
#define GP_ERR_GRID_CAPACITY 7   /* BridgeError::GridCapacity (bridge.rs:46-47, 469-474) */
#define GP_ERR_PANEL_CAPACITY 8  /* BridgeError::PanelCapacity (bridge.rs:48-49, 381-386, 410-415) */
#define GP_ERR_SINGULAR 9        /* cap::SolveError::Singular (cap/solver.rs:39-40) */
#define GP_ERR_EMPTY 10          /* cap::SolveError::Empty (cap/solver.rs:37-38) */

/* ------------------------------------------------------------ context
 * Stands in for the panic-contained lazy GPU context of the backend crate
 * (crates/backend/src/gpu.rs:264-284) and gpu_ready() (crates/backend/src/lib.rs:50-59).
 * A context owns one HIP device, one stream and its device memory.  Calls on
 * one context are serialized internally; use one context per thread for
 * concurrency.  gp_ctx_create never aborts: with no usable device it returns
 * GP_ERR_NO_DEVICE and sets *out to NULL. */
typedef struct gp_ctx gp_ctx;

int gp_ctx_create(gp_ctx **out, uint32_t device_ordinal);
void gp_ctx_destroy(gp_ctx *ctx);
/* Message of the last failing call on ctx (or of the last failing
 * gp_ctx_create on this thread when ctx is NULL).  Never NULL. */
const char *gp_last_error(const gp_ctx *ctx);
/* 1 if a gfx950 device is usable now, else 0 (gpu_ready(), backend/src/lib.rs:50-59). */
int gp_device_ready(void);
/* Blocks until every queued operation on the context's device has finished
 * (no reference counterpart; benchmark timing brackets).  Every solve entry
 * point already returns only after its own work is complete. */
int gp_device_synchronize(gp_ctx *ctx);
/* Device name, arch and CU count as text. */
int gp_device_info(gp_ctx *ctx, char *buf, size_t len);
/* Library version string (static storage). */
const char *gp_version(void);

/* ------------------------------------------------------------ panels
 * The solve-time panel set (PanelSet, integrals/panel.rs:78-119): flat
 * quadrilateral panels, 4 vertices each, x,y,z per vertex, metres.  A
 * triangle repeats its last vertex (the repeated vertex adds a zero-length
 * edge and a zero-area fan triangle, so its integrals are unchanged).
 * conductor[i] in [0, n_conductors). */
typedef struct {
  uint32_t n_panels;
  uint32_t n_conductors;
  const double *vertices;    /* n_panels * 12 */
  const uint32_t *conductor; /* n_panels */
} gp_panels;

enum { GP_PREC_FP64 = 0, GP_PREC_MIXED = 1 };
enum { GP_METHOD_AUTO = 0, GP_METHOD_DENSE = 1, GP_METHOD_FMM = 2 };
/* M2L kernels: FP64 on matrix cores (FP64 VALU when the RHS block is not a
 * multiple of 16), FP32 on the VALU, BF16 inputs with FP32 accumulation on
 * matrix cores, FP64 forced onto the VALU, and BF16-rounded operands on the
 * FP32 VALU (the same rounding as GP_M2L_BF16 through independent kernels,
 * a cross-check of the matrix-core path; inner operator only). */
enum { GP_M2L_FP64 = 0, GP_M2L_FP32 = 1, GP_M2L_BF16 = 2, GP_M2L_FP64_VALU = 3,
       GP_M2L_BF16_VALU = 4 };
/* Near-field values of the inner operator of single-RHS mixed solves: FP32,
 * or FP16 scaled per row by the row's largest magnitude (FP32 scale). */
enum { GP_NEAR_FP32 = 0, GP_NEAR_FP16 = 1 };

/* Solver options.  gp_cap_opts_default() fills the reference's own
 * constants: the dense/FMM rule of bridge.rs:397-404 (dense below 500
 * panels, else FMM order 4, depth 4, tol 1e-8), the CPQR-truncated M2L
 * operators of bbfmm.rs:132-174, GMRES restart 200 and 50 restarts
 * (cap/fmm_solver.rs:164), FP64 throughout. */
typedef struct {
  uint32_t precision;    /* GP_PREC_FP64 | GP_PREC_MIXED */
  uint32_t method;       /* GP_METHOD_AUTO (reference rule) | DENSE | FMM */
  uint32_t fmm_order;    /* Chebyshev order n (reference: 4) */
  uint32_t max_level;    /* octree depth; 0 = chosen from leaf_target */
  uint32_t leaf_target;  /* panel-weighted mean leaf population when max_level == 0 */
  uint32_t restart;      /* GMRES(m) restart length */
  uint32_t max_restarts; /* GMRES restart cycles */
  uint32_t rhs_block;    /* right-hand sides solved together (0 = all) */
  double tol;            /* relative residual target per right-hand side */
  double inner_tol;      /* GP_PREC_MIXED: inner GMRES relative tolerance */
  uint32_t max_refine;   /* GP_PREC_MIXED: outer FP64 refinement steps */
  uint32_t m2l_inner;    /* GP_PREC_MIXED inner operator: GP_M2L_BF16 (default), _FP32 or _BF16_VALU */
  uint32_t m2l_outer;    /* FP64 operator (FP64 solve, mixed outer residual): GP_M2L_FP64 or _FP64_VALU */
  uint32_t strict;       /* 1: non-convergence returns GP_ERR_NOT_CONVERGED */
  uint32_t dense_limit;  /* AUTO: dense below this many panels (reference 500) */
  uint32_t rhs_first;    /* solve conductor columns [rhs_first, rhs_first + rhs_count) */
  uint32_t rhs_count;    /* 0 = all columns; unsolved columns of C are left zero */
  uint32_t inner_passes; /* GP_PREC_MIXED: Gram-Schmidt passes of the inner GMRES, 1 or 2 */
  uint32_t inner_basis;  /* GP_PREC_MIXED: inner Krylov basis storage, GP_BASIS_FP32 | _FP64 */
  uint32_t m2l_rank;     /* GP_PREC_MIXED inner M2L of a block of RHS: 0 = full n^3; r > 0
                            (default 32) projects the interaction matrices on their leading
                            singular vectors (r rounded up to 32, 64 or 128; no projection
                            when that reaches n^3 rounded up to 32, i.e. below order 4) and
                            runs split-BF16 products (hi*hi + hi*lo + lo*hi).  Single-RHS
                            solves and the FP64 operator are never projected. */
  uint32_t step_graphs;  /* GP_PREC_MIXED: 1 (default) = from the second refinement step on,
                            each inner GMRES step (operator apply, orthogonalisation, Givens
                            update) runs as a HIP graph captured on first use and replayed;
                            a batch (gp_cap_solve_batched with skip_converged) replays four
                            steps per graph and reads the convergence flags once per graph;
                            0 = kernels launched one by one.  Off while profiling. */
  uint32_t m2l_compress; /* 1 (default) = the reference's truncated M2L operators: every far-field
                            interaction matrix compressed by column-pivoted QR, truncated where the
                            remaining squared column norm falls below eps = max(1e-3 * 0.1^n,
                            1e-14) of the largest (bbfmm.rs:132-174, linalg.rs:277-381), and
                            applied in factored form, U (V^T x), as the reference applies it, by
                            the single-RHS FP64 tiles (the dense product U V^T for blocks of RHS
                            and where the autotuner prefers it); 0 = the full matrices.  The
                            adjoint solves of gp_cap_jacobian always apply the full matrices. */
  uint32_t fuse_leaf;    /* 1 (default) = FP32 blocks of RHS at orders <= 5 run the finest L2L and
                            L2P as one kernel; 0 = two launches (the same sums). */
  uint32_t skip_converged; /* gp_cap_solve_batched: 1 (default) = the operator applies of a GMRES
                            step skip the problems whose system has converged (their basis
                            vector is zero; the others' results are unchanged bit for bit);
                            0 = every problem in every step. */
  uint32_t batch_groups; /* gp_cap_solve_batched: the FMM problems are solved in this many groups
                            (each its own stacked operator and GMRES), and the host set-up of a
                            group (octrees, near-field and interaction lists) runs on a second
                            thread while the device solves the group before it; the first
                            group's set-up overlaps the dense problems' solve.  Group k takes
                            a share of the FMM panels proportional to 1.5^k (a set-up costs
                            about 0.6 of the solve of the same panels, so each set-up hides
                            behind the previous group's solve).  0 (default) = 1 below 2
                            million FMM panels, else 4; 1 = one group, every launch from the
                            calling thread; results agree with one group to the solve
                            tolerance. */
  uint32_t near_inner;   /* GP_PREC_MIXED with one RHS per problem (gp_cap_solve_batched, or
                            rhs_block 1): the inner operator's near-field values, GP_NEAR_FP32
                            (default) or GP_NEAR_FP16 (each row scaled by its largest magnitude
                            and stored as FP16, 11 significant bits per entry: half the memory
                            of FP32, the same answer; the single-RHS product is bound by its
                            per-leaf load latency, not by the value bytes, and measured slower
                            with FP16 values).  Blocks of RHS always keep FP32 values; the outer
                            residual is FP64 either way. */
  uint32_t inner_order;  /* GP_PREC_MIXED: Chebyshev order of the inner operator's far field
                            (P2M, M2M, M2L, L2L, L2P of the FP32 / BF16 operator, on the same
                            tree, near field and interaction lists as the FP64 operator, its
                            M2L matrices truncated by the same CPQR rule at that order); 0
                            (default) = fmm_order.  Must be in [2, fmm_order].  The FP64 outer
                            residual keeps fmm_order, so the converged answer is that of the
                            fmm_order operator whatever the inner order: a lower order only
                            changes how far each inner solve moves the iterate (the refinement
                            contraction is bounded by inner_tol plus the inner operator's
                            relative error). */
  uint32_t ground;       /* GP_GROUND_NONE (default): the panels as given in free space, the
                            reference's model (its substrate is a meshed finite plate among the
                            panels, bridge.rs:346-379).  GP_GROUND_IMAGE: a perfectly conducting
                            infinite plane z = ground_z by the method of images, the Green's
                            function 1/|x - y| - 1/|x - y*| with y* = (y.x, y.y, 2 ground_z -
                            y.z); every vertex must lie above the plane; C is the conductors'
                            Maxwell matrix with the plane as the grounded reference.  Dense
                            path: each entry adds the mirror panel's integral; FMM path: the
                            mirror panels join the octree as explicit panels (2N in all) with
                            antisymmetric right-hand sides.  gp_solve_capacitance replaces its
                            substrate plate by the plane at the plate's height (ground_z is then
                            ignored).  Not available for the dielectric solves and
                            gp_cap_jacobian. */
  double ground_z;       /* GP_GROUND_IMAGE: the plane's height, metres */
  uint32_t m2l_tile_order; /* single-RHS M2L target tiles (64 targets of one octant): 0 (default)
                            = longest interaction lists first, 1 = box (Morton) order within
                            the octant; the same sums either way, only which targets share a
                            tile and which tiles run together change. */
} gp_cap_opts;
enum { GP_GROUND_NONE = 0, GP_GROUND_IMAGE = 1 };
enum { GP_BASIS_FP32 = 0, GP_BASIS_FP64 = 1 };

void gp_cap_opts_default(gp_cap_opts *opts);

/* Per-solve report (the reference has none; BackendTelemetry,
 * crates/backend/src/lib.rs:75-84, is the closest). */
typedef struct {
  uint32_t n_panels, n_conductors, method_used, precision_used;
  uint32_t iterations;    /* GMRES iterations (inner, summed over refinement), max over RHS */
  uint32_t matvecs;       /* operator applications (each covers all RHS of a block) */
  uint32_t refine_steps;  /* outer FP64 refinement steps (mixed) */
  uint32_t converged;     /* 1 if every RHS reached tol */
  uint32_t max_level, n_leaves, n_boxes, fmm_order;
  uint64_t near_nnz, m2l_pairs;
  double final_rel_residual; /* max over RHS of ||b - A x|| / ||b|| (FP64 operator) */
  double seconds_total, seconds_setup, seconds_solve;
  double seconds_near, seconds_far, seconds_orth; /* time in near-field, far-field, GMRES kernels */
  double seconds_upload;  /* host-to-device copies of the panel data (vertices, centroids,
                             areas, conductor ids), part of seconds_total */
} gp_cap_report;

/* Maxwell capacitance matrix of a panel set.  Stands in for cap::solve
 * (crates/pex/src/quasistatic/cap/solver.rs:100-134) and
 * cap::fmm_solver::solve (cap/fmm_solver.rs:147-172), i.e. for the method
 * selection and solve of bridge.rs:397-408 before the eps_r scaling.
 * C: n_conductors^2 doubles, farads, vacuum.  report may be NULL. */
int gp_cap_solve(gp_ctx *ctx, const gp_panels *panels, const gp_cap_opts *opts, double *C,
                 gp_cap_report *report);

/* The per-net solves of one layout in one call: what the reference runs as
 * one solve_capacitance (bridge.rs:307-443 -> cap::solve /
 * cap::fmm_solver::solve) per net.  Each problem is independent (its own
 * conductors, e.g. one net and its substrate plate); problems below
 * dense_limit (GP_METHOD_AUTO) take the dense path, all others are solved
 * together as one stacked block-diagonal FMM operator with one GMRES per
 * problem in lockstep.  Columns [rhs_first, rhs_first + rhs_count) of each
 * problem's C are solved (FMM problems; dense ones solve every column).
 * C: the problems' n_conductors^2 blocks back to back, farads, vacuum.
 * reports: n_problems entries or NULL (times are those of the whole call). */
int gp_cap_solve_batched(gp_ctx *ctx, const gp_panels *problems, uint32_t n_problems,
                         const gp_cap_opts *opts, double *C, gp_cap_report *reports);

/* Geometric parameters for sensitivities.  Parameter p moves the panels
 * panel[param_off[p] .. param_off[p+1]) (caller indices, each at most once
 * per parameter) with vertex velocities dverts (12 doubles per entry, the
 * d vertex / d theta of the panel's 4 vertices, metres per metre of theta);
 * all other panels stay.  A width change stretches the side and cross
 * faces, a spacing or height change translates a conductor. */
typedef struct {
  uint32_t n_params;
  const uint32_t *param_off;  /* n_params + 1 */
  const uint32_t *panel;      /* param_off[n_params] entries */
  const double *dverts;       /* 12 * param_off[n_params] */
} gp_param_map;

/* C and its Jacobian dC/dtheta (no reference counterpart; the reference
 * has no adjoint: the forward half of a forward/backward pair modelled on
 * torch.autograd.Function, whose backward is dL/dtheta_p = sum_ij
 * dL/dC_ij J[p][i][j]).  One forward solve P sigma_j = e_j and one adjoint
 * solve P^T lambda_i = a_i (a_i: areas on conductor i) per conductor, then
 *   dC_ij/dtheta = sum_(k on i) (da_k/dtheta) sigma_kj
 *                  - sum_(k,l) lambda_ki (dP_kl/dtheta) sigma_lj
 * with dP_kl the exact forward-mode derivative of the closed-form integral
 * (collocation point and panel both moving) over every entry a moved panel
 * touches.  Method per opts (dense below dense_limit, else FMM with the
 * transposed operator for the adjoint).  C: n_conductors^2 (F, vacuum);
 * J: n_params * n_conductors^2 (F/m).  The dielectric k scales C (bridge.rs:
 * 405-407): dC/dk = C / k needs no solve.  report may be NULL. */
int gp_cap_jacobian(gp_ctx *ctx, const gp_panels *panels, const gp_cap_opts *opts,
                    const gp_param_map *params, double *C, double *J, gp_cap_report *report);

/* The precorrected FMM operator apply y = P x of FmmCapOperator::apply
 * (cap/fmm_solver.rs:105-130): exact-Wilton near tiles plus the far field of
 * the panel charges area * x, built as gp_cap_solve builds it for
 * GP_METHOD_FMM (opts: fmm_order, max_level / leaf_target).  precision
 * GP_PREC_FP64 applies the FP64 operator, GP_PREC_MIXED the FP32 one.
 * X, Y: n_panels x nr row-major (panel-major), SI units (V per C/m^2). */
int gp_fmm_apply(gp_ctx *ctx, const gp_panels *panels, const gp_cap_opts *opts, uint32_t nr,
                 uint32_t precision, const double *X, double *Y);

/* One M2L interaction matrix of the Chebyshev bbFMM as the FMM operator
 * applies it (host only: no context, no device).  G[m * n^3 + l] =
 * 1 / |x_m - x_l - 2 o| over the tensor Chebyshev nodes of order `order`
 * (2..6) on the reference box for the far-field box offset o (components in
 * [-3, 3], at least one of magnitude 2 or 3), as precompute_m2l builds it
 * (bbfmm.rs:132-166); with compressed = 1 its truncation U V^T by
 * LowRankApprox::compress (linalg.rs:277-381) at eps = max(1e-3 * 0.1^n,
 * 1e-14) (bbfmm.rs:150), whose rank goes to *rank (may be NULL).
 * G: order^6 doubles. */
int gp_m2l_operator(uint32_t order, int32_t o0, int32_t o1, int32_t o2, uint32_t compressed,
                    double *G, uint32_t *rank);

/* The factors of that truncation, as the reference keeps and applies them
 * (LowRankApprox u and vt, linalg.rs:277-381; y += u (vt x), bbfmm.rs:132-174)
 * and the FP64 FMM operator's factored M2L kernel reads them (host only):
 * U[m * n^3 + s] and Vt[s * n^3 + l] for s < *rank, zero beyond.  U, Vt:
 * order^6 doubles each. */
int gp_m2l_factors(uint32_t order, int32_t o0, int32_t o1, int32_t o2, double *U, double *Vt,
                   uint32_t *rank);

/* ------------------------------------------------------------ fixed workspace
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid allocation): the dense path of gp_cap_solve (opts->method =
 * GP_METHOD_DENSE in FP64: collocation assembly, one-work-group LU, charge
 * accumulation) run entirely in a fixed per-device workspace of
 * GP_NOALLOC_SCRATCH_BYTES, which the library holds from the first
 * gp_ctx_create on a device until the process ends.  The call makes no
 * hipMalloc, hipFree or allocating library call.  Inputs and outputs as
 * gp_cap_solve; C equals its dense result.  At most GP_NOALLOC_MAX_PANELS
 * panels (the reference's dense rule is below 500) and
 * GP_NOALLOC_MAX_CONDUCTORS conductors, else GP_ERR_INVALID_INPUT.  Calls on
 * one device share the workspace, which keeps the last call's contents
 * (every region a call reads it first writes); the library serializes these
 * calls within a process, and calls that overlapped on one device would race. */
#define GP_NOALLOC_MAX_PANELS 512
#define GP_NOALLOC_MAX_CONDUCTORS 64
#define GP_NOALLOC_SCRATCH_BYTES 2726656
int gp_cap_solve_dense_noalloc(gp_ctx *ctx, const gp_panels *panels, double *C,
                               gp_cap_report *report);

/* Kernel profiler (no reference counterpart; the reference has no timers).
 * Enabled, every device kernel launch of this context is bracketed by HIP
 * events on its stream.  gp_profile_report writes a JSON object
 * {"kernel": {"launches": n, "seconds": s, "flops": f, "bytes": b}, ...}
 * with the summed event durations and algorithmic flops / bytes since the
 * last report (or enable), then resets the counters. */
int gp_profile_enable(gp_ctx *ctx, int enable);
int gp_profile_report(gp_ctx *ctx, char *json, size_t len);

/* Kernel-variant autotuner (no reference counterpart; the context's
 * autotune cache).  Tunable kernels (near-field products, M2L work-group
 * orders) time their variants on the first launch of each shape and keep the
 * fastest for the context's lifetime.  gp_tune_force(ctx, v) pins every
 * tunable kernel to variant v (clamped to its count) and clears the cache;
 * v < 0 returns to tuning.  gp_tune_report writes a JSON list
 * [{"kernel": name, "key": [k0, k1, k2], "variant": v, "ms": [t0, ...]}, ...]. */
int gp_tune_force(gp_ctx *ctx, int variant);
int gp_tune_report(gp_ctx *ctx, char *json, size_t len);

/* ------------------------------------------------------------ layout bridge
 * One (net, layer) volume group of bridge.rs:58-65, 131-141: the rectangles
 * of one net on one layer, extruded from height_nm to height_nm+thickness_nm.
 * rects: n_rects * 4 int32 (x0, y0, x1, y1) in DBU. */
typedef struct {
  uint32_t net;
  uint32_t n_rects;
  double height_nm;
  double thickness_nm;
  double dielectric_k;
  const int32_t *rects;
} gp_volume_group;

/* Panelization of bridge.rs:307-386 + 445-652 (union grid, greedy-merged
 * top/bottom faces, exposed sidewalls, substrate plate).  Groups must come
 * in the reference's BTreeMap order, i.e. sorted by (net, layer).
 * refine >= 1 splits every reference panel refine x refine (1 = the
 * reference mesh).  max_panels 0 = the reference cap MAX_PANELS = 20000.
 * Two-call protocol: with vertices == NULL only *n_panels is written. */
int gp_mesh_groups(const gp_volume_group *groups, uint32_t n_groups, double xy_scale,
                   uint32_t refine, uint32_t max_panels, uint32_t *n_panels, double *vertices,
                   uint32_t *conductor, double *panel_eps);

/* solve_capacitance (bridge.rs:307-443): mesh, solve, scale by eps_r
 * (homogeneous) or by per-panel eps (inhomogeneous: the reference's
 * conductor-role dielectric solve, dense below dense_limit panels and
 * GP_ERR_PANEL_CAPACITY above, as bridge.rs:409-415; opts->method =
 * GP_METHOD_FMM solves it with the FMM operator at any size), and return
 * cap_af = C[i][i] * 1e18 per net in ascending net order.  No groups: GP_OK
 * with *n_nets = 0 (bridge.rs:311-313).
 * nets_out / cap_af_out: room for the number of distinct nets.
 * C_out (optional): (n_nets+1)^2 farads with eps applied, substrate last. */
int gp_solve_capacitance(gp_ctx *ctx, const gp_volume_group *groups, uint32_t n_groups,
                         double xy_scale, uint32_t refine, uint32_t max_panels,
                         const gp_cap_opts *opts, uint32_t *n_nets, uint32_t *nets_out,
                         double *cap_af_out, double *C_out, gp_cap_report *report);

/* ------------------------------------------------------------ dielectrics
 * Panel roles of cap::dielectric::Problem (cap/dielectric.rs:10-14):
 * a conductor panel (conductor id, eps_surrounding: the relative
 * permittivity of the medium around it, weighting its charge), or a
 * dielectric-interface panel (eps_out / eps_in: relative permittivities on
 * the outer / inner side; reference: a point on the inner side, which
 * orients the panel's normal outward, dielectric.rs:39-46). */
enum { GP_ROLE_CONDUCTOR = 0, GP_ROLE_DIELECTRIC = 1 };
typedef struct {
  uint32_t kind;
  uint32_t conductor;
  double eps_surrounding;
  double eps_out;
  double eps_in;
  double reference[3];
} gp_panel_role;

/* cap::solve_dielectric (cap/dielectric.rs:47-151): conductor rows of
 * potential coefficients, interface rows of normal fields (7-point rule,
 * integrals/panel.rs:213-253) with the jump (eps_out + eps_in) /
 * (2 eps0 (eps_out - eps_in)) on the diagonal, dense FP64 LU on the device.
 * roles: one per panel.  C: n_conductors^2 farads, permittivities applied. */
int gp_cap_solve_dielectric(gp_ctx *ctx, const gp_panels *panels, const gp_panel_role *roles,
                            double *C, gp_cap_report *report);

/* The problem of gp_cap_solve_dielectric on the FMM operator, for sizes the
 * dense LU cannot reach (no counterpart in the reference, which solves it
 * dense only; FP64).  The rows are those of cap::solve_dielectric
 * (cap/dielectric.rs:76-151) over the near field of the FMM (each leaf and
 * its neighbour leaves); the far field is the centroid-charge field of the
 * conductor FMM: its potential for conductor rows and minus its normal
 * derivative (the differentiated local expansion) for interface rows.
 * Jacobi-preconditioned GMRES per conductor (opts: fmm_order, max_level or
 * leaf_target, tol, restart, max_restarts, m2l_compress; precision must be
 * GP_PREC_FP64).  C: n_conductors^2 farads, permittivities applied. */
int gp_cap_solve_dielectric_fmm(gp_ctx *ctx, const gp_panels *panels, const gp_panel_role *roles,
                                const gp_cap_opts *opts, double *C, gp_cap_report *report);

/* ------------------------------------------------------------ FastCap input
 * FastCap "quickif" geometry text (.qui), as cap::geometry::parse
 * (crates/pex/src/quasistatic/cap/geometry.rs:56-135) reads it: Q (quad)
 * and T (triangle) panel lines, N name lines, comment lines starting with
 * 0, * or %; conductor numbers interned to 0-based indices in first-seen
 * order; a C include or any other line is GP_ERR_INVALID_INPUT with
 * "line <n>: ..." in gp_last_error(NULL).  Coordinates are multiplied by
 * scale.  The panels come in the gp_panels layout (a triangle repeats its
 * last vertex), ready for gp_cap_solve.
 * Two-call protocol: with vertices == NULL only *n_panels, *n_conductors and
 * *names_len (bytes of the names, each NUL-terminated, in conductor order)
 * are written; the second call fills every non-NULL output (names needs
 * *names_len bytes, n_vertices gets 3 or 4 per panel). */
int gp_qui_parse(const char *text, size_t text_len, double scale, uint32_t *n_panels,
                 uint32_t *n_conductors, size_t *names_len, double *vertices,
                 uint32_t *conductor, uint8_t *n_vertices, char *names);

/* ------------------------------------------------------------ DC resistance
 * One layout polygon for the resistance half of extract_quasistatic
 * (bridge.rs:97-129, 180-299).  xy: the ring of n_vertices points (x, y
 * int32 DBU, first point not repeated), as GeometryStore holds it.  has_pex:
 * the polygon's layer has deck.pex parameters (else it is skipped, as
 * bridge.rs:99-101); sheet_res_ohm_sq / via_res_ohm are those parameters;
 * volume_conductor is is_volume_conductor(params) (bridge.rs:170-177);
 * thickness_nm / height_nm are the layer's ProcessStack values. */
typedef struct {
  uint32_t net;
  uint32_t n_vertices;
  const int32_t *xy;
  uint32_t has_pex;
  uint32_t volume_conductor;
  double sheet_res_ohm_sq;
  double via_res_ohm;
  double thickness_nm;
  double height_nm;
} gp_res_polygon;

/* r_ohm per net (NetParasitics.r_ohm, analytical/mod.rs:353-358) of
 * extract_quasistatic: the polygons' via resistances in polygon order, then
 * the port resistance of every sheet-resistance polygon, which the
 * reference obtains from solve_resistance -> henry::solve_with(Direct) at
 * f = 0 (bridge.rs:180-299, henry/solver.rs:336-584) and this library
 * evaluates in closed form, operation for operation (each polygon is an
 * isolated one-filament segment, so the nodal system is diagonal).  Nets
 * come in ascending order; only nets the reference's map would hold appear.
 * Errors: GP_ERR_INVALID_INPUT (the reference's UnsupportedGeometry /
 * InvalidProcessValue / InvalidDatabaseUnit messages), GP_ERR_NON_FINITE.
 * Two-call protocol: with nets_out == NULL only *n_nets is written. */
int gp_solve_resistance(const gp_res_polygon *polys, uint32_t n_polys, double dbu_nm,
                        uint32_t *n_nets, uint32_t *nets_out, double *r_ohm_out);

#ifdef __cplusplus
}
#endif
#endif /* GPURIFY_H */
