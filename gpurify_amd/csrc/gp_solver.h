// Host orchestration of the capacitance solve.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gpurify.h"
#include "gp_common.h"
#include "gp_tree.h"

struct gp_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // uploads and set-up kernels of a batch's next group of problems, issued
  // by a second host thread while `stream` solves the current group
  hipStream_t setup_stream = nullptr;
  // the near-field assembly of that next group, so the set-up's later
  // uploads (each synchronising setup_stream) need not wait for it
  hipStream_t asm_stream = nullptr;
  std::mutex mu;
  std::string last_error;
  std::string device_name;
  int cu_count = 0;
  gp::Prof prof;
  gp::Tuner tuner;
  // rocBLAS handle of the large dense dielectric LU (rocSOLVER), created on
  // first use on this context's stream; released by gp_ctx_destroy.
  std::shared_ptr<void> blas;
};

namespace gp {

// Maxwell C (ncond x ncond, farads, vacuum) of g.  eps_panel, when given,
// weights each panel's charge by its surrounding permittivity (the bridge's
// inhomogeneous conductor-role solve; dense or FMM, the operator is the
// same potential-coefficient matrix).
// cap::solve_dielectric (cap/dielectric.rs:47-151), dense FP64 on the device.
void solve_dense_roles(gp_ctx *ctx, const PanelGeom &g, const gp_panel_role *roles, double *C,
                       gp_cap_report *rep, const double *ground_z = nullptr);
void cap_solve(gp_ctx *ctx, const PanelGeom &g, const gp_cap_opts &opts, double *C,
               gp_cap_report *rep, const double *eps_panel);

// cap::solve_dielectric's problem (cap/dielectric.rs:76-151) on the FMM
// operator (no reference counterpart; FP64): near tiles with the interface
// rows' normal-field integrals and jump terms, far field of the centroid
// charges (potential for conductor rows, minus its normal derivative for
// interface rows).  C: n_conductors^2 farads, permittivities applied.
void solve_dielectric_fmm(gp_ctx *ctx, const PanelGeom &g, const gp_panel_role *roles,
                          const gp_cap_opts &opts, double *C, gp_cap_report *rep);

// Maxwell C of each independent problem of a batch, C[p] ncond_p x ncond_p.
// Problems below the dense limit (method AUTO) or all (DENSE) take the dense
// path one by one; the FMM problems are solved together (stacked operator,
// lockstep per-problem GMRES).  reps: one per problem, or null.
void cap_solve_batched(gp_ctx *ctx, const std::vector<PanelGeom> &gs, const gp_cap_opts &opts,
                       const std::vector<double *> &C, gp_cap_report *reps);

// Geometric parameters: parameter p moves the caller's panels
// panel[off[p] .. off[p+1]) with vertex velocities dv (12 per entry, metres
// per metre of theta).
struct ParamMap {
  int n = 0;
  const uint32_t *off = nullptr, *panel = nullptr;
  const double *dv = nullptr;
};
// C (farads) and J = dC/dtheta (n x ncond x ncond, F/m) by one forward and
// one adjoint (transposed) solve per conductor and forward-mode derivatives
// of the closed-form integrals of every entry a moved panel touches.
void cap_jacobian(gp_ctx *ctx, const PanelGeom &g, const gp_cap_opts &opts, const ParamMap &pm,
                  double *C, double *J, gp_cap_report *rep);

// Apply the precorrected FMM operator (FP64, no preconditioner) to nr
// columns X (n x nr, caller order), y = P X, farads^-1 m^-2 units of P.
// Test entry used to compare the device operator with the oracle's.
void fmm_apply_host(gp_ctx *ctx, const PanelGeom &g, const gp_cap_opts &opts, int nr,
                    const double *X, double *Y, int precision);

}  // namespace gp
