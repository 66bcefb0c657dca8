// The dense solve in a fixed per-device workspace (no device allocation).
#pragma once

#include "../../include/gpurify.h"

struct gp_ctx;

namespace gp {

// Takes the address of the fixed workspace on `device` (which must be the
// current device) once per device; later calls return at once.  A failure
// is kept and reported by the entry point.
void noalloc_init(int device);

// gp_cap_solve's dense path (FP64 assembly, LU, accumulation) in the fixed
// workspace; same result.  At most GP_NOALLOC_MAX_PANELS panels and
// GP_NOALLOC_MAX_CONDUCTORS conductors.
void cap_solve_dense_noalloc(gp_ctx *ctx, const gp_panels &in, double *C, gp_cap_report *rep);

}  // namespace gp
