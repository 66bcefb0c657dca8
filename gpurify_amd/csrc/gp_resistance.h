// DC resistance half of extract_quasistatic (bridge.rs:180-299 through the
// henry nodal solve at f = 0).
#pragma once
#include <cstdint>
#include <map>

#include "../../include/gpurify.h"

namespace gp {

// Equivalent rectangle of a Manhattan polygon (analytical/mod.rs:577-628);
// throws Error(GP_ERR_INVALID_INPUT) with the reference's message.
void rectilinear_metrics(const int32_t *xy, int n, uint32_t poly, double &length_nm,
                         double &width_nm);
// The port resistance of one sheet-resistance polygon (ohm).
double segment_dc_ohm(const int32_t *xy, int n, uint32_t poly, double dbu_nm,
                      double sheet_res_ohm_sq, double thickness_nm, double height_nm);
// r_ohm per net over a layout's polygons (via resistances, then ports).
void solve_resistance(const gp_res_polygon *polys, uint32_t n_polys, double dbu_nm,
                      std::map<uint32_t, double> &r_of_net);

}  // namespace gp
