// Layout -> panel bridge (bridge.rs:307-652).
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "../../include/gpurify.h"

namespace gp {

struct Meshed {
  std::vector<uint32_t> nets;  // ascending net ids; conductor i = nets[i], substrate last
  std::vector<double> verts;   // n * 12, metres
  std::vector<uint32_t> cond;
  std::vector<double> eps;     // surrounding dielectric per panel
  double ground_z = 0.0;       // height of the substrate plate (bridge.rs:353), metres
  size_t n_plate = 0;          // the plate's panels, the last ones
};

// Throws gp::Error with the BridgeError class of the failure.
void mesh_groups(const gp_volume_group *groups, uint32_t n_groups, double xy_scale, int refine,
                 size_t max_panels, Meshed &m);

}  // namespace gp
