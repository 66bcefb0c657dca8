// FastCap quickif (.qui) geometry reader (cap/geometry.rs:56-135).
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace gp {

struct QuiGeometry {
  std::vector<double> verts;          // 12 per panel (a triangle repeats its last vertex)
  std::vector<uint32_t> cond;         // 0-based conductor index, first-seen order
  std::vector<uint8_t> n_vertices;    // 4 (Q) or 3 (T)
  std::vector<std::string> names;     // per conductor index
};

// Throws gp::Error(GP_ERR_INVALID_INPUT, "line <n>: ...") on a malformed line.
void parse_qui(const char *text, size_t len, double scale, QuiGeometry &geo);

}  // namespace gp
