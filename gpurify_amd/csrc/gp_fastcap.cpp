// FastCap "quickif" (.qui) geometry reader: the panel input format of
// crates/pex/src/quasistatic/cap/geometry.rs:56-135, feeding gp_cap_solve.
//
// Grammar (per trimmed, non-empty line; the line's FIRST CHARACTER selects
// the kind, as geometry.rs:82-88 does, so "Qx 1 ..." is a Q line):
//   0 / * / %  title or comment                   (geometry.rs:88)
//   Q c x1 y1 z1 ... x4 y4 z4   quadrilateral       (geometry.rs:89-104)
//   T c x1 y1 z1 ... x3 y3 z3   triangle            (geometry.rs:105-119)
//   N c name   display name of conductor number c   (geometry.rs:120-129)
//   C ...      sub-file include: an error           (geometry.rs:130-132)
//   anything else: an error naming the line         (geometry.rs:133)
// Conductor numbers are signed 64-bit integers interned to 0-based indices
// in first-seen order (geometry.rs:62-77); a conductor is named by the last
// N line for its number seen before it was interned, else "cond<c>", and an
// N line after interning renames it.  Tokens after the needed numbers are
// ignored.  Coordinates are multiplied by `scale`.
#include "gp_fastcap.h"

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <string_view>

#include "gp_common.h"
#include "../../include/gpurify.h"

namespace gp {
namespace {

[[noreturn]] void fail_at(size_t lineno, const std::string &msg) {
  throw Error(GP_ERR_INVALID_INPUT, "line " + std::to_string(lineno) + ": " + msg);
}

bool ieq(std::string_view a, const char *b) {
  if (a.size() != std::strlen(b)) return false;
  for (size_t i = 0; i < a.size(); ++i)
    if (std::tolower((unsigned char)a[i]) != b[i]) return false;
  return true;
}

// The decimal forms a Rust f64 parse accepts: [+-] then digits with an
// optional fraction (at least one digit on either side of the point) and an
// optional exponent [eE][+-]digits, or inf / infinity / nan in any case.
// (No hexadecimal forms, unlike strtod.)  The value is then strtod's, which
// like Rust's parser rounds correctly.
bool parse_f64(std::string_view t, double &v) {
  size_t i = 0;
  if (i < t.size() && (t[i] == '+' || t[i] == '-')) ++i;
  const std::string_view body = t.substr(i);
  if (ieq(body, "inf") || ieq(body, "infinity") || ieq(body, "nan")) {
    v = ieq(body, "nan") ? NAN : (t[0] == '-' ? -INFINITY : INFINITY);
    return true;
  }
  size_t d0 = 0, d1 = 0;
  while (i < t.size() && std::isdigit((unsigned char)t[i])) ++i, ++d0;
  if (i < t.size() && t[i] == '.') {
    ++i;
    while (i < t.size() && std::isdigit((unsigned char)t[i])) ++i, ++d1;
  }
  if (d0 + d1 == 0) return false;
  if (i < t.size() && (t[i] == 'e' || t[i] == 'E')) {
    ++i;
    if (i < t.size() && (t[i] == '+' || t[i] == '-')) ++i;
    size_t de = 0;
    while (i < t.size() && std::isdigit((unsigned char)t[i])) ++i, ++de;
    if (de == 0) return false;
  }
  if (i != t.size()) return false;
  // strtod on a NUL-terminated copy (the text need not be terminated);
  // overflow gives +-inf, as Rust's parse does
  char buf[128];
  if (t.size() < sizeof(buf)) {
    std::memcpy(buf, t.data(), t.size());
    buf[t.size()] = 0;
    v = std::strtod(buf, nullptr);
  } else {
    v = std::strtod(std::string(t).c_str(), nullptr);
  }
  return true;
}

// A Rust i64 parse: [+-] then one or more digits, in range.
bool parse_i64(std::string_view t, int64_t &v) {
  size_t i = 0;
  const bool neg = !t.empty() && t[0] == '-';
  if (!t.empty() && (t[0] == '+' || t[0] == '-')) ++i;
  if (i == t.size()) return false;
  uint64_t mag = 0;
  const uint64_t lim = neg ? (uint64_t)INT64_MAX + 1 : (uint64_t)INT64_MAX;
  for (; i < t.size(); ++i) {
    if (!std::isdigit((unsigned char)t[i])) return false;
    const uint64_t d = (uint64_t)(t[i] - '0');
    if (mag > (lim - d) / 10) return false;
    mag = mag * 10 + d;
  }
  v = neg ? (int64_t)(0 - mag) : (int64_t)mag;
  return true;
}

}  // namespace

void parse_qui(const char *text, size_t len, double scale, QuiGeometry &geo) {
  geo = QuiGeometry();
  std::map<int64_t, uint32_t> idx_of;
  std::map<int64_t, std::string> names_by_number;
  auto intern = [&](int64_t num) -> uint32_t {
    auto it = idx_of.find(num);
    if (it != idx_of.end()) return it->second;
    const uint32_t i = (uint32_t)geo.names.size();
    idx_of[num] = i;
    auto nm = names_by_number.find(num);
    geo.names.push_back(nm != names_by_number.end() ? nm->second : "cond" + std::to_string(num));
    return i;
  };
  size_t pos = 0, lineno = 0;
  std::vector<std::string_view> tok;
  while (pos < len) {
    // str::lines: split at '\n', one trailing '\r' removed with it
    size_t end = pos;
    while (end < len && text[end] != '\n') ++end;
    std::string_view raw(text + pos, end - pos);
    pos = end + 1;
    ++lineno;
    if (!raw.empty() && raw.back() == '\r') raw.remove_suffix(1);
    tok.clear();
    for (size_t i = 0; i < raw.size();) {
      while (i < raw.size() && std::isspace((unsigned char)raw[i])) ++i;
      const size_t j = i;
      while (i < raw.size() && !std::isspace((unsigned char)raw[i])) ++i;
      if (i > j) tok.push_back(raw.substr(j, i - j));
    }
    if (tok.empty()) continue;
    const char first = tok[0][0];
    switch (first) {
      case '0':
      case '*':
      case '%':
        continue;
      case 'Q':
      case 'q':
      case 'T':
      case 't': {
        const bool quad = first == 'Q' || first == 'q';
        const int nnum = quad ? 12 : 9;
        int64_t c;
        if (tok.size() < 2 || !parse_i64(tok[1], c))
          fail_at(lineno, std::string(quad ? "Q" : "T") + ": missing conductor number");
        if (tok.size() - 2 < (size_t)nnum)
          fail_at(lineno, "expected " + std::to_string(nnum) + " numbers, got " +
                              std::to_string(tok.size() - 2));
        double f[12];
        for (int k = 0; k < nnum; ++k)
          if (!parse_f64(tok[2 + k], f[k]))
            fail_at(lineno, "bad number '" + std::string(tok[2 + k]) + "'");
        const uint32_t ci = intern(c);
        for (int k = 0; k < 12; ++k)
          geo.verts.push_back(f[k < nnum ? k : k - 3] * scale);  // a triangle repeats v3
        geo.cond.push_back(ci);
        geo.n_vertices.push_back(quad ? 4 : 3);
        break;
      }
      case 'N':
      case 'n': {
        int64_t c;
        if (tok.size() >= 3 && parse_i64(tok[1], c)) {
          names_by_number[c] = std::string(tok[2]);
          auto it = idx_of.find(c);
          if (it != idx_of.end()) geo.names[it->second] = std::string(tok[2]);
        }
        break;
      }
      case 'C':
      case 'c':
        fail_at(lineno, "C (sub-file include) not yet supported");
      default:
        // the trimmed line, as geometry.rs:133 prints it
        {
          size_t a = 0, b = raw.size();
          while (a < b && std::isspace((unsigned char)raw[a])) ++a;
          while (b > a && std::isspace((unsigned char)raw[b - 1])) --b;
          fail_at(lineno, "unrecognized line: " + std::string(raw.substr(a, b - a)));
        }
    }
  }
}

}  // namespace gp
