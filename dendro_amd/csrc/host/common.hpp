// Shared host-side definitions of libdendrohip: error reporting and Morton arithmetic.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstdarg>
#include <string>
#include <stdexcept>
#include <vector>

#include "dendro_hip.h"

namespace dgh {

// ---------------------------------------------------------------- errors
// Library code throws dgh::Error; every extern "C" entry point catches it, stores the
// message for dgh_last_error and returns the code.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

[[noreturn]] void fail(int code, const char* fmt, ...);
void set_last_error(const std::string& msg);

template <class F>
int guarded(F&& f) {
  try {
    f();
    return DGH_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return DGH_ERR_NOMEM;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return DGH_ERR_ARG;
  }
}

// ---------------------------------------------------------------- Morton order
// An octant is (m, lev): m is the Morton index of its anchor at the finest depth of the
// tree (bits interleaved z,y,x from most to least significant, so the child number of a
// level-l octant in a depth-D tree is (m >> 3(D-l)) & 7 = 4z+2y+x), lev its level.
// Sorting by (m, lev) is the reference's Morton order with ancestors before descendants
// (include/oct/TreeNode.txx:160-180).  21 bits per axis.
inline uint64_t spread3(uint32_t v) {
  uint64_t x = v & 0x1fffff;
  x = (x | (x << 32)) & 0x1f00000000ffffULL;
  x = (x | (x << 16)) & 0x1f0000ff0000ffULL;
  x = (x | (x << 8)) & 0x100f00f00f00f00fULL;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ULL;
  x = (x | (x << 2)) & 0x1249249249249249ULL;
  return x;
}
inline uint32_t compact3(uint64_t x) {
  x &= 0x1249249249249249ULL;
  x = (x | (x >> 2)) & 0x10c30c30c30c30c3ULL;
  x = (x | (x >> 4)) & 0x100f00f00f00f00fULL;
  x = (x | (x >> 8)) & 0x1f0000ff0000ffULL;
  x = (x | (x >> 16)) & 0x1f00000000ffffULL;
  x = (x | (x >> 32)) & 0x1fffffULL;
  return static_cast<uint32_t>(x);
}
inline uint64_t morton(uint32_t x, uint32_t y, uint32_t z) {
  return spread3(x) | (spread3(y) << 1) | (spread3(z) << 2);
}
inline void demorton(uint64_t m, uint32_t& x, uint32_t& y, uint32_t& z) {
  x = compact3(m);
  y = compact3(m >> 1);
  z = compact3(m >> 2);
}

struct Oct {
  uint64_t m;       // Morton index of the anchor
  uint8_t lev;      // level
  uint8_t flags;    // OCT_NODE | OCT_PSEUDO
  bool operator<(const Oct& o) const { return m < o.m || (m == o.m && lev < o.lev); }
  bool operator==(const Oct& o) const { return m == o.m && lev == o.lev; }
};
constexpr uint8_t OCT_NODE = 128;    // anchor is a regular (non-hanging) node
constexpr uint8_t OCT_PSEUDO = 64;   // positive-boundary pseudo-octant

// Morton span (number of finest cells) of a level-l octant in a depth-D tree.
inline uint64_t span(uint32_t D, uint32_t l) { return uint64_t(1) << (3 * (D - l)); }
inline uint32_t side(uint32_t D, uint32_t l) { return uint32_t(1) << (D - l); }
inline uint32_t child_number(const Oct& o, uint32_t D) {
  return static_cast<uint32_t>((o.m >> (3 * (D - o.lev))) & 7);
}

// ---------------------------------------------------------------- element-type table
// Hanging masks of the 18 element types of each child number, in eType order.  This table
// fixes the eType index of the stencil tables, so it must equal the reference's
// (include/oda/cnumEtypes.h:15-229, Matlab/cNumBasedHangingMasks.m).
extern const uint8_t kEtypeMask[8][18];
int etype_of(int child, int mask);   // -1 if invalid

}  // namespace dgh
