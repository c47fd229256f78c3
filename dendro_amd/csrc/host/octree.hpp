// Host-side octree and DA construction of libdendrohip.
#pragma once

#include <vector>

#include "common.hpp"

namespace dgh {

// ---- octrees (octree.cpp); octants at the tree's max depth D
std::vector<Oct> octree_from_points(const double* xyz, uint64_t n, uint32_t D, uint32_t max_pts,
                                    bool inc_corner);
std::vector<Oct> balance(const std::vector<Oct>& in, uint32_t D, bool inc_corner);
std::vector<Oct> coarsen(const std::vector<Oct>& in, uint32_t D);
bool is_complete_linear(const std::vector<Oct>& v, uint32_t D);

// ---- stencils (stencils.cpp)
void generate_stencils(double* lap2, double* mass2, double* rmat2, double* rmat1);
void generate_transfer_weights(double* W /* [144][27][8] */);
void generate_shfn(double* S /* [144][8][27] */);
void generate_graddiv(double* GD /* [144][24][24] */);

// ---- one DA level (da.cpp)
// Octants of the DA live at depth Dda = D + 1 (the octree is embedded one level deeper so
// that positive-boundary pseudo-octants fit, src/oct/OctFunctions.C:636-700); the anchors
// keep their coordinates, levels grow by one.
struct DALevel {
  uint32_t D = 0;                    // octree max depth; DA max depth is D + 1
  std::vector<Oct> octs;             // elements [0, n_elem) then pseudo-octants, DA levels
  uint32_t n_elem = 0, n_node = 0;
  std::vector<uint32_t> node_of;     // node index of each octant, or kNoNode
  std::vector<uint32_t> nlist;       // [n_elem][8]
  std::vector<uint8_t> child, hnmask, lev;
  uint32_t n_hanging = 0;
  uint32_t max_depth() const { return D + 1; }
  // index of the octant whose region holds the finest cell (x,y,z), or -1
  int64_t find_cell(uint32_t x, uint32_t y, uint32_t z, bool elements_only) const;
};
constexpr uint32_t kNoNode = 0xffffffffu;

void build_da(const std::vector<Oct>& tree, uint32_t D, DALevel& da);

// Fine node -> (coarse element, position code) map between two DA levels of one hierarchy.
void build_transfer(const DALevel& coarse, const DALevel& fine, std::vector<uint32_t>& celem,
                    std::vector<uint8_t>& pos);

// ---- owner-computes tiling of one level (tiles.cpp)
struct TileRec {
  uint32_t elem_off;   // first record of the tile
  uint32_t n_elem;     // records (own + halo elements)
  uint32_t node_off;   // first entry of the tile's node table
  uint32_t n_touch;    // nodes touched by the records (owned first)
  uint32_t n_own;      // nodes owned (stored) by the tile
  uint32_t pad[3];
};
struct TilePlan {
  uint32_t tile_elems = 0;          // element cap per tile
  uint32_t budget = 0;              // cap on n_own + n_touch per tile (0: none)
  uint32_t own_cap = 0;             // with budget: n_own <= own_cap, n_touch <= budget - own_cap
  std::vector<TileRec> tiles;
  std::vector<uint16_t> lid;        // [R][8] tile-local node ids, vertex order k^child
  std::vector<uint16_t> meta;       // [R] child-0 stencil type | level << 5
  std::vector<uint32_t> rec_elem;   // [R] element of each record
  std::vector<uint32_t> node_tab;   // global node ids, per tile: owned then others
  uint32_t max_touch = 0, max_own = 0, max_elem = 0;
  uint32_t max_slots = 0;           // max over tiles of n_own + n_touch
};
// Tiles of at most T consecutive elements; with budget > 0 a tile also ends before its
// owned plus touched node count would exceed the budget (tiles then vary in length), and with
// own_cap > 0 before it owns more than own_cap nodes or touches more than budget - own_cap
// (the owned-node sums then sit in a fixed-size LDS region ahead of the staged x).
void build_tiles(uint32_t E, uint32_t N, const uint32_t* nl, const uint8_t* child, const uint8_t* hn,
                 const uint8_t* lev, uint32_t T, TilePlan& plan, uint32_t budget = 0, uint32_t own_cap = 0);
// Tile boundaries alone: bounds[t] = first element of tile t, bounds.back() = E.
void tile_bounds(uint32_t E, uint32_t N, const uint32_t* nl, uint32_t T, uint32_t budget,
                 std::vector<uint32_t>& bounds, uint32_t own_cap = 0);

// Matching of coarse elements to fine elements (same element or 8 children), in order:
// first fine element of each coarse element and whether it is the same element.
void match_levels(const DALevel& coarse, const DALevel& fine, std::vector<uint32_t>& first_fine,
                  std::vector<uint8_t>& same);

}  // namespace dgh
