// Tiles of the fused flux / draining / update kernel: host-built static
// tables (swe_tiles.cpp) and their device view (swe_fused.hip).
#pragma once

#include <stdint.h>

#include <vector>

namespace swe {

#ifndef SWE_TILE_CELLS
#define SWE_TILE_CELLS 256
#endif
constexpr int kTileCells = SWE_TILE_CELLS;      // cells per tile (a build-time knob: 256 or 512)
constexpr int kTileMaxLocal = 2 * kTileCells;  // tile cells + ring
constexpr int kTileMaxEval = 3 * kTileCells;   // edge evaluations per tile (LDS flux slots)
static_assert(kTileCells == 256 || kTileCells == 512, "tile size");
constexpr uint32_t kNbWall = 0xffffu, kNbFree = 0xfffeu, kNbOut = 0xfffdu;

// one evaluated edge: its cells and nodes (one 16-byte load)
struct alignas(16) EdgeRec {
  int et0, et1, ep0, ep1;
};

struct TileTables {
  int64_t ntiles = 0;
  std::vector<uint32_t> code;       // [3][T] tile cells: slot << 16 | neighbour code
  std::vector<int> ring_off;        // [ntiles + 1]
  std::vector<int> ring_cell;       // ring cells (global ids), tile by tile
  std::vector<uint32_t> ring_code;  // [ring entry][3] codes of the ring cells
  std::vector<int> ev_off;          // [ntiles + 1]
  std::vector<uint16_t> ev;         // evaluations: local cell << 2 | k
  std::vector<EdgeRec> evrec;       // per evaluation: the edge's et0, et1, ep0, ep1
  std::vector<int> evedge;          // per evaluation: the edge
  int max_eval = 0, max_ring = 0;
};

// false when a tile exceeds the kernel's LDS capacity (a mesh far from a
// cell-local order) or the tables are inconsistent: the fused kernel is
// then not used
bool build_tiles(int64_t T, const int* tt, const int* te, const uint32_t* cflag, const int* et0, const int* et1,
                 const int* ep0, const int* ep1, TileTables& out);

struct DevTiles {
  const uint32_t* code;
  const int* ring_off;
  const int* ring_cell;
  const uint32_t* ring_code;
  const int* ev_off;
  const EdgeRec* evrec;  // et0, et1, ep0, ep1 of each evaluated edge
  const int* evedge;     // the edge of each evaluation (for the fluxes a storing launch writes)
  int64_t ntiles;
};

}  // namespace swe
