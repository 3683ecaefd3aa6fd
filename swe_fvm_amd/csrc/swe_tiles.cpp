// Static tables of the fused flux / draining / update kernel (swe_fused.hip).
//
// The cells are cut into tiles of kTileCells consecutive (Morton-ordered)
// cells.  A tile's kernel needs the fluxes of every edge of its own cells
// (the update) and of every edge of their outside neighbours, the tile's
// ring (their draining time steps, which the update of a cell next to them
// reads).  Its local cells are the tile's own cells (local ids 0 ..
// kTileCells - 1) followed by the ring (kTileCells + j, in order of first
// appearance).  Every edge of a local cell is evaluated once per tile, by
// its et0 cell when that cell is local, else by its local cell; the
// evaluations are numbered in order (local cell, then k) and their number
// is the LDS slot of that edge's flux.  Per (local cell, k) the tables hold
//   code = slot << 16 | nb
// with nb the local id of the neighbour across edge k, or kNbWall /
// kNbFree (boundary edge of that tag) or kNbOut (a ring cell's neighbour
// outside the tile: its global id is tt's).  Each evaluation also records
// its edge's index: a storing launch writes an edge's flux from the one
// evaluation whose et0 cell is a tile cell (the edge's et0 cell evaluates
// it in that cell's own tile, and no other tile holds that cell).
#include <algorithm>
#include <cstdint>
#include <vector>

#include "swe_tiles.h"

namespace swe {

bool build_tiles(int64_t T, const int* tt, const int* te, const uint32_t* cflag, const int* et0, const int* et1,
                 const int* ep0, const int* ep1, TileTables& out) {
  const int64_t nt = (T + kTileCells - 1) / kTileCells;
  out = TileTables{};
  out.ntiles = nt;
  out.code.assign(size_t(3 * T), 0u);
  out.ring_off.assign(size_t(nt) + 1, 0);
  out.ev_off.assign(size_t(nt) + 1, 0);
  std::vector<int> loc(size_t(T), -1);  // local id of a cell in the current tile
  std::vector<int> ring;
  std::vector<uint32_t> codes;          // [local][k] of the current tile
  std::vector<int> slot;                // [local][k]: evaluation slot
  for (int64_t t = 0; t < nt; ++t) {
    const int64_t c0 = t * kTileCells, c1 = std::min<int64_t>(T, c0 + kTileCells);
    const int nc = int(c1 - c0);
    ring.clear();
    for (int64_t i = c0; i < c1; ++i) loc[size_t(i)] = int(i - c0);
    for (int64_t i = c0; i < c1; ++i)
      for (int k = 0; k < 3; ++k) {
        const int y = tt[size_t(k * T + i)];
        if (y >= 0 && loc[size_t(y)] < 0) {
          loc[size_t(y)] = kTileCells + int(ring.size());
          ring.push_back(y);
        }
      }
    const int nl = kTileCells + int(ring.size());
    auto global = [&](int l) -> int64_t { return l < kTileCells ? c0 + l : ring[size_t(l - kTileCells)]; };
    auto local_ok = [&](int l) { return l < nc || l >= kTileCells; };
    // neighbour codes and evaluators
    codes.assign(size_t(3 * nl), 0u);
    slot.assign(size_t(3 * nl), -1);
    int nev = 0;
    for (int l = 0; l < nl; ++l) {
      if (!local_ok(l)) continue;
      const int64_t x = global(l);
      for (int k = 0; k < 3; ++k) {
        const int y = tt[size_t(k * T + x)];
        uint32_t nb;
        bool evaluator;
        if (y < 0) {
          nb = et1[te[size_t(k * T + x)]] == -1 ? kNbWall : kNbFree;
          evaluator = true;
        } else if (loc[size_t(y)] >= 0) {
          nb = uint32_t(loc[size_t(y)]);
          evaluator = (cflag[x] >> k) & 1u;  // the et0 side of a local pair
        } else {
          nb = kNbOut;  // ring cell's outside neighbour
          evaluator = true;
        }
        codes[size_t(3 * l + k)] = nb;
        if (evaluator) {
          slot[size_t(3 * l + k)] = nev++;
          out.ev.push_back(uint16_t((l << 2) | k));
          const int e = te[size_t(k * T + x)];
          out.evrec.push_back(EdgeRec{et0[e], et1[e], ep0[e], ep1[e]});
          out.evedge.push_back(e);
        }
      }
    }
    // the other side of every local pair reads its twin's slot
    for (int l = 0; l < nl; ++l) {
      if (!local_ok(l)) continue;
      const int64_t x = global(l);
      for (int k = 0; k < 3; ++k) {
        if (slot[size_t(3 * l + k)] >= 0) continue;
        const int ly = int(codes[size_t(3 * l + k)]);
        const int64_t y = global(ly);
        const int e = te[size_t(k * T + x)];
        for (int q = 0; q < 3; ++q)
          if (te[size_t(q * T + y)] == e) slot[size_t(3 * l + k)] = slot[size_t(3 * ly + q)];
        if (slot[size_t(3 * l + k)] < 0) return false;  // inconsistent tables
      }
    }
    if (nev > kTileMaxEval || nl > kTileMaxLocal) return false;
    out.max_eval = std::max(out.max_eval, nev);
    out.max_ring = std::max(out.max_ring, int(ring.size()));
    for (int l = 0; l < nl; ++l) {
      if (!local_ok(l)) continue;
      for (int k = 0; k < 3; ++k) {
        const uint32_t c = (uint32_t(slot[size_t(3 * l + k)]) << 16) | codes[size_t(3 * l + k)];
        if (l < kTileCells) out.code[size_t(k * T + c0 + l)] = c;
        else out.ring_code.push_back(c);  // [ring entry][k], 3 per entry
      }
    }
    for (int y : ring) out.ring_cell.push_back(y);
    out.ring_off[size_t(t) + 1] = int(out.ring_cell.size());
    out.ev_off[size_t(t) + 1] = int(out.ev.size());
    for (int64_t i = c0; i < c1; ++i) loc[size_t(i)] = -1;
    for (int y : ring) loc[size_t(y)] = -1;
  }
  return true;
}

}  // namespace swe
