// Host check of the fused kernel's tile tables (swe_tiles.cpp): replays the
// kernel's index arithmetic (swe_fused.hip) for every tile and verifies
// that every index stays in bounds and names the right edge and cells:
//   - every evaluation's (cell, k) is a local cell and its edge's et0 / et1
//     cells are the f / t the kernel derives;
//   - each edge is evaluated once per tile;
//   - every (local cell, k) code points at the slot of its own edge and its
//     neighbour code at the cell across that edge (or the boundary tag);
//   - LDS capacities hold;
//   - every edge has exactly one evaluation whose et0 cell is a tile cell
//     (the flux a storing launch writes), and each evaluation's edge index
//     is its edge.
// Built with g++ into a shared library by tests/test_tiles_host.py.
#include <cstdint>
#include <cstdio>
#include <set>
#include <string>
#include <vector>

#include "../../swe_fvm_amd/csrc/swe_tiles.h"

using namespace swe;

extern "C" int tiles_check(int64_t T, const int* tt, const int* te, const uint32_t* cflag, const int* et0,
                           const int* et1, const int* ep0, const int* ep1, int64_t E, int64_t* stats, char* msg,
                           int msglen) {
  auto bad = [&](const std::string& m) {
    std::snprintf(msg, size_t(msglen), "%s", m.c_str());
    return 1;
  };
  TileTables P;
  if (!build_tiles(T, tt, te, cflag, et0, et1, ep0, ep1, P)) return bad("build_tiles refused the mesh");
  if (P.code.size() != size_t(3 * T)) return bad("code size");
  int64_t evals = 0, ring = 0;
  std::vector<int> owners(size_t(E), 0);  // per edge: tiles whose storing launch writes its flux
  std::vector<int64_t> owned;
  for (int64_t tile = 0; tile < P.ntiles; ++tile) {
    const int64_t c0 = tile * kTileCells;
    const int nc = int(T - c0 < kTileCells ? T - c0 : kTileCells);
    const int rg0 = P.ring_off[size_t(tile)], nr = P.ring_off[size_t(tile) + 1] - rg0;
    const int e0 = P.ev_off[size_t(tile)], ne = P.ev_off[size_t(tile) + 1] - e0;
    if (nr < 0 || ne < 0 || ne > kTileMaxEval || kTileCells + nr > kTileMaxLocal) return bad("capacity");
    if (size_t(rg0 + nr) > P.ring_cell.size() || size_t(3 * (rg0 + nr)) > P.ring_code.size() ||
        size_t(e0 + ne) > P.ev.size() || P.evrec.size() != P.ev.size() || P.evedge.size() != P.ev.size())
      return bad("table range");
    const int nl = kTileCells + nr;
    auto valid_local = [&](int l) { return (l >= 0 && l < nc) || (l >= kTileCells && l < nl); };
    auto global = [&](int l) -> int64_t {
      return l < kTileCells ? c0 + l : P.ring_cell[size_t(rg0 + l - kTileCells)];
    };
    auto code_of = [&](int l, int64_t g, int k) -> uint32_t {
      return l < kTileCells ? P.code[size_t(k * T + g)] : P.ring_code[size_t(3 * (rg0 + l - kTileCells) + k)];
    };
    std::vector<int64_t> slot_edge(size_t(ne), -1);
    std::set<int64_t> seen;
    for (int j = 0; j < ne; ++j) {
      const uint32_t ev = P.ev[size_t(e0 + j)];
      const int l = int(ev >> 2), k = int(ev & 3u);
      if (k > 2 || !valid_local(l)) return bad("evaluation names no local cell");
      const int64_t x = global(l);
      if (x < 0 || x >= T) return bad("evaluation cell out of range");
      const uint32_t nb = code_of(l, x, k) & 0xffffu;
      int64_t y, t;
      if (nb == kNbWall) { y = -1; t = -1; }
      else if (nb == kNbFree) { y = -1; t = -2; }
      else {
        if (nb != kNbOut && !valid_local(int(nb))) return bad("neighbour code out of range");
        y = nb == kNbOut ? tt[size_t(k * T + x)] : global(int(nb));
        t = y;
      }
      if (nb == kNbOut && l < kTileCells) return bad("tile cell with an outside neighbour");
      const bool own0 = (cflag[x] >> k) & 1u;
      const int64_t f = own0 ? x : y;
      if (!own0) t = x;
      const int64_t e = te[size_t(k * T + x)];
      if (e < 0 || e >= E) return bad("edge out of range");
      if (f < 0 || f >= T || f != et0[e]) return bad("f is not the edge's et0");
      if (t != et1[e]) return bad("t is not the edge's et1");
      if (!seen.insert(e).second) return bad("edge evaluated twice in a tile");
      const EdgeRec& rec = P.evrec[size_t(e0 + j)];
      if (rec.et0 != f || rec.et1 != et1[e] || rec.ep0 < 0 || rec.ep1 < 0) return bad("edge record");
      if (P.evedge[size_t(e0 + j)] != e) return bad("evaluation's edge index");
      // the evaluation a storing launch writes: its et0 cell is a tile cell
      if (f >= c0 && f < c0 + nc) owned.push_back(e);
      slot_edge[size_t(j)] = e;
    }
    for (int l = 0; l < nl; ++l) {
      if (!valid_local(l)) continue;
      const int64_t x = global(l);
      for (int k = 0; k < 3; ++k) {
        const uint32_t c = code_of(l, x, k);
        const int s = int(c >> 16);
        const uint32_t nb = c & 0xffffu;
        if (s < 0 || s >= ne) return bad("slot out of range");
        if (slot_edge[size_t(s)] != te[size_t(k * T + x)]) return bad("slot of another edge");
        const int n = tt[size_t(k * T + x)];
        if (nb == kNbWall || nb == kNbFree) {
          if (n >= 0) return bad("boundary code on an interior edge");
          if ((nb == kNbWall) != (et1[te[size_t(k * T + x)]] == -1)) return bad("boundary tag");
        } else if (nb == kNbOut) {
          if (l < kTileCells || n < 0) return bad("outside code");
        } else if (global(int(nb)) != n) {
          return bad("neighbour code names another cell");
        }
      }
    }
    evals += ne;
    ring += nr;
    for (int64_t e : owned) owners[size_t(e)]++;
    owned.clear();
  }
  for (int64_t e = 0; e < E; ++e)
    if (owners[size_t(e)] != 1) return bad("an edge's flux is written by " + std::to_string(owners[size_t(e)]) + " tiles");
  stats[0] = P.ntiles;
  stats[1] = evals;
  stats[2] = ring;
  stats[3] = P.max_eval;
  stats[4] = P.max_ring;
  return 0;
}
