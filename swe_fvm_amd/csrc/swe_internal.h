// Host-side internal types shared by the mesh and solver translation units.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

namespace swe {

// Owning copy of the reference's Topology tables (include/TriangMesh.h:55-64),
// int64 row-major: ep, et are E x 2; tp, te, tt are T x 3.
struct Topo {
  int64_t V = 0, E = 0, T = 0;
  std::vector<int64_t> ep, et, tp, te, tt;
  // partner edge of every PERIODIC (-3) boundary edge, -1 elsewhere; empty
  // when the mesh has no periodic edges
  std::vector<int64_t> pe;
};

int fail(int code, const std::string& msg);
int build_topology(int64_t V, int64_t T, const int64_t* tri, int64_t nb, const int64_t* bnd,
                   Topo& m);

}  // namespace swe

struct swe_topology {
  swe::Topo m;
};
