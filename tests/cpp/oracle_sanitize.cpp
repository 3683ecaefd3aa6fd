// Runs the oracle (oracle/swe_oracle.cpp, linked in) under AddressSanitizer
// and UndefinedBehaviorSanitizer on a small wet/dry mesh: every scheme, flux,
// wave speed and option for a few CFL steps, plus the integrals.  Exit code
// 0 when the run is clean; the sanitizers abort on the first error.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

extern "C" {
struct orc_topo;
struct Solver;
orc_topo* orc_topology_build(int64_t num_nodes, int64_t T, const int64_t* tri, int64_t nb, const int64_t* bnd);
int64_t orc_topology_num_edges(const orc_topo* h);
void orc_topology_copy(const orc_topo* h, int64_t* ep, int64_t* et, int64_t* tp, int64_t* te, int64_t* tt);
void orc_topology_free(orc_topo* h);
Solver* orc_create(int64_t V, const double* xyz, int64_t E, const int64_t* ep, const int64_t* et, int64_t T,
                   const int64_t* tp, const int64_t* te, const int64_t* tt, int flux, int ws, double cor,
                   double tau, int opts, const int64_t* pe);
void orc_destroy(Solver* s);
void orc_set_state(Solver* s, const double* prim);
void orc_get_state(const Solver* s, double* prim);
void orc_run(Solver* s, int scheme, double dt, double cfl, int64_t nsteps, double* dts);
void orc_integrals(Solver* s, double* out, double* per_cell);
}

int main() {
  const int n = 12;  // n x n squares, two CCW triangles each, on [0, 1]^2
  const int64_t V = (n + 1) * (n + 1), T = 2 * n * n;
  std::vector<double> xyz(3 * V);
  for (int j = 0; j <= n; ++j)
    for (int i = 0; i <= n; ++i) {
      int64_t k = j * (n + 1) + i;
      double x = double(i) / n, y = double(j) / n;
      xyz[3 * k] = x;
      xyz[3 * k + 1] = y;
      xyz[3 * k + 2] = 0.9 * std::exp(-((x - 0.6) * (x - 0.6) + (y - 0.4) * (y - 0.4)) / 0.02);
    }
  std::vector<int64_t> tri, bnd;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      int64_t a = j * (n + 1) + i, b = a + 1, c = a + n + 1, d = c + 1;
      tri.insert(tri.end(), {a, b, d, a, d, c});
    }
  for (int i = 0; i < n; ++i) {
    bnd.insert(bnd.end(), {int64_t(i), int64_t(i + 1)});
    bnd.insert(bnd.end(), {int64_t(n * (n + 1) + i), int64_t(n * (n + 1) + i + 1)});
    bnd.insert(bnd.end(), {int64_t(i * (n + 1)), int64_t((i + 1) * (n + 1))});
    bnd.insert(bnd.end(), {int64_t(i * (n + 1) + n), int64_t((i + 1) * (n + 1) + n)});
  }
  orc_topo* topo = orc_topology_build(V, T, tri.data(), int64_t(bnd.size() / 2), bnd.data());
  const int64_t E = orc_topology_num_edges(topo);
  std::vector<int64_t> ep(2 * E), et(2 * E), tp(3 * T), te(3 * T), tt(3 * T);
  orc_topology_copy(topo, ep.data(), et.data(), tp.data(), te.data(), tt.data());
  // a dam, a dry pocket and the emerging bump: dry, partially and fully wet cells
  std::vector<double> prim(3 * T);
  for (int64_t t = 0; t < T; ++t) {
    double cx = 0., cy = 0., bz = 0.;
    for (int k = 0; k < 3; ++k) {
      cx += xyz[3 * tp[3 * t + k]] / 3.;
      cy += xyz[3 * tp[3 * t + k] + 1] / 3.;
      bz += xyz[3 * tp[3 * t + k] + 2] / 3.;
    }
    double level = (cx < 0.3) ? 0.8 : 0.5;
    double h = (std::fabs(cx - 0.8) < 0.08 && std::fabs(cy - 0.8) < 0.08) ? 0. : std::fmax(0., level - bz);
    prim[3 * t] = bz + h;
    prim[3 * t + 1] = h > 0 ? 0.1 : 0.;
    prim[3 * t + 2] = h > 0 ? -0.05 : 0.;
  }
  int runs = 0;
  const int opts_list[] = {0, 1, 2, 4, 8, 16 | 32};
  for (int flux = 0; flux < 2; ++flux)
    for (int ws = 0; ws < 3; ++ws)
      for (int opts : opts_list)
        for (int scheme = 0; scheme < 3; ++scheme) {
          Solver* s = orc_create(V, xyz.data(), E, ep.data(), et.data(), T, tp.data(), te.data(), tt.data(),
                                 flux, ws, 0.3, 0.5, opts, nullptr);
          if (!s) return 2;
          orc_set_state(s, prim.data());
          double dts[4];
          orc_run(s, scheme, 0.0, 0.15, 4, dts);
          double I[3];
          orc_integrals(s, I, nullptr);
          std::vector<double> out(3 * T);
          orc_get_state(s, out.data());
          for (double v : out)
            if (!std::isfinite(v)) return 3;
          orc_destroy(s);
          ++runs;
        }
  orc_topology_free(topo);
  std::printf("%d clean runs\n", runs);
  return 0;
}
