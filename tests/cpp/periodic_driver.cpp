// A channel periodic in x written against the drop-in facade:
// StructTriangMesh{12, 8, 0.125} with a bathymetry that is periodic in x,
// the left and right wall edges paired by translation (Topology::SetPeriodic,
// matched by their midpoints' y, which both sides compute alike), a surface
// hump near the seam moving across it, Domain{mesh}, Coriolis 0.1,
// HLLC<Einfeldt>, ten Solvers::SSPRK2 steps of 2e-3.  Writes the node
// heights (z.dat), the pairs (pairs.dat) and w u v per cell before (out0.dat)
// and after (out1.dat) the steps with 17 digits.
#include <swe_fvm.hpp>

#include <cmath>
#include <cstdio>
#include <vector>

static void dump(const SpaceDisc& sd, const std::string& path) {
  FILE* f = std::fopen(path.c_str(), "w");
  const auto& vol = sd.GetVolField();
  for (Idx i = 0; i < sd.GetDomain().GetTopology().NumTriangles(); i++)
    std::fprintf(f, "%.17g %.17g %.17g\n", vol.w(i), vol.u(i), vol.v(i));
  std::fclose(f);
}

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s out_prefix\n", argv[0]);
    return 2;
  }
  const std::string p = argv[1];
  try {
    const size_t ni = 12, nj = 8;
    const double h = 0.125, X = double(ni) * h;
    StructTriangMesh mesh{ni, nj, h};
    PointArray& g = mesh.Geometry();
    const double pi = std::acos(-1.);
    for (Idx n = 0; n < g.cols(); ++n) g(2, n) = 0.05 * std::cos(2. * pi * g(0, n) / X) * std::sin(pi * g(1, n));
    Topology& m = mesh.GetTopology();
    std::vector<Idx> left, right;
    for (Idx e = 0; e < m.NumEdges(); ++e) {
      if (!m.IsEdgeBoundary(e)) continue;
      const auto ep = m.EdgePoints(e);
      if (g(0, ep[0]) == 0. && g(0, ep[1]) == 0.) left.push_back(e);
      if (g(0, ep[0]) == X && g(0, ep[1]) == X) right.push_back(e);
    }
    if (left.size() != nj || right.size() != nj) throw DomainError("side edges not found");
    const Idx npairs = Idx(nj);
    EdgeTagArray pairs(npairs);
    Idx k = 0;
    for (Idx a : left) {
      const auto pa = m.EdgePoints(a);
      const double ya = 0.5 * (g(1, pa[0]) + g(1, pa[1]));
      for (Idx b : right) {
        const auto pb = m.EdgePoints(b);
        if (0.5 * (g(1, pb[0]) + g(1, pb[1])) == ya) {
          pairs(k, 0) = a;
          pairs(k, 1) = b;
          ++k;
        }
      }
    }
    if (k != Idx(nj)) throw DomainError("unmatched side edges");
    m.SetPeriodic(pairs);
    const std::vector<Idx> partner = m.PeriodicPartners();
    for (Idx q = 0; q < k; ++q)
      if (partner[size_t(pairs(q, 0))] != pairs(q, 1) || m.EdgeTriangs(pairs(q, 0))[1] != Idx(Boundaries::PERIODIC))
        throw DomainError("pairing not recorded");

    Domain b{mesh};
    VolumeField v0{b, size_t(m.NumTriangles())};
    for (Idx i = 0; i < m.NumTriangles(); ++i) {
      const Point t = b.T(i);
      const double r2 = (t[0] - 0.1) * (t[0] - 0.1) + (t[1] - 0.5) * (t[1] - 0.5);
      v0.set_prim(i, Array<3>{1. + 0.1 * std::exp(-r2 / 0.02), 0.2, 0.});
    }
    SpaceDisc sd{Fluxes::HLLC<Wavespeeds::Einfeldt>, b, v0, 0.1};
    TimeDisc td{&sd};
    dump(sd, p + "out0.dat");
    for (int s = 0; s < 10; ++s) Solvers::SSPRK2(&td, 2e-3);
    dump(sd, p + "out1.dat");
    FILE* fz = std::fopen((p + "z.dat").c_str(), "w");
    for (Idx n = 0; n < g.cols(); ++n) std::fprintf(fz, "%.17g\n", g(2, n));
    std::fclose(fz);
    FILE* fp = std::fopen((p + "pairs.dat").c_str(), "w");
    for (Idx q = 0; q < k; ++q) std::fprintf(fp, "%lld %lld\n", (long long)pairs(q, 0), (long long)pairs(q, 1));
    std::fclose(fp);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
