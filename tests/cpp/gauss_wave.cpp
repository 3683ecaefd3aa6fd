// testGaussWave (reference examples/Main.cpp:172-195) written against the
// drop-in facade: bowl mesh, flat bottom, w = 1 + exp(-5 |T - (4, 4)|^2),
// SpaceDisc{Fluxes::HLL<Wavespeeds::Einfeldt>, ...}, one Euler step.
// Prints w hu hv per cell before and after the step with 17 digits.
#include <swe_fvm.hpp>

#include <cmath>
#include <cstdio>

static void dumpFields(const SpaceDisc& sd, FILE* out) {
  const auto& vol = sd.GetVolField();
  const auto& m = sd.GetDomain().GetTopology();
  for (Idx i = 0; i < m.NumTriangles(); i++)
    std::fprintf(out, "%.17g\t%.17g\t%.17g\n", vol.w(i), vol.hu(i), vol.hv(i));
}

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s bowl.msh out_prefix\n", argv[0]);
    return 2;
  }
  try {
    PointArray geometry;
    Topology m = Topology::read_msh(argv[1], geometry);
    for (Idx i = 0; i < geometry.cols(); i++) geometry(2, i) = 0.;
    Domain b(geometry, m);

    VolumeField v0{b, size_t(m.NumTriangles())};
    for (Idx i = 0; i < m.NumTriangles(); ++i) {
      Point t = b.T(i);
      double r = (t[0] - 4.) * (t[0] - 4.) + (t[1] - 4.) * (t[1] - 4.);
      v0.set_prim(i, Array<3>{1. + std::exp(-5. * r), 0., 0.});
    }

    SpaceDisc sd{Fluxes::HLL<Wavespeeds::Einfeldt>, b, v0};
    TimeDisc td{&sd};

    double dt = 0.001;
    std::string p = argv[2];
    FILE* f0 = std::fopen((p + "out0.dat").c_str(), "w");
    dumpFields(sd, f0);
    std::fclose(f0);
    Solvers::Euler(&td, dt);
    FILE* f1 = std::fopen((p + "out1.dat").c_str(), "w");
    dumpFields(sd, f1);
    std::fclose(f1);
    std::printf("cells %lld edges %lld r %.17g cfl_dt %.17g\n", (long long)m.NumTriangles(),
                (long long)m.NumEdges(), sd.GetMinLenToWavespeed(), td.CFLdt());
  } catch (const DomainError& e) {
    std::fprintf(stderr, "Domain error: %s\n", e.what());
    return 1;
  } catch (const SolverError& e) {
    std::fprintf(stderr, "Solver error: %s\n", e.what());
    return 1;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
