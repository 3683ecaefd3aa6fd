// A driver written against the drop-in facade (include/swe_fvm.hpp) in the
// style of the reference's own drivers and solvers: it fills the initial
// field through the assigners (v0.prim(i) = ..., cons(i) *= ..., /= ...),
// steps by writing every cell through GetVolField().cons(i) with
// TimeDisc::RHS (the update loops of src/Solvers.cpp:5-59: Euler with
// `+=`, SSP-RK2 with U0 and the stage state), reads the reconstructions
// (MUSCLObject::Reconstruct*), the geometry getters of Domain and the
// per-edge fluxes Fluxes::HLLC<W>(sd, e, from, to, &r), and runs the same
// steps on a second solver through Solvers::*.  Everything is written to
// text files for tests/test_cpp_facade.py, which compares it with the
// oracle.
#include <swe_fvm.hpp>

#include <cstdio>
#include <string>

static void dump_state(const SpaceDisc& sd, const std::string& path) {
  FILE* f = std::fopen(path.c_str(), "w");
  const VolumeField& vol = sd.GetVolField();
  for (Idx i = 0; i < Idx(vol.Size()); ++i) std::fprintf(f, "%.17g %.17g %.17g\n", vol.w(i), vol.u(i), vol.v(i));
  std::fclose(f);
}

// Euler: every cell's conserved state += RHS (stage-start draining steps)
static void euler_by_cells(TimeDisc* td, double h) {
  SpaceDisc* s = td->GetSpaceDisc();
  const Idx n = s->GetDomain().GetTopology().NumTriangles();
  s->ComputeInterfaceValues();
  s->ComputeFluxes();
  for (Idx c = 0; c < n; ++c) s->GetVolField().cons(c) += td->RHS(c, h);
}

// SSP-RK2 in Shu-Osher form, cell by cell
static void ssprk2_by_cells(TimeDisc* td, double h) {
  SpaceDisc* s = td->GetSpaceDisc();
  const Idx n = s->GetDomain().GetTopology().NumTriangles();
  s->ComputeInterfaceValues();
  s->ComputeFluxes();
  const VolumeField start = s->GetVolField();
  for (Idx c = 0; c < n; ++c) s->GetVolField().cons(c) = start.cons(c) + td->RHS(c, h);
  s->ComputeInterfaceValues();
  s->ComputeFluxes();
  const VolumeField& mid = s->GetVolField();
  for (Idx c = 0; c < n; ++c)
    s->GetVolField().cons(c) = 0.5 * start.cons(c) + 0.5 * mid.cons(c) + td->RHS(c, 0.5 * h);
}

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s mesh.msh out_prefix\n", argv[0]);
    return 2;
  }
  const std::string p = argv[2];
  try {
    PointArray geometry;
    Topology m = Topology::read_msh(argv[1], geometry);
    for (Idx i = 0; i < geometry.cols(); ++i) {
      const double x = geometry(0, i) - 4., y = geometry(1, i) - 4.;
      geometry(2, i) = 0.02 * (x * x + 0.5 * y * y) - 0.25;
    }
    Domain b(geometry, m);
    const Idx T = m.NumTriangles();
    {
      FILE* f = std::fopen((p + "z.dat").c_str(), "w");
      for (Idx i = 0; i < geometry.cols(); ++i) std::fprintf(f, "%.17g\n", geometry(2, i));
      std::fclose(f);
    }

    // initial field through the assigners: dry clamp below ground, thin
    // films (desingularised velocities), and the compound writes
    VolumeField v0{b, size_t(T)};
    for (Idx i = 0; i < T; ++i) {
      const Point c = b.T(i);
      const double level = 0.12 - 0.01 * (c[0] - 4.);
      v0.prim(i) = Array<3>{level, 0.05 * (c[1] - 4.), -0.03 * (c[0] - 4.)};
      if (i % 7 == 3) v0.cons(i) *= 0.5;
      if (i % 11 == 5) v0.cons(i) /= 4.;
      if (i % 13 == 1) v0.prim(i) += Array<3>{0.001, 0.01, 0.};
      if (i % 17 == 2) v0.prim(i) -= Array<3>{0.0, 0.0, 0.02};
    }
    {
      FILE* f = std::fopen((p + "init.dat").c_str(), "w");
      for (Idx i = 0; i < T; ++i) std::fprintf(f, "%.17g %.17g %.17g\n", v0.w(i), v0.u(i), v0.v(i));
      std::fclose(f);
    }
    bool refused = false;
    try {
      v0.cons(0) /= 0.;
    } catch (const std::runtime_error&) {
      refused = true;
    }

    SpaceDisc sd{Fluxes::HLLC<Wavespeeds::Einfeldt>, b, v0, 0.3};
    TimeDisc td{&sd};

    // reconstructions of one cell of each class, and geometry getters
    FILE* fr = std::fopen((p + "recon.dat").c_str(), "w");
    Idx found[3] = {-1, -1, -1};
    for (Idx i = 0; i < T; ++i) {
      const int k = sd.IsDryCell(i) ? 0 : (sd.IsPartWetCell(i) ? 1 : 2);
      if (found[k] < 0 && !(k == 2 && m.IsTriangleBoundary(i))) found[k] = i;
    }
    for (Idx i : found) {
      if (i < 0) continue;
      const MUSCLObject::MUSCL r[4] = {sd.ReconstructDryCell(i), sd.ReconstructPartWetCell1(i),
                                       sd.ReconstructFullWetCell(i), sd.ReconstructPartWetCell2(i)};
      for (int k = 0; k < 4; ++k) {
        if (k == 2 && m.IsTriangleBoundary(i)) continue;
        const Array<3> o = r[k].AtOrigin();
        const Matrix32d g = r[k].Gradient(b.T(i));
        const Array<3> at = r[k].AtPoint(b.E(m.TriangEdges(i)[0]));
        std::fprintf(fr, "%lld %d %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", (long long)i, k, o[0], o[1],
                     o[2], g(0, 0), g(0, 1), at[0], at[1], at[2]);
      }
    }
    std::fclose(fr);
    FILE* fg = std::fopen((p + "geom.dat").c_str(), "w");
    for (Idx e = 0; e < m.NumEdges(); e += 97) {
      const Idx t0 = m.EdgeTriangs(e)[0];
      const Vector2d tg = b.Tang(e, t0), nm = b.Norm(e, t0), sl = b.TriangSlope(t0);
      const Point c = b.C(e);
      const Point ct = b.T(t0);
      std::fprintf(fg, "%lld %lld %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", (long long)e,
                   (long long)t0, tg[0], tg[1], nm[0], nm[1], sl[0], sl[1], c[0], c[1], c[2],
                   b.Z(t0, ct[0], ct[1]));
    }
    std::fclose(fg);

    // per-edge fluxes in both orientations, with the min-updated r
    sd.ComputeInterfaceValues();
    sd.ComputeFluxes();
    FILE* ff = std::fopen((p + "flux.dat").c_str(), "w");
    for (Idx e = 0; e < m.NumEdges(); e += 53) {
      const auto et = m.EdgeTriangs(e);
      double r = 1.;
      const Array<3> a = Fluxes::HLLC<Wavespeeds::Einfeldt>(&sd, e, et[0], et[1], &r);
      const double ra = r;
      Array<3> bflux;
      if (et[1] >= 0) bflux = Fluxes::HLL<Wavespeeds::Davis>(&sd, e, et[1], et[0], &r);
      std::fprintf(ff, "%lld %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", (long long)e, a[0], a[1], a[2], ra,
                   bflux[0], bflux[1], bflux[2], r);
    }
    std::fclose(ff);
    const Array<3> elem = ElemFlux(Vector2d{0.6, 0.8}, Array<3>{2., 1., -0.5});
    std::printf("elemflux %.17g %.17g %.17g\n", elem[0], elem[1], elem[2]);

    // steps: cell by cell through the assigners, and through Solvers::*
    const double dt = 2e-3;
    euler_by_cells(&td, dt);
    euler_by_cells(&td, dt);
    ssprk2_by_cells(&td, dt);
    ssprk2_by_cells(&td, dt);
    dump_state(sd, p + "cells.dat");

    SpaceDisc sd2{Fluxes::HLLC<Wavespeeds::Einfeldt>, b, v0, 0.3};
    TimeDisc td2{&sd2};
    Solvers::Euler(&td2, dt);
    Solvers::Euler(&td2, dt);
    Solvers::SSPRK2(&td2, dt);
    Solvers::SSPRK2(&td2, dt);
    dump_state(sd2, p + "solvers.dat");
    std::printf("cells %lld refused_div0 %d\n", (long long)T, refused ? 1 : 0);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
