// Host-only parts of the facade (no device): the assigners' write path
// (include/Assigners.h:4-50), the Domain geometry getters and ElemFlux on a
// two-triangle mesh.  Prints one labelled line per check for
// tests/test_cpp_facade.py.
#include <swe_fvm.hpp>

#include <cstdio>

int main() {
  try {
    PointArray g(4);
    const double xyz[4][3] = {{0., 0., -1.}, {2., 0., -0.5}, {0., 1., -0.75}, {2., 1.5, -0.25}};
    for (Idx i = 0; i < 4; ++i)
      for (int r = 0; r < 3; ++r) g(r, i) = xyz[i][r];
    TriangTagArray tri{{0, 1, 2}, {1, 3, 2}};
    EdgeTagArray bnd{{0, 1}, {1, 3}, {3, 2}, {2, 0}};
    Topology m = Topology::build(4, tri, bnd);
    Domain b(g, m);
    VolumeField v(b, 2);
    v.prim(0) = Array<3>{-0.5, 1., 2.};           // dry: clamped to the centroid bathymetry
    v.prim(1) = Array<3>{b.T(1)[2] + 5e-4, 1., 2.};  // thin film: desingularised velocity
    const Array<3> p1 = v.prim(1);
    std::printf("prim %.17g %.17g %.17g %.17g %.17g %.17g\n", v.w(0), v.u(0), v.v(0), p1[0], p1[1], p1[2]);
    v.cons(1) = Array<3>{0.5, 0.25, -0.125};
    v.cons(1) += Array<3>{0.25, 0.25, 0.};
    v.cons(1) *= 2.;
    v.cons(1) /= 4.;
    v.cons(1) -= Array<3>{0., 0.125, 0.};
    std::printf("cons %.17g %.17g %.17g\n", v.w(1), v.u(1), v.v(1));
    int thrown = 0;
    try { v.cons(1) /= 0.; } catch (const std::runtime_error&) { thrown = 1; }
    std::printf("div0 %d\n", thrown);
    for (Idx e = 0; e < m.NumEdges(); ++e) {
      const Idx t0 = m.EdgeTriangs(e)[0];
      const Vector2d t = b.Tang(e, t0), n = b.Norm(e, t0);
      const Point c = b.C(e);
      std::printf("edge %lld %lld %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", (long long)e, (long long)t0, t[0], t[1],
                  n[0], n[1], c[0], c[1], c[2]);
    }
    for (Idx t = 0; t < 2; ++t) {
      const Vector2d s = b.TriangSlope(t);
      std::printf("slope %lld %.17g %.17g %.17g\n", (long long)t, s[0], s[1], b.Z(t, 1., 0.5));
    }
    std::printf("extent %.17g %.17g %.17g %.17g\n", b.MinX(), b.MaxX(), b.MinY(), b.MaxY());
    const Array<3> f = ElemFlux(Vector2d{0.6, 0.8}, Array<3>{2., 1., -0.5});
    const Array<3> f0 = ElemFlux(Vector2d{0.6, 0.8}, Array<3>{0., 1., -0.5});
    std::printf("elem %.17g %.17g %.17g %.17g\n", f[0], f[1], f[2], f0[0] + f0[1] + f0[2]);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
