// ORACLE — test infrastructure only.
//
// A scalar C++17 CPU restatement of SWE_FVM's explicit finite-volume step
// (no Eigen).  It is what the HIP product path's results are compared
// against; it is imported only by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py.  The product library never links it.
//
// Every function cites the reference file:line it restates.  Where the
// reference HEAD is broken, the documented semantic decision is applied:
//   D1  Domain::E(e) = midpoint of the two edge nodes (x, y, z)
//       (reference stub: src/Bathymetry.cpp:29-31).
//   D2  ReconstructFullWetCell takes the plane gradient of w through the
//       neighbour points (reference passes bathymetry: src/MUSCLObject.cpp:45-64).
//       opts bit 0 restores the HEAD-literal gradient of z.
//   D3  hp_m = wp_m - z_m   (src/MUSCLObject.cpp:67-68).
//   D4  ReconstructPartWetCell2 writes the z row of its plane points
//       (src/MUSCLObject.cpp:143,158-159,180); opts bit 1 keeps the HEAD
//       transposed writes (they are lost).
//   D5  draining time step reads the stage-start state (Jacobi);
//       opts bit 2 restores the literal in-place read (src/TimeDisc.cpp:47,64).
//   D6  pass 2 of ComputeInterfaceValues reads the pass-1 node maximum;
//       opts bit 3 restores the literal order-dependent read.
//   D8  fabs, never the integer abs (src/PointOperations.cpp:13,32).
//   D9  Einfeldt Roe velocity (cl*ul + cl*ur)/(cl+cr), bug-compatible
//       (src/Fluxes.cpp:22).
//   D10 HLLC CFL denominator max(tol, max(al, ar)) (include/Fluxes.h:89).
//   D11 SOLID_WALL (-1) as the reference, FREE_FLOW (-2) as a transmissive
//       edge (declared, unimplemented in the reference); CUSTOM is
//       rejected.
//   D18 PERIODIC (-3, declared but never paired in the reference): edges
//       paired by a partner table pe; the pair's flux is evaluated once,
//       from its lower edge e (normal outward of et0(e), left state the
//       slot of et0(e) at E(e), right state the slot of et0(pe[e]) at
//       E(pe[e]), edge bathymetry of e, CFL lengths of both cells over
//       L(e)), and the partner edge carries its negation; across the pair
//       the draining time step of the neighbour is the partner cell's.
//   D12 CFL mode: dt = cfl * r with r from the stage-1 flux pass, r <= 1.
//
// Floating point: built with -ffp-contract=off; every expression is written
// in the evaluation order stated next to it.  Where the reference's order
// is fixed by Eigen internals (the centroid sum, the 2x2 partial-pivot LU)
// the order chosen here is stated and is parity-unpinned at the ulp level.
//
// cbrt and log2 are replaced by deterministic restatements built from IEEE
// basic operations (Newton iteration for the cube root; exact exponent
// extraction for (int)log2), so that any implementation following the same
// steps reproduces the same bits.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace orc {

using Idx = int64_t;
constexpr double tol = 1e-13;                         // include/Includes.h:30
constexpr double kInf = std::numeric_limits<double>::infinity();

enum Opt : int {
  OPT_HEAD_GRADIENT = 1,      // D2 literal
  OPT_HEAD_PARTWET2 = 2,      // D4 literal
  OPT_SEQ_DRAINING = 4,       // D5 literal
  OPT_SEQ_NODEMAX = 8,        // D6 literal
  OPT_FRICTION = 16,          // linear friction -tau (hu, hv) (tau is unused in the reference)
  OPT_DRAIN_LENGTH = 32,      // D16 corrected: outflow F0 * L in the draining dt
  OPT_EXACT_BALANCE = 64,     // D19: groupings that keep a lake at rest to the bit
};
// Boundaries (include/Includes.h:21): SOLID_WALL = -1 (src/SpaceDisc.cpp:66-69)
// and FREE_FLOW = -2, declared but unimplemented in the reference: here the
// outside state equals the inside edge state (a transmissive edge), so the
// fluxer sees R = L and the CFL length uses the inside cell only.
constexpr Idx SOLID_WALL = -1, FREE_FLOW = -2, PERIODIC = -3;

enum Flux : int { FLUX_HLL = 0, FLUX_HLLC = 1 };
enum Wave : int { WS_RUSANOV = 0, WS_DAVIS = 1, WS_EINFELDT = 2 };
enum Scheme : int { SCHEME_EULER = 0, SCHEME_SSPRK2 = 1, SCHEME_SSPRK3 = 2 };

// std::min / std::max semantics: min(a,b) = (b < a) ? b : a.
static inline double dmin(double a, double b) { return (b < a) ? b : a; }
static inline double dmax(double a, double b) { return (a < b) ? b : a; }

// IsWet: include/Bathymetry.h:5-8
static inline bool IsWet(double h) { return h > 1e-12; }

// Deterministic cube root (replaces libm cbrt at src/MUSCLObject.cpp:103).
// x = m * 2^(3q), m in [0.5, 4); linear seed on [0.5, 4); five Newton steps
// y <- y - (y^3 - m) / (3 y^2); result y * 2^q.
static double cbrt_det(double x) {
  if (x == 0.0 || !std::isfinite(x)) return x;
  if (x < 0.0) return -cbrt_det(-x);
  int e;
  double m = std::frexp(x, &e);  // x = m 2^e, m in [0.5, 1)
  int r = e % 3;
  if (r < 0) r += 3;
  int q = (e - r) / 3;
  m = std::ldexp(m, r);          // m in [0.5, 4)
  double y = 0.7937005259840998 + (m - 0.5) * 0.22679491924311227;
  for (int it = 0; it < 5; ++it) {
    double y2 = y * y;
    y = y - (y2 * y - m) / (3.0 * y2);
  }
  return std::ldexp(y, q);
}

// (int)log2(d) for d > 0, from the exponent bits (src/PointOperations.cpp:33).
static int ilog2_trunc(double d) {
  int e;
  double m = std::frexp(d, &e);  // d = m 2^e, m in [0.5, 1): floor(log2 d) = e - 1
  int fl = e - 1;
  if (d >= 1.0) return fl;
  return (m == 0.5) ? fl : fl + 1;  // truncation toward zero for d < 1
}

// CubicPoly: include/CubicPolyMath.h:6-19 ; x^3 + b x^2 + c x + d
struct CubicPoly {
  double b, c, d;
  double operator()(double x) const { return x * x * x + b * x * x + c * x + d; }
};

// Bisection: src/PointOperations.cpp:26-40
static double Bisection(const CubicPoly& f, double xmin, double xmax, int accuracy = 50) {
  double fmin_ = f(xmin), fmax_ = f(xmax);
  if (std::signbit(fmin_) == std::signbit(fmax_))
    return (std::fabs(fmin_) < std::fabs(fmax_)) ? xmin : xmax;
  double d = xmax - xmin;
  int n = (d > 0.0) ? accuracy + ilog2_trunc(d) : -1;
  double x = 0.5 * (xmin + xmax);
  for (int i = 0; i <= n; i++) {
    x = 0.5 * (xmin + xmax);
    double fx = f(x);
    if (std::signbit(fmin_) != std::signbit(fx)) {
      xmax = x;
    } else {
      xmin = x;
      fmin_ = fx;
    }
  }
  return x;
}

struct V3 { double x, y, z; };

// Gradient: src/PointOperations.cpp:42-48.  Plane through three (x, y, z)
// points; 2x2 partial-pivot LU of [d1x d1y; d2x d2y] g = [d1z; d2z].
static void Gradient(const V3& p0, const V3& p1, const V3& p2, double g[2]) {
  double a00 = p1.x - p0.x, a01 = p1.y - p0.y, b0 = p1.z - p0.z;
  double a10 = p2.x - p0.x, a11 = p2.y - p0.y, b1 = p2.z - p0.z;
  if (std::fabs(a10) > std::fabs(a00)) {
    std::swap(a00, a10); std::swap(a01, a11); std::swap(b0, b1);
  }
  double l = a10 / a00;
  double u11 = a11 - l * a01;
  double y1 = b1 - l * b0;
  double x1 = y1 / u11;
  double x0 = (b0 - a01 * x1) / a00;
  g[0] = x0;
  g[1] = x1;
}

// ------------------------------------------------------------------ mesh --

struct Mesh {
  Idx V = 0, E = 0, T = 0;
  std::vector<double> px, py, pz;
  std::vector<Idx> ep, et, tp, te, tt;  // row-major E x 2, E x 2, T x 3 x3
  std::vector<Idx> pe;                  // partner of each PERIODIC edge (D18), -1 elsewhere
};

// Topology table builder — conventions of the reference's mesh reader,
// pinned by notebooks/topology.dat (written by examples/Main.cpp:146-165):
// boundary edges first in the given line order; the remaining edges numbered
// by first appearance over cells, k = 0..2, edge (tp[k], tp[k+1]);
// ep0 < ep1; te[k] = edge (tp[k], tp[k+1]); interior et = (later, earlier);
// boundary et = (cell, -1); tt[k] = neighbour across te[k], -1 on the wall.
static void BuildTopology(Mesh& m, const Idx* tri, Idx T, const Idx* bnd, Idx nb) {
  m.T = T;
  std::unordered_map<uint64_t, Idx> map;
  map.reserve(size_t(2 * (T + nb) + 16));
  auto key = [](Idx a, Idx b) {
    Idx lo = std::min(a, b), hi = std::max(a, b);
    return (uint64_t(lo) << 32) | uint64_t(hi);
  };
  m.ep.clear(); m.et.clear();
  for (Idx k = 0; k < nb; ++k) {
    Idx a = bnd[2 * k], b = bnd[2 * k + 1];
    uint64_t kk = key(a, b);
    if (map.count(kk)) continue;
    map[kk] = Idx(m.ep.size() / 2);
    m.ep.push_back(std::min(a, b)); m.ep.push_back(std::max(a, b));
    m.et.push_back(-1); m.et.push_back(-1);
  }
  m.tp.assign(tri, tri + 3 * T);
  m.te.assign(size_t(3 * T), -1);
  m.tt.assign(size_t(3 * T), -1);
  for (Idx t = 0; t < T; ++t) {
    for (int k = 0; k < 3; ++k) {
      Idx a = tri[3 * t + k], b = tri[3 * t + (k + 1) % 3];
      uint64_t kk = key(a, b);
      auto it = map.find(kk);
      Idx e;
      if (it == map.end()) {
        e = Idx(m.ep.size() / 2);
        map.emplace(kk, e);
        m.ep.push_back(std::min(a, b)); m.ep.push_back(std::max(a, b));
        m.et.push_back(-1); m.et.push_back(-1);
      } else {
        e = it->second;
      }
      m.te[3 * t + k] = e;
      if (m.et[2 * e] == -1) {
        m.et[2 * e] = t;
      } else {
        m.et[2 * e + 1] = m.et[2 * e];
        m.et[2 * e] = t;
      }
    }
  }
  m.E = Idx(m.ep.size() / 2);
  for (Idx t = 0; t < T; ++t)
    for (int k = 0; k < 3; ++k) {
      Idx e = m.te[3 * t + k];
      Idx a = m.et[2 * e], b = m.et[2 * e + 1];
      m.tt[3 * t + k] = (a == t) ? b : a;
    }
}

// --------------------------------------------------------------- solver --

struct MUSCL {  // include/MUSCLObject.h:15-55 ; only grad row 0 is ever non-zero
  double o[3];
  double g[2];
  Idx i;
};

struct Solver {
  Mesh m;
  int flux = FLUX_HLLC, ws = WS_EINFELDT, opts = 0;
  double cor = 0.0, tau = 0.0;
  std::vector<double> vol;      // T x 3 primitive (w, u, v)
  std::vector<double> edg;      // 2E x 3 edge slots (w, u, v)
  std::vector<double> src;      // 2E x 3 source slots (row 0 unused)
  std::vector<double> f;        // E x 3 fluxes
  std::vector<double> max_wp;   // V
  std::vector<double> max_wp1;  // pass-1 snapshot (D6)
  double r = 1.0;
  const double* drain = nullptr;  // state used by ComputeDrainingDt

  // ---- Domain (src/Bathymetry.cpp) ----
  V3 P(Idx n) const { return {m.px[n], m.py[n], m.pz[n]}; }
  // T(t): src/Bathymetry.cpp:24-27 ; ((p0 c + p1 c) + p2 c), c = 1/3
  V3 Tc(Idx t) const {
    const double c = 1. / 3.;
    V3 a = P(m.tp[3 * t]), b = P(m.tp[3 * t + 1]), d = P(m.tp[3 * t + 2]);
    return {(a.x * c + b.x * c) + d.x * c, (a.y * c + b.y * c) + d.y * c,
            (a.z * c + b.z * c) + d.z * c};
  }
  // E(e): D1 midpoint
  V3 Ec(Idx e) const {
    V3 a = P(m.ep[2 * e]), b = P(m.ep[2 * e + 1]);
    return {0.5 * (a.x + b.x), 0.5 * (a.y + b.y), 0.5 * (a.z + b.z)};
  }
  // L(e): src/Bathymetry.cpp:82-85, Len at src/PointOperations.cpp:4-6
  double L(Idx e) const {
    V3 a = P(m.ep[2 * e]), b = P(m.ep[2 * e + 1]);
    return std::sqrt((a.x - b.x) * (a.x - b.x) + (a.y - b.y) * (a.y - b.y));
  }
  // Area(t): src/Bathymetry.cpp:87-90, TriangArea at src/PointOperations.cpp:12-14
  double Area(Idx t) const {
    V3 a = P(m.tp[3 * t]), b = P(m.tp[3 * t + 1]), c = P(m.tp[3 * t + 2]);
    double ux = b.x - a.x, uy = b.y - a.y, vx = c.x - a.x, vy = c.y - a.y;
    return 0.5 * std::fabs(ux * vy - vx * uy);
  }
  // Tang(e, t): src/Bathymetry.cpp:69-76 (outward from t)
  void Tang(Idx e, Idx t, double tn[2]) const {
    V3 a = P(m.ep[2 * e]), b = P(m.ep[2 * e + 1]);
    double l = L(e);
    double tx = (b.x - a.x) / l, ty = (b.y - a.y) / l;
    V3 c = Tc(t);
    double dx = c.x - a.x, dy = c.y - a.y;
    if (dx * ty - tx * dy > 0.) { tx = -tx; ty = -ty; }
    tn[0] = tx; tn[1] = ty;
  }
  // Norm(e, t): src/Bathymetry.cpp:78-80 ; (t_y, -t_x)
  void Norm(Idx e, Idx t, double n[2]) const {
    double tn[2];
    Tang(e, t, tn);
    n[0] = tn[1]; n[1] = -tn[0];
  }
  // TriangSlope: src/Bathymetry.cpp:10-13
  void TriangSlope(Idx t, double g[2]) const {
    Gradient(P(m.tp[3 * t]), P(m.tp[3 * t + 1]), P(m.tp[3 * t + 2]), g);
  }

  // ---- topology predicates (src/TriangMesh.cpp:5-23) ----
  bool IsEdgeBoundary(Idx e) const { return m.et[2 * e + 1] < 0; }
  bool IsTriangleBoundary(Idx t) const {
    return IsEdgeBoundary(m.te[3 * t]) || IsEdgeBoundary(m.te[3 * t + 1]) ||
           IsEdgeBoundary(m.te[3 * t + 2]);
  }

  // ---- value fields (include/ValueField.h, src/ValueField.cpp) ----
  double bcell(Idx t) const { return Tc(t).z; }                       // :8-10
  double bedge(Idx e) const { return 0.5 * (m.pz[m.ep[2 * e]] + m.pz[m.ep[2 * e + 1]]); }  // :13-15 + D1
  static Idx Slot(Idx e, Idx from, Idx to) { return 2 * e + (Idx)(from < to); }  // include/ValueField.h:70-75
  double w(Idx t) const { return vol[3 * t]; }
  double u(Idx t) const { return vol[3 * t + 1]; }
  double v(Idx t) const { return vol[3 * t + 2]; }
  double h(Idx t) const { return vol[3 * t] - bcell(t); }

  // PrimAssigner::operator= (src/Assigners.cpp:8-20)
  static void PrimAssign(double* col, double b, const double rhs[3]) {
    double hh = rhs[0] - b;
    if (!IsWet(hh)) { col[0] = b; col[1] = 0.; col[2] = 0.; return; }
    col[0] = rhs[0]; col[1] = rhs[1]; col[2] = rhs[2];
    if (hh < 1e-3) {
      double fac = std::sqrt(2.) * hh / std::sqrt(hh * hh + 1e-6);
      col[1] *= fac; col[2] *= fac;
    }
  }
  // ConsAssigner::operator= (src/Assigners.cpp:28-44)
  static void ConsAssign(double* col, double b, const double rhs[3]) {
    double hh = rhs[0];
    if (!IsWet(hh)) { col[0] = b; col[1] = 0.; col[2] = 0.; return; }
    double ih;
    if (hh < 1e-3) ih = std::sqrt(2.) * hh / std::sqrt(hh * hh * hh * hh + 1e-12);
    else ih = 1. / hh;
    col[0] = hh + b; col[1] = rhs[1] * ih; col[2] = rhs[2] * ih;
  }

  // ---- classification (src/MUSCLObject.cpp:13-29) ----
  bool IsDryCell(Idx i) const { return !IsWet(h(i)); }
  bool IsFullWetCell(Idx i) const {
    if (IsTriangleBoundary(i)) return false;
    double z0 = m.pz[m.tp[3 * i]], z1 = m.pz[m.tp[3 * i + 1]], z2 = m.pz[m.tp[3 * i + 2]];
    double max_bp = dmax(dmax(z0, z1), z2);
    return max_bp < w(i);
  }
  bool IsPartWetCell(Idx i) const { return !IsDryCell(i) && !IsFullWetCell(i); }

  // MUSCL::AtPoint (include/MUSCLObject.h:26-33)
  void AtPoint(const MUSCL& mu, const V3& p, double out[3]) const {
    V3 c = Tc(mu.i);
    double dx = p.x - c.x, dy = p.y - c.y;
    double a0 = mu.o[0] + (mu.g[0] * dx + mu.g[1] * dy);
    if ((a0 - p.z) >= 0) {
      out[0] = a0; out[1] = mu.o[1]; out[2] = mu.o[2];
    } else {
      out[0] = p.z; out[1] = 0.; out[2] = 0.;
    }
  }

  // ReconstructDryCell: src/MUSCLObject.cpp:31-36
  MUSCL ReconstructDryCell(Idx i) const {
    MUSCL mu{{bcell(i), 0., 0.}, {0., 0.}, i};
    TriangSlope(i, mu.g);
    return mu;
  }

  // ReconstructPartWetCell1: src/MUSCLObject.cpp:86-112
  MUSCL ReconstructPartWetCell1(Idx i) const {
    double z0 = m.pz[m.tp[3 * i]], z1 = m.pz[m.tp[3 * i + 1]], z2 = m.pz[m.tp[3 * i + 2]];
    double b13 = dmax(dmax(z0, z1), z2);
    double b23 = dmin(dmin(z0, z1), z2);
    double b12 = 3. * bcell(i) - b23 - b13;
    double b_delim = b12 + (1. / 3.) * (b13 - b12) * (b13 - b12) / (b13 - b23);
    double wi = w(i), w_rec;
    if (wi >= b13) {
      w_rec = wi;
    } else if (wi <= b_delim) {
      w_rec = b23 + cbrt_det(3. * h(i) * (b13 - b23) * (b12 - b23));
    } else {
      double a = -3. * b13;
      double b = 3. * (b12 * b13 + b13 * b23 - b12 * b23);
      double c = (b13 - b23) * (3. * h(i) * (b13 - b12) - b12 * (b12 + b23)) - b23 * b23 * b13;
      w_rec = Bisection(CubicPoly{a, b, c}, b12, b13);
    }
    return MUSCL{{w_rec, u(i), v(i)}, {0., 0.}, i};
  }

  // ReconstructFullWetCell: src/MUSCLObject.cpp:38-84 (D2, D3)
  MUSCL ReconstructFullWetCell(Idx i) const {
    const Idx* ip = &m.tp[3 * i];
    const Idx* ie = &m.te[3 * i];
    const Idx* it = &m.tt[3 * i];
    V3 pts[3];
    for (int k = 0; k < 3; ++k) {
      Idx j = it[k];
      if (IsFullWetCell(j)) {
        V3 c = Tc(j);
        pts[k] = {c.x, c.y, (opts & OPT_HEAD_GRADIENT) ? c.z : w(j)};
      } else if (IsDryCell(j)) {
        return MUSCL{{w(i), u(i), v(i)}, {0., 0.}, i};
      } else {
        MUSCL mj = ReconstructPartWetCell1(j);
        V3 pt = Ec(ie[k]);
        double a[3];
        AtPoint(mj, pt, a);
        double val = 0.5 * (w(i) + a[0]);
        pts[k] = {pt.x, pt.y, (opts & OPT_HEAD_GRADIENT) ? pt.z : val};
      }
    }
    double g[2];
    Gradient(pts[0], pts[1], pts[2], g);

    // vertex positivity: wp_m = g . (P_m - centroid)_xy + w_i with the
    // (I - 1/3) weighting of src/MUSCLObject.cpp:66
    V3 P0 = P(ip[0]), P1 = P(ip[1]), P2 = P(ip[2]);
    const double c23 = 1. - 1. / 3., c13 = -1. / 3.;
    V3 Pm[3] = {P0, P1, P2};
    bool all_wet = true;
    for (int mm = 0; mm < 3; ++mm) {
      double dxm = 0., dym = 0.;
      double wx[3], wy[3];
      for (int k = 0; k < 3; ++k) {
        double cf = (k == mm) ? c23 : c13;
        wx[k] = Pm[k].x * cf; wy[k] = Pm[k].y * cf;
      }
      dxm = (wx[0] + wx[1]) + wx[2];
      dym = (wy[0] + wy[1]) + wy[2];
      double wp = (g[0] * dxm + g[1] * dym) + w(i);
      double hp = wp - Pm[mm].z;
      if (!IsWet(hp)) all_wet = false;
    }
    if (!all_wet) { g[0] = 0.; g[1] = 0.; }

    // TVD limiter on w (rows u, v have zero gradient): src/MUSCLObject.cpp:74-81
    double tvd = 1.;
    V3 ci = Tc(i);
    for (int k = 0; k < 3; ++k) {
      double wk = w(it[k]);
      double vtmin = dmin(w(i), wk), vtmax = dmax(w(i), wk);
      V3 pe = Ec(ie[k]);
      double vek = w(i) + (g[0] * (pe.x - ci.x) + g[1] * (pe.y - ci.y));
      if (!((vtmin <= vek) && (vek <= vtmax))) tvd = 0.;
    }
    return MUSCL{{w(i), u(i), v(i)}, {tvd * g[0], tvd * g[1]}, i};
  }

  // ReconstructPartWetCell2: src/MUSCLObject.cpp:114-191 (D4, D6)
  MUSCL ReconstructPartWetCell2(Idx i, const std::vector<double>& mwp) const {
    Idx ip[3] = {m.tp[3 * i], m.tp[3 * i + 1], m.tp[3 * i + 2]};
    if (m.pz[ip[0]] > m.pz[ip[1]]) std::swap(ip[0], ip[1]);
    if (m.pz[ip[1]] > m.pz[ip[2]]) std::swap(ip[1], ip[2]);
    if (m.pz[ip[0]] > m.pz[ip[1]]) std::swap(ip[0], ip[1]);
    double b23 = m.pz[ip[0]], b12 = m.pz[ip[1]], b13 = m.pz[ip[2]];
    if ((w(i) > b13) || (b13 - b23 < tol)) return ReconstructPartWetCell1(i);

    double w23 = mwp[ip[0]];
    double h23 = w23 - b23;
    double ratio_b = (b12 - b23) / (b13 - b23);
    double hd1 = 1. / 3. * h23 * ratio_b;
    double hd2 = 1. / 3. * h23 * (2. * b13 - b12 - b23) / (b13 - b23);
    double hi = h(i);
    double ratio_h = hi / h23;
    const bool head = (opts & OPT_HEAD_PARTWET2) != 0;

    V3 Q0 = P(ip[0]), Q1 = P(ip[1]), Q2 = P(ip[2]);
    V3 p0 = Q0, p1, p2;
    if (!head) p0.z = w23;

    if (hi <= hd1) {  // one wet vertex
      double k2 = std::sqrt(3. * ratio_h / ratio_b);
      p1 = {k2 * Q1.x + (1. - k2) * Q0.x, k2 * Q1.y + (1. - k2) * Q0.y, k2 * Q1.z + (1. - k2) * Q0.z};
      double k3 = std::sqrt(3. * ratio_h * ratio_b);
      p2 = {k3 * Q2.x + (1. - k3) * Q0.x, k3 * Q2.y + (1. - k3) * Q0.y, k3 * Q2.z + (1. - k3) * Q0.z};
    } else if (hi >= hd2) {  // three wet vertices
      double delta_w = 1.5 * (hi - hd2);
      p1 = Q1;
      if (!head) { p1.z += delta_w; p1.z += (1. - ratio_b) * h23; }
      p2 = Q2;
      p2.z += delta_w;
    } else {  // two wet vertices
      double alpha = 3. * ratio_h;
      double beta = (b13 - b12) / (b13 - b23);
      double k1 = 1. - Bisection(CubicPoly{0., (alpha - 3.) / beta, (1. + beta - alpha) / (beta * beta)}, 0., 1.);
      if (k1 < tol) return ReconstructPartWetCell1(i);
      double k3 = 1. - beta * (1. - k1);
      double bp1 = k1 * b13 + (1. - k1) * b12;
      p1 = Q1;
      if (!head) p1.z = b12 + (k1 / k3) * beta * h23;
      p2 = {k1 * Q2.x + (1. - k1) * Q1.x, k1 * Q2.y + (1. - k1) * Q1.y, k1 * Q2.z + (1. - k1) * Q1.z};
      p2.z = bp1;
    }
    MUSCL mu{{0., u(i), v(i)}, {0., 0.}, i};
    Gradient(p0, p1, p2, mu.g);
    V3 c = Tc(i);
    mu.o[0] = w23 + (mu.g[0] * (c.x - Q0.x) + mu.g[1] * (c.y - Q0.y));
    return mu;
  }

  // SpaceDisc::UpdateInterfaceValues: src/SpaceDisc.cpp:15-31
  void UpdateInterfaceValues(const MUSCL& mu) {
    Idx i = mu.i;
    for (int k = 0; k < 3; ++k) {
      Idx node = m.tp[3 * i + k], e = m.te[3 * i + k], nb = m.tt[3 * i + k];
      double a[3];
      AtPoint(mu, P(node), a);
      max_wp[node] = dmax(max_wp[node], a[0]);
      V3 pe = Ec(e);
      AtPoint(mu, pe, a);
      Idx s = Slot(e, i, nb);
      PrimAssign(&edg[3 * s], pe.z, a);
      double ue = edg[3 * s + 1], ve = edg[3 * s + 2];
      src[3 * s + 1] = mu.g[0];
      src[3 * s + 2] = mu.g[1];
      src[3 * s + 1] += cor * (-ve);
      src[3 * s + 2] += cor * ue;
    }
  }

  // SpaceDisc::ComputeInterfaceValues: src/SpaceDisc.cpp:33-52
  void ComputeInterfaceValues() {
    max_wp = m.pz;
    for (Idx i = 0; i < m.T; ++i) {
      if (IsDryCell(i)) UpdateInterfaceValues(ReconstructDryCell(i));
      else if (!IsFullWetCell(i)) UpdateInterfaceValues(ReconstructPartWetCell1(i));
      else UpdateInterfaceValues(ReconstructFullWetCell(i));
    }
    if (!(opts & OPT_SEQ_NODEMAX)) max_wp1 = max_wp;
    for (Idx i = 0; i < m.T; ++i) {
      if (IsPartWetCell(i))
        UpdateInterfaceValues(ReconstructPartWetCell2(i, (opts & OPT_SEQ_NODEMAX) ? max_wp : max_wp1));
    }
  }

  // TriangAverage<k, n> (include/PointOperations.h:24-44): mean of f over
  // the n^2 sub-triangle centroids of (p0, p1, p2), points carried in 3-D.
  template <class F>
  void TriangAverage2(int n, const V3& p0, const V3& p1, const V3& p2, F f, double out[2]) const {
    const double hs = 1. / n;
    const V3 di = {hs * (p1.x - p0.x), hs * (p1.y - p0.y), hs * (p1.z - p0.z)};
    const V3 dj = {hs * (p2.x - p0.x), hs * (p2.y - p0.y), hs * (p2.z - p0.z)};
    const double c = 1. / 3.;
    const V3 dt = {c * (di.x + dj.x), c * (di.y + dj.y), c * (di.z + dj.z)};
    double sum[2] = {0., 0.};
    V3 pi = p0;
    auto acc = [&](const V3& q) {
      double r[2];
      f(q, r);
      sum[0] += hs * r[0];
      sum[1] += hs * r[1];
    };
    for (int i = 0; i < n; ++i) {
      V3 pt = {pi.x + dt.x, pi.y + dt.y, pi.z + dt.z};
      for (int j = 0; j < n - i - 1; ++j) {
        acc(pt);
        acc({pt.x + dt.x, pt.y + dt.y, pt.z + dt.z});
        pt = {pt.x + dj.x, pt.y + dj.y, pt.z + dj.z};
      }
      acc(pt);
      pi = {pi.x + di.x, pi.y + di.y, pi.z + di.z};
    }
    out[0] = hs * sum[0];
    out[1] = hs * sum[1];
  }

  // Per-cell terms of ComputeIntegrals (src/SpaceDisc.cpp:77-103, commented
  // out in the reference): area * (h, 1/2 h |u|^2, 1/2 h^2 + h b).  Full-wet
  // cells use the cell average; partially wet cells add area * h and the
  // TriangAverage<2,10> of the energies of their pass-2 reconstruction (the
  // bathymetry at a quadrature point is the point's interpolated z); dry
  // cells add nothing.  Call after ComputeInterfaceValues.
  void CellIntegrals(Idx i, double out[3]) const {
    out[0] = out[1] = out[2] = 0.;
    if (IsDryCell(i)) return;
    const double A = Area(i);
    if (IsFullWetCell(i)) {
      const double hh = h(i), b = bcell(i), uu = u(i), vv = v(i);
      out[0] = A * hh;
      out[1] = A * (0.5 * hh * (uu * uu + vv * vv));
      out[2] = A * (0.5 * hh * hh + hh * b);
      return;
    }
    out[0] = A * h(i);
    const MUSCL mu = ReconstructPartWetCell2(i, (opts & OPT_SEQ_NODEMAX) ? max_wp : max_wp1);
    double e[2];
    TriangAverage2(10, P(m.tp[3 * i]), P(m.tp[3 * i + 1]), P(m.tp[3 * i + 2]),
                   [&](const V3& q, double r[2]) {
                     double z[3];
                     AtPoint(mu, q, z);
                     const double hq = z[0] - q.z;
                     r[0] = 0.5 * hq * (z[1] * z[1] + z[2] * z[2]);
                     r[1] = 0.5 * hq * hq + hq * q.z;
                   }, e);
    out[1] = A * e[0];
    out[2] = A * e[1];
  }

  // ElemFlux: include/SpaceDisc.h:4-18
  static void ElemFlux(const double n[2], const double U[3], double out[3]) {
    out[0] = 0.; out[1] = 0.; out[2] = 0.;
    double hh = U[0];
    if (IsWet(hh)) {
      double hveln = U[1] * n[0] + U[2] * n[1];
      out[0] = hveln;
      out[1] = (hveln / hh) * U[1] + (0.5 * hh * hh) * n[0];
      out[2] = (hveln / hh) * U[2] + (0.5 * hh * hh) * n[1];
    }
  }

  // Wavespeeds: src/Fluxes.cpp:5-26
  void Wavespeed(double ul, double hl, double ur, double hr, double a[2]) const {
    double cl = std::sqrt(hl), cr = std::sqrt(hr);
    if (ws == WS_RUSANOV) {
      double ap = dmax(std::fabs(ul) + cl, std::fabs(ur) + cr);
      a[0] = -ap; a[1] = ap;
    } else if (ws == WS_DAVIS) {
      a[0] = dmin(ul - cl, ur - cr); a[1] = dmax(ul + cl, ur + cr);
    } else {
      double uRoe = (cl * ul + cl * ur) / (cl + cr);  // D9
      double cRoe = std::sqrt(0.5 * (hl + hr));
      a[0] = dmin(ul - cl, uRoe - cRoe); a[1] = dmax(ur + cr, uRoe + cRoe);
    }
  }

  // edge slot state in conservative form (ValueField::cons, include/ValueField.h:34,39-41)
  void SlotCons(Idx e, Idx s, double U[3]) const {
    double hh = edg[3 * s] - bedge(e);
    U[0] = hh; U[1] = hh * edg[3 * s + 1]; U[2] = hh * edg[3 * s + 2];
  }

  // Fluxes::HLL<W>: include/Fluxes.h:14-54
  // sr_p >= 0: the right state's slot and to_p the cell whose area enters
  // the CFL length (a periodic pair, D18)
  void HLL(Idx e, Idx from, Idx to, double* rr, double F[3], Idx sr_p = -1, Idx to_p = -1) const {
    double n[2];
    Norm(e, from, n);
    Idx sl = Slot(e, from, to), sr = sr_p >= 0 ? sr_p : (to < 0 ? sl : Slot(e, to, from));
    if (to_p >= 0) to = to_p;
    double be = bedge(e);
    double ul = edg[3 * sl + 1] * n[0] + edg[3 * sl + 2] * n[1];
    double hl = edg[3 * sl] - be;
    double ur = edg[3 * sr + 1] * n[0] + edg[3 * sr + 2] * n[1];
    double hr = edg[3 * sr] - be;
    F[0] = F[1] = F[2] = 0.;
    if (hl + hr <= 1e-10) return;
    double a[2];
    Wavespeed(ul, hl, ur, hr, a);
    double al = dmin(0., a[0]), ar = dmax(0., a[1]);
    if (ar - al <= 1e-10) return;
    if (rr) {
      double le = L(e);
      double dl = 2. * Area(from) / le, dr = to < 0 ? dl : 2. * Area(to) / le;
      double c = std::fabs(cor);
      double lw = dmin(dl, dr) / (c + dmax(-al, ar));
      *rr = dmin(*rr, lw);
    }
    double Ul[3], Ur[3], Fl[3], Fr[3];
    SlotCons(e, sl, Ul); SlotCons(e, sr, Ur);
    ElemFlux(n, Ul, Fl); ElemFlux(n, Ur, Fr);
    if (opts & OPT_EXACT_BALANCE) {
      // D19: the same flux as Fl + al (ar (Ur - Ul) - (Fr - Fl)) / (ar - al),
      // which is Fl exactly when the two states are equal
      for (int c = 0; c < 3; ++c)
        F[c] = Fl[c] + al * (ar * (Ur[c] - Ul[c]) - (Fr[c] - Fl[c])) / (ar - al);
      return;
    }
    for (int c = 0; c < 3; ++c)
      F[c] = (ar * Fl[c] - al * Fr[c] + (al * ar) * (Ur[c] - Ul[c])) / (ar - al);
  }

  // Fluxes::HLLC<W>: include/Fluxes.h:56-111
  void HLLC(Idx e, Idx from, Idx to, double* rr, double F[3], Idx sr_p = -1, Idx to_p = -1) const {
    double t[2], n[2];
    Tang(e, from, t);
    n[0] = t[1]; n[1] = -t[0];
    Idx sl = Slot(e, from, to), sr = sr_p >= 0 ? sr_p : (to < 0 ? sl : Slot(e, to, from));
    if (to_p >= 0) to = to_p;
    double be = bedge(e);
    double ul = edg[3 * sl + 1] * n[0] + edg[3 * sl + 2] * n[1];
    double vl = edg[3 * sl + 1] * t[0] + edg[3 * sl + 2] * t[1];
    double hl = edg[3 * sl] - be;
    double ur = edg[3 * sr + 1] * n[0] + edg[3 * sr + 2] * n[1];
    double vr = edg[3 * sr + 1] * t[0] + edg[3 * sr + 2] * t[1];
    double hr = edg[3 * sr] - be;
    F[0] = F[1] = F[2] = 0.;
    if (hl + hr <= 1e-10) return;
    double a[2];
    Wavespeed(ul, hl, ur, hr, a);
    double al = a[0], ar = a[1];
    double ustar = (ar - ur) * hr * ur - (al - ul) * hl * ul + 0.5 * (hl * hl - hr * hr);
    ustar /= (hr * (ar - ur) - hl * (al - ul));
    if (rr) {
      double le = L(e);
      double dl = 2. * Area(from) / le, dr = to < 0 ? dl : 2. * Area(to) / le;
      double c = std::fabs(cor);
      double lw = dmin(dl, dr) / (c + dmax(tol, dmax(al, ar)));
      *rr = dmin(*rr, lw);
    }
    double Ul[3], Ur[3], Fe[3];
    SlotCons(e, sl, Ul); SlotCons(e, sr, Ur);
    if (ustar <= 0) {
      double us = vr * t[0] + ustar * t[1];
      double vs = vr * n[0] + ustar * n[1];
      // D19: the ratio first, so that at rest (u* = ur = 0) hs is hr exactly
      double hs = (opts & OPT_EXACT_BALANCE) ? hr * ((ar - ur) / (ar - ustar)) : hr * (ar - ur) / (ar - ustar);
      double Us[3] = {hs, hs * us, hs * vs};
      ElemFlux(n, Ur, Fe);
      double s = dmax(0., ar);
      for (int c = 0; c < 3; ++c) F[c] = Fe[c] + s * (Us[c] - Ur[c]);
    } else {
      double us = vl * t[0] + ustar * t[1];
      double vs = vl * n[0] + ustar * n[1];
      double hs = (opts & OPT_EXACT_BALANCE) ? hl * ((al - ul) / (al - ustar)) : hl * (al - ul) / (al - ustar);
      double Us[3] = {hs, hs * us, hs * vs};
      ElemFlux(n, Ul, Fe);
      double s = dmin(0., al);
      for (int c = 0; c < 3; ++c) F[c] = Fe[c] + s * (Us[c] - Ul[c]);
    }
  }

  // SpaceDisc::ComputeFluxes: src/SpaceDisc.cpp:54-74 (D11)
  void ComputeFluxes() {
    r = 1.;
    for (Idx e = 0; e < m.E; ++e) {
      Idx lf = m.et[2 * e], lt = m.et[2 * e + 1];
      double* F = &f[3 * e];
      if (lt == SOLID_WALL) {
        double n[2];
        Norm(e, lf, n);
        // D19: with OPT_EXACT_BALANCE the inside edge depth, the depth of
        // the RHS's pressure correction, instead of the cell mean
        const double hw = (opts & OPT_EXACT_BALANCE) ? edg[3 * Slot(e, lf, lt)] - bedge(e) : h(lf);
        double U[3] = {hw, 0., 0.};
        ElemFlux(n, U, F);
      } else if (lt == PERIODIC) {  // D18: the pair's lower edge, its negation on the other
        const Idx q = m.pe[e], p = std::min(e, q), s2 = std::max(e, q);
        const Idx fp = m.et[2 * p], fs = m.et[2 * s2];
        // slots of boundary edges: Slot(e, cell, -1) = 2 e
        if (flux == FLUX_HLL) HLL(p, fp, PERIODIC, &r, F, 2 * s2, fs);
        else HLLC(p, fp, PERIODIC, &r, F, 2 * s2, fs);
        if (e != p) { F[0] = -F[0]; F[1] = -F[1]; F[2] = -F[2]; }
      } else if (flux == FLUX_HLL) {
        HLL(e, lf, lt, &r, F);
      } else {
        HLLC(e, lf, lt, &r, F);
      }
    }
  }

  // TimeDisc::ComputeDrainingDt: src/TimeDisc.cpp:43-66 (D5: reads `drain`)
  double ComputeDrainingDt(Idx i) const {
    if (i < 0) return kInf;
    double hi = drain[3 * i] - bcell(i);
    if (!IsWet(hi)) return 0.;
    double sum = 0.;
    for (int k = 0; k < 3; ++k) {
      Idx e = m.te[3 * i + k];
      double fe = f[3 * e];
      double out = (i == m.et[2 * e]) ? fe : -fe;
      // D16: the reference adds the flux per unit length; the outflow
      // through the edge is F0 * L (OPT_DRAIN_LENGTH)
      if (opts & OPT_DRAIN_LENGTH) out = out * L(e);
      sum += dmax(0., out);
    }
    return sum > tol ? Area(i) * hi / sum : kInf;
  }

  // TimeDisc::RHS: src/TimeDisc.cpp:3-41
  void RHS(Idx i, double dt, double res[3]) const {
    res[0] = res[1] = res[2] = 0.;
    double i_area = 1. / Area(i);
    double dti = ComputeDrainingDt(i);
    for (int k = 0; k < 3; ++k) {
      Idx e = m.te[3 * i + k], nb = m.tt[3 * i + k];
      int sgn = (m.et[2 * e] == i) ? 1 : -1;
      // D18: across a periodic pair the neighbour is the partner edge's cell
      const Idx nbd = (nb < 0 && m.et[2 * e + 1] == PERIODIC) ? m.et[2 * m.pe[e]] : nb;
      double dtik = ComputeDrainingDt(nbd);
      const double* F = &f[3 * e];
      double dtk = (sgn * F[0]) > 0. ? dmin(dt, dti) : dmin(dt, dtik);
      double c_ek = i_area * L(e);
      Idx s = Slot(e, i, nb);
      double h_ek = edg[3 * s] - bedge(e);
      if (opts & OPT_EXACT_BALANCE) {
        // D19: the flux minus the cell's own hydrostatic edge pressure,
        // zero to the bit where the edge states are equal and at rest
        double n[2];
        Norm(e, i, n);
        double q = 0.5 * h_ek * h_ek;
        double g0 = sgn * F[0], g1 = sgn * F[1] - n[0] * q, g2 = sgn * F[2] - n[1] * q;
        res[0] -= dtk * (c_ek * g0);
        res[1] -= dtk * (c_ek * g1);
        res[2] -= dtk * (c_ek * g2);
        double cs = dt * (1. / 3.);
        res[1] -= cs * src[3 * s + 1] * h_ek;
        res[2] -= cs * src[3 * s + 2] * h_ek;
        continue;
      }
      double cf = dtk * sgn * c_ek;
      res[0] -= cf * F[0];
      res[1] -= cf * F[1];
      res[2] -= cf * F[2];
      double cs = dt * (1. / 3.);
      res[1] -= cs * src[3 * s + 1] * h_ek;
      res[2] -= cs * src[3 * s + 2] * h_ek;
      double n[2];
      Norm(e, i, n);
      double q = 0.5 * h_ek * h_ek;
      res[1] += dtk * (n[0] * c_ek * q);
      res[2] += dtk * (n[1] * c_ek * q);
    }
    if ((opts & OPT_FRICTION) && tau != 0.) {
      const double hh = h(i);
      const bool dry = IsDryCell(i);
      const double uu = dry ? 0. : u(i), vv = dry ? 0. : v(i);
      res[1] -= dt * (tau * (hh * uu));
      res[2] -= dt * (tau * (hh * vv));
    }
  }

  void cons_of(const std::vector<double>& st, Idx i, double U[3]) const {
    double hh = st[3 * i] - bcell(i);
    U[0] = hh; U[1] = hh * st[3 * i + 1]; U[2] = hh * st[3 * i + 2];
  }

  // One update loop: vol(i) <- a0 * U0(i) + a1 * Ucur(i) + RHS(i, dts)
  // (src/Solvers.cpp:11-13, 23-25, 30-32, 42-58)
  void UpdateLoop(const std::vector<double>* U0, double a0, double a1, double dts) {
    std::vector<double> cur = vol;  // stage-start copy
    drain = (opts & OPT_SEQ_DRAINING) ? vol.data() : cur.data();
    for (Idx i = 0; i < m.T; ++i) {
      double R[3], Uc[3], rhs[3];
      RHS(i, dts, R);
      cons_of(cur, i, Uc);
      if (U0) {
        double Ua[3];
        cons_of(*U0, i, Ua);
        // D19: with OPT_EXACT_BALANCE a convex combination is written
        // Uc + a0 (Ua - Uc), which is Uc to the bit where Ua = Uc
        if ((opts & OPT_EXACT_BALANCE) && a0 + a1 == 1.)
          for (int c = 0; c < 3; ++c) rhs[c] = (Uc[c] + a0 * (Ua[c] - Uc[c])) + R[c];
        else
          for (int c = 0; c < 3; ++c) rhs[c] = (a0 * Ua[c] + a1 * Uc[c]) + R[c];
      } else {
        for (int c = 0; c < 3; ++c) rhs[c] = Uc[c] + R[c];
      }
      ConsAssign(&vol[3 * i], bcell(i), rhs);
    }
    drain = nullptr;
  }

  // Solvers: src/Solvers.cpp:5-59 ; dt <= 0 selects CFL mode (D12) with
  // dt = cfl * r from the first stage's flux pass.  Returns the dt used.
  double Step(int scheme, double dt, double cfl) {
    ComputeInterfaceValues();
    ComputeFluxes();
    if (dt <= 0.) dt = cfl * r;
    std::vector<double> U0 = vol;
    UpdateLoop(nullptr, 0., 0., dt);
    if (scheme == SCHEME_EULER) return dt;
    ComputeInterfaceValues();
    ComputeFluxes();
    if (scheme == SCHEME_SSPRK2) {
      UpdateLoop(&U0, 0.5, 0.5, 0.5 * dt);
      return dt;
    }
    UpdateLoop(&U0, 0.75, 0.25, 0.25 * dt);
    ComputeInterfaceValues();
    ComputeFluxes();
    UpdateLoop(&U0, 1. / 3., 2. / 3., (2. / 3.) * dt);
    return dt;
  }
};

}  // namespace orc

using namespace orc;

// ------------------------------------------------------------------ C ABI --
extern "C" {

struct orc_topo { Mesh m; };

// Build reference-convention tables from a triangle list and boundary lines.
orc_topo* orc_topology_build(int64_t num_nodes, int64_t T, const int64_t* tri,
                             int64_t nb, const int64_t* bnd) {
  auto* h = new orc_topo;
  h->m.V = num_nodes;
  BuildTopology(h->m, tri, T, bnd, nb);
  return h;
}
int64_t orc_topology_num_edges(const orc_topo* h) { return h->m.E; }
void orc_topology_copy(const orc_topo* h, int64_t* ep, int64_t* et, int64_t* tp,
                       int64_t* te, int64_t* tt) {
  const Mesh& m = h->m;
  std::memcpy(ep, m.ep.data(), sizeof(int64_t) * 2 * m.E);
  std::memcpy(et, m.et.data(), sizeof(int64_t) * 2 * m.E);
  std::memcpy(tp, m.tp.data(), sizeof(int64_t) * 3 * m.T);
  std::memcpy(te, m.te.data(), sizeof(int64_t) * 3 * m.T);
  std::memcpy(tt, m.tt.data(), sizeof(int64_t) * 3 * m.T);
}
void orc_topology_free(orc_topo* h) { delete h; }

// Solver on given tables.  xyz: V x 3 row-major (x, y, z).
Solver* orc_create(int64_t V, const double* xyz, int64_t E, const int64_t* ep,
                   const int64_t* et, int64_t T, const int64_t* tp, const int64_t* te,
                   const int64_t* tt, int flux, int ws, double cor, double tau, int opts,
                   const int64_t* pe) {
  for (int64_t e = 0; e < E; ++e) {
    if (et[2 * e + 1] < PERIODIC) return nullptr;  // D11: CUSTOM is rejected
    if (et[2 * e + 1] == PERIODIC) {               // D18: paired both ways
      if (!pe || pe[e] < 0 || pe[e] >= E || pe[e] == e || et[2 * pe[e] + 1] != PERIODIC || pe[pe[e]] != e)
        return nullptr;
    }
  }
  auto* s = new Solver;
  Mesh& m = s->m;
  m.V = V; m.E = E; m.T = T;
  m.px.resize(V); m.py.resize(V); m.pz.resize(V);
  for (int64_t i = 0; i < V; ++i) {
    m.px[i] = xyz[3 * i]; m.py[i] = xyz[3 * i + 1]; m.pz[i] = xyz[3 * i + 2];
  }
  m.ep.assign(ep, ep + 2 * E); m.et.assign(et, et + 2 * E);
  m.tp.assign(tp, tp + 3 * T); m.te.assign(te, te + 3 * T); m.tt.assign(tt, tt + 3 * T);
  if (pe) m.pe.assign(pe, pe + E);
  s->flux = flux; s->ws = ws; s->cor = cor; s->tau = tau; s->opts = opts;
  s->vol.assign(3 * T, 0.);
  s->edg.assign(6 * E, 0.);
  s->src.assign(6 * E, 0.);
  s->f.assign(3 * E, 0.);
  s->max_wp = m.pz;
  return s;
}
void orc_destroy(Solver* s) { delete s; }
void orc_set_state(Solver* s, const double* prim) { std::memcpy(s->vol.data(), prim, sizeof(double) * s->vol.size()); }
void orc_get_state(const Solver* s, double* prim) { std::memcpy(prim, s->vol.data(), sizeof(double) * s->vol.size()); }
void orc_compute_interface_values(Solver* s) { s->ComputeInterfaceValues(); }
double orc_compute_fluxes(Solver* s) { s->ComputeFluxes(); return s->r; }
void orc_get_edge_field(const Solver* s, double* out) { std::memcpy(out, s->edg.data(), sizeof(double) * s->edg.size()); }
void orc_get_src_field(const Solver* s, double* out) { std::memcpy(out, s->src.data(), sizeof(double) * s->src.size()); }
void orc_get_fluxes(const Solver* s, double* out) { std::memcpy(out, s->f.data(), sizeof(double) * s->f.size()); }
void orc_get_max_wp(const Solver* s, double* out) { std::memcpy(out, s->max_wp.data(), sizeof(double) * s->max_wp.size()); }
double orc_min_len_to_wavespeed(const Solver* s) { return s->r; }

// Per-cell classification: 0 dry, 1 partially wet, 2 fully wet.
void orc_classify(const Solver* s, int8_t* out) {
  for (int64_t i = 0; i < s->m.T; ++i)
    out[i] = s->IsDryCell(i) ? 0 : (s->IsFullWetCell(i) ? 2 : 1);
}

// Reconstruction of one cell (kind: 0 dry, 1 partwet1, 2 fullwet, 3 partwet2
// using the current max_wp, 4 partwet2 using the pass-1 node maximum, as the
// step does).  out = origin (w, u, v), grad (gx, gy).
void orc_reconstruct(const Solver* s, int64_t i, int kind, double* out) {
  MUSCL mu;
  if (kind == 0) mu = s->ReconstructDryCell(i);
  else if (kind == 1) mu = s->ReconstructPartWetCell1(i);
  else if (kind == 2) mu = s->ReconstructFullWetCell(i);
  else if (kind == 3) mu = s->ReconstructPartWetCell2(i, s->max_wp);
  else mu = s->ReconstructPartWetCell2(i, s->max_wp1);
  out[0] = mu.o[0]; out[1] = mu.o[1]; out[2] = mu.o[2]; out[3] = mu.g[0]; out[4] = mu.g[1];
}

// Mass, kinetic and potential energy (sums over cells in index order) and,
// optionally, the per-cell terms (3 per cell); recomputes the interface
// values of the current state first.
void orc_integrals(Solver* s, double* out, double* per_cell) {
  s->ComputeInterfaceValues();
  double tot[3] = {0., 0., 0.};
  for (int64_t i = 0; i < s->m.T; ++i) {
    double c[3];
    s->CellIntegrals(i, c);
    for (int k = 0; k < 3; ++k) tot[k] += c[k];
    if (per_cell) for (int k = 0; k < 3; ++k) per_cell[3 * i + k] = c[k];
  }
  for (int k = 0; k < 3; ++k) out[k] = tot[k];
}

// TimeDisc::RHS with draining dt read from the current state.
void orc_rhs(Solver* s, int64_t i, double dt, double* out) {
  s->drain = s->vol.data();
  s->RHS(i, dt, out);
  s->drain = nullptr;
}

// nsteps steps; dt > 0 fixed, dt <= 0 CFL mode.  dts (optional) receives each dt.
void orc_run(Solver* s, int scheme, double dt, double cfl, int64_t nsteps, double* dts) {
  for (int64_t k = 0; k < nsteps; ++k) {
    double d = s->Step(scheme, dt, cfl);
    if (dts) dts[k] = d;
  }
}

// Wall-clock seconds of nsteps steps (steady_clock), for the CPU baseline.
double orc_time_run(Solver* s, int scheme, double dt, double cfl, int64_t nsteps) {
  auto t0 = std::chrono::steady_clock::now();
  for (int64_t k = 0; k < nsteps; ++k) s->Step(scheme, dt, cfl);
  auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

// Scalar helpers exposed for pinning tests.
double orc_cbrt(double x) { return cbrt_det(x); }
int orc_ilog2_trunc(double x) { return ilog2_trunc(x); }
double orc_bisection_cubic(double b, double c, double d, double xmin, double xmax) {
  return Bisection(CubicPoly{b, c, d}, xmin, xmax);
}
void orc_gradient(const double* p /*3x3 rows = points*/, double* g) {
  Gradient({p[0], p[1], p[2]}, {p[3], p[4], p[5]}, {p[6], p[7], p[8]}, g);
}
// Numerical flux of one edge through the solver's own ComputeFluxes /
// Solver::HLL / Solver::HLLC (src/SpaceDisc.cpp:54-74, include/Fluxes.h:14-111)
// on a two-cell mesh: xyz holds 4 nodes (x, y, z rows) A, B (the edge), C
// (third vertex of the `from` cell, triangle A B C) and D (third vertex of
// the `to` cell, triangle B A D); both must be counter-clockwise.  The edge
// slots are set to the given primitive states Lp (seen from `from`) and Rp;
// to_tag < 0 drops the `to` cell and tags the edge (SOLID_WALL, FREE_FLOW),
// and a wall's flux then reads the `from` cell's average Lp.  Returns the
// flux, the normal Norm(e, from) and the CFL length r of that pass.
int orc_flux_edge(int flux, int ws, double cor, const double* xyz, const double* Lp,
                  const double* Rp, int to_tag, double* F, double* n_out, double* r_out) {
  Solver s;
  s.flux = flux; s.ws = ws; s.cor = cor;
  Mesh& m = s.m;
  const bool inner = to_tag >= 0;
  m.V = inner ? 4 : 3;
  for (Idx i = 0; i < m.V; ++i) {
    m.px.push_back(xyz[3 * i]); m.py.push_back(xyz[3 * i + 1]); m.pz.push_back(xyz[3 * i + 2]);
  }
  // the later cell of an interior edge is its et0 (the `from` cell): list
  // B A D first
  const Idx tri2[6] = {1, 0, 3, 0, 1, 2}, tri1[3] = {0, 1, 2};
  const Idx bnd2[8] = {1, 3, 3, 0, 1, 2, 2, 0}, bnd1[6] = {0, 1, 1, 2, 2, 0};
  if (inner) BuildTopology(m, tri2, 2, bnd2, 4);
  else BuildTopology(m, tri1, 1, bnd1, 3);
  Idx e = -1;
  for (Idx q = 0; q < m.E; ++q)
    if (std::min(m.ep[2 * q], m.ep[2 * q + 1]) == 0 && std::max(m.ep[2 * q], m.ep[2 * q + 1]) == 1) e = q;
  const Idx from = inner ? 1 : 0, to = inner ? 0 : to_tag;
  if (e < 0 || m.et[2 * e] != from || (inner && m.et[2 * e + 1] != to)) return -1;
  if (!inner) {
    if (to_tag < FREE_FLOW) return -1;
    m.et[2 * e + 1] = to_tag;
    m.tt[3 * from + 0] = to_tag;  // edge (A, B) is local edge 0 of A B C
  }
  s.vol.assign(3 * m.T, 0.);
  s.edg.assign(6 * m.E, 0.);
  s.src.assign(6 * m.E, 0.);
  s.f.assign(3 * m.E, 0.);
  for (int c = 0; c < 3; ++c) s.vol[3 * from + c] = Lp[c];
  if (inner)
    for (int c = 0; c < 3; ++c) s.vol[3 * to + c] = Rp[c];
  const Idx sl = Solver::Slot(e, from, to), sr = inner ? Solver::Slot(e, to, from) : sl;
  for (int c = 0; c < 3; ++c) {
    s.edg[3 * sl + c] = Lp[c];
    if (inner) s.edg[3 * sr + c] = Rp[c];
  }
  s.ComputeFluxes();
  for (int c = 0; c < 3; ++c) F[c] = s.f[3 * e + c];
  s.Norm(e, from, n_out);
  *r_out = s.r;
  return 0;
}
// Solver::HLL / HLLC of edge e seen from `from` towards `to` on the edge
// slots of the last ComputeInterfaceValues (include/Fluxes.h:14-111);
// *r = min(*r, candidate) as there.
void orc_edge_flux(Solver* s, int flux, int ws, int64_t e, int64_t from, int64_t to, double* r, double* F) {
  const int ws0 = s->ws;
  s->ws = ws;
  if (flux == FLUX_HLL) s->HLL(e, from, to, r, F);
  else s->HLLC(e, from, to, r, F);
  s->ws = ws0;
}
void orc_wavespeed(int ws, double ul, double hl, double ur, double hr, double* a) {
  Solver s;
  s.ws = ws;
  s.Wavespeed(ul, hl, ur, hr, a);
}

}  // extern "C"
