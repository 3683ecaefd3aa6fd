// swe_fvm.hpp — header-only C++17 facade with the names and signatures of
// SWE_FVM's operator interface, over the C ABI of swe_fvm.h.
//
// Replaces the reference headers include/{Includes,Exceptions,TriangMesh,
// Bathymetry,ValueField,SpaceDisc,Fluxes,TimeDisc,Solvers}.h for callers of
// the explicit step (examples/Main.cpp:172-195 style).  Eigen is not used:
// Array<k> and Storage<k> are small value types with the accessors the
// drivers use (operator[], operator(), col()).  All numerics of
// ComputeInterfaceValues, ComputeFluxes, RHS and the Solvers run on the GPU;
// a fluxer that is not one of Fluxes::HLL/HLLC<Wavespeeds::*> is rejected
// with SolverError (there is no host path).
#pragma once

#include <stdint.h>

#include <array>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "swe_fvm.h"

// ------------------------------------------------------- include/Includes.h
using Idx = int64_t;
enum class Boundaries : Idx { SOLID_WALL = -1, FREE_FLOW = -2, PERIODIC = -3, CUSTOM = -4 };
constexpr inline double tol = 1e-13;

template <size_t k>
struct Array {
  double v[k] = {};
  Array() = default;
  Array(std::initializer_list<double> l) {
    size_t i = 0;
    for (double x : l) if (i < k) v[i++] = x;
  }
  static Array Zero() { return Array(); }
  double& operator[](size_t i) { return v[i]; }
  double operator[](size_t i) const { return v[i]; }
  double& operator()(size_t i) { return v[i]; }
  double operator()(size_t i) const { return v[i]; }
  Array operator+(const Array& o) const { Array r; for (size_t i = 0; i < k; ++i) r.v[i] = v[i] + o.v[i]; return r; }
  Array operator-(const Array& o) const { Array r; for (size_t i = 0; i < k; ++i) r.v[i] = v[i] - o.v[i]; return r; }
  Array operator*(double s) const { Array r; for (size_t i = 0; i < k; ++i) r.v[i] = v[i] * s; return r; }
  friend Array operator*(double s, const Array& a) { return a * s; }
};

// k x N column-major storage (Eigen::Array<double, k, Dynamic>)
template <size_t k>
struct Storage {
  std::vector<double> d;
  Storage() = default;
  explicit Storage(Idx cols) : d(size_t(cols) * k, 0.0) {}
  Idx cols() const { return Idx(d.size() / k); }
  void resize(Idx cols) { d.assign(size_t(cols) * k, 0.0); }
  double& operator()(Idx r, Idx c) { return d[size_t(c) * k + size_t(r)]; }
  double operator()(Idx r, Idx c) const { return d[size_t(c) * k + size_t(r)]; }
  Array<k> col(Idx c) const { Array<k> a; std::memcpy(a.v, &d[size_t(c) * k], sizeof(a.v)); return a; }
  void set_col(Idx c, const Array<k>& a) { std::memcpy(&d[size_t(c) * k], a.v, sizeof(a.v)); }
  double* data() { return d.data(); }
  const double* data() const { return d.data(); }
};

using Point = Array<3>;
using PointArray = Storage<3>;

// ----------------------------------------------------- include/Exceptions.h
struct DomainError : std::runtime_error { using std::runtime_error::runtime_error; };
struct SolverError : std::runtime_error { using std::runtime_error::runtime_error; };

namespace swe_detail {
inline void check(int rc) {
  if (rc == SWE_OK) return;
  std::string msg = swe_last_error();
  if (rc == SWE_ERR_DOMAIN || rc == SWE_ERR_ARG) throw DomainError(msg);
  if (rc == SWE_ERR_SOLVER) throw SolverError(msg);
  throw std::runtime_error(msg);
}
}  // namespace swe_detail

// ----------------------------------------------------- include/TriangMesh.h
using NodeTag = Idx;
using EdgeTag = std::array<Idx, 2>;
using TriangTag = std::array<Idx, 3>;

// row-major int64 table with C columns (Eigen::Array<Idx, Dynamic, C, RowMajor>)
template <int C>
struct TagArray {
  std::vector<Idx> d;
  TagArray() = default;
  explicit TagArray(Idx rows) : d(size_t(rows) * C, 0) {}
  TagArray(std::initializer_list<std::array<Idx, C>> rows) {
    for (auto& r : rows) d.insert(d.end(), r.begin(), r.end());
  }
  Idx rows() const { return Idx(d.size() / C); }
  Idx& operator()(Idx r, Idx c) { return d[size_t(r) * C + size_t(c)]; }
  Idx operator()(Idx r, Idx c) const { return d[size_t(r) * C + size_t(c)]; }
  const Idx* data() const { return d.data(); }
  Idx* data() { return d.data(); }
};
using EdgeTagArray = TagArray<2>;
using TriangTagArray = TagArray<3>;

// Owning Topology (the reference holds references to caller arrays,
// include/TriangMesh.h:57-64; this one copies them).
struct Topology {
  static Topology create(NodeTag numNodes, const EdgeTagArray& edgeNodes, const EdgeTagArray& edgeElements,
                         const TriangTagArray& elementNodes, const TriangTagArray& elementEdges,
                         const TriangTagArray& elementNeighbours) {
    return Topology(numNodes, edgeNodes, edgeElements, elementNodes, elementEdges, elementNeighbours);
  }
  Topology(NodeTag numNodes, const EdgeTagArray& edgeNodes, const EdgeTagArray& edgeElements,
           const TriangTagArray& elementNodes, const TriangTagArray& elementEdges,
           const TriangTagArray& elementNeighbours) {
    if (edgeNodes.rows() != edgeElements.rows() || elementNodes.rows() != elementEdges.rows() ||
        elementNodes.rows() != elementNeighbours.rows())
      throw DomainError("table sizes disagree");
    swe_topology* t = nullptr;
    swe_detail::check(swe_topology_create(numNodes, edgeNodes.rows(), edgeNodes.data(), edgeElements.data(),
                                          elementNodes.rows(), elementNodes.data(), elementEdges.data(),
                                          elementNeighbours.data(), &t));
    adopt(t);
  }
  // tables by the reference mesh reader's conventions
  static Topology build(NodeTag numNodes, const TriangTagArray& triangles, const EdgeTagArray& boundaryLines) {
    swe_topology* t = nullptr;
    swe_detail::check(swe_topology_build(numNodes, triangles.rows(), triangles.data(), boundaryLines.rows(),
                                         boundaryLines.data(), &t));
    return Topology(t);
  }
  // Gmsh 4.1 reader; geometry receives node (x, y, z)
  static Topology read_msh(const std::string& path, PointArray& geometry) {
    swe_topology* t = nullptr;
    double* xyz = nullptr;
    int64_t nn = 0;
    swe_detail::check(swe_topology_read_msh(path.c_str(), &xyz, &nn, &t));
    geometry.resize(nn);
    std::memcpy(geometry.data(), xyz, sizeof(double) * size_t(3 * nn));
    swe_free(xyz);
    return Topology(t);
  }

  NodeTag NumNodes() const { return swe_topology_num_nodes(h_.get()); }
  NodeTag NumEdges() const { return swe_topology_num_edges(h_.get()); }
  NodeTag NumTriangles() const { return swe_topology_num_elements(h_.get()); }
  EdgeTag EdgePoints(NodeTag i) const { return {ep_(i, 0), ep_(i, 1)}; }
  EdgeTag EdgeTriangs(NodeTag i) const { return {et_(i, 0), et_(i, 1)}; }
  bool IsEdgeBoundary(NodeTag i) const { return et_(i, 1) < 0; }
  TriangTag TriangPoints(NodeTag i) const { return {tp_(i, 0), tp_(i, 1), tp_(i, 2)}; }
  TriangTag TriangEdges(NodeTag i) const { return {te_(i, 0), te_(i, 1), te_(i, 2)}; }
  TriangTag TriangTriangs(NodeTag i) const { return {tt_(i, 0), tt_(i, 1), tt_(i, 2)}; }
  bool IsTriangleBoundary(NodeTag i) const {
    return IsEdgeBoundary(te_(i, 0)) || IsEdgeBoundary(te_(i, 1)) || IsEdgeBoundary(te_(i, 2));
  }
  const swe_topology* handle() const { return h_.get(); }

 private:
  explicit Topology(swe_topology* t) { adopt(t); }
  void adopt(swe_topology* t) {
    h_ = std::shared_ptr<swe_topology>(t, swe_topology_destroy);
    ep_ = EdgeTagArray(NumEdges()); et_ = EdgeTagArray(NumEdges());
    tp_ = TriangTagArray(NumTriangles()); te_ = TriangTagArray(NumTriangles()); tt_ = TriangTagArray(NumTriangles());
    swe_detail::check(swe_topology_copy(t, ep_.data(), et_.data(), tp_.data(), te_.data(), tt_.data()));
  }
  std::shared_ptr<swe_topology> h_;
  EdgeTagArray ep_, et_;
  TriangTagArray tp_, te_, tt_;
};

// ------------------------------------------------------ include/Bathymetry.h
constexpr inline bool IsWet(double h) noexcept { return h > 1e-12; }
inline Point DryState(double b) noexcept { return {b, 0., 0.}; }

// Node geometry (x, y, bathymetry z) on a topology (owning copy).
struct Domain {
  Domain(const PointArray& geometry, const Topology& topology) : m_geometry(geometry), m_topology(topology) {
    if (geometry.cols() != topology.NumNodes()) throw DomainError("geometry and topology disagree");
  }
  const Topology& GetTopology() const noexcept { return m_topology; }
  const PointArray& GetGeometry() const { return m_geometry; }
  size_t Size() const { return size_t(m_geometry.cols()); }
  Point P(NodeTag p) const { return m_geometry.col(p); }
  // Domain::T (src/Bathymetry.cpp:24-27)
  Point T(NodeTag t) const {
    auto tp = m_topology.TriangPoints(t);
    const double c = 1. / 3.;
    Point a = P(tp[0]), b = P(tp[1]), d = P(tp[2]);
    return {(a[0] * c + b[0] * c) + d[0] * c, (a[1] * c + b[1] * c) + d[1] * c, (a[2] * c + b[2] * c) + d[2] * c};
  }
  // edge midpoint
  Point E(NodeTag e) const {
    auto ep = m_topology.EdgePoints(e);
    Point a = P(ep[0]), b = P(ep[1]);
    return {0.5 * (a[0] + b[0]), 0.5 * (a[1] + b[1]), 0.5 * (a[2] + b[2])};
  }
  double L(NodeTag e) const {
    auto ep = m_topology.EdgePoints(e);
    Point a = P(ep[0]), b = P(ep[1]);
    return std::sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]));
  }
  double Area(NodeTag t) const {
    auto tp = m_topology.TriangPoints(t);
    Point a = P(tp[0]), b = P(tp[1]), c = P(tp[2]);
    return 0.5 * std::fabs((b[0] - a[0]) * (c[1] - a[1]) - (c[0] - a[0]) * (b[1] - a[1]));
  }

 private:
  PointArray m_geometry;
  Topology m_topology;
};

// ------------------------------------------------------ include/ValueField.h
// Primitive (w, u, v) per cell; b = centroid bathymetry.
struct VolumeField {
  VolumeField(const Domain& b, size_t size = 0) : m_str(Idx(size)), m_b(&b) {}
  size_t Size() const { return size_t(m_str.cols()); }
  void Resize(size_t size) { m_str.resize(Idx(size)); }
  double b(Idx i) const { return m_b->T(i)[2]; }
  Array<3> prim(Idx i) const { return m_str.col(i); }
  void set_prim(Idx i, const Array<3>& p) { m_str.set_col(i, p); }
  double w(Idx i) const { return m_str(0, i); }
  double u(Idx i) const { return m_str(1, i); }
  double v(Idx i) const { return m_str(2, i); }
  double h(Idx i) const { return w(i) - b(i); }
  double hu(Idx i) const { return h(i) * u(i); }
  double hv(Idx i) const { return h(i) * v(i); }
  Array<3> cons(Idx i) const { return {h(i), hu(i), hv(i)}; }
  Storage<3>& storage() { return m_str; }
  const Storage<3>& storage() const { return m_str; }

 private:
  Storage<3> m_str;
  const Domain* m_b;
};

// Edge slots 2e + (from < to) (include/ValueField.h:70-75); b = midpoint z.
struct EdgeField {
  EdgeField() = default;
  EdgeField(const Domain* b, Idx slots) : m_str(slots), m_b(b) {}
  static Idx Id(Idx e, Idx from, Idx to) { return 2 * e + Idx(from < to); }
  double b(Idx e, Idx, Idx) const { return m_b->E(e)[2]; }
  Array<3> prim(Idx e, Idx f, Idx t) const { return m_str.col(Id(e, f, t)); }
  double w(Idx e, Idx f, Idx t) const { return m_str(0, Id(e, f, t)); }
  double u(Idx e, Idx f, Idx t) const { return m_str(1, Id(e, f, t)); }
  double v(Idx e, Idx f, Idx t) const { return m_str(2, Id(e, f, t)); }
  double h(Idx e, Idx f, Idx t) const { return w(e, f, t) - b(e, f, t); }
  Array<2> vel(Idx e, Idx f, Idx t) const { return {u(e, f, t), v(e, f, t)}; }
  Storage<3>& storage() { return m_str; }

 private:
  Storage<3> m_str;
  const Domain* m_b = nullptr;
};

// ------------------------------------------------- include/Fluxes.h (decl)
struct SpaceDisc;
namespace Wavespeeds {
// src/Fluxes.cpp:5-26; as template arguments these select the device estimate
inline Array<2> Rusanov(double ul, double hl, double ur, double hr) {
  double cl = std::sqrt(hl), cr = std::sqrt(hr);
  double a = std::fabs(ul) + cl, b = std::fabs(ur) + cr;
  double ap = (a < b) ? b : a;
  return {-ap, ap};
}
inline Array<2> Davis(double ul, double hl, double ur, double hr) {
  double cl = std::sqrt(hl), cr = std::sqrt(hr);
  double a = ul - cl, b = ur - cr, c = ul + cl, d = ur + cr;
  return {(b < a) ? b : a, (c < d) ? d : c};
}
inline Array<2> Einfeldt(double ul, double hl, double ur, double hr) {
  double cl = std::sqrt(hl), cr = std::sqrt(hr);
  double uRoe = (cl * ul + cl * ur) / (cl + cr);
  double cRoe = std::sqrt(0.5 * (hl + hr));
  double a = ul - cl, b = uRoe - cRoe, c = ur + cr, d = uRoe + cRoe;
  return {(b < a) ? b : a, (c < d) ? d : c};
}
}  // namespace Wavespeeds
using WavespeedFn = Array<2> (*)(double, double, double, double);
namespace swe_detail {
inline int wave_code(WavespeedFn w) {
  if (w == &Wavespeeds::Rusanov) return SWE_WAVE_RUSANOV;
  if (w == &Wavespeeds::Davis) return SWE_WAVE_DAVIS;
  if (w == &Wavespeeds::Einfeldt) return SWE_WAVE_EINFELDT;
  return -1;
}
}  // namespace swe_detail
namespace Fluxes {
template <WavespeedFn W>
Array<3> HLL(SpaceDisc* const sd, Idx e, Idx from, Idx to, double* r);
template <WavespeedFn W>
Array<3> HLLC(SpaceDisc* const sd, Idx e, Idx from, Idx to, double* r);
}  // namespace Fluxes

// ------------------------------------------------------- include/SpaceDisc.h
struct SpaceDisc {
  using Fluxer = std::function<Array<3>(SpaceDisc* const, Idx, Idx, Idx, double*)>;
  using FluxFn = Array<3> (*)(SpaceDisc* const, Idx, Idx, Idx, double*);

  SpaceDisc(const Fluxer& fluxer, const Domain& b, const VolumeField& v0, double cor = 0, double tau = 0,
            int options = 0, int device = 0)
      : m_b(b), m_vol(v0), m_cor(cor), m_tau(tau) {
    int flux = -1, ws = -1;
    if (const FluxFn* f = fluxer.target<FluxFn>()) identify(*f, flux, ws);
    if (flux < 0) throw SolverError("unsupported fluxer: use Fluxes::HLL<W> or Fluxes::HLLC<W>");
    swe_solver* s = nullptr;
    swe_detail::check(swe_create(b.GetTopology().handle(), b.GetGeometry().data(), flux, ws, cor, tau, options,
                                 device, &s));
    m_h = std::shared_ptr<swe_solver>(s, swe_destroy);
    swe_detail::check(swe_set_state(s, v0.storage().data()));
  }

  // VolumeField on the host: reading downloads, writing through the
  // mutable accessor is uploaded before the next device operation.
  const VolumeField& GetVolField() const { pull(); return m_vol; }
  VolumeField& GetVolField() { pull(); m_dirty = true; return m_vol; }
  const Domain& GetDomain() const noexcept { return m_b; }

  bool IsDryCell(Idx i) const { return cls(i) == 0; }
  bool IsFullWetCell(Idx i) const { return cls(i) == 2; }
  bool IsPartWetCell(Idx i) const { return cls(i) == 1; }

  const EdgeField& GetEdgField() const {
    push();
    m_edg = EdgeField(&m_b, 2 * m_b.GetTopology().NumEdges());
    swe_detail::check(swe_get_edge_field(m_h.get(), m_edg.storage().data()));
    return m_edg;
  }
  const EdgeField& GetSrcField() const {
    push();
    m_src = EdgeField(&m_b, 2 * m_b.GetTopology().NumEdges());
    swe_detail::check(swe_get_src_field(m_h.get(), m_src.storage().data()));
    return m_src;
  }
  const Storage<3>& GetFluxes() const {
    push();
    m_f.resize(m_b.GetTopology().NumEdges());
    swe_detail::check(swe_get_fluxes(m_h.get(), m_f.data()));
    return m_f;
  }
  double GetTau() const noexcept { return m_tau; }
  double GetCor() const noexcept { return m_cor; }
  double GetMinLenToWavespeed() const {
    double r = 0.;
    swe_detail::check(swe_min_len_to_wavespeed(m_h.get(), &r));
    return r;
  }

  void ComputeInterfaceValues() {
    push();
    swe_detail::check(swe_compute_interface_values(m_h.get()));
    m_cls.clear();
  }
  void ComputeFluxes() {
    push();
    swe_detail::check(swe_compute_fluxes(m_h.get(), nullptr));
  }

  // nsteps steps on the device (dt > 0 fixed, dt <= 0 CFL mode)
  void Run(int scheme, double dt, Idx nsteps, double cfl = 0.15) {
    push();
    swe_detail::check(swe_run(m_h.get(), scheme, dt, cfl, nsteps));
    swe_detail::check(swe_synchronize(m_h.get()));
    m_stale = true;
    m_cls.clear();
  }
  swe_solver* handle() const { return m_h.get(); }

 private:
  template <WavespeedFn W>
  static bool is_flux(FluxFn f, int& flux, int& ws) {
    if (f == &Fluxes::HLL<W>) { flux = SWE_FLUX_HLL; ws = swe_detail::wave_code(W); return true; }
    if (f == &Fluxes::HLLC<W>) { flux = SWE_FLUX_HLLC; ws = swe_detail::wave_code(W); return true; }
    return false;
  }
  static void identify(FluxFn f, int& flux, int& ws) {
    is_flux<&Wavespeeds::Rusanov>(f, flux, ws) || is_flux<&Wavespeeds::Davis>(f, flux, ws) ||
        is_flux<&Wavespeeds::Einfeldt>(f, flux, ws);
  }
  int cls(Idx i) const {
    if (m_cls.empty()) {
      push();
      m_cls.resize(size_t(m_b.GetTopology().NumTriangles()));
      swe_detail::check(swe_compute_interface_values(m_h.get()));
      swe_detail::check(swe_get_cell_class(m_h.get(), m_cls.data()));
    }
    return m_cls[size_t(i)];
  }
  void pull() const {
    push();
    if (m_stale) {
      swe_detail::check(swe_get_state(m_h.get(), m_vol.storage().data()));
      m_stale = false;
    }
  }
  void push() const {
    if (m_dirty) {
      swe_detail::check(swe_set_state(m_h.get(), m_vol.storage().data()));
      m_dirty = false;
      m_cls.clear();
    }
  }

  const Domain& m_b;
  mutable VolumeField m_vol;
  mutable EdgeField m_edg, m_src;
  mutable Storage<3> m_f;
  mutable std::vector<int8_t> m_cls;
  mutable bool m_dirty = false, m_stale = false;
  std::shared_ptr<swe_solver> m_h;
  double m_cor, m_tau;
};

// ------------------------------------------------------- include/Fluxes.h
namespace Fluxes {
// Flux of edge e from the last ComputeFluxes (the device evaluates every
// edge; from/to follow the edge's orientation et0 -> et1).
template <WavespeedFn W>
Array<3> HLL(SpaceDisc* const sd, Idx e, Idx, Idx, double* r) {
  if (r) *r = sd->GetMinLenToWavespeed();
  return sd->GetFluxes().col(e);
}
template <WavespeedFn W>
Array<3> HLLC(SpaceDisc* const sd, Idx e, Idx, Idx, double* r) {
  if (r) *r = sd->GetMinLenToWavespeed();
  return sd->GetFluxes().col(e);
}
}  // namespace Fluxes

// -------------------------------------------------------- include/TimeDisc.h
struct TimeDisc {
  explicit TimeDisc(SpaceDisc* sd = nullptr) : m_sd(sd) {}
  SpaceDisc* GetSpaceDisc() { return m_sd; }
  const SpaceDisc* GetSpaceDisc() const { return m_sd; }
  void SetSpaceDisc(SpaceDisc* sd) { m_sd = sd; }
  double CFLdt() const { return m_constCFL * m_sd->GetMinLenToWavespeed(); }
  Array<3> RHS(Idx i, double dt) const {
    Array<3> out;
    swe_detail::check(swe_rhs(m_sd->handle(), i, dt, out.v));
    return out;
  }

 protected:
  SpaceDisc* m_sd;
  const double m_constCFL = 0.15;  // include/TimeDisc.h:22
};

// --------------------------------------------------------- include/Solvers.h
namespace Solvers {
inline void Euler(TimeDisc* const td, double dt) { td->GetSpaceDisc()->Run(SWE_SCHEME_EULER, dt, 1); }
inline void SSPRK2(TimeDisc* const td, double dt) { td->GetSpaceDisc()->Run(SWE_SCHEME_SSPRK2, dt, 1); }
inline void SSPRK3(TimeDisc* const td, double dt) { td->GetSpaceDisc()->Run(SWE_SCHEME_SSPRK3, dt, 1); }
}  // namespace Solvers
