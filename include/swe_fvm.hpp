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
  Array operator/(double s) const { Array r; for (size_t i = 0; i < k; ++i) r.v[i] = v[i] / s; return r; }
  Array operator-() const { Array r; for (size_t i = 0; i < k; ++i) r.v[i] = -v[i]; return r; }
  Array& operator+=(const Array& o) { for (size_t i = 0; i < k; ++i) v[i] += o.v[i]; return *this; }
  Array& operator-=(const Array& o) { for (size_t i = 0; i < k; ++i) v[i] -= o.v[i]; return *this; }
  Array& operator*=(double s) { for (size_t i = 0; i < k; ++i) v[i] *= s; return *this; }
  friend Array operator*(double s, const Array& a) { return a * s; }
  // dot product summed in index order (Eigen's order for fixed sizes <= 3)
  double dot(const Array& o) const { double d = v[0] * o.v[0]; for (size_t i = 1; i < k; ++i) d += v[i] * o.v[i]; return d; }
  static constexpr size_t size() { return k; }
};

// 2-vectors of the geometry getters (Eigen::Vector2d in the reference)
using Vector2d = Array<2>;

// 3 x 2 gradient of a reconstruction (Eigen::Matrix32d): rows w, u, v
struct Matrix32d {
  double m[3][2] = {{0., 0.}, {0., 0.}, {0., 0.}};
  static Matrix32d Zero() { return Matrix32d(); }
  double& operator()(int r, int c) { return m[r][c]; }
  double operator()(int r, int c) const { return m[r][c]; }
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
  // a Topology owning the tables of a handle from the C ABI
  static Topology adopt_handle(swe_topology* t) { return Topology(t); }
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
  // Boundary edges pairs(k, 0) and pairs(k, 1) become PERIODIC partners
  // (swe_topology_set_periodic; the reference declares the tag only,
  // include/Includes.h:21).  Copies of this Topology share its tables and
  // see the pairing; solvers built before keep the tables they copied.
  void SetPeriodic(const EdgeTagArray& pairs) {
    swe_detail::check(swe_topology_set_periodic(h_.get(), pairs.rows(), pairs.data()));
    swe_detail::check(swe_topology_copy(h_.get(), ep_.data(), et_.data(), tp_.data(), te_.data(), tt_.data()));
  }
  // the partner of each edge (-1: none)
  std::vector<Idx> PeriodicPartners() const {
    std::vector<Idx> p(static_cast<size_t>(NumEdges()));
    swe_detail::check(swe_topology_periodic(h_.get(), p.data()));
    return p;
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

// ------------------------------------------------- include/PointOperations.h
// src/PointOperations.cpp:4-24, 42-48
inline double Len(const Point& a, const Point& b) noexcept {
  return std::sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]));
}
inline double Det(const Point& a, const Point& b) noexcept { return a[0] * b[1] - b[0] * a[1]; }
inline double TriangArea(const Point& a, const Point& b, const Point& c) noexcept {
  return 0.5 * std::fabs(Det(b - a, c - a));
}
// intersection of the lines a1 a2 and b1 b2; SolverError when the
// determinant of their directions is below tol in magnitude (the reference
// tests den < tol, which also refuses every pair crossing the other way
// round, half of a mesh's interior edges in Domain::C)
inline Point Intersection(const Point& a1, const Point& a2, const Point& b1, const Point& b2) {
  const double den = Det(a2 - a1, b2 - b1);
  if (std::fabs(den) < tol) throw SolverError("Intersection: the lines are (nearly) parallel");
  return a1 + (a2 - a1) * Det(b1 - a1, b2 - b1) / den;
}
// plane through three points (x, y, value): the 2 x 2 system of the
// differences to the first point, solved by partial-pivot LU
inline Vector2d Gradient(const Point& p0, const Point& p1, const Point& p2) {
  double a00 = p1[0] - p0[0], a01 = p1[1] - p0[1], b0 = p1[2] - p0[2];
  double a10 = p2[0] - p0[0], a11 = p2[1] - p0[1], b1 = p2[2] - p0[2];
  if (std::fabs(a10) > std::fabs(a00)) {
    std::swap(a00, a10); std::swap(a01, a11); std::swap(b0, b1);
  }
  const double l = a10 / a00;
  const double u11 = a11 - l * a01;
  const double y1 = b1 - l * b0;
  const double x1 = y1 / u11;
  const double x0 = (b0 - a01 * x1) / a00;
  return {x0, x1};
}

// ------------------------------------------------------ include/Bathymetry.h
constexpr inline bool IsWet(double h) noexcept { return h > 1e-12; }
inline Point DryState(double b) noexcept { return {b, 0., 0.}; }

// Node geometry (x, y, bathymetry z) on a topology (owning copy).
// ----------------------------------------------- include/StructTriangMesh.h
// StructTriangMesh(ni, nj, h) (include/StructTriangMesh.h:4-15): ni x nj
// squares of side h from (x0, y0), four CCW triangles per square around a
// centre node (swe_mesh_struct).  The reference's class derives from a
// TriangMesh holding the nodes, which HEAD no longer has; here the mesh owns
// its node geometry (z = 0, to be set by the caller) and its Topology, and a
// Domain is built from it.
struct StructTriangMesh {
  StructTriangMesh(size_t ni, size_t nj, double h, double x0 = 0., double y0 = 0.) : m_ni(ni), m_nj(nj) {
    swe_topology* t = nullptr;
    double* xyz = nullptr;
    int64_t nn = 0;
    swe_detail::check(swe_mesh_struct(int64_t(ni), int64_t(nj), h, x0, y0, &xyz, &nn, &t));
    m_geometry.resize(nn);
    std::memcpy(m_geometry.data(), xyz, sizeof(double) * size_t(3 * nn));
    swe_free(xyz);
    m_topology = std::make_shared<Topology>(Topology::adopt_handle(t));
  }
  size_t Ni() const { return m_ni; }
  size_t Nj() const { return m_nj; }
  PointArray& Geometry() { return m_geometry; }
  const PointArray& Geometry() const { return m_geometry; }
  Topology& GetTopology() { return *m_topology; }
  const Topology& GetTopology() const { return *m_topology; }

 private:
  size_t m_ni, m_nj;
  PointArray m_geometry;
  std::shared_ptr<Topology> m_topology;
};

struct Domain {
  Domain(const PointArray& geometry, const Topology& topology) : m_geometry(geometry), m_topology(topology) {
    if (geometry.cols() != topology.NumNodes()) throw DomainError("geometry and topology disagree");
  }
  // Domain{StructTriangMesh{n, n, l / n}} as the reference's drivers write it
  explicit Domain(const StructTriangMesh& m) : Domain(m.Geometry(), m.GetTopology()) {}
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
  double MinX() const { return extreme(0, true); }
  double MaxX() const { return extreme(0, false); }
  double MinY() const { return extreme(1, true); }
  double MaxY() const { return extreme(1, false); }
  // Domain::Tang (src/Bathymetry.cpp:69-76): unit edge direction, turned
  // so that cell t lies on its right (Norm then points out of t)
  Vector2d Tang(NodeTag e, NodeTag t) const {
    auto ep = m_topology.EdgePoints(e);
    const Point a = P(ep[0]), b = P(ep[1]);
    const double l = L(e);
    double tx = (b[0] - a[0]) / l, ty = (b[1] - a[1]) / l;
    const Point c = T(t);
    if ((c[0] - a[0]) * ty - tx * (c[1] - a[1]) > 0.) { tx = -tx; ty = -ty; }
    return {tx, ty};
  }
  // Domain::Norm (src/Bathymetry.cpp:78-80): (t_y, -t_x), outward of t
  Vector2d Norm(NodeTag e, NodeTag t) const {
    const Vector2d tn = Tang(e, t);
    return {tn[1], -tn[0]};
  }
  // Domain::TriangSlope (src/Bathymetry.cpp:10-13): bathymetry gradient
  Vector2d TriangSlope(NodeTag t) const {
    auto tp = m_topology.TriangPoints(t);
    return Gradient(P(tp[0]), P(tp[1]), P(tp[2]));
  }
  // Domain::Z (src/Bathymetry.cpp:15-18): slope . ((x, y) - first vertex)
  double Z(NodeTag t, double x, double y) const {
    auto tp = m_topology.TriangPoints(t);
    const Point p0 = P(tp[0]);
    return TriangSlope(t).dot(Vector2d{x - p0[0], y - p0[1]});
  }
  // Domain::C (src/Bathymetry.cpp:33-40): the edge midpoint on a boundary
  // edge, else where the line between the two centroids crosses the edge
  Point C(NodeTag e) const {
    if (m_topology.IsEdgeBoundary(e)) return E(e);
    auto ep = m_topology.EdgePoints(e);
    auto et = m_topology.EdgeTriangs(e);
    return Intersection(P(ep[0]), P(ep[1]), T(et[0]), T(et[1]));
  }

 private:
  double extreme(int row, bool lo) const {
    double r = m_geometry(row, 0);
    for (Idx i = 1; i < m_geometry.cols(); ++i) {
      const double x = m_geometry(row, i);
      if (lo ? (x < r) : (r < x)) r = x;
    }
    return r;
  }
  PointArray m_geometry;
  Topology m_topology;
};

// ------------------------------------------------------- include/Assigners.h
// Writes through a value field in primitive or conservative form with the
// reference's dry clamp and small-depth velocity desingularisation
// (include/Assigners.h:4-50, src/Assigners.cpp:8-44).  b is the
// bathymetry under the written column.
template <class Derived, size_t k>
struct BaseAssigner {
  BaseAssigner(Storage<k>* field, double b, Idx col) : m_field(field), m_b(b), m_col(col) {}
  void operator+=(const Array<k>& rhs) { self() = self().Get() + rhs; }
  void operator-=(const Array<k>& rhs) { self() = self().Get() - rhs; }
  void operator*=(double scalar) { self() = self().Get() * scalar; }
  void operator/=(double scalar) {
    if (scalar == 0.) throw std::runtime_error("value field write divided by zero");
    self() *= (1. / scalar);
  }
  operator Array<k>() const { return static_cast<const Derived&>(*this).Get(); }

 protected:
  Storage<k>* const m_field;
  double m_b;
  Idx m_col;

 private:
  Derived& self() { return static_cast<Derived&>(*this); }
};

struct PrimAssigner : BaseAssigner<PrimAssigner, 3> {
  using BaseAssigner<PrimAssigner, 3>::BaseAssigner;
  using BaseAssigner<PrimAssigner, 3>::operator+=;
  Array<3> Get() const { return m_field->col(m_col); }
  void operator=(const Array<3>& rhs) {
    const double h = rhs[0] - m_b;
    if (!IsWet(h)) { m_field->set_col(m_col, DryState(m_b)); return; }
    Array<3> c = rhs;
    if (h < 1e-3) {
      const double f = std::sqrt(2.) * h / std::sqrt(h * h + 1e-6);
      c[1] *= f;
      c[2] *= f;
    }
    m_field->set_col(m_col, c);
  }
};

struct ConsAssigner : BaseAssigner<ConsAssigner, 3> {
  using BaseAssigner<ConsAssigner, 3>::BaseAssigner;
  Array<3> Get() const {
    Array<3> c = m_field->col(m_col);
    const double h = c[0] - m_b;
    return {h, c[1] * h, c[2] * h};
  }
  void operator=(const Array<3>& rhs) {
    const double h = rhs[0];
    if (!IsWet(h)) { m_field->set_col(m_col, DryState(m_b)); return; }
    double ih;
    if (h < 1e-3) ih = std::sqrt(2.) * h / std::sqrt(h * h * h * h + 1e-12);
    else ih = 1. / h;
    m_field->set_col(m_col, Array<3>{h + m_b, rhs[1] * ih, rhs[2] * ih});
  }
};

// ------------------------------------------------------ include/ValueField.h
// Primitive (w, u, v) per cell; b = centroid bathymetry.  The const
// accessors read; prim(i) / cons(i) on a mutable field return assigners, so
// `field.cons(i) = U` and `field.prim(i) += dW` write as in the reference.
struct VolumeField {
  VolumeField(const Domain& b, size_t size = 0) : m_str(Idx(size)), m_b(&b) {}
  size_t Size() const { return size_t(m_str.cols()); }
  void Resize(size_t size) { m_str.resize(Idx(size)); }
  double b(Idx i) const { return m_b->T(i)[2]; }
  Array<3> prim(Idx i) const { return m_str.col(i); }
  Array<3> cons(Idx i) const { return {h(i), hu(i), hv(i)}; }
  PrimAssigner prim(Idx i) { return PrimAssigner(&m_str, b(i), i); }
  ConsAssigner cons(Idx i) { return ConsAssigner(&m_str, b(i), i); }
  void set_prim(Idx i, const Array<3>& p) { m_str.set_col(i, p); }
  double w(Idx i) const { return m_str(0, i); }
  double u(Idx i) const { return m_str(1, i); }
  double v(Idx i) const { return m_str(2, i); }
  double& u(Idx i) { return m_str(1, i); }
  double& v(Idx i) { return m_str(2, i); }
  Array<2> vel(Idx i) const { return {u(i), v(i)}; }
  double h(Idx i) const { return w(i) - b(i); }
  double hu(Idx i) const { return h(i) * u(i); }
  double hv(Idx i) const { return h(i) * v(i); }
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
  Array<3> cons(Idx e, Idx f, Idx t) const { return {h(e, f, t), hu(e, f, t), hv(e, f, t)}; }
  PrimAssigner prim(Idx e, Idx f, Idx t) { return PrimAssigner(&m_str, b(e, f, t), Id(e, f, t)); }
  ConsAssigner cons(Idx e, Idx f, Idx t) { return ConsAssigner(&m_str, b(e, f, t), Id(e, f, t)); }
  double w(Idx e, Idx f, Idx t) const { return m_str(0, Id(e, f, t)); }
  double u(Idx e, Idx f, Idx t) const { return m_str(1, Id(e, f, t)); }
  double v(Idx e, Idx f, Idx t) const { return m_str(2, Id(e, f, t)); }
  double h(Idx e, Idx f, Idx t) const { return w(e, f, t) - b(e, f, t); }
  double hu(Idx e, Idx f, Idx t) const { return h(e, f, t) * u(e, f, t); }
  double hv(Idx e, Idx f, Idx t) const { return h(e, f, t) * v(e, f, t); }
  Array<2> vel(Idx e, Idx f, Idx t) const { return {u(e, f, t), v(e, f, t)}; }
  Storage<3>& storage() { return m_str; }

 private:
  Storage<3> m_str;
  const Domain* m_b = nullptr;
};

// -------------------------------------------------------- ElemFlux (SpaceDisc.h)
// physical flux of U = (h, hu, hv) through normal n (include/SpaceDisc.h:4-18)
inline Array<3> ElemFlux(const Vector2d& n, const Array<3>& U) {
  Array<3> res;
  const double h = U[0];
  if (IsWet(h)) {
    const double hveln = U[1] * n[0] + U[2] * n[1];
    res[0] = hveln;
    res[1] = (hveln / h) * U[1] + (0.5 * h * h) * n[0];
    res[2] = (hveln / h) * U[2] + (0.5 * h * h) * n[1];
  }
  return res;
}

// ------------------------------------------------------ include/MUSCLObject.h
struct MUSCLObject {
  // A cell's reconstruction (include/MUSCLObject.h:15-55): value at the
  // centroid and gradient; AtPoint clamps to the dry state below ground.
  struct MUSCL {
    MUSCL(const Domain& b, const Array<3>& at_origin, const Matrix32d& grad, Idx i)
        : m_b(&b), m_at_origin(at_origin), m_grad(grad), m_i(i) {}
    Idx OriginId() const noexcept { return m_i; }
    Array<3> AtOrigin() const noexcept { return m_at_origin; }
    Array<3> AtPoint(const Point& p) const {
      const Point c = m_b->T(m_i);
      const double dx = p[0] - c[0], dy = p[1] - c[1];
      Array<3> a;
      for (int r = 0; r < 3; ++r) a[size_t(r)] = m_at_origin[size_t(r)] + (m_grad(r, 0) * dx + m_grad(r, 1) * dy);
      if ((a[0] - p[2]) >= 0) return a;
      return DryState(p[2]);
    }
    Matrix32d dryGradient() const {
      Matrix32d g;
      const Vector2d s = m_b->TriangSlope(m_i);
      g(0, 0) = s[0];
      g(0, 1) = s[1];
      return g;
    }
    Matrix32d Gradient(const Point& p) const {
      if (AtPoint(p)[0] - p[2] >= 0) return m_grad;
      return dryGradient();
    }

   private:
    const Domain* m_b;
    Array<3> m_at_origin;
    Matrix32d m_grad;
    Idx m_i;
  };
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
struct SpaceDisc : MUSCLObject {
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

  // VolumeField on the host: reading downloads what the device changed;
  // writing through the mutable accessor (field.cons(i) = ..., prim(i) =
  // ..., u(i) = ...) is uploaded before the next ComputeInterfaceValues,
  // ComputeFluxes, Run or field getter.  TimeDisc::RHS does not upload: it
  // evaluates on the state of the last interface values, so a loop that
  // writes cells one by one between RHS calls (src/Solvers.cpp:5-59) gets
  // every draining time step from the stage-start state (the Jacobi
  // reading of src/TimeDisc.cpp:43-66).
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

  // MUSCLObject::Reconstruct* (src/MUSCLObject.cpp:31-191) of cell i of the
  // current state, evaluated on the device by the step's own code (the
  // interface values are recomputed first)
  MUSCL ReconstructDryCell(Idx i) const { return reconstruct(i, 0); }
  MUSCL ReconstructPartWetCell1(Idx i) const { return reconstruct(i, 1); }
  MUSCL ReconstructFullWetCell(Idx i) const { return reconstruct(i, 2); }
  MUSCL ReconstructPartWetCell2(Idx i) const { return reconstruct(i, 3); }

 private:
  MUSCL reconstruct(Idx i, int kind) const {
    push();
    double o[5];
    swe_detail::check(swe_reconstruct(m_h.get(), i, kind, o));
    Matrix32d g;
    g(0, 0) = o[3];
    g(0, 1) = o[4];
    return MUSCL(m_b, Array<3>{o[0], o[1], o[2]}, g, i);
  }
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
// Flux of edge e seen from cell `from` towards `to` with the edge slots of
// the last ComputeInterfaceValues, evaluated on the device by the step's
// own flux code; *r (if given) takes the min with this edge's
// length-to-wavespeed candidate (include/Fluxes.h:14-111).  ComputeFluxes
// evaluates every edge on the device itself.
template <WavespeedFn W>
Array<3> HLL(SpaceDisc* const sd, Idx e, Idx from, Idx to, double* r) {
  Array<3> out;
  swe_detail::check(swe_edge_flux(sd->handle(), SWE_FLUX_HLL, swe_detail::wave_code(W), e, from, to, r, out.v));
  return out;
}
template <WavespeedFn W>
Array<3> HLLC(SpaceDisc* const sd, Idx e, Idx from, Idx to, double* r) {
  Array<3> out;
  swe_detail::check(swe_edge_flux(sd->handle(), SWE_FLUX_HLLC, swe_detail::wave_code(W), e, from, to, r, out.v));
  return out;
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
