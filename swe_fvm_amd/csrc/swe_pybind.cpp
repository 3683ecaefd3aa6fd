// Python module SWE_FVM: the reference's pybind binding of Topology
// (pybind/Topology.cpp:8-30) with the same constructor, Topology.create and
// num_nodes / num_edges / num_elements, over the C ABI's owning topology
// (the reference's Topology keeps references to the caller's arrays, which
// dangle once the temporaries of the call are gone).  Also the mesh sources
// of the C ABI (Gmsh reader, StructTriangMesh, seeded jittered Delaunay), the
// tables back as numpy arrays, and Solver: the device solver with its
// adjoint and tangent-linear runs, on which the torch.autograd.Function
// swe_fvm_amd.adjoint.ShallowWaterRun works.  Host code only.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>

#include <memory>
#include <stdexcept>
#include <string>

#include "swe_fvm.h"

namespace py = pybind11;
using namespace pybind11::literals;

namespace {

using I64 = py::array_t<int64_t, py::array::c_style>;

void check(int rc) {
  if (rc == SWE_OK) return;
  std::string msg = swe_last_error();
  if (rc == SWE_ERR_ARG || rc == SWE_ERR_DOMAIN) throw py::value_error(msg);
  throw std::runtime_error(msg);
}

const int64_t* rows(const I64& a, int64_t cols, const char* name, int64_t* n) {
  if (a.ndim() != 2 || a.shape(1) != cols)
    throw py::value_error(std::string(name) + " must be an N x " + std::to_string(cols) + " int64 array");
  *n = a.shape(0);
  return a.data();
}

struct Topology {
  std::shared_ptr<swe_topology> h;

  explicit Topology(swe_topology* t) : h(t, swe_topology_destroy) {}

  // Topology(numNodes, edgeNodes, edgeElements, elementNodes, elementEdges,
  //          elementNeighbours) (include/TriangMesh.h:16-40)
  Topology(int64_t num_nodes, const I64& edge_nodes, const I64& edge_elements, const I64& element_nodes,
           const I64& element_edges, const I64& element_neighbours) {
    int64_t E, E2, T, T2, T3;
    const int64_t* ep = rows(edge_nodes, 2, "edgeNodes", &E);
    const int64_t* et = rows(edge_elements, 2, "edgeElements", &E2);
    const int64_t* tp = rows(element_nodes, 3, "elementNodes", &T);
    const int64_t* te = rows(element_edges, 3, "elementEdges", &T2);
    const int64_t* tt = rows(element_neighbours, 3, "elementNeighbours", &T3);
    if (E2 != E || T2 != T || T3 != T) throw py::value_error("table sizes disagree");
    swe_topology* t = nullptr;
    check(swe_topology_create(num_nodes, E, ep, et, T, tp, te, tt, &t));
    h.reset(t, swe_topology_destroy);
  }

  int64_t num_nodes() const { return swe_topology_num_nodes(h.get()); }
  int64_t num_edges() const { return swe_topology_num_edges(h.get()); }
  int64_t num_elements() const { return swe_topology_num_elements(h.get()); }

  py::dict tables() const {
    const py::ssize_t E = num_edges(), T = num_elements();
    I64 ep({E, py::ssize_t(2)}), et({E, py::ssize_t(2)});
    I64 tp({T, py::ssize_t(3)}), te({T, py::ssize_t(3)}), tt({T, py::ssize_t(3)});
    check(swe_topology_copy(h.get(), ep.mutable_data(), et.mutable_data(), tp.mutable_data(),
                            te.mutable_data(), tt.mutable_data()));
    return py::dict("edge_nodes"_a = ep, "edge_elements"_a = et, "element_nodes"_a = tp,
                    "element_edges"_a = te, "element_neighbours"_a = tt);
  }
};

using F64 = py::array_t<double, py::array::c_style | py::array::forcecast>;

const double* sized(const F64& a, int64_t n, const char* name) {
  if (a.size() != n) throw py::value_error(std::string(name) + " must hold " + std::to_string(n) + " values");
  return a.data();
}

int scheme_code(const std::string& s) {
  if (s == "Euler") return SWE_SCHEME_EULER;
  if (s == "SSPRK2") return SWE_SCHEME_SSPRK2;
  if (s == "SSPRK3") return SWE_SCHEME_SSPRK3;
  throw py::value_error("scheme must be Euler, SSPRK2 or SSPRK3");
}

// SpaceDisc (include/SpaceDisc.h:24) on the device, for differentiation
// from Python: the method names are those adjoint.ShallowWaterRun (a
// torch.autograd.Function) calls, so the Function runs on this object as on
// swe_fvm_amd.SpaceDisc.  Arrays are float64 numpy arrays on the host; the
// GIL is released while the device works.
struct Solver {
  std::shared_ptr<swe_solver> h;
  Topology topo;  // the solver keeps its own tables; held for the caller's convenience
  int64_t T = 0, E = 0, V = 0;

  Solver(const Topology& t, const F64& xyz, const std::string& flux, const std::string& wavespeed, double cor,
         double tau, int options, int device)
      : topo(t) {
    const int f = flux == "HLL" ? SWE_FLUX_HLL : flux == "HLLC" ? SWE_FLUX_HLLC : -1;
    const int w = wavespeed == "Rusanov" ? SWE_WAVE_RUSANOV
                  : wavespeed == "Davis" ? SWE_WAVE_DAVIS
                  : wavespeed == "Einfeldt" ? SWE_WAVE_EINFELDT : -1;
    if (f < 0) throw py::value_error("flux must be HLL or HLLC");
    if (w < 0) throw py::value_error("wavespeed must be Rusanov, Davis or Einfeldt");
    const double* g = sized(xyz, 3 * t.num_nodes(), "xyz (V x 3)");
    swe_solver* s = nullptr;
    check(swe_create(t.h.get(), g, f, w, cor, tau, options, device, &s));
    h.reset(s, swe_destroy);
    T = swe_num_cells(s);
    E = swe_num_edges(s);
    V = swe_num_nodes(s);
  }

  F64 cells3() const { return F64({py::ssize_t(T), py::ssize_t(3)}); }

  void set_state(const F64& prim) { check(swe_set_state(h.get(), sized(prim, 3 * T, "prim (T x 3)"))); }
  F64 state() const {
    F64 out = cells3();
    check(swe_get_state(h.get(), out.mutable_data()));
    return out;
  }
  void set_bathymetry(const F64& z) { check(swe_set_bathymetry(h.get(), sized(z, V, "z (V)"))); }
  void run(const std::string& scheme, int64_t nsteps, double dt, double cfl) {
    const int sc = scheme_code(scheme);
    int rc;
    {
      py::gil_scoped_release nogil;
      rc = swe_run(h.get(), sc, dt, cfl, nsteps);
      if (rc == SWE_OK) rc = swe_synchronize(h.get());
    }
    check(rc);
  }
  // step sizes of the last run (swe_get_dts)
  F64 dts(int64_t nsteps) const {
    F64 out({py::ssize_t(nsteps)});
    check(swe_get_dts(h.get(), out.mutable_data(), nsteps));
    return out;
  }
  const double* steps(const py::object& dts, int64_t nsteps, F64& keep) const {
    if (dts.is_none()) return nullptr;
    keep = dts.cast<F64>();
    return sized(keep, nsteps, "dts (nsteps)");
  }
  py::tuple adjoint_run(const F64& final_bar, int64_t nsteps, const std::string& scheme, const py::object& dt,
                        const py::object& dts, int64_t checkpoint_every, bool want_z) {
    const int sc = scheme_code(scheme);
    const double* fb = sized(final_bar, 3 * T, "final_bar (T x 3)");
    F64 keep, s0 = cells3(), ms(2);
    const double* d = steps(dts, nsteps, keep);
    const double step = dt.is_none() ? 0.0 : dt.cast<double>();
    py::object zb = py::none();
    F64 z;
    if (want_z) z = F64(V);
    double* zp = want_z ? z.mutable_data() : nullptr;
    double* sp = s0.mutable_data();
    double* mp = ms.mutable_data();
    int rc;
    {
      py::gil_scoped_release nogil;
      rc = swe_adjoint_run(h.get(), sc, step, d, nsteps, checkpoint_every, fb, sp, zp, mp);
    }
    check(rc);
    if (want_z) zb = z;
    return py::make_tuple(s0, zb, ms);
  }
  py::tuple tangent_run(const F64& state0_dot, int64_t nsteps, const py::object& z_dot, const std::string& scheme,
                        const py::object& dt, const py::object& dts) {
    const int sc = scheme_code(scheme);
    const double* sd = sized(state0_dot, 3 * T, "state0_dot (T x 3)");
    F64 keep, zkeep, out = cells3();
    const double* d = steps(dts, nsteps, keep);
    const double* zd = nullptr;
    if (!z_dot.is_none()) {
      zkeep = z_dot.cast<F64>();
      zd = sized(zkeep, V, "z_dot (V)");
    }
    const double step = dt.is_none() ? 0.0 : dt.cast<double>();
    double ms = 0.0;
    double* op = out.mutable_data();
    int rc;
    {
      py::gil_scoped_release nogil;
      rc = swe_tangent_run(h.get(), sc, step, d, nsteps, sd, zd, op, &ms);
    }
    check(rc);
    return py::make_tuple(out, ms);
  }
};

// (xyz V x 3, Topology) from a C-ABI mesh source
py::tuple mesh_result(int rc, double* xyz, int64_t nn, swe_topology* t) {
  check(rc);
  py::array_t<double, py::array::c_style> g({py::ssize_t(nn), py::ssize_t(3)});
  std::copy(xyz, xyz + 3 * nn, g.mutable_data());
  swe_free(xyz);
  return py::make_tuple(g, Topology(t));
}

}  // namespace

PYBIND11_MODULE(SWE_FVM, m) {
  m.doc() = "SWE_FVM mesh topology (owning tables), mesh sources and the device solver with its adjoint";
  py::class_<Topology>(m, "Topology")
      .def(py::init<int64_t, const I64&, const I64&, const I64&, const I64&, const I64&>(), "numNodes"_a,
           "edgeNodes"_a.noconvert(), "edgeElements"_a.noconvert(), "elementNodes"_a.noconvert(),
           "elementEdges"_a.noconvert(), "elementNeighbours"_a.noconvert())
      .def_static(
          "create",
          [](int64_t n, const I64& a, const I64& b, const I64& c, const I64& d, const I64& e) {
            return Topology(n, a, b, c, d, e);
          },
          "numNodes"_a.noconvert(), "edgeNodes"_a.noconvert(), "edgeElements"_a.noconvert(),
          "elementNodes"_a.noconvert(), "elementEdges"_a.noconvert(), "elementNeighbours"_a.noconvert())
      .def("num_nodes", &Topology::num_nodes)
      .def("num_edges", &Topology::num_edges)
      .def("num_elements", &Topology::num_elements)
      .def("tables", &Topology::tables, "edge and element tables as int64 arrays")
      .def_static(
          "build",
          [](int64_t num_nodes, const I64& tri, const I64& bnd) {
            int64_t T, B;
            const int64_t* t = rows(tri, 3, "triangles", &T);
            const int64_t* b = rows(bnd, 2, "boundary lines", &B);
            swe_topology* out = nullptr;
            check(swe_topology_build(num_nodes, T, t, B, b, &out));
            return Topology(out);
          },
          "numNodes"_a, "triangles"_a, "boundaryLines"_a,
          "tables of the reference's mesh reader from CCW triangles and boundary lines");
  m.def(
      "read_msh",
      [](const std::string& path) {
        double* xyz = nullptr;
        int64_t nn = 0;
        swe_topology* t = nullptr;
        int rc = swe_topology_read_msh(path.c_str(), &xyz, &nn, &t);
        return mesh_result(rc, xyz, nn, t);
      },
      "path"_a, "Gmsh 4.1 ASCII mesh -> (xyz, Topology)");
  m.def(
      "struct_mesh",
      [](int64_t ni, int64_t nj, double h, double x0, double y0) {
        double* xyz = nullptr;
        int64_t nn = 0;
        swe_topology* t = nullptr;
        int rc = swe_mesh_struct(ni, nj, h, x0, y0, &xyz, &nn, &t);
        return mesh_result(rc, xyz, nn, t);
      },
      "ni"_a, "nj"_a, "h"_a, "x0"_a = 0.0, "y0"_a = 0.0, "StructTriangMesh(ni, nj, h) -> (xyz, Topology)");
  m.def(
      "jittered_mesh",
      [](int64_t n, double jitter, uint64_t seed, double x0, double y0, double L) {
        double* xyz = nullptr;
        int64_t nn = 0;
        swe_topology* t = nullptr;
        int rc = swe_mesh_jittered(n, jitter, seed, x0, y0, L, &xyz, &nn, &t);
        return mesh_result(rc, xyz, nn, t);
      },
      "n"_a, "jitter"_a = 0.25, "seed"_a = 1, "x0"_a = 0.0, "y0"_a = 0.0, "L"_a = 1.0,
      "seeded Delaunay of a jittered grid, Morton order -> (xyz, Topology)");
  py::class_<Solver>(m, "Solver", py::dynamic_attr(),
                     "SpaceDisc on a device with run, adjoint_run and tangent_run; the object "
                     "swe_fvm_amd.adjoint.ShallowWaterRun (torch.autograd.Function) works on")
      .def(py::init<const Topology&, const F64&, const std::string&, const std::string&, double, double, int, int>(),
           "topology"_a, "xyz"_a, "flux"_a = "HLLC", "wavespeed"_a = "Einfeldt", "cor"_a = 0.0, "tau"_a = 0.0,
           "options"_a = 0, "device"_a = 0,
           "SpaceDisc(Fluxes::<flux><Wavespeeds::<wavespeed>>, Domain(xyz, topology), cor, tau) on a device")
      .def_readonly("T", &Solver::T)
      .def_readonly("E", &Solver::E)
      .def_readonly("V", &Solver::V)
      .def("SetVolField", &Solver::set_state, "prim"_a, "primitive (w, u, v) per cell, T x 3")
      .def("GetVolField", &Solver::state)
      .def("set_bathymetry", &Solver::set_bathymetry, "z"_a, "node bathymetry (V); the state is kept")
      .def("run", &Solver::run, "scheme"_a = "SSPRK2", "nsteps"_a = 1, "dt"_a = 1e-3, "cfl"_a = 0.15,
           "nsteps steps; dt > 0 fixed, dt <= 0 CFL mode")
      .def("dts", &Solver::dts, "nsteps"_a, "the step sizes of the last run (swe_get_dts)")
      .def("adjoint_run", &Solver::adjoint_run, "final_bar"_a, "nsteps"_a, "scheme"_a = "SSPRK2",
           "dt"_a = py::none(), "dts"_a = py::none(), "checkpoint_every"_a = 50, "want_z"_a = true,
           "swe_adjoint_run: (adjoint of the initial state T x 3, of the node bathymetry V or None, "
           "[forward sweep ms, reverse sweep ms])")
      .def("tangent_run", &Solver::tangent_run, "state0_dot"_a, "nsteps"_a, "z_dot"_a = py::none(),
           "scheme"_a = "SSPRK2", "dt"_a = py::none(), "dts"_a = py::none(),
           "swe_tangent_run: (tangent of the final state T x 3, ms)");
}
