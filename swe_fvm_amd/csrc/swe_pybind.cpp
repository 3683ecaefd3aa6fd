// Python module SWE_FVM: the reference's pybind binding of Topology
// (pybind/Topology.cpp:8-30) with the same constructor, Topology.create and
// num_nodes / num_edges / num_elements, over the C ABI's owning topology
// (the reference's Topology keeps references to the caller's arrays, which
// dangle once the temporaries of the call are gone).  Also the mesh sources
// of the C ABI (Gmsh reader, StructTriangMesh, seeded jittered Delaunay) and
// the tables back as numpy arrays.  Host code only.
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
  m.doc() = "SWE_FVM mesh topology (owning tables) and mesh sources";
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
}
