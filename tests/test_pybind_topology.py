"""The Python module SWE_FVM (pybind/Topology.cpp:8-30 restated over the C
ABI): the reference notebook's one-triangle Topology.create (which printed
garbage counts through dangling references, notebooks/Untitled1.ipynb),
the noconvert argument rules, and the mesh sources against the reference's
topology.dat fixture."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(HERE, "swe_fvm_amd"))
SWE_FVM = pytest.importorskip("SWE_FVM", reason="build the module with __graft_entry__.build()")


def one_triangle():
    return (3, np.array([[0, 1], [0, 2], [1, 2]]), np.array([[0, -1], [0, -1], [0, -1]]),
            np.array([[0, 1, 2]]), np.array([[0, 1, 2]]), np.array([[-1, -1, -1]]))


def test_notebook_one_triangle_counts():
    args = one_triangle()
    mesh = SWE_FVM.Topology.create(*args)
    del args  # the tables are owned: dropping the caller's arrays is safe
    assert (mesh.num_nodes(), mesh.num_edges(), mesh.num_elements()) == (3, 3, 1)
    t = mesh.tables()
    assert np.array_equal(t["edge_nodes"], [[0, 1], [0, 2], [1, 2]])
    assert np.array_equal(t["element_neighbours"], [[-1, -1, -1]])
    m2 = SWE_FVM.Topology(*one_triangle())
    assert m2.num_edges() == 3


def test_noconvert_and_bad_tables_rejected():
    n, ep, et, tp, te, tt = one_triangle()
    with pytest.raises(TypeError):
        SWE_FVM.Topology.create(n, ep.astype(np.float64), et, tp, te, tt)
    with pytest.raises(TypeError):
        SWE_FVM.Topology.create(n, ep.astype(np.int32), et, tp, te, tt)
    with pytest.raises(ValueError):
        SWE_FVM.Topology.create(n, ep, et, np.array([[0, 1, 3]]), te, tt)  # node out of range
    SWE_FVM.Topology.create(n, ep, et, tp, te, np.array([[-2, -1, -1]]))  # a tag: accepted here
    with pytest.raises(ValueError):
        SWE_FVM.Topology.create(n, ep[:2], et, tp, te, tt)


def test_read_msh_matches_reference_tables(golden):
    xyz, topo = SWE_FVM.read_msh(golden.bowl_path())
    ref = golden.topology()
    t = topo.tables()
    assert np.array_equal(t["edge_nodes"], ref["ep"]) and np.array_equal(t["edge_elements"], ref["et"])
    assert np.array_equal(t["element_nodes"], ref["tp"]) and np.array_equal(t["element_edges"], ref["te"])
    assert np.array_equal(t["element_neighbours"], ref["tt"])
    assert xyz.shape == (topo.num_nodes(), 3)
    xb, tri, bnd = golden.bowl()
    rebuilt = SWE_FVM.Topology.build(xb.shape[0], tri, bnd)
    assert np.array_equal(rebuilt.tables()["edge_nodes"], ref["ep"])


def test_mesh_sources_sizes():
    xyz, topo = SWE_FVM.struct_mesh(100, 100, 0.04, -2.0, -2.0)
    assert (topo.num_elements(), topo.num_nodes(), topo.num_edges()) == (40000, 20201, 60200)
    assert xyz.min(0)[0] == -2.0 and abs(xyz.max(0)[0] - 2.0) < 1e-12
    xyz, topo = SWE_FVM.jittered_mesh(20, 0.25, 7)
    assert (topo.num_elements(), topo.num_nodes()) == (800, 441)
    assert topo.num_edges() == 441 + 800 - 1


def test_solver_binding_checks_its_arguments_before_the_device():
    """SWE_FVM.Solver (the device solver for the autograd Function) exports
    the Function's methods and rejects bad arguments with ValueError before
    any device call; without a device it fails with RuntimeError."""
    for name in ("SetVolField", "GetVolField", "set_bathymetry", "run", "adjoint_run", "tangent_run"):
        assert hasattr(SWE_FVM.Solver, name)
    xyz, topo = SWE_FVM.struct_mesh(4, 4, 0.25)
    with pytest.raises(ValueError, match="flux"):
        SWE_FVM.Solver(topo, xyz, "HLLX", "Einfeldt")
    with pytest.raises(ValueError, match="wavespeed"):
        SWE_FVM.Solver(topo, xyz, "HLL", "Roe")
    with pytest.raises(ValueError, match="xyz"):
        SWE_FVM.Solver(topo, xyz[:3], "HLLC", "Einfeldt")
    import torch
    if not torch.cuda.is_available():
        with pytest.raises(RuntimeError):
            SWE_FVM.Solver(topo, xyz, "HLLC", "Einfeldt")
