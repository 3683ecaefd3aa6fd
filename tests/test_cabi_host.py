"""The C-ABI library on the CPU: it loads, exports every declared symbol, and
its host-side mesh code (table builder, Gmsh reader, generators) is exact.
No compute call reaches a device here."""
import ctypes
import json
import os
import re

import numpy as np
import pytest

import swe_fvm_amd as swe
from swe_fvm_amd import _lib
from oracle import oracle as O
from tests import _common as C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(swe_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    names = declared_symbols()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n
    bound = {s[0] for s in _lib.SIGNATURES}
    assert names == bound, names ^ bound


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(ImportError):
        _lib.load()


def test_read_msh_reproduces_reference_tables(golden):
    """The library's Gmsh reader + table builder on examples/bowl.msh gives
    notebooks/topology.dat bit for bit."""
    xyz, topo = swe.read_msh(golden.bowl_path())
    ref = golden.topology()
    got = topo.tables()
    for k in ("ep", "et", "tp", "te", "tt"):
        assert np.array_equal(got[k], ref[k]), k
    mxyz, _, _ = golden.bowl()
    assert np.array_equal(xyz, mxyz)


def test_topology_create_one_triangle():
    """notebooks/Untitled1.ipynb: the reference prints 3, 44070616,
    140721212595328 (dangling references); the owning tables give 3, 3, 1."""
    one = json.load(open(os.path.join(C.GOLDEN, "one_triangle.json")))
    t = swe.Topology.create(one["numNodes"], np.array(one["edgeNodes"]), np.array(one["edgeElements"]),
                            np.array(one["elementNodes"]), np.array(one["elementEdges"]),
                            np.array(one["elementNeighbours"]))
    del one["edgeNodes"]  # the library copied the tables
    assert (t.num_nodes(), t.num_edges(), t.num_elements()) == (3, 3, 1)


def test_builder_matches_oracle_on_generated_meshes():
    rng = np.random.default_rng(0)
    for n, seed in [(7, 1), (20, 2), (33, 3)]:
        xyz, topo = swe.jittered_mesh(n, 0.25, seed)
        t = topo.tables()
        # rebuild from the raw triangles and the wall lines
        walls = t["ep"][t["et"][:, 1] < 0]
        again = O.build_topology(xyz.shape[0], t["tp"], walls)
        for k in ("ep", "et", "tp", "te", "tt"):
            assert np.array_equal(again[k], t[k]), (n, k)
        perm = rng.permutation(t["tp"].shape[0])
        a = swe.Topology.build(xyz.shape[0], t["tp"][perm], walls).tables()
        b = O.build_topology(xyz.shape[0], t["tp"][perm], walls)
        for k in ("ep", "et", "tp", "te", "tt"):
            assert np.array_equal(a[k], b[k]), (n, k)


def _ccw(xyz, tp):
    P = xyz
    return (P[tp[:, 1], 0] - P[tp[:, 0], 0]) * (P[tp[:, 2], 1] - P[tp[:, 0], 1]) - \
           (P[tp[:, 2], 0] - P[tp[:, 0], 0]) * (P[tp[:, 1], 1] - P[tp[:, 0], 1])


@pytest.mark.parametrize("ni,nj", [(1, 1), (3, 2), (10, 10)])
def test_struct_mesh_invariants(ni, nj):
    """StructTriangMesh (include/StructTriangMesh.h:4-15): 4 triangles per
    square around a centre node."""
    h = 0.25
    xyz, topo = swe.StructTriangMesh(ni, nj, h, -1.0, 2.0)
    t = topo.tables()
    V, E, T = topo.NumNodes(), topo.NumEdges(), topo.NumTriangles()
    assert V == (ni + 1) * (nj + 1) + ni * nj
    assert T == 4 * ni * nj
    assert E == ni * (nj + 1) + (ni + 1) * nj + 4 * ni * nj == V + T - 1
    assert int(np.sum(t["et"][:, 1] < 0)) == 2 * (ni + nj)
    area = 0.5 * _ccw(xyz, t["tp"])
    assert np.all(area > 0) and np.allclose(area, h * h / 4)
    assert np.allclose(area.sum(), ni * nj * h * h)
    # mirror symmetry of the node set about the domain centre
    cx, cy = -1.0 + ni * h / 2, 2.0 + nj * h / 2
    pts = {(round(x, 9), round(y, 9)) for x, y in xyz[:, :2]}
    assert pts == {(round(2 * cx - x, 9), round(y, 9)) for x, y in xyz[:, :2]}
    assert pts == {(round(x, 9), round(2 * cy - y, 9)) for x, y in xyz[:, :2]}


def test_jittered_mesh_invariants():
    xyz, topo = swe.jittered_mesh(40, 0.25, 11)
    t = topo.tables()
    tp = t["tp"]
    assert np.all(_ccw(xyz, tp) > 0)
    assert topo.NumEdges() == topo.NumNodes() + topo.NumTriangles() - 1
    assert np.isclose(0.5 * _ccw(xyz, tp).sum(), 1.0)
    # locally Delaunay across every interior edge
    et = t["et"]
    for e in np.where(et[:, 1] >= 0)[0]:
        a, b = et[e]
        d = list(set(tp[b]) - set(tp[a]))[0]
        A, B, Cc = xyz[tp[a], :2]
        D = xyz[d, :2]
        m = np.array([[A[0] - D[0], A[1] - D[1], (A - D) @ (A - D)],
                      [B[0] - D[0], B[1] - D[1], (B - D) @ (B - D)],
                      [Cc[0] - D[0], Cc[1] - D[1], (Cc - D) @ (Cc - D)]])
        assert np.linalg.det(m) <= 1e-14
    # Morton order: neighbours are close in index
    tt = t["tt"]
    gap = np.abs(tt - np.arange(tp.shape[0])[:, None])[tt >= 0]
    assert np.median(gap) < 40
    # seeded: identical tables and coordinates
    xyz2, topo2 = swe.jittered_mesh(40, 0.25, 11)
    assert np.array_equal(xyz, xyz2) and np.array_equal(topo2.tables()["te"], t["te"])


def test_solver_needs_a_device_or_rejects_bad_tables():
    """Unsupported boundary tags are rejected before any device work
    (src/SpaceDisc.cpp:66-72 implements SOLID_WALL; FREE_FLOW is added here,
    PERIODIC and CUSTOM are not); without a device creation fails loudly
    instead of falling back to the CPU."""
    xyz, topo = swe.jittered_mesh(4, 0.0, 1)
    t = topo.tables()
    et = t["et"].copy()
    et[np.where(et[:, 1] < 0)[0][0], 1] = -3  # PERIODIC
    bad = swe.Topology.create(topo.NumNodes(), t["ep"], et, t["tp"], t["te"], t["tt"])
    fl = swe.Fluxes.HLLC[swe.Wavespeeds.Einfeldt]
    prim = np.zeros((topo.NumTriangles(), 3))
    with pytest.raises(swe.DomainError):
        swe.SpaceDisc(fl, swe.Domain(xyz, bad), prim)
    tt = t["tt"].copy()
    tt[0, 0] = (tt[0, 0] + 1) % topo.NumTriangles()
    with pytest.raises(ValueError):
        swe.SpaceDisc(fl, swe.Domain(xyz, swe.Topology.create(topo.NumNodes(), t["ep"], t["et"], t["tp"], t["te"], tt)), prim)
    try:
        import torch  # noqa: F401
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if not has_gpu:
        with pytest.raises(swe.DeviceError):
            swe.SpaceDisc(fl, swe.Domain(xyz, topo), prim)


def test_unknown_fluxer_is_rejected():
    xyz, topo = swe.jittered_mesh(4, 0.0, 1)
    with pytest.raises(swe.SolverError):
        swe.SpaceDisc(lambda *a: None, swe.Domain(xyz, topo), np.zeros((topo.NumTriangles(), 3)))


def test_empty_and_out_of_range_tables_are_rejected_before_the_device():
    """Degenerate inputs fail with a clear error on the host, with or
    without a device: a mesh without triangles, node / edge / element
    indices out of range, ragged tables, a degenerate triangle."""
    xyz, topo = swe.jittered_mesh(4, 0.0, 1)
    t = topo.tables()
    V = topo.NumNodes()
    fl = swe.Fluxes.HLLC[swe.Wavespeeds.Einfeldt]
    z2, z3 = np.zeros((0, 2), np.int64), np.zeros((0, 3), np.int64)
    empty = swe.Topology.create(V, z2, z2, z3, z3, z3)
    assert empty.NumTriangles() == 0
    with pytest.raises(ValueError, match="without triangles"):
        swe.SpaceDisc(fl, swe.Domain(xyz, empty), np.zeros((0, 3)))
    from swe_fvm_amd.ensemble import Ensemble
    with pytest.raises(ValueError, match="without triangles"):
        Ensemble(fl, [swe.Domain(xyz, topo), swe.Domain(xyz, empty)],
                 [np.zeros((topo.NumTriangles(), 3)), np.zeros((0, 3))])
    for key, row, col, value, msg in [("tp", 0, 0, V + 5, "element node"), ("te", 0, 0, -1, "element edge"),
                                      ("ep", 0, 1, V, "edge node")]:
        tab = {k: v.copy() for k, v in t.items()}
        tab[key][row, col] = value
        with pytest.raises(swe.DomainError, match=msg):
            swe.Topology.create(V, tab["ep"], tab["et"], tab["tp"], tab["te"], tab["tt"])
    with pytest.raises(ValueError):
        swe.Topology.create(V, t["ep"][:-1], t["et"], t["tp"], t["te"], t["tt"])
    tri = t["tp"].copy()
    tri[0, 1] = tri[0, 0]
    with pytest.raises(ValueError):
        swe.Topology.build(V, tri, np.zeros((0, 2), np.int64))


def test_boundary_line_off_the_triangulation_is_rejected(golden):
    """A boundary line that is no triangle's edge would become an edge
    without a cell: the builder refuses it; a line along an interior edge is
    an ordinary interior edge numbered among the boundary edges; a mesh
    file whose counts or end markers are wrong is an error, not a partial
    mesh."""
    tri = np.array([[0, 1, 2], [0, 2, 3]], np.int64)  # two triangles of a square
    with pytest.raises(swe.DomainError, match="not an edge of any triangle"):
        swe.Topology.build(4, tri, np.array([[0, 1], [1, 3]], np.int64))
    t = swe.Topology.build(4, tri, np.array([[0, 2]], np.int64)).tables()
    assert list(t["ep"][0]) == [0, 2] and t["et"][0, 1] >= 0  # interior, numbered first
    text = open(golden.bowl_path()).read()
    path = golden.bowl_path() + ".cut.msh"
    with open(path, "w") as f:
        f.write(text[: text.index("$EndElements")])
    with pytest.raises(ValueError, match="truncated"):
        swe.read_msh(path)
    os.remove(path)


def test_corrupted_tables_fail_cleanly_before_the_device():
    """400 seeded single-entry corruptions of a small mesh's five tables
    (indices out of range, boundary tags, another element or node): each
    ends in a clear error from the table checks of Topology.create or of the
    solver, or, where the tables stay valid (a wall retagged free flow, the
    same value), in the solver's creation; none crashes the process."""
    xyz, topo = swe.jittered_mesh(5, 0.25, 3)
    t = topo.tables()
    V, T, E = topo.NumNodes(), topo.NumTriangles(), topo.NumEdges()
    fl = swe.Fluxes.HLLC[swe.Wavespeeds.Einfeldt]
    rng = np.random.default_rng(11)
    rejected = 0
    for it in range(400):
        tab = {k: v.copy() for k, v in t.items()}
        key = ["ep", "et", "tp", "te", "tt"][it % 5]
        r, c = int(rng.integers(0, tab[key].shape[0])), int(rng.integers(0, tab[key].shape[1]))
        tab[key][r, c] = int(rng.choice([-4, -3, -2, -1, 0, 1, 5, V - 1, V, T, E, 2 ** 31 - 1]))
        prim = np.zeros((T, 3))
        prim[:, 0] = 1.0
        try:
            tp = swe.Topology.create(V, tab["ep"], tab["et"], tab["tp"], tab["te"], tab["tt"])
            swe.SpaceDisc(fl, swe.Domain(xyz, tp), prim)
        except (swe.DomainError, ValueError) as e:
            rejected += 1
            assert str(e)
        except swe.core.DeviceError:
            pass  # valid tables; no device on this host
    assert rejected >= 380
