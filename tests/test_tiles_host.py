"""Tile tables of the fused flux-draining-update kernel, checked on the host:
tests/cpp/tiles_check.cpp replays the kernel's index arithmetic over every
tile (bounds, edge identities, one evaluation per edge and tile, slots and
neighbour codes) on generated, structured and file meshes, with walls and
free-flow edges; a mesh in random cell order is refused (the fused kernel
is then not used)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import swe_fvm_amd as swe

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module", params=[256, 512], ids=["tile256", "tile512"])
def checker(tmp_path_factory, request):
    """The checker built with the tables' tile size (the build-time knob
    SWE_TILE_CELLS of the library)."""
    out = str(tmp_path_factory.mktemp("tiles") / "libtiles_check.so")
    cmd = ["g++", "-std=c++17", "-O2", "-shared", "-fPIC", "-Wall", "-Wextra", f"-DSWE_TILE_CELLS={request.param}",
           os.path.join(ROOT, "tests", "cpp", "tiles_check.cpp"),
           os.path.join(ROOT, "swe_fvm_amd", "csrc", "swe_tiles.cpp"), "-o", out]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    L = ctypes.CDLL(out)
    L.tiles_check.restype = ctypes.c_int
    L.tiles_check.argtypes = [ctypes.c_int64] + [ctypes.c_void_p] * 7 + [ctypes.c_int64, ctypes.c_void_p,
                                                                          ctypes.c_char_p, ctypes.c_int]
    return L


def soa(topo):
    t = topo.tables()
    T = t["tp"].shape[0]
    tt = np.ascontiguousarray(t["tt"].T.reshape(-1), dtype=np.int32)
    te = np.ascontiguousarray(t["te"].T.reshape(-1), dtype=np.int32)
    et0 = np.ascontiguousarray(t["et"][:, 0], dtype=np.int32)
    et1 = np.ascontiguousarray(t["et"][:, 1], dtype=np.int32)
    ep0 = np.ascontiguousarray(t["ep"][:, 0], dtype=np.int32)
    ep1 = np.ascontiguousarray(t["ep"][:, 1], dtype=np.int32)
    fl = np.zeros(T, np.uint32)
    for k in range(3):
        fl |= (et0[t["te"][:, k]] == np.arange(T)).astype(np.uint32) << k
    return T, tt, te, fl, et0, et1, ep0, ep1


def run(checker, topo):
    T, tt, te, fl, et0, et1, ep0, ep1 = soa(topo)
    stats = np.zeros(5, np.int64)
    msg = ctypes.create_string_buffer(256)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    rc = checker.tiles_check(T, p(tt), p(te), p(fl), p(et0), p(et1), p(ep0), p(ep1), et0.shape[0], p(stats), msg,
                             256)
    return rc, msg.value.decode(), stats


@pytest.mark.parametrize("n,seed", [(6, 1), (40, 2), (171, 3)])
def test_tiles_of_generated_meshes(checker, n, seed):
    xyz, topo = swe.jittered_mesh(n, 0.25, seed)
    rc, msg, st = run(checker, topo)
    assert rc == 0, msg
    T = topo.NumTriangles()
    print(f"T={T}: tiles {st[0]}, evaluations per cell {st[1] / T:.3f}, ring per tile {st[2] / st[0]:.1f}, "
          f"max evaluations {st[3]}, max ring {st[4]}")
    assert st[1] >= 1.5 * T - 1  # every edge at least once


def test_tiles_with_free_flow_edges(checker):
    xyz, topo = swe.jittered_mesh(50, 0.25, 4)
    t = topo.tables()
    b = np.nonzero(t["et"][:, 1] < 0)[0]
    rc, msg, _ = run(checker, topo.with_boundary_tag(b[::3], swe.FREE_FLOW))
    assert rc == 0, msg


def test_tiles_of_structured_and_file_meshes(checker, golden):
    xyz, topo = swe.StructTriangMesh(30, 30, 0.04)
    rc, msg, _ = run(checker, topo)
    assert rc == 0, msg
    # the reference's bowl.msh in file order, and Morton-reordered
    xyz, topo = swe.read_msh(golden.bowl_path())
    rc, msg, _ = run(checker, topo)
    assert rc == 0 or "refused" in msg, msg
    dom, _, _ = swe.morton_order(swe.Domain(xyz, topo))
    rc, msg, _ = run(checker, dom.topology)
    assert rc == 0, msg


def test_random_cell_order_is_refused(checker):
    xyz, topo = swe.jittered_mesh(120, 0.25, 5)
    rng = np.random.default_rng(0)
    t = topo.tables()
    perm = rng.permutation(t["tp"].shape[0])
    tri = t["tp"][perm]
    bnd = t["ep"][t["et"][:, 1] < 0]
    shuffled = swe.Topology.build(xyz.shape[0], tri, bnd)
    rc, msg, _ = run(checker, shuffled)
    assert rc == 1 and "refused" in msg
