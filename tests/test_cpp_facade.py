"""The C++ drop-in facade (include/swe_fvm.hpp): compiles with g++ against
the C ABI; on the GPU, a testGaussWave driver written against it matches
the oracle bit for bit."""
import os
import subprocess

import numpy as np
import pytest

from tests import _common as C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "gauss_wave.cpp")
LIBDIR = os.path.join(ROOT, "swe_fvm_amd")


def compile_driver(out):
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"), SRC,
           "-o", out, "-L", LIBDIR, "-lswe_fvm_amd", f"-Wl,-rpath,{LIBDIR}"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    return out


def test_facade_compiles(tmp_path):
    compile_driver(str(tmp_path / "gauss_wave"))


@pytest.mark.gpu
def test_facade_gauss_wave_matches_oracle(tmp_path, golden):
    from oracle import oracle as O
    exe = compile_driver(str(tmp_path / "gauss_wave"))
    res = subprocess.run([exe, golden.bowl_path(), str(tmp_path) + "/"], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stderr
    print(res.stdout)
    out0 = np.loadtxt(tmp_path / "out0.dat")
    out1 = np.loadtxt(tmp_path / "out1.dat")
    xyz, tri, bnd = golden.bowl()
    topo = O.build_topology(xyz.shape[0], tri, bnd)
    xyz = xyz.copy()
    xyz[:, 2] = 0.0
    # the driver evaluates exp with the C library; numpy's exp may differ in
    # the last bit, so the oracle starts from the driver's own initial state
    prim = C.gauss_wave_state(xyz, topo["tp"])
    assert np.max(np.abs(out0[:, 0] - prim[:, 0])) <= 4e-16
    assert np.all(out0[:, 1:] == 0.0)
    prim = np.stack([out0[:, 0], np.zeros(len(out0)), np.zeros(len(out0))], axis=1)
    s = O.OracleSolver(xyz, topo, "HLL", "Einfeldt")
    s.set_state(prim)
    s.run("Euler", 1, dt=1e-3)
    st = s.state()
    ref = np.stack([st[:, 0], st[:, 0] * st[:, 1], st[:, 0] * st[:, 2]], axis=1)
    assert np.array_equal(out1, ref)


DRIVER = os.path.join(ROOT, "tests", "cpp", "assigner_driver.cpp")


def compile_assigner_driver(out):
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"), DRIVER,
           "-o", out, "-L", LIBDIR, "-lswe_fvm_amd", f"-Wl,-rpath,{LIBDIR}"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    return out


def test_assigner_driver_compiles(tmp_path):
    compile_assigner_driver(str(tmp_path / "assigner_driver"))


def _host_assigners():
    """PrimAssigner / ConsAssigner (src/Assigners.cpp:8-44) restated in
    numpy: the reference of the facade's host writes."""
    def prim(b, rhs):
        h = rhs[0] - b
        if not h > 1e-12:
            return np.array([b, 0.0, 0.0])
        c = np.array(rhs, dtype=np.float64)
        if h < 1e-3:
            f = np.sqrt(2.0) * h / np.sqrt(h * h + 1e-6)
            c[1] *= f
            c[2] *= f
        return c

    def cons_get(b, c):
        h = c[0] - b
        return np.array([h, c[1] * h, c[2] * h])

    def cons_set(b, rhs):
        h = rhs[0]
        if not h > 1e-12:
            return np.array([b, 0.0, 0.0])
        ih = np.sqrt(2.0) * h / np.sqrt(h * h * h * h + 1e-12) if h < 1e-3 else 1.0 / h
        return np.array([h + b, rhs[1] * ih, rhs[2] * ih])
    return prim, cons_get, cons_set


@pytest.mark.gpu
def test_assigner_driver_matches_oracle(tmp_path, golden):
    """The facade's write path, reconstruction, geometry and per-edge flux
    getters and both stepping routes against the oracle, bit for bit."""
    from oracle import oracle as O
    exe = compile_assigner_driver(str(tmp_path / "assigner_driver"))
    res = subprocess.run([exe, golden.bowl_path(), str(tmp_path) + "/"], capture_output=True, text=True,
                         timeout=600)
    assert res.returncode == 0, res.stderr
    print(res.stdout)
    assert "refused_div0 1" in res.stdout
    xyz, tri, bnd = golden.bowl()
    topo = O.build_topology(xyz.shape[0], tri, bnd)
    xyz = xyz.copy()
    xyz[:, 2] = np.loadtxt(tmp_path / "z.dat")
    tp = topo["tp"]
    c3 = 1.0 / 3.0
    cen = (xyz[tp[:, 0]] * c3 + xyz[tp[:, 1]] * c3) + xyz[tp[:, 2]] * c3
    bc = cen[:, 2]
    init = np.loadtxt(tmp_path / "init.dat")
    # the initial writes, replayed with the restated assigners
    prim, cons_get, cons_set = _host_assigners()
    T = tp.shape[0]
    for i in range(T):
        c = cen[i]
        st = prim(bc[i], [0.12 - 0.01 * (c[0] - 4.0), 0.05 * (c[1] - 4.0), -0.03 * (c[0] - 4.0)])
        if i % 7 == 3:
            st = cons_set(bc[i], cons_get(bc[i], st) * 0.5)
        if i % 11 == 5:
            st = cons_set(bc[i], cons_get(bc[i], st) * (1.0 / 4.0))
        if i % 13 == 1:
            st = prim(bc[i], st + np.array([0.001, 0.01, 0.0]))
        if i % 17 == 2:
            st = prim(bc[i], st - np.array([0.0, 0.0, 0.02]))
        assert np.array_equal(st, init[i]), i
    assert np.any(init[:, 0] == bc) and np.any((init[:, 0] - bc > 0) & (init[:, 0] - bc < 1e-3))
    # both stepping routes equal the oracle
    s = O.OracleSolver(xyz, topo, "HLLC", "Einfeldt", cor=0.3)
    s.set_state(init)
    s.run("Euler", 2, dt=2e-3)
    s.run("SSPRK2", 2, dt=2e-3)
    ref = s.state()
    for name in ("cells.dat", "solvers.dat"):
        got = np.loadtxt(tmp_path / name)
        print(name, "bitwise", np.array_equal(got, ref), "max diff", np.max(np.abs(got - ref)))
        assert np.array_equal(got, ref), name
    # reconstructions (pass 2 with the pass-1 node maximum: oracle kind 4)
    o = O.OracleSolver(xyz, topo, "HLLC", "Einfeldt", cor=0.3)
    o.set_state(init)
    o.compute_interface_values()
    rec = np.loadtxt(tmp_path / "recon.dat", ndmin=2)
    assert len(rec) >= 8
    for row in rec:
        i, k = int(row[0]), int(row[1])
        want = o.reconstruct(i, 4 if k == 3 else k)
        # PartWet2 of a dry cell divides 0 by 0 in both (h23 = 0)
        assert np.array_equal(row[2:5], want[:3], equal_nan=True), (i, k)
        # Gradient(T(i)) is the reconstruction's unless it is dry there
        dry_at_c = want[0] - bc[i] < 0
        if not dry_at_c:
            assert np.array_equal(row[5:7], want[3:5], equal_nan=True), (i, k)
    # per-edge fluxes in both orientations and the min-updated r
    o.compute_fluxes()
    fl = np.loadtxt(tmp_path / "flux.dat", ndmin=2)
    et = topo["et"]
    for row in fl:
        e = int(row[0])
        F, r = o.edge_flux("HLLC", "Einfeldt", e, et[e, 0], et[e, 1], 1.0)
        assert np.array_equal(row[1:4], F) and row[4] == r, e
        if et[e, 1] >= 0:
            G, r2 = o.edge_flux("HLL", "Davis", e, et[e, 1], et[e, 0], r)
            assert np.array_equal(row[5:8], G) and row[8] == r2, e
    # geometry getters against their definitions
    geo = np.loadtxt(tmp_path / "geom.dat", ndmin=2)
    ep = topo["ep"]
    for row in geo:
        e, t0 = int(row[0]), int(row[1])
        a, b = xyz[ep[e, 0]], xyz[ep[e, 1]]
        L = np.sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]))
        tx, ty = (b[0] - a[0]) / L, (b[1] - a[1]) / L
        if (cen[t0, 0] - a[0]) * ty - tx * (cen[t0, 1] - a[1]) > 0:
            tx, ty = -tx, -ty
        assert row[2] == tx and row[3] == ty and row[4] == ty and row[5] == -tx
        g = O.gradient(xyz[tp[t0]])
        assert row[6] == g[0] and row[7] == g[1]
        if et[e, 1] < 0:
            assert np.array_equal(row[8:11], 0.5 * (a + b))
        else:
            # the crossing of the centroid line lies on the edge
            cpt = row[8:11]
            lam = ((cpt[0] - a[0]) * (b[0] - a[0]) + (cpt[1] - a[1]) * (b[1] - a[1])) / (L * L)
            assert -1e-12 <= lam <= 1 + 1e-12
            assert abs((cpt[0] - a[0]) * (b[1] - a[1]) - (cpt[1] - a[1]) * (b[0] - a[0])) <= 1e-12
        p0 = xyz[tp[t0, 0]]
        assert abs(row[11] - (g[0] * (cen[t0, 0] - p0[0]) + g[1] * (cen[t0, 1] - p0[1]))) <= 1e-14


def test_facade_host_parts(tmp_path):
    """Assigners, Domain getters and ElemFlux of the facade (host code, no
    device) against their definitions (src/Assigners.cpp:8-44,
    src/Bathymetry.cpp:10-80, include/SpaceDisc.h:4-18)."""
    exe = str(tmp_path / "facade_host")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "facade_host.cpp"), "-o", exe, "-L", LIBDIR, "-lswe_fvm_amd",
           f"-Wl,-rpath,{LIBDIR}"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    res = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert res.returncode == 0, res.stderr
    lines = {}
    for ln in res.stdout.splitlines():
        key, *vals = ln.split()
        lines.setdefault(key, []).append([float(x) for x in vals])
    prim, cons_get, cons_set = _host_assigners()
    xyz = np.array([[0., 0., -1.], [2., 0., -0.5], [0., 1., -0.75], [2., 1.5, -0.25]])
    c3 = 1.0 / 3.0
    bc0 = (xyz[0, 2] * c3 + xyz[1, 2] * c3) + xyz[2, 2] * c3
    bc1 = (xyz[1, 2] * c3 + xyz[3, 2] * c3) + xyz[2, 2] * c3
    p = lines["prim"][0]
    assert p[:3] == list(prim(bc0, [-0.5, 1.0, 2.0]))
    assert p[3:] == list(prim(bc1, [bc1 + 5e-4, 1.0, 2.0]))
    assert 0 < p[4] < 1.0  # desingularised
    st = cons_set(bc1, [0.5, 0.25, -0.125])
    st = cons_set(bc1, cons_get(bc1, st) + np.array([0.25, 0.25, 0.0]))
    st = cons_set(bc1, cons_get(bc1, st) * 2.0)
    st = cons_set(bc1, cons_get(bc1, st) * (1.0 / 4.0))
    st = cons_set(bc1, cons_get(bc1, st) - np.array([0.0, 0.125, 0.0]))
    assert lines["cons"][0] == list(st)
    assert lines["div0"][0] == [1.0]
    cen = [np.array([(xyz[a, 0] * c3 + xyz[b_, 0] * c3) + xyz[c, 0] * c3,
                     (xyz[a, 1] * c3 + xyz[b_, 1] * c3) + xyz[c, 1] * c3]) for a, b_, c in ((0, 1, 2), (1, 3, 2))]
    for row in lines["edge"]:
        t0 = int(row[1])
        tx, ty, nx, ny = row[2:6]
        assert abs(np.hypot(tx, ty) - 1.0) < 1e-15 and (nx, ny) == (ty, -tx)
        # the normal points away from the cell: its centroid lies behind the edge
        mid = row[6:8]
        assert (cen[t0] - mid) @ np.array([nx, ny]) < 0
    # the interior edge's C is where the centroid line crosses it
    interior = [r for r in lines["edge"] if abs(r[6] - 1.0) < 0.9 and abs(r[7] - 0.5) < 0.6]
    assert interior
    for row in lines["slope"]:
        t = int(row[0])
        pts = xyz[[0, 1, 2]] if t == 0 else xyz[[1, 3, 2]]
        A = np.array([pts[1, :2] - pts[0, :2], pts[2, :2] - pts[0, :2]])
        g = np.linalg.solve(A, [pts[1, 2] - pts[0, 2], pts[2, 2] - pts[0, 2]])
        assert np.allclose(row[1:3], g, rtol=1e-14, atol=1e-15)
        assert abs(row[3] - (g[0] * (1.0 - pts[0, 0]) + g[1] * (0.5 - pts[0, 1]))) < 1e-15
    assert lines["extent"][0] == [0.0, 2.0, 0.0, 1.5]
    e = lines["elem"][0]
    hveln = 1.0 * 0.6 + (-0.5) * 0.8
    assert e[0] == hveln and e[1] == (hveln / 2.0) * 1.0 + 2.0 * 0.6 and e[2] == (hveln / 2.0) * -0.5 + 2.0 * 0.8
    assert e[3] == 0.0


PERIODIC_DRIVER = os.path.join(ROOT, "tests", "cpp", "periodic_driver.cpp")


def compile_periodic_driver(out):
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"), PERIODIC_DRIVER,
           "-o", out, "-L", LIBDIR, "-lswe_fvm_amd", f"-Wl,-rpath,{LIBDIR}"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    return out


def test_periodic_driver_compiles(tmp_path):
    """StructTriangMesh, Domain{mesh} and Topology::SetPeriodic compile
    against the C ABI."""
    compile_periodic_driver(str(tmp_path / "periodic_driver"))


@pytest.mark.gpu
def test_periodic_driver_matches_oracle(tmp_path):
    """The facade's StructTriangMesh and periodic pairing on the device: the
    driver's channel (pairs found in C++ by translation) after ten SSP-RK2
    steps equals the oracle on the same tables bit for bit; the mesh and
    pairs equal the Python ones."""
    import swe_fvm_amd as swe
    from oracle import oracle as O
    exe = compile_periodic_driver(str(tmp_path / "periodic_driver"))
    res = subprocess.run([exe, str(tmp_path) + "/"], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    z = np.loadtxt(tmp_path / "z.dat")
    pairs = np.loadtxt(tmp_path / "pairs.dat", dtype=np.int64, ndmin=2)
    out0, out1 = np.loadtxt(tmp_path / "out0.dat"), np.loadtxt(tmp_path / "out1.dat")
    xyz, topo = swe.StructTriangMesh(12, 8, 0.125)
    xyz = xyz.copy()
    xyz[:, 2] = z
    py_pairs = swe.periodic_pairs(xyz, topo, (1.5, 0.0))
    assert {tuple(sorted(p)) for p in pairs.tolist()} == {tuple(sorted(p)) for p in py_pairs.tolist()}
    per = topo.with_periodic(pairs)
    s = O.OracleSolver(xyz, per.tables(), "HLLC", "Einfeldt", cor=0.1)
    s.set_state(out0)
    s.run("SSPRK2", 10, dt=2e-3)
    st = s.state()
    print(f"periodic facade driver: bitwise={np.array_equal(out1, st)}", flush=True)
    assert np.array_equal(out1, st)
    assert np.max(np.abs(out1 - out0)) > 1e-4  # the hump moved
