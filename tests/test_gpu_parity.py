"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bar (floating point): h, hu, hv within 1e-11 relative after N steps; the
kernels follow the oracle's operation order under -ffp-contract=off, so most
cases agree bit for bit, which is reported.  Integer outputs (cell classes)
must be identical.
"""
import numpy as np
import pytest

import swe_fvm_amd as swe
from oracle import oracle as O
from tests import _common as C

pytestmark = pytest.mark.gpu

TOL = 1e-11

FLUXES = [("HLL", "Rusanov"), ("HLL", "Davis"), ("HLL", "Einfeldt"),
          ("HLLC", "Rusanov"), ("HLLC", "Davis"), ("HLLC", "Einfeldt")]


def make_pair(xyz, topo, prim, flux="HLLC", wave="Einfeldt", cor=0.0, opts=0):
    fl = getattr(swe.Fluxes, flux)[getattr(swe.Wavespeeds, wave)]
    sd = swe.SpaceDisc(fl, swe.Domain(xyz, topo), prim, cor=cor, options=opts)
    orc = O.OracleSolver(xyz, topo.tables(), flux=flux, wave=wave, cor=cor, opts=opts)
    orc.set_state(prim)
    return sd, orc


def cons(prim, xyz, tp):
    h = prim[:, 0] - C.cell_bath(xyz, tp)
    return np.stack([h, h * prim[:, 1], h * prim[:, 2]], axis=1)


def assert_state_close(sd, orc, xyz, tp, tol=TOL, what=""):
    g = sd.GetVolField()
    r = orc.state()
    d = C.max_rel_diff(cons(g, xyz, tp), cons(r, xyz, tp))
    bitwise = np.array_equal(g, r)
    print(f"{what}: max rel diff {d:.3e} bitwise={bitwise}", flush=True)
    assert d <= tol, f"{what}: {d}"
    return bitwise


def test_bowl_gauss_wave_euler(golden):
    """testGaussWave (examples/Main.cpp:172-195): one Euler step, HLL<Einfeldt>."""
    xyz, topo = swe.read_msh(golden.bowl_path())
    xyz[:, 2] = 0.0
    tp = topo.tables()["tp"]
    prim = C.gauss_wave_state(xyz, tp)
    for opts in (0, swe.OPT_HEAD_GRADIENT):
        sd, orc = make_pair(xyz, topo, prim, "HLL", "Einfeldt", opts=opts)
        swe.Solvers.Euler(swe.TimeDisc(sd), 1e-3)
        orc.run("Euler", 1, dt=1e-3)
        assert_state_close(sd, orc, xyz, tp, what=f"bowl Euler opts={opts}")


def test_basic_msh_matches_oracle(golden):
    """notebooks/basic.msh (four named walls) read by the library: a hump
    over a sloping bottom with a dry corner, SSP-RK2 in CFL mode, both
    fluxes, against the oracle."""
    xyz, topo = swe.read_msh(golden.msh_path("basic.msh.gz"))
    xyz[:, 2] = 0.6 * (xyz[:, 0] + xyz[:, 1] + 2.0) / 4.0
    tp = topo.tables()["tp"]
    c = xyz[tp].mean(1)
    w = 0.8 + 0.2 * np.exp(-8.0 * (c[:, 0] ** 2 + c[:, 1] ** 2))
    prim = np.stack([np.maximum(w, C.cell_bath(xyz, tp)), 0.05 * np.ones(len(c)), np.zeros(len(c))], 1)
    for flux in ("HLL", "HLLC"):
        sd, orc = make_pair(xyz, topo, prim, flux, "Einfeldt")
        sd.run("SSPRK2", 40, dt=0.0, cfl=0.3)
        orc.run("SSPRK2", 40, dt=0.0, cfl=0.3)
        assert_state_close(sd, orc, xyz, tp, what=f"basic.msh {flux}")


@pytest.mark.parametrize("flux,wave", FLUXES)
def test_interface_values_and_fluxes(flux, wave):
    """ComputeInterfaceValues + ComputeFluxes getters on a wet/dry mesh."""
    xyz, topo, prim = C.wet_dry_mesh(20, seed=5)
    tp = topo.tables()["tp"]
    sd, orc = make_pair(xyz, topo, prim, flux, wave, cor=0.3)
    sd.ComputeInterfaceValues()
    orc.compute_interface_values()
    cls = sd.CellClass()
    assert np.array_equal(cls, orc.classify())
    assert (cls == 0).any() and (cls == 1).any() and (cls == 2).any()
    edg, oedg = sd.GetEdgField(), orc.edge_field()
    assert C.max_rel_diff(edg, oedg) <= TOL
    src, osrc = sd.GetSrcField(), orc.src_field()
    assert C.max_rel_diff(src[:, 1:], osrc[:, 1:]) <= TOL
    assert C.max_rel_diff(sd.GetMaxWp(), orc.max_wp()) <= TOL
    sd.ComputeFluxes()
    r = orc.compute_fluxes()
    assert C.max_rel_diff(sd.GetFluxes(), orc.fluxes()) <= TOL
    assert abs(sd.GetMinLenToWavespeed() - r) <= TOL * abs(r)
    td = swe.TimeDisc(sd)
    assert abs(td.CFLdt() - 0.15 * r) <= TOL * r
    for i in list(range(0, tp.shape[0], max(1, tp.shape[0] // 37))):
        a, b = td.RHS(i, 1e-3), orc.rhs(i, 1e-3)
        assert np.max(np.abs(a - b)) <= TOL * max(1.0, np.max(np.abs(b))), i


@pytest.mark.parametrize("scheme", ["Euler", "SSPRK2", "SSPRK3"])
@pytest.mark.parametrize("flux,wave", FLUXES)
def test_steps_wet_dry(scheme, flux, wave):
    """N steps on a wet/dry mesh in CFL mode; every scheme and flux."""
    xyz, topo, prim = C.wet_dry_mesh(24, seed=11)
    tp = topo.tables()["tp"]
    sd, orc = make_pair(xyz, topo, prim, flux, wave, cor=0.1)
    nsteps = 6
    sd.run(scheme, nsteps, dt=0.0, cfl=0.15)
    odts = orc.run(scheme, nsteps, dt=0.0, cfl=0.15)
    gdts = sd.dts(nsteps)
    assert np.max(np.abs(gdts - odts) / odts) <= TOL
    assert_state_close(sd, orc, xyz, tp, what=f"{scheme} {flux}<{wave}>")


@pytest.mark.parametrize("opts", [0, swe.OPT_HEAD_GRADIENT, swe.OPT_HEAD_PARTWET2])
def test_head_compat_options(opts):
    xyz, topo, prim = C.wet_dry_mesh(24, seed=2)
    tp = topo.tables()["tp"]
    sd, orc = make_pair(xyz, topo, prim, "HLLC", "Einfeldt", opts=opts)
    sd.run("SSPRK2", 5, dt=2e-3)
    orc.run("SSPRK2", 5, dt=2e-3)
    assert_state_close(sd, orc, xyz, tp, what=f"opts={opts}")


def test_thacker_bowl_short():
    """Thacker paraboloid on StructTriangMesh (4 triangles per square)."""
    from swe_fvm_amd import workloads as W
    dom, prim, th = W.thacker(n=24, quad=8)
    xyz, topo = dom.geometry, dom.topology
    tp = topo.tables()["tp"]
    sd, orc = make_pair(xyz, topo, prim, "HLLC", "Einfeldt")
    sd.run("SSPRK2", 20, dt=0.0, cfl=0.15)
    orc.run("SSPRK2", 20, dt=0.0, cfl=0.15)
    assert_state_close(sd, orc, xyz, tp, what="thacker")


def test_lake_at_rest_gpu():
    """Well-balancedness with emerging islands, 300 steps: |w - 1| <= 1e-13 on
    fully wet cells; |hu|, |hv| stay at round-off (<= 1e-12): the flux and the
    pressure correction of TimeDisc::RHS (src/TimeDisc.cpp:31,35) group the
    same product differently, so each step leaves ~1e-16 that accumulates as
    a random walk; the oracle reproduces these bits exactly."""
    from swe_fvm_amd import workloads as W
    dom, prim = W.lake_at_rest(n=120, seed=4)
    xyz, topo = dom.geometry, dom.topology
    tp = topo.tables()["tp"]
    fl = swe.Fluxes.HLLC[swe.Wavespeeds.Einfeldt]
    sd = swe.SpaceDisc(fl, dom, prim)
    sd.run("SSPRK2", 300, dt=0.0, cfl=0.15)
    st = sd.GetVolField()
    h = st[:, 0] - C.cell_bath(xyz, tp)
    full = np.max(xyz[tp, 2], axis=1) < 1.0
    assert np.max(np.abs(st[full, 0] - 1.0)) <= 1e-13
    mom = max(np.max(np.abs(h * st[:, 1])), np.max(np.abs(h * st[:, 2])))
    print(f"lake at rest 300 steps: max|w-1| {np.max(np.abs(st[full, 0] - 1.0)):.2e} max|hu|,|hv| {mom:.2e}")
    assert mom <= 1e-12


def test_mid_size_against_oracle():
    """~125k cells, dam break with dry pockets, SSPRK2 CFL mode, 10 steps."""
    from swe_fvm_amd import workloads as W
    dom, prim = W.dam_break(n=250, seed=7)
    xyz, topo = dom.geometry, dom.topology
    tp = topo.tables()["tp"]
    sd, orc = make_pair(xyz, topo, prim, "HLLC", "Einfeldt")
    sd.run("SSPRK2", 10, dt=0.0, cfl=0.15)
    orc.run("SSPRK2", 10, dt=0.0, cfl=0.15)
    assert_state_close(sd, orc, xyz, tp, what="dam break n=250")


def test_large_properties_determinism_mass_positivity():
    """2M cells: two runs bitwise identical, mass conserved, h >= 0."""
    from swe_fvm_amd import workloads as W
    dom, prim = W.dam_break(n=1000, seed=3)
    xyz, topo = dom.geometry, dom.topology
    tp = topo.tables()["tp"]
    fl = swe.Fluxes.HLLC[swe.Wavespeeds.Einfeldt]
    m0 = C.mass(xyz, tp, prim)
    outs = []
    for _ in range(2):
        sd = swe.SpaceDisc(fl, dom, prim)
        sd.run("SSPRK2", 20, dt=0.0, cfl=0.15)
        outs.append(sd.GetVolField())
        del sd
    assert np.array_equal(outs[0], outs[1])
    m1 = C.mass(xyz, tp, outs[0])
    assert abs(m1 - m0) <= 1e-12 * abs(m0)
    h = outs[0][:, 0] - C.cell_bath(xyz, tp)
    assert np.min(h) >= 0.0


def test_full_size_dam_break_paths_agree():
    """The benchmark's workload at full size (dam break, n = 2828: 15,995,168
    cells), 6 SSP-RK2 steps in CFL mode.  The tile kernels (the default path)
    and the separate flux, draining and update kernels (each pinned to the
    oracle at the sizes above) give the same state bit for bit, and so does
    the AoSoA-64 state layout; mass is conserved to 1e-12 relative and every
    depth stays >= 0."""
    import gc

    from swe_fvm_amd import workloads as W
    print("[full size] building the 16M-cell dam break", flush=True)
    dom, prim = W.dam_break(n=2828, seed=3)
    fl = swe.Fluxes.HLLC[swe.Wavespeeds.Einfeldt]
    outs = {}
    for name, fused, layout in (("tiles", True, 0), ("separate", False, 0), ("tiles_aosoa64", True, 64)):
        sd = swe.SpaceDisc(fl, dom, prim)
        assert sd.set_fused(fused) == fused  # the tiles are built on this mesh
        sd.set_layout(layout)
        sd.run("SSPRK2", 6, dt=0.0, cfl=0.15)
        outs[name] = sd.GetVolField()
        del sd
        gc.collect()
        print(f"[full size] {name} done", flush=True)
    assert np.array_equal(outs["tiles"], outs["separate"])
    assert np.array_equal(outs["tiles"], outs["tiles_aosoa64"])
    assert not np.array_equal(outs["tiles"], prim)  # the steps moved the state
    xyz, tp = dom.geometry, dom.topology.tables()["tp"]
    m0, m1 = C.mass(xyz, tp, prim), C.mass(xyz, tp, outs["tiles"])
    assert abs(m1 - m0) <= 1e-12 * abs(m0)
    assert np.min(outs["tiles"][:, 0] - C.cell_bath(xyz, tp)) >= 0.0


@pytest.mark.parametrize("scheme", ["Euler", "SSPRK2", "SSPRK3"])
@pytest.mark.parametrize("flux,wave", [("HLL", "Einfeldt"), ("HLLC", "Einfeldt"), ("HLLC", "Davis")])
def test_exact_balance_matches_oracle(scheme, flux, wave):
    """SWE_OPT_EXACT_BALANCE on a wet/dry mesh with Coriolis: CFL steps (the
    flux kernel, then the tile kernel reading its fluxes, for the first
    stage) and fixed-dt steps (the fused tile kernel) against the oracle,
    bit for bit."""
    xyz, topo, prim = C.wet_dry_mesh(24, seed=11)
    tp = topo.tables()["tp"]
    for dt in (0.0, 2e-3):
        sd, orc = make_pair(xyz, topo, prim, flux, wave, cor=0.1, opts=swe.OPT_EXACT_BALANCE)
        sd.run(scheme, 6, dt=dt, cfl=0.15)
        orc.run(scheme, 6, dt=dt, cfl=0.15)
        assert assert_state_close(sd, orc, xyz, tp, what=f"exact balance {scheme} {flux}<{wave}> dt={dt}")
