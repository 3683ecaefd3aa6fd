"""PERIODIC boundaries (include/Includes.h:21: declared, never paired in the
reference; paired here by Topology.with_periodic): the oracle's definition
is pinned by closed forms (a uniform stream in a doubly periodic square
stays uniform, mass is conserved to round-off, the two edges of a pair carry
opposite fluxes bit for bit, a periodic pair of a structured mesh carries
the flux an interior edge would), and the GPU matches the oracle bit for
bit."""
import numpy as np
import pytest

import swe_fvm_amd as swe
from oracle import oracle as O
from tests import _common as C


def periodic_mesh(n, seed, x=True, y=True, z=None):
    xyz, topo = swe.jittered_mesh(n, 0.2, seed)
    if z is not None:
        xyz[:, 2] = z(xyz[:, 0], xyz[:, 1])
    pairs = []
    if x:
        pairs.append(swe.periodic_pairs(xyz, topo, (1.0, 0.0)))
    if y:
        pairs.append(swe.periodic_pairs(xyz, topo, (0.0, 1.0)))
    return xyz, topo.with_periodic(np.concatenate(pairs)), topo


def test_pairs_cover_the_sides():
    n = 12
    xyz, per, plain = periodic_mesh(n, 1)
    t = per.tables()
    assert "pe" in t
    paired = np.nonzero(t["pe"] >= 0)[0]
    assert paired.shape[0] == 4 * n  # every boundary edge of the square
    assert np.all(t["et"][paired, 1] == swe.PERIODIC)
    assert np.all(t["pe"][t["pe"][paired]] == paired)
    # the partner's midpoint is the edge's, shifted by one period
    ep = t["ep"]
    mid = 0.5 * (xyz[ep[:, 0], :2] + xyz[ep[:, 1], :2])
    d = mid[t["pe"][paired]] - mid[paired]
    assert np.all(np.isclose(np.abs(d).max(1), 1.0) & np.isclose(np.abs(d).min(1), 0.0))
    # tags survive a Morton reordering and a stacking
    dom, _, _ = swe.morton_order(swe.Domain(xyz, per))
    t2 = dom.topology.tables()
    assert np.count_nonzero(t2["pe"] >= 0) == 4 * n
    big, off = swe.stack_domains([swe.Domain(xyz, per), swe.Domain(xyz, plain)])
    t3 = big.topology.tables()
    assert np.count_nonzero(t3["pe"] >= 0) == 4 * n


def test_unpaired_periodic_tag_is_refused():
    xyz, topo = swe.jittered_mesh(6, 0.2, 2)
    t = topo.tables()
    b = np.nonzero(t["et"][:, 1] < 0)[0]
    with pytest.raises(swe.DomainError):
        topo.with_boundary_tag(b[:2], swe.PERIODIC)
    # tables that carry the tag without a partner table: refused by the
    # solver before any device is touched, and by the oracle
    t["et"][b[:2], 1] = swe.PERIODIC
    bad = swe.Topology.create(xyz.shape[0], t["ep"], t["et"], t["tp"], t["te"], t["tt"])
    with pytest.raises(swe.DomainError):
        swe.SpaceDisc(swe.Fluxes.HLLC[swe.Wavespeeds.Einfeldt], swe.Domain(xyz, bad),
                      np.tile([1.0, 0.0, 0.0], (bad.NumTriangles(), 1)), device=0)
    with pytest.raises(ValueError):
        O.OracleSolver(xyz, t)
    # a partner that is not periodic itself
    pairs = swe.periodic_pairs(xyz, topo, (1.0, 0.0))
    t2 = topo.with_periodic(pairs).tables()
    t2["pe"][pairs[0, 1]] = -1
    with pytest.raises(ValueError):
        O.OracleSolver(xyz, t2)


@pytest.mark.parametrize("opts", [0, O.OPT_DRAIN_LENGTH])
@pytest.mark.parametrize("flux,wave", [("HLL", "Einfeldt"), ("HLLC", "Einfeldt"), ("HLLC", "Davis")])
def test_uniform_stream_in_a_periodic_square(flux, wave, opts):
    """With every boundary edge periodic a uniform stream over a flat bottom
    has nowhere to go: it stays uniform (to round-off) and keeps its mass."""
    xyz, per, _ = periodic_mesh(10, 3, z=lambda x, y: 0.0 * x)
    T = per.NumTriangles()
    prim = np.tile([1.0, 0.3, -0.2], (T, 1))
    s = O.OracleSolver(xyz, per.tables(), flux, wave, opts=opts)
    s.set_state(prim)
    s.run("SSPRK2", 20, dt=1e-3)
    assert np.max(np.abs(s.state() - prim)) < 1e-13


def _mass(xyz, tp, prim):
    return float(np.sum(C.cell_area(xyz, tp) * (prim[:, 0] - C.cell_bath(xyz, tp))))


@pytest.mark.parametrize("scheme", ["Euler", "SSPRK2", "SSPRK3"])
def test_mass_conserved_and_pair_fluxes_opposite(scheme):
    """A hump crossing the periodic sides over a periodic bottom: mass is
    conserved to round-off (walls only on y in the channel case) and the two
    edges of every pair carry opposite fluxes, bit for bit."""
    z = lambda x, y: 0.05 * np.cos(2 * np.pi * x) * np.cos(2 * np.pi * y)
    for x_per, y_per in ((True, True), (True, False)):
        xyz, per, _ = periodic_mesh(16, 4, x=x_per, y=y_per, z=z)
        t = per.tables()
        tp = t["tp"]
        cen = xyz[tp].mean(1)
        bc = C.cell_bath(xyz, tp)
        r2 = (cen[:, 0] - 0.05) ** 2 + (cen[:, 1] - 0.5) ** 2
        prim = np.stack([bc + 0.5 + 0.2 * np.exp(-r2 / 0.01), 0.1 + 0 * r2, 0 * r2], 1)
        s = O.OracleSolver(xyz, t, "HLLC", "Einfeldt", cor=0.2)
        s.set_state(prim)
        m0 = _mass(xyz, tp, prim)
        s.run(scheme, 30, dt=0.0, cfl=0.15)
        st = s.state()
        assert abs(_mass(xyz, tp, st) - m0) <= 1e-13 * m0
        assert np.max(np.abs(st - prim)) > 1e-3  # the hump moved
        s.compute_interface_values()
        s.compute_fluxes()
        F = s.fluxes()
        p = np.nonzero(t["pe"] >= 0)[0]
        assert np.array_equal(F[p], -F[t["pe"][p]])


def test_periodic_pair_carries_the_interior_flux():
    """On a structured mesh, the strip between two columns of cells across
    the periodic seam sees the same edge states as an interior edge between
    the same columns of a mesh one column wider: HLL<Davis> (symmetric)
    fluxes through the pairs equal those interior fluxes at round-off."""
    n = 8
    xyz, topo = swe.StructTriangMesh(n, n, 1.0 / n)
    t = topo.tables()
    pairs = swe.periodic_pairs(xyz, topo, (1.0, 0.0))
    assert pairs.shape[0] == n
    per = topo.with_periodic(pairs)
    T = per.NumTriangles()
    rng = np.random.default_rng(5)
    cen = xyz[t["tp"]].mean(1)
    prim = np.stack([1.0 + 0.1 * np.cos(2 * np.pi * cen[:, 0]) + 0.01 * rng.random(T),
                     0.2 * np.sin(2 * np.pi * cen[:, 0]), 0.1 + 0 * cen[:, 0]], 1)
    s = O.OracleSolver(xyz, per.tables(), "HLL", "Davis")
    s.set_state(prim)
    s.compute_interface_values()
    s.compute_fluxes()
    F = s.fluxes()
    edg = s.edge_field()
    ep = t["ep"]
    for a, b in pairs:
        p = min(a, b)
        q = max(a, b)
        # left state: et0(p)'s slot at p; right state: et0(q)'s slot at q
        L, R = edg[2 * p], edg[2 * q]
        P0, P1 = xyz[ep[p, 0]], xyz[ep[p, 1]]
        cp = cen[t["et"][p, 0]]
        tx, ty = (P1[:2] - P0[:2]) / np.hypot(*(P1[:2] - P0[:2]))
        if (cp[0] - P0[0]) * ty - tx * (cp[1] - P0[1]) > 0:
            tx, ty = -tx, -ty
        nrm = np.array([ty, -tx])
        hl, hr = L[0], R[0]  # flat bottom
        ul, ur = L[1:] @ nrm, R[1:] @ nrm
        cl, cr = np.sqrt(hl), np.sqrt(hr)
        al, ar = min(0.0, min(ul - cl, ur - cr)), max(0.0, max(ul + cl, ur + cr))
        U = lambda h, v: np.array([h, h * v[0], h * v[1]])
        Fp = lambda h, v: np.array([h * (v @ nrm), h * (v @ nrm) * v[0] + 0.5 * h * h * nrm[0],
                                    h * (v @ nrm) * v[1] + 0.5 * h * h * nrm[1]])
        want = (ar * Fp(hl, L[1:]) - al * Fp(hr, R[1:]) + al * ar * (U(hr, R[1:]) - U(hl, L[1:]))) / (ar - al)
        assert np.allclose(F[p], want, rtol=1e-13, atol=1e-15)
        assert np.array_equal(F[q], -F[p])


# ------------------------------------------------------------------ GPU --

@pytest.mark.gpu
@pytest.mark.parametrize("scheme", ["Euler", "SSPRK2", "SSPRK3"])
@pytest.mark.parametrize("flux,wave", [("HLL", "Davis"), ("HLLC", "Einfeldt")])
def test_gpu_periodic_matches_oracle(scheme, flux, wave):
    """A wet/dry doubly periodic mesh (bumps, dry pockets, Coriolis, CFL
    steps) and a channel (periodic in x, free flow and walls in y)."""
    z = lambda x, y: 0.3 * np.exp(-((x - 0.4) ** 2 + (y - 0.6) ** 2) / 0.02) + 0.02 * np.cos(2 * np.pi * x)
    for x_per, y_per in ((True, True), (True, False)):
        xyz, per, _ = periodic_mesh(40, 7, x=x_per, y=y_per, z=z)
        if not y_per:
            t = per.tables()
            walls = np.nonzero(t["et"][:, 1] == swe.SOLID_WALL)[0]
            per = per.with_boundary_tag(walls[::2], swe.FREE_FLOW)
        t = per.tables()
        tp = t["tp"]
        bc = C.cell_bath(xyz, tp)
        rng = np.random.default_rng(8)
        T = tp.shape[0]
        h = np.maximum(0.0, 0.25 - bc)
        h[rng.random(T) < 0.05] = 0.0
        prim = np.stack([bc + h, 0.2 * (h > 0), -0.1 * (h > 0)], 1)
        fl = getattr(swe.Fluxes, flux)[getattr(swe.Wavespeeds, wave)]
        sd = swe.SpaceDisc(fl, swe.Domain(xyz, per), prim, cor=0.3)
        assert not sd.set_fused(True)  # the fused tiles hold no partners
        orc = O.OracleSolver(xyz, t, flux, wave, cor=0.3)
        orc.set_state(prim)
        sd.run(scheme, 12, dt=0.0, cfl=0.15)
        dts_o = orc.run(scheme, 12, dt=0.0, cfl=0.15)
        g, r = sd.GetVolField(), orc.state()
        print(f"periodic x={x_per} y={y_per} {scheme} {flux}<{wave}>: bitwise={np.array_equal(g, r)}", flush=True)
        assert np.array_equal(g, r)
        assert np.array_equal(sd.dts(12), dts_o)
        sd.ComputeInterfaceValues()
        sd.ComputeFluxes()
        orc.compute_interface_values()
        rr = orc.compute_fluxes()
        assert np.array_equal(sd.GetFluxes(), orc.fluxes())
        assert sd.GetMinLenToWavespeed() == rr


@pytest.mark.gpu
def test_gpu_periodic_ensemble_and_adjoint_refusal():
    from swe_fvm_amd.ensemble import Ensemble
    z = lambda x, y: 0.05 * np.cos(2 * np.pi * x)
    mem = []
    for k, n in enumerate((10, 17, 13)):
        xyz, per, plain = periodic_mesh(n, 20 + k, z=z)
        topo = per if k != 1 else plain
        tp = topo.tables()["tp"]
        bc = C.cell_bath(xyz, tp)
        cen = xyz[tp].mean(1)
        prim = np.stack([bc + 0.4 + 0.1 * np.exp(-((cen[:, 0] - 0.5) ** 2) / 0.01), 0.1 + 0 * bc, 0 * bc], 1)
        mem.append((swe.Domain(xyz, topo), prim))
    fl = swe.Fluxes.HLLC[swe.Wavespeeds.Einfeldt]
    ens = Ensemble(fl, [d for d, _ in mem], [p for _, p in mem])
    ens.run("SSPRK2", cfl=0.15, t_end=1e9, max_steps=8, graph=True)
    got = ens.states()
    for k, (dom, prim) in enumerate(mem):
        sd = swe.SpaceDisc(fl, dom, prim)
        sd.run("SSPRK2", 8, dt=0.0, cfl=0.15)
        assert np.array_equal(got[k], sd.GetVolField()), k
    sd = swe.SpaceDisc(fl, mem[0][0], mem[0][1])
    with pytest.raises(ValueError):
        sd.adjoint_run(np.ones_like(mem[0][1]), 2, dt=1e-3)
