"""FREE_FLOW boundary edges and linear friction (include/Includes.h:21 and
include/SpaceDisc.h:47: declared but unimplemented in the reference).  The
oracle's definitions are pinned by closed forms (a uniform stream leaves
through free-flow edges unchanged; friction decays it by the SSP-RK2
amplification factor of -tau), and the GPU matches the oracle bit for bit."""
import numpy as np
import pytest

import swe_fvm_amd as swe
from oracle import oracle as O
from tests import _common as C


def _flat(n, seed):
    xyz, topo = swe.jittered_mesh(n, 0.2, seed)
    xyz[:, 2] = 0.0
    return xyz, topo


def _tagged(topo, tag):
    t = topo.tables()
    b = np.nonzero(t["et"][:, 1] < 0)[0]
    return topo.with_boundary_tag(b, tag)


def _uniform(topo, w=1.0, u=0.3, v=-0.2):
    T = topo.NumTriangles()
    return np.tile([w, u, v], (T, 1)).astype(np.float64)


@pytest.mark.parametrize("flux,wave", [("HLL", "Einfeldt"), ("HLLC", "Einfeldt"), ("HLLC", "Rusanov")])
def test_uniform_stream_passes_free_flow_edges(flux, wave):
    """With the draining dt from the outflow F0 * L, a uniform stream leaves
    through free-flow edges unchanged; walls reflect it.  The reference's
    draining dt (the sum of F0 alone, src/TimeDisc.cpp:56-62) drops below the
    CFL step in wet cells and disturbs even this stream (D16)."""
    xyz, topo = _flat(12, 3)
    prim = _uniform(topo)
    ff = _tagged(topo, swe.FREE_FLOW)
    a = O.OracleSolver(xyz, ff.tables(), flux, wave, opts=O.OPT_DRAIN_LENGTH)
    a.set_state(prim)
    a.run("SSPRK2", 20, dt=0.0, cfl=0.15)
    assert np.max(np.abs(a.state() - prim)) < 1e-13
    b = O.OracleSolver(xyz, topo.tables(), flux, wave, opts=O.OPT_DRAIN_LENGTH)
    b.set_state(prim)
    b.run("SSPRK2", 20, dt=0.0, cfl=0.15)
    assert np.max(np.abs(b.state() - prim)) > 1e-3
    c = O.OracleSolver(xyz, ff.tables(), flux, wave)
    c.set_state(prim)
    c.run("SSPRK2", 20, dt=0.0, cfl=0.15)
    assert np.max(np.abs(c.state() - prim)) > 1e-4


def test_free_flow_flux_is_the_physical_flux():
    """On a free-flow edge the fluxer sees R = L: HLL and HLLC reduce to the
    physical flux (q.n, (q.n/h) q + h^2/2 n) of the inside edge state."""
    xyz, topo = _flat(6, 1)
    rng = np.random.default_rng(2)
    T = topo.NumTriangles()
    prim = np.stack([1.0 + 0.1 * rng.random(T), 0.2 * rng.standard_normal(T), 0.2 * rng.standard_normal(T)], 1)
    ff = _tagged(topo, swe.FREE_FLOW)
    t = ff.tables()
    for flux in ("HLL", "HLLC"):
        s = O.OracleSolver(xyz, t, flux, "Einfeldt")
        s.set_state(prim)
        s.compute_interface_values()
        s.compute_fluxes()
        F, edg = s.fluxes(), s.edge_field()
        for e in np.nonzero(t["et"][:, 1] < 0)[0][:10]:
            f = t["et"][e, 0]
            a, b = xyz[t["ep"][e, 0]], xyz[t["ep"][e, 1]]
            L = np.hypot(*(b - a)[:2])
            tx, ty = (b - a)[:2] / L
            cen = xyz[t["tp"][f]].mean(0)
            if (cen[0] - a[0]) * ty - tx * (cen[1] - a[1]) > 0:
                tx, ty = -tx, -ty
            n = np.array([ty, -tx])
            w_, u_, v_ = edg[2 * e]
            h = w_  # flat bottom
            qn = h * (u_ * n[0] + v_ * n[1])
            want = [qn, qn * u_ + 0.5 * h * h * n[0], qn * v_ + 0.5 * h * h * n[1]]
            assert np.allclose(F[e], want, rtol=1e-13, atol=1e-15), (flux, e)


def test_friction_decays_a_uniform_stream():
    """With OPT_FRICTION a uniform stream through free-flow edges decays by
    the SSP-RK2 factor 1 - k + k^2/2 (k = tau dt) per step; without the bit
    tau has no effect, as in the reference."""
    xyz, topo = _flat(10, 5)
    ff = _tagged(topo, swe.FREE_FLOW).tables()
    prim = _uniform(topo)
    tau, dt, n = 0.7, 2e-3, 30
    s = O.OracleSolver(xyz, ff, "HLLC", "Einfeldt", tau=tau, opts=O.OPT_FRICTION | O.OPT_DRAIN_LENGTH)
    s.set_state(prim)
    s.run("SSPRK2", n, dt=dt)
    k = tau * dt
    g = (1 - k + 0.5 * k * k) ** n
    st = s.state()
    assert np.allclose(st[:, 1], 0.3 * g, rtol=1e-12) and np.allclose(st[:, 2], -0.2 * g, rtol=1e-12)
    assert np.allclose(st[:, 0], 1.0, atol=1e-13)
    a = O.OracleSolver(xyz, ff, "HLLC", "Einfeldt", tau=tau, opts=O.OPT_DRAIN_LENGTH)
    b = O.OracleSolver(xyz, ff, "HLLC", "Einfeldt", tau=0.0, opts=O.OPT_DRAIN_LENGTH)
    for x in (a, b):
        x.set_state(prim)
        x.run("SSPRK2", 5, dt=dt)
    assert np.array_equal(a.state(), b.state())


def _coriolis_map(scheme, k):
    """Amplification matrix of one step on (hu, hv) for the source
    cor (h v, -h u) alone: M = I + k J with k = cor dt, J = [[0, 1], [-1, 0]];
    Shu-Osher stages (src/Solvers.cpp:5-59) combine it into M (Euler),
    (I + M^2) / 2 (SSPRK2) and I / 3 + M / 2 + M^3 / 6 (SSPRK3)."""
    I = np.eye(2)
    M = I + k * np.array([[0.0, 1.0], [-1.0, 0.0]])
    if scheme == "Euler":
        return M
    if scheme == "SSPRK2":
        return 0.5 * (I + M @ M)
    return I / 3 + 0.5 * M + (M @ M @ M) / 6


def _coriolis_case():
    xyz, topo = _flat(10, 7)
    ff = _tagged(topo, swe.FREE_FLOW)
    return xyz, ff, _uniform(topo)


@pytest.mark.parametrize("opts", [0, O.OPT_DRAIN_LENGTH])
@pytest.mark.parametrize("scheme", ["Euler", "SSPRK2", "SSPRK3"])
def test_coriolis_rotates_a_uniform_stream(scheme, opts):
    """The Coriolis source (src/SpaceDisc.cpp:29, src/TimeDisc.cpp:34) on a
    uniform stream over a flat bottom through free-flow edges: the fluxes
    cancel, the depth stays put and (u, v) follows the closed-form map of
    each scheme to 1e-13; the opposite sign of cor rotates the other way."""
    xyz, ff, prim = _coriolis_case()
    cor, dt, n = 0.9, 2e-3, 25
    s = O.OracleSolver(xyz, ff.tables(), "HLLC", "Einfeldt", cor=cor, opts=opts)
    s.set_state(prim)
    s.run(scheme, n, dt=dt)
    st = s.state()
    want = np.linalg.matrix_power(_coriolis_map(scheme, cor * dt), n) @ prim[0, 1:]
    assert np.allclose(st[:, 0], 1.0, rtol=0, atol=1e-13)
    assert np.max(np.abs(st[:, 1:] - want)) <= 1e-13 * np.abs(want).max()
    wrong = np.linalg.matrix_power(_coriolis_map(scheme, -cor * dt), n) @ prim[0, 1:]
    assert np.max(np.abs(st[:, 1:] - wrong)) > 1e-3


@pytest.mark.gpu
@pytest.mark.parametrize("scheme", ["Euler", "SSPRK2", "SSPRK3"])
def test_gpu_coriolis_rotates_a_uniform_stream(scheme):
    xyz, ff, prim = _coriolis_case()
    cor, dt, n = 0.9, 2e-3, 25
    fl = swe.Fluxes.HLLC[swe.Wavespeeds.Einfeldt]
    sd = swe.SpaceDisc(fl, swe.Domain(xyz, ff), prim, cor=cor)
    sd.run(scheme, n, dt=dt)
    g = sd.GetVolField()
    want = np.linalg.matrix_power(_coriolis_map(scheme, cor * dt), n) @ prim[0, 1:]
    print(f"{scheme}: max |(u, v) - closed form| {np.max(np.abs(g[:, 1:] - want)):.2e}", flush=True)
    assert np.allclose(g[:, 0], 1.0, rtol=0, atol=1e-13)
    assert np.max(np.abs(g[:, 1:] - want)) <= 1e-13 * np.abs(want).max()
    orc = O.OracleSolver(xyz, ff.tables(), "HLLC", "Einfeldt", cor=cor)
    orc.set_state(prim)
    orc.run(scheme, n, dt=dt)
    assert np.array_equal(g, orc.state())


@pytest.mark.gpu
@pytest.mark.parametrize("drain_len", [False, True])
@pytest.mark.parametrize("scheme", ["Euler", "SSPRK2", "SSPRK3"])
@pytest.mark.parametrize("flux,wave", [("HLL", "Davis"), ("HLLC", "Einfeldt")])
def test_gpu_free_flow_and_friction_match_oracle(scheme, flux, wave, drain_len):
    xyz, topo, prim = C.wet_dry_mesh(30, seed=6)
    t = topo.tables()
    b = np.nonzero(t["et"][:, 1] < 0)[0]
    mixed = topo.with_boundary_tag(b[::2], swe.FREE_FLOW)  # every other boundary edge
    fl = getattr(swe.Fluxes, flux)[getattr(swe.Wavespeeds, wave)]
    opts = swe.OPT_FRICTION | (swe.OPT_DRAIN_LENGTH if drain_len else 0)
    sd = swe.SpaceDisc(fl, swe.Domain(xyz, mixed), prim, cor=0.4, tau=0.8, options=opts)
    orc = O.OracleSolver(xyz, mixed.tables(), flux, wave, cor=0.4, tau=0.8, opts=opts)
    orc.set_state(prim)
    sd.run(scheme, 12, dt=0.0, cfl=0.15)
    orc.run(scheme, 12, dt=0.0, cfl=0.15)
    g, r = sd.GetVolField(), orc.state()
    print(f"{scheme} {flux}<{wave}> free flow + friction, drain length {drain_len}: "
          f"bitwise={np.array_equal(g, r)}", flush=True)
    tp = t["tp"]
    h = lambda p: np.stack([p[:, 0] - C.cell_bath(xyz, tp), p[:, 1], p[:, 2]], 1)
    assert C.max_rel_diff(h(g), h(r)) <= 1e-11
    # getters on the same tables: fluxes, the CFL length, one RHS
    sd.ComputeInterfaceValues()
    sd.ComputeFluxes()
    orc.compute_interface_values()
    rr = orc.compute_fluxes()
    assert C.max_rel_diff(sd.GetFluxes(), orc.fluxes()) <= 1e-11
    assert abs(sd.GetMinLenToWavespeed() - rr) <= 1e-11 * rr
    i = int(np.nonzero((C.cell_area(xyz, tp) > 0) & (g[:, 1] != 0))[0][0])
    td = swe.TimeDisc(sd)
    assert np.allclose(td.RHS(i, 1e-3), orc.rhs(i, 1e-3), rtol=1e-11, atol=1e-18)
