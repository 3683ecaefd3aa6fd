"""The adjoint of the step (swe_adjoint_run) against central finite
differences of the oracle (the CPU restatement of the reference solver) on
a 72-cell mesh and of the GPU forward (itself bit-identical to the oracle)
on larger ones, in random directions of the initial state and of the node
bathymetry; against the tangent-linear build by duality; the checkpoint
interval does not change the result; the HEAD options are refused; the
torch.autograd.Function gives the same gradients through either binding.
Finite differences must not cross the limiter's all-or-nothing decisions:
the perturbations are ~1e-8 against level differences ~1e-3."""
import numpy as np
import pytest

import swe_fvm_amd as swe
from swe_fvm_amd import adjoint as AD
from tests import _common as C

pytestmark = pytest.mark.gpu


def wet_case(n=16, seed=3, amp=0.3):
    """Deep water over bumps, every cell wet above its highest vertex."""
    xyz, topo = swe.jittered_mesh(n, 0.25, seed)
    rng = np.random.default_rng(seed)
    x, y = xyz[:, 0], xyz[:, 1]
    z = np.zeros_like(x)
    for _ in range(4):
        cx, cy = rng.uniform(0.2, 0.8, 2)
        z += amp * rng.uniform(0.3, 1.0) * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * 0.1 ** 2))
    xyz[:, 2] = z
    dom = swe.Domain(xyz, topo)
    c = dom.centroids()
    # a tilt keeps neighbouring levels apart, so the finite differences
    # below do not cross the limiter's all-or-nothing decisions
    w = 1.0 + 0.03 * c[:, 0] + 0.05 * np.exp(-((c[:, 0] - 0.4) ** 2 + (c[:, 1] - 0.5) ** 2) / 0.02)
    u = 0.1 * np.sin(2 * c[:, 1])
    v = -0.05 * np.cos(3 * c[:, 0])
    return dom, np.stack([w, u, v], 1)


def directional_fd(sd, prim0, z, nsteps, fn, dp, dz, scheme, dt, eps):
    """Central difference at eps and at eps / 4; None when they disagree,
    i.e. when the perturbation crossed a limiter or wave-speed switch (the
    step is piecewise smooth) and the difference measures the jump."""
    def J(p, zz):
        sd.set_bathymetry(zz)
        sd.SetVolField(p)
        sd.run(scheme, nsteps, dt=dt)
        return fn(sd.GetVolField())[0]

    def cd(e):
        return (J(prim0 + e * dp, z + e * dz) - J(prim0 - e * dp, z - e * dz)) / (2 * e)
    a, b = cd(eps), cd(eps / 4)
    sd.set_bathymetry(z)
    return a if abs(a - b) <= 1e-5 * abs(a) + 1e-12 else None


@pytest.mark.parametrize("scheme,nsteps,flux", [("Euler", 1, "HLLC"), ("SSPRK2", 12, "HLLC"),
                                                ("SSPRK3", 6, "HLL"), ("SSPRK2", 12, "HLL")])
def test_adjoint_matches_finite_differences(scheme, nsteps, flux):
    dom, prim0 = wet_case()
    fl = getattr(swe.Fluxes, flux)[swe.Wavespeeds.Einfeldt]
    sd = swe.SpaceDisc(fl, dom, prim0, cor=0.3)
    z = dom.geometry[:, 2].copy()
    rng = np.random.default_rng(7)
    cvec = rng.standard_normal(prim0.shape)

    def fn(p):  # a linear functional of the whole final state
        return float(np.sum(cvec * p)), cvec

    dt = 2e-3
    J, s0, zb, ms = AD.run_and_adjoint(sd, prim0, z, nsteps, fn, scheme=scheme, dt=dt, checkpoint_every=5)
    checked = 0
    for trial in range(4):
        dp = rng.standard_normal(prim0.shape) * np.array([1e-2, 1e-2, 1e-2])
        dz = rng.standard_normal(z.shape) * 1e-2 if trial > 0 else np.zeros_like(z)
        fd = directional_fd(sd, prim0, z, nsteps, fn, dp, dz, scheme, dt, 1e-6)
        ad = float(np.sum(s0 * dp) + np.sum(zb * dz))
        if fd is None:
            print(f"{scheme} {flux} x{nsteps} trial {trial}: the difference crossed a switch", flush=True)
            continue
        print(f"{scheme} {flux} x{nsteps} trial {trial}: adjoint {ad:.12e} fd {fd:.12e} "
              f"rel {abs(ad - fd) / abs(fd):.2e}", flush=True)
        assert abs(ad - fd) <= 1e-5 * abs(fd) + 1e-12
        checked += 1
    assert checked >= 2


@pytest.mark.parametrize("scheme,nsteps,flux", [("Euler", 1, "HLLC"), ("SSPRK2", 3, "HLL"),
                                                ("SSPRK3", 2, "HLLC")])
def test_adjoint_matches_differences_of_the_oracle(scheme, nsteps, flux):
    """The GPU adjoint against directional central differences of the oracle
    (the CPU restatement of the reference solver), in state and bathymetry
    directions, on a 72-cell mesh."""
    from oracle import oracle as O
    dom, prim0 = wet_case(n=6, seed=11)
    assert prim0.shape[0] == 72
    fl = getattr(swe.Fluxes, flux)[swe.Wavespeeds.Einfeldt]
    sd = swe.SpaceDisc(fl, dom, prim0, cor=0.3)
    xyz0 = dom.geometry.copy()
    z = xyz0[:, 2].copy()
    tables = dom.topology.tables()
    rng = np.random.default_rng(17)
    cvec = rng.standard_normal(prim0.shape)
    dt = 2e-3
    J, s0, zb, _ = AD.run_and_adjoint(sd, prim0, z, nsteps, lambda p: (float(np.sum(cvec * p)), cvec),
                                      scheme=scheme, dt=dt, checkpoint_every=2)

    def J_oracle(p, zz):
        xyz = xyz0.copy()
        xyz[:, 2] = zz
        o = O.OracleSolver(xyz, tables, flux, "Einfeldt", cor=0.3)
        o.set_state(p)
        o.run(scheme, nsteps, dt=dt)
        return float(np.sum(cvec * o.state()))

    assert abs(J_oracle(prim0, z) - J) <= 1e-13 * abs(J)  # the forward is the oracle's
    checked = 0
    for trial in range(4):
        dp = rng.standard_normal(prim0.shape) * 1e-2
        dz = rng.standard_normal(z.shape) * 1e-2 if trial > 0 else np.zeros_like(z)
        cd = [(J_oracle(prim0 + e * dp, z + e * dz) - J_oracle(prim0 - e * dp, z - e * dz)) / (2 * e)
              for e in (1e-6, 2.5e-7)]
        if abs(cd[0] - cd[1]) > 1e-5 * abs(cd[0]) + 1e-12:
            print(f"{scheme} {flux} trial {trial}: the difference crossed a switch", flush=True)
            continue
        ad = float(np.sum(s0 * dp) + np.sum(zb * dz))
        print(f"{scheme} {flux} x{nsteps} trial {trial}: adjoint {ad:.12e} oracle fd {cd[0]:.12e} "
              f"rel {abs(ad - cd[0]) / abs(cd[0]):.2e}", flush=True)
        assert abs(ad - cd[0]) <= 1e-5 * abs(cd[0]) + 1e-12
        checked += 1
    assert checked >= 2


def test_gauge_loss_gradient_and_checkpoint_invariance():
    dom, prim0 = wet_case(n=20, seed=5)
    fl = swe.Fluxes.HLLC[swe.Wavespeeds.Einfeldt]
    sd = swe.SpaceDisc(fl, dom, prim0)
    z = dom.geometry[:, 2].copy()
    rng = np.random.default_rng(1)
    gauges = rng.choice(prim0.shape[0], 16, replace=False)
    fn = AD.gauge_misfit(gauges, prim0[gauges, 0] + 0.01)
    outs = []
    for ck in (1, 7, 40):
        J, s0, zb, ms = AD.run_and_adjoint(sd, prim0, z, 40, fn, dt=2e-3, checkpoint_every=ck)
        outs.append((J, s0, zb))
    for o in outs[1:]:
        assert o[0] == outs[0][0] and np.array_equal(o[1], outs[0][1]) and np.array_equal(o[2], outs[0][2])
    J, s0, zb = outs[0]
    checked = 0
    for trial in range(3):
        dp = np.zeros_like(prim0)
        dp[:, 0] = rng.standard_normal(prim0.shape[0]) * 1e-2
        dz = rng.standard_normal(z.shape) * 1e-2
        fd = directional_fd(sd, prim0, z, 40, fn, dp, dz, "SSPRK2", 2e-3, 1e-6)
        ad = float(np.sum(s0 * dp) + np.sum(zb * dz))
        if fd is None:
            continue
        print(f"gauge loss trial {trial}: adjoint {ad:.12e} fd {fd:.12e}", flush=True)
        assert abs(ad - fd) <= 1e-5 * abs(fd) + 1e-12
        checked += 1
    assert checked >= 2


def draining_dt(sd, dom, prim, length=True):
    """Per-cell draining dt h A / sum of outflows F0 L (with OPT_DRAIN_LENGTH;
    F0 alone otherwise, as the reference) from the GPU's fluxes of prim."""
    sd.SetVolField(prim)
    sd.ComputeInterfaceValues()
    sd.ComputeFluxes()
    F0 = sd.GetFluxes()[:, 0]
    t = dom.topology.tables()
    xyz = dom.geometry
    L = np.hypot(*(xyz[t["ep"][:, 1], :2] - xyz[t["ep"][:, 0], :2]).T) if length else np.ones(len(F0))
    out = np.zeros(prim.shape[0])
    for side, sg in ((0, 1.0), (1, -1.0)):
        c = t["et"][:, side]
        m = c >= 0
        np.add.at(out, c[m], np.maximum(0.0, sg * F0[m] * L[m]))
    h = prim[:, 0] - xyz[t["tp"], 2].mean(1)
    with np.errstate(divide="ignore"):
        return np.where(out > 1e-12, h * dom.areas() / out, np.inf)


@pytest.mark.parametrize("scheme", ["Euler", "SSPRK2"])
def test_adjoint_through_active_draining_dt(scheme):
    """Shallow water with a rough velocity field at CFL 0.8: the draining dt
    of some cells is below the stage dt, so the flux and depth terms of its
    adjoint are exercised (OPT_DRAIN_LENGTH)."""
    xyz, topo = swe.jittered_mesh(16, 0.25, 4)
    rng = np.random.default_rng(4)
    xyz[:, 2] = 0.01 * rng.random(xyz.shape[0])
    dom = swe.Domain(xyz, topo)
    c = dom.centroids()
    T = c.shape[0]
    prim0 = np.stack([0.06 + 0.01 * c[:, 0], rng.uniform(-0.4, 0.4, T), rng.uniform(-0.4, 0.4, T)], 1)
    fl = swe.Fluxes.HLLC[swe.Wavespeeds.Einfeldt]
    sd = swe.SpaceDisc(fl, dom, prim0, options=swe.OPT_DRAIN_LENGTH)
    sd.SetVolField(prim0)
    sd.ComputeInterfaceValues()
    sd.ComputeFluxes()
    dt = 0.8 * sd.GetMinLenToWavespeed()
    active = int(np.sum(draining_dt(sd, dom, prim0) < dt))
    print(f"{scheme}: dt {dt:.3e}, {active} of {T} cells with an active draining dt", flush=True)
    assert active > 0
    z = dom.geometry[:, 2].copy()
    wvec = rng.standard_normal(prim0.shape)

    def fn(p):
        return float(np.sum(wvec * p)), wvec

    nsteps = 2
    J, s0, zb, ms = AD.run_and_adjoint(sd, prim0, z, nsteps, fn, scheme=scheme, dt=dt, checkpoint_every=1)
    checked = 0
    for trial in range(4):
        dp = rng.standard_normal(prim0.shape) * np.array([1e-3, 1e-2, 1e-2])
        dz = rng.standard_normal(z.shape) * 1e-3 if trial > 0 else np.zeros_like(z)
        fd = directional_fd(sd, prim0, z, nsteps, fn, dp, dz, scheme, dt, 1e-6)
        ad = float(np.sum(s0 * dp) + np.sum(zb * dz))
        if fd is None:
            print(f"draining {scheme} trial {trial}: the difference crossed a switch", flush=True)
            continue
        print(f"draining {scheme} trial {trial}: adjoint {ad:.12e} fd {fd:.12e} "
              f"rel {abs(ad - fd) / abs(fd):.2e}", flush=True)
        assert abs(ad - fd) <= 1e-5 * abs(fd) + 1e-12
        checked += 1
    assert checked >= 2


@pytest.mark.parametrize("scheme,flux,opts", [("Euler", "HLLC", 0), ("SSPRK2", "HLLC", 0), ("SSPRK3", "HLL", 0),
                                               ("SSPRK2", "HLL", swe.OPT_FRICTION),
                                               ("SSPRK2", "HLLC", swe.OPT_DRAIN_LENGTH)])
def test_tangent_linear_duality(scheme, flux, opts):
    """The dual-number build of the stage kernels (swe_tangent_run): its
    values are the forward's bit for bit, its tangent agrees with central
    differences, and with the adjoint it satisfies the duality
    <ybar, J xdot> = <J^T ybar, xdot> to round-off."""
    dom, prim0 = wet_case(n=16, seed=3)
    fl = getattr(swe.Fluxes, flux)[swe.Wavespeeds.Einfeldt]
    sd = swe.SpaceDisc(fl, dom, prim0, cor=0.3, tau=0.5, options=opts)
    z = dom.geometry[:, 2].copy()
    rng = np.random.default_rng(11)
    n, dt = 8, 2e-3
    sd.SetVolField(prim0)
    sd.run(scheme, n, dt=dt)
    ref = sd.GetVolField()
    for trial in range(3):
        pdot = rng.standard_normal(prim0.shape) * np.array([1e-2, 1e-2, 1e-2])
        zdot = rng.standard_normal(z.shape) * 1e-2 if trial > 0 else None
        sd.SetVolField(prim0)
        tdot, _ = sd.tangent_run(pdot, n, z_dot=zdot, scheme=scheme, dt=dt)
        assert np.array_equal(sd.GetVolField(), ref)
        ybar = rng.standard_normal(prim0.shape)
        sd.SetVolField(prim0)
        s0, zb, _ = sd.adjoint_run(ybar, n, scheme=scheme, dt=dt, checkpoint_every=3)
        lhs = float(np.sum(ybar * tdot))
        rhs = float(np.sum(s0 * pdot) + (np.sum(zb * zdot) if zdot is not None else 0.0))
        scale = float(np.sum(np.abs(ybar * tdot)))
        print(f"{scheme} {flux} opts {opts} trial {trial}: <ybar, J xdot> {lhs:.15e} <J^T ybar, xdot> {rhs:.15e} "
              f"diff/scale {abs(lhs - rhs) / scale:.1e}", flush=True)
        assert abs(lhs - rhs) <= 1e-12 * scale
        # the tangent against central differences of the forward
        fn = lambda p, c=ybar: (float(np.sum(c * p)), c)
        fd = directional_fd(sd, prim0, z, n, fn, pdot, zdot if zdot is not None else np.zeros_like(z),
                            scheme, dt, 1e-6)
        if fd is not None:
            assert abs(fd - lhs) <= 1e-5 * abs(fd) + 1e-12


def test_duality_on_the_inversion_workload():
    """The bathymetry-inversion workload at 1,800 cells with a swirling
    initial flow, 100 SSP-RK2 steps at CFL ~0.2: with the reference's
    draining dt (outflow F0 without edge length) the draining dt is below the
    stage dt in many moving cells (as in the 250k-cell inversion, whose cells
    are 140 times smaller), so its adjoint is exercised broadly; duality with
    the tangent-linear build holds."""
    from swe_fvm_amd import workloads as W
    dom, prim0, z_true, gauges = W.adjoint_inversion(n=30, n_gauges=16)
    c0 = dom.centroids()
    prim0[:, 1] = 0.3 * np.sin(2 * np.pi * c0[:, 1])
    prim0[:, 2] = 0.3 * np.cos(2 * np.pi * c0[:, 0])
    sd = swe.SpaceDisc(swe.Fluxes.HLLC[swe.Wavespeeds.Einfeldt], dom, prim0)
    z = dom.geometry[:, 2].copy()
    n, dt = 100, 5e-3
    sd.SetVolField(prim0)
    sd.run("SSPRK2", n // 2, dt=dt)
    mid = sd.GetVolField()
    active = int(np.sum(draining_dt(sd, dom, mid, length=False) < dt))
    print(f"inversion workload: {active} of {mid.shape[0]} cells with an active draining dt", flush=True)
    assert active > mid.shape[0] // 10
    rng = np.random.default_rng(5)
    c = dom.centroids()
    fn = AD.gauge_misfit(gauges, prim0[gauges, 0] + 1e-3)
    J, s0, zb, _ = AD.run_and_adjoint(sd, prim0, z, n, fn, dt=dt, checkpoint_every=40)
    ybar = fn(sd.GetVolField())[1]
    for _ in range(3):
        pdot = np.zeros_like(prim0)
        pdot[:, 0] = W.smooth_field(c[:, 0], c[:, 1], rng, 1e-2)
        zdot = W.smooth_field(dom.geometry[:, 0], dom.geometry[:, 1], rng, 1e-2)
        sd.SetVolField(prim0)
        tdot, _ = sd.tangent_run(pdot, n, z_dot=zdot, dt=dt)
        lhs = float(np.sum(ybar * tdot))
        rhs = float(np.sum(s0 * pdot) + np.sum(zb * zdot))
        scale = float(np.sum(np.abs(ybar * tdot)))
        print(f"inversion duality: {lhs:.15e} {rhs:.15e} diff/scale {abs(lhs - rhs) / scale:.1e}", flush=True)
        assert abs(lhs - rhs) <= 1e-12 * scale


def test_duality_with_free_flow_edges():
    """Half the boundary edges FREE_FLOW (R = L across them; the adjoint adds
    the outside slot's adjoint to the inside one): duality with the
    tangent-linear build, state and bathymetry directions, a flow leaving
    through the open edges."""
    dom0, prim0 = wet_case(n=14, seed=8)
    t = dom0.topology.tables()
    b = np.nonzero(t["et"][:, 1] < 0)[0]
    dom = swe.Domain(dom0.geometry, dom0.topology.with_boundary_tag(b[::2], swe.FREE_FLOW))
    prim0 = prim0.copy()
    prim0[:, 1] += 0.2
    sd = swe.SpaceDisc(swe.Fluxes.HLLC[swe.Wavespeeds.Einfeldt], dom, prim0, cor=0.2)
    z = dom.geometry[:, 2].copy()
    rng = np.random.default_rng(4)
    n, dt = 10, 2e-3
    ybar = rng.standard_normal(prim0.shape)
    sd.SetVolField(prim0)
    s0, zb, _ = sd.adjoint_run(ybar, n, dt=dt, checkpoint_every=4)
    for _ in range(2):
        pdot = rng.standard_normal(prim0.shape) * 1e-2
        zdot = rng.standard_normal(z.shape) * 1e-2
        sd.SetVolField(prim0)
        tdot, _ = sd.tangent_run(pdot, n, z_dot=zdot, dt=dt)
        lhs = float(np.sum(ybar * tdot))
        rhs = float(np.sum(s0 * pdot) + np.sum(zb * zdot))
        scale = float(np.sum(np.abs(ybar * tdot)))
        print(f"free flow duality: {lhs:.15e} {rhs:.15e} diff/scale {abs(lhs - rhs) / scale:.1e}", flush=True)
        assert abs(lhs - rhs) <= 1e-12 * scale


@pytest.mark.parametrize("flux", ["HLL", "HLLC"])
def test_duality_dam_break_onto_a_dry_bed(flux):
    """A flat-bottom dam break onto a dry bed stays in the supported regime
    (cells are dry or fully wet) and exercises the dry-cell reverse (level =
    bathymetry, dry neighbours zero the gradient), the velocity
    desingularisation of thin fronts (h < 1e-3) and the zero draining dt of
    dry cells: the tangent-linear values equal the forward's bit for bit and
    the duality holds."""
    xyz, topo = swe.jittered_mesh(16, 0.25, 6)
    xyz[:, 2] = 0.0
    dom = swe.Domain(xyz, topo)
    c = dom.centroids()
    T = c.shape[0]
    prim0 = np.stack([np.where(c[:, 0] < 0.4, 1.0, 0.0), np.zeros(T), np.zeros(T)], 1)
    sd = swe.SpaceDisc(getattr(swe.Fluxes, flux)[swe.Wavespeeds.Einfeldt], dom, prim0)
    z = dom.geometry[:, 2].copy()
    n, dt = 20, 2e-3
    sd.SetVolField(prim0)
    sd.run("SSPRK2", n, dt=dt)
    ref = sd.GetVolField()
    h = ref[:, 0] - C.cell_bath(xyz, topo.tables()["tp"])
    thin = int(np.sum((h > 1e-12) & (h < 1e-3)))
    dry = int(np.sum(h <= 1e-12))
    print(f"{flux}: {dry} dry cells, {thin} thin-front cells after {n} steps", flush=True)
    assert dry > 0 and thin > 0
    rng = np.random.default_rng(2)
    ybar = rng.standard_normal(prim0.shape)
    sd.SetVolField(prim0)
    s0, zb, _ = sd.adjoint_run(ybar, n, dt=dt, checkpoint_every=7)
    for _ in range(2):
        pdot = rng.standard_normal(prim0.shape) * 1e-2
        zdot = rng.standard_normal(z.shape) * 1e-3
        sd.SetVolField(prim0)
        tdot, _ = sd.tangent_run(pdot, n, z_dot=zdot, dt=dt)
        assert np.array_equal(sd.GetVolField(), ref)
        lhs = float(np.sum(ybar * tdot))
        rhs = float(np.sum(s0 * pdot) + np.sum(zb * zdot))
        scale = float(np.sum(np.abs(ybar * tdot)))
        print(f"{flux} dry bed duality: {lhs:.15e} {rhs:.15e} diff/scale {abs(lhs - rhs) / scale:.1e}", flush=True)
        assert abs(lhs - rhs) <= 1e-12 * scale


def test_stacked_batch_equals_members():
    """A batch (stack_domains) gives each member's forward and adjoint bit
    for bit: the kernels never mix members."""
    (da, pa), (db, pb) = wet_case(n=12, seed=3), wet_case(n=10, seed=4)
    big, off = swe.stack_domains([da, db])
    fl = swe.Fluxes.HLLC[swe.Wavespeeds.Einfeldt]
    prim = np.concatenate([pa, pb])
    z = big.geometry[:, 2].copy()
    rng = np.random.default_rng(2)
    cvec = rng.standard_normal(prim.shape)
    fn = lambda p, c=cvec: (float(np.sum(c * p)), c)
    sd = swe.SpaceDisc(fl, big, prim, cor=0.2)
    J, s0, zb, _ = AD.run_and_adjoint(sd, prim, z, 6, fn, dt=2e-3, checkpoint_every=4)
    fin = sd.GetVolField()
    Ta, Va = pa.shape[0], da.geometry.shape[0]
    for dom, p0, cs, vs in ((da, pa, slice(0, Ta), slice(0, Va)), (db, pb, slice(Ta, None), slice(Va, None))):
        s1 = swe.SpaceDisc(fl, dom, p0, cor=0.2)
        c1 = cvec[cs]
        _, a0, az, _ = AD.run_and_adjoint(s1, p0, dom.geometry[:, 2].copy(), 6,
                                          lambda p, c=c1: (float(np.sum(c * p)), c), dt=2e-3, checkpoint_every=4)
        assert np.array_equal(s1.GetVolField(), fin[cs])
        assert np.array_equal(a0, s0[cs]) and np.array_equal(az, zb[vs])


def shoreline_case(n=20, seed=2):
    """Emerging bumps under a sloping, non-flat surface: dry, partially wet
    (below their highest vertex: PartWet1 cube-root and bisection levels,
    PartWet2 planes) and fully wet cells.  The surface is nowhere flat (on a
    flat surface every gradient is zero and the TVD check sits on its
    boundary, a kink that finite differences average over).  Returns (domain,
    prim, mask of dry cells, mask of their nodes): perturbing a dry cell's
    level or its bathymetry moves it off the wet/dry kink h = 0."""
    xyz, topo, _ = C.wet_dry_mesh(n, seed=seed)
    dom = swe.Domain(xyz, topo)
    tp = topo.tables()["tp"]
    c = dom.centroids()
    bc = C.cell_bath(xyz, tp)
    level = 0.8 + 0.03 * np.sin(4 * c[:, 0] + 0.3) * np.cos(3 * c[:, 1]) + 0.1 * c[:, 0]
    h = np.maximum(0.0, level - bc)
    rng = np.random.default_rng(seed + 1)
    prim = np.stack([bc + h, rng.normal(0, 0.1, len(h)) * (h > 0), rng.normal(0, 0.1, len(h)) * (h > 0)], 1)
    dry = h <= 1e-12
    dry_nodes = np.zeros(xyz.shape[0], bool)
    dry_nodes[tp[dry].ravel()] = True
    return dom, prim, dry, dry_nodes


@pytest.mark.parametrize("scheme", ["Euler", "SSPRK2"])
def test_tangent_through_shorelines(scheme):
    """Through shorelines (partially wet cells below their highest vertex:
    PartWet1's cube-root and bisection levels, PartWet2's planes through the
    node maxima): the tangent-linear values equal the forward's bit for bit
    and the tangent agrees with central differences of the forward."""
    dom, prim0, dry, dry_nodes = shoreline_case()
    xyz, tp = dom.geometry, dom.topology.tables()["tp"]
    sd = swe.SpaceDisc(swe.Fluxes.HLLC[swe.Wavespeeds.Einfeldt], dom, prim0)
    sd.ComputeInterfaceValues()
    cls = sd.CellClass()
    below = int(np.sum((cls == 1) & (prim0[:, 0] < xyz[tp, 2].max(1)) & ~dry))
    print(f"{below} partially wet cells below their highest vertex, {int(dry.sum())} dry", flush=True)
    assert below > 10
    z = dom.geometry[:, 2].copy()
    n, dt = 4, 1e-3
    sd.SetVolField(prim0)
    sd.run(scheme, n, dt=dt)
    ref = sd.GetVolField()
    rng = np.random.default_rng(12)
    ybar = rng.standard_normal(prim0.shape)
    fn = lambda p, c=ybar: (float(np.sum(c * p)), c)
    checked = 0
    for trial in range(4):
        pdot = rng.standard_normal(prim0.shape) * 1e-2
        pdot[dry] = 0.0
        zdot = rng.standard_normal(z.shape) * 1e-2 if trial % 2 else np.zeros_like(z)
        zdot[dry_nodes] = 0.0
        sd.SetVolField(prim0)
        tdot, _ = sd.tangent_run(pdot, n, z_dot=zdot, scheme=scheme, dt=dt)
        assert np.array_equal(sd.GetVolField(), ref)
        lhs = float(np.sum(ybar * tdot))
        fd = directional_fd(sd, prim0, z, n, fn, pdot, zdot, scheme, dt, 1e-6)
        if fd is None:
            print(f"shoreline {scheme} trial {trial}: the difference crossed a switch", flush=True)
            continue
        print(f"shoreline {scheme} trial {trial}: tangent {lhs:.12e} fd {fd:.12e} "
              f"rel {abs(lhs - fd) / abs(fd):.2e}", flush=True)
        assert abs(lhs - fd) <= 1e-5 * abs(fd) + 1e-12
        checked += 1
    assert checked >= 2


@pytest.mark.parametrize("scheme,flux", [("Euler", "HLLC"), ("SSPRK2", "HLLC"), ("SSPRK3", "HLL")])
def test_adjoint_through_shorelines(scheme, flux):
    """The adjoint through partially wet cells below their highest vertex
    (PartWet1/PartWet2 local Jacobians, the node maxima routed to the cells
    attaining them, neighbours' cube-root and bisection levels): duality with
    the tangent-linear build (itself checked against central differences in
    test_tangent_through_shorelines), state and bathymetry directions."""
    dom, prim0, dry, dry_nodes = shoreline_case()
    fl = getattr(swe.Fluxes, flux)[swe.Wavespeeds.Einfeldt]
    sd = swe.SpaceDisc(fl, dom, prim0, cor=0.2)
    z = dom.geometry[:, 2].copy()
    n, dt = 6, 1e-3
    rng = np.random.default_rng(21)
    ybar = rng.standard_normal(prim0.shape)
    sd.SetVolField(prim0)
    s0, zb, _ = sd.adjoint_run(ybar, n, scheme=scheme, dt=dt, checkpoint_every=4)
    for _ in range(3):
        pdot = rng.standard_normal(prim0.shape) * 1e-2
        zdot = rng.standard_normal(z.shape) * 1e-2
        sd.SetVolField(prim0)
        tdot, _ = sd.tangent_run(pdot, n, z_dot=zdot, scheme=scheme, dt=dt)
        lhs = float(np.sum(ybar * tdot))
        rhs = float(np.sum(s0 * pdot) + np.sum(zb * zdot))
        scale = float(np.sum(np.abs(ybar * tdot)))
        print(f"shoreline {scheme} {flux} duality: {lhs:.15e} {rhs:.15e} diff/scale {abs(lhs - rhs) / scale:.1e}",
              flush=True)
        assert abs(lhs - rhs) <= 1e-12 * scale


@pytest.mark.parametrize("opts", [swe.OPT_HEAD_GRADIENT, swe.OPT_EXACT_BALANCE])
def test_adjoint_refuses_head_options(opts):
    """The HEAD-compatibility and exact-balance groupings are not
    differentiated: the adjoint and tangent runs refuse them."""
    xyz, topo, prim = C.wet_dry_mesh(12, seed=2)
    sd = swe.SpaceDisc(swe.Fluxes.HLLC[swe.Wavespeeds.Einfeldt], swe.Domain(xyz, topo), prim,
                       options=opts)
    with pytest.raises(ValueError):
        sd.adjoint_run(np.ones_like(prim), 2, dt=1e-3)
    with pytest.raises(ValueError):
        sd.tangent_run(np.ones_like(prim), 2, dt=1e-3)


def test_autograd_function_matches():
    torch = pytest.importorskip("torch")
    dom, prim0 = wet_case(n=12, seed=9)
    sd = swe.SpaceDisc(swe.Fluxes.HLLC[swe.Wavespeeds.Einfeldt], dom, prim0)
    z = dom.geometry[:, 2].copy()
    p = torch.tensor(prim0, dtype=torch.float64, requires_grad=True)
    zt = torch.tensor(z, dtype=torch.float64, requires_grad=True)
    out = AD.ShallowWaterRun.apply(p, zt, sd, 10, "SSPRK2", 2e-3, 4)
    loss = (out[:, 0] ** 2).sum()
    loss.backward()
    fn = lambda q: (float(np.sum(q[:, 0] ** 2)), np.stack([2 * q[:, 0], 0 * q[:, 0], 0 * q[:, 0]], 1))
    J, s0, zb, ms = AD.run_and_adjoint(sd, prim0, z, 10, fn, dt=2e-3, checkpoint_every=4)
    assert abs(loss.item() - J) <= 1e-12 * abs(J)
    assert np.allclose(p.grad.numpy(), s0, rtol=1e-12, atol=1e-15)
    assert np.allclose(zt.grad.numpy(), zb, rtol=1e-12, atol=1e-15)


def test_autograd_function_through_the_pybind_solver():
    """The autograd Function on the pybind module's Solver: loss and
    gradients equal those through swe_fvm_amd.SpaceDisc bit for bit (both
    call the same C ABI), and the Solver's tangent satisfies the duality
    with its adjoint."""
    torch = pytest.importorskip("torch")
    from tests import _common as C
    SWE_FVM = C.swe_fvm_module()
    dom, prim0 = wet_case(n=12, seed=9)
    xyz_p, topo_p = SWE_FVM.jittered_mesh(12, 0.25, 9)
    assert np.array_equal(xyz_p[:, :2], dom.geometry[:, :2])
    ps = SWE_FVM.Solver(topo_p, dom.geometry, "HLLC", "Einfeldt")
    sd = swe.SpaceDisc(swe.Fluxes.HLLC[swe.Wavespeeds.Einfeldt], dom, prim0)
    z = dom.geometry[:, 2].copy()
    res = []
    for solver in (sd, ps):
        p = torch.tensor(prim0, dtype=torch.float64, requires_grad=True)
        zt = torch.tensor(z, dtype=torch.float64, requires_grad=True)
        out = AD.ShallowWaterRun.apply(p, zt, solver, 10, "SSPRK2", 2e-3, 4)
        loss = (out[:, 0] ** 2).sum()
        loss.backward()
        res.append((loss.item(), p.grad.numpy().copy(), zt.grad.numpy().copy()))
    assert res[0][0] == res[1][0]
    assert np.array_equal(res[0][1], res[1][1]) and np.array_equal(res[0][2], res[1][2])
    rng = np.random.default_rng(4)
    dp, dz = rng.standard_normal(prim0.shape), rng.standard_normal(z.shape)
    ps.set_bathymetry(z)
    ps.SetVolField(prim0)
    fin_dot, _ = ps.tangent_run(dp, 10, z_dot=dz, scheme="SSPRK2", dt=2e-3)
    fin = ps.GetVolField()
    ybar = np.zeros_like(fin)
    ybar[:, 0] = 2 * fin[:, 0]
    lhs = float(np.sum(ybar * fin_dot))
    rhs = float(np.sum(res[1][1] * dp) + np.sum(res[1][2] * dz))
    assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + abs(rhs))


def test_adjoint_is_independent_of_the_tuning():
    """The recomputation inside the adjoint sweep launches with the solver's
    tuning (work items, workgroup sizes): every choice gives the same
    gradients bit for bit."""
    dom, prim0 = wet_case(n=14, seed=21)
    z = dom.geometry[:, 2].copy()
    rng = np.random.default_rng(2)
    cvec = rng.standard_normal(prim0.shape)
    res = []
    for tune in [(2, 1, 256, 256), (1, 2, 512, 512), (4, 1, 512, 256)]:
        sd = swe.SpaceDisc(swe.Fluxes.HLLC[swe.Wavespeeds.Davis], dom, prim0)
        sd.set_tuning(*tune)
        J, s0, zb, _ = AD.run_and_adjoint(sd, prim0, z, 8, lambda p: (float(np.sum(cvec * p)), cvec),
                                          scheme="SSPRK3", dt=2e-3, checkpoint_every=3)
        res.append((J, s0, zb))
    for J, s0, zb in res[1:]:
        assert J == res[0][0]
        assert np.array_equal(s0, res[0][1]) and np.array_equal(zb, res[0][2])


def test_gradient_descent_lowers_the_gauge_misfit():
    """The use the adjoint is for: steepest descent on the node bathymetry of
    a small bathymetry inversion.  Each backtracking step is accepted only
    under the Armijo condition J(z - s g) <= J(z) - 0.1 s |g|^2 (a tenth of
    the decrease the gradient predicts to first order), and the misfit falls
    at every iteration (steepest descent stalls soon on this ill-posed
    problem: 40 gauges, ~600 nodes)."""
    from swe_fvm_amd import workloads as W
    dom, prim, z_true, gauges = W.adjoint_inversion(n=24, seed=5, n_gauges=40)
    sd = swe.SpaceDisc(swe.Fluxes.HLLC[swe.Wavespeeds.Einfeldt], dom, prim)
    N, dt = 200, 2e-3
    sd.set_bathymetry(z_true)
    sd.SetVolField(prim)
    sd.run("SSPRK2", N, dt=dt)
    fn = AD.gauge_misfit(gauges, sd.GetVolField()[gauges, 0].copy())

    def misfit(zz):
        sd.set_bathymetry(zz)
        sd.SetVolField(prim)
        sd.run("SSPRK2", N, dt=dt)
        return fn(sd.GetVolField())[0]

    z = dom.geometry[:, 2].copy()
    J, _, zb, _ = AD.run_and_adjoint(sd, prim, z, N, fn, dt=dt, checkpoint_every=50)
    Js = [J]
    for it in range(3):
        g2 = float(np.sum(zb * zb))
        step = 0.1 / np.max(np.abs(zb))  # at most 0.1 of bathymetry per node
        while True:
            zn = z - step * zb
            Jn = misfit(zn)
            if Jn <= J - 0.1 * step * g2:
                break
            step *= 0.5
            assert step * np.max(np.abs(zb)) > 1e-7, f"no Armijo step at iteration {it}"
        z = zn
        J, _, zb, _ = AD.run_and_adjoint(sd, prim, z, N, fn, dt=dt, checkpoint_every=50)
        assert abs(J - Jn) <= 1e-12 * Jn  # the adjoint's forward is the run's
        Js.append(J)
        print(f"descent iteration {it}: step {step:.3e} misfit {J:.6e}", flush=True)
    assert all(b < a for a, b in zip(Js, Js[1:]))


def test_cell_bathymetry_gradient_matches_differences_of_the_oracle():
    """Per-cell bathymetry b with node heights z = M b (z_n the mean of b
    over the cells at node n): the adjoint's per-cell gradient M^T dJ/dz
    against directional central differences of the oracle in b, together
    with the initial free surface (eta = w) direction."""
    from oracle import oracle as O
    dom, prim0 = wet_case(n=6, seed=13)
    tables = dom.topology.tables()
    cells = AD.CellBathymetry(tables["tp"], dom.geometry.shape[0])
    xyz0 = dom.geometry.copy()
    b0 = xyz0[tables["tp"], 2].mean(1)
    fl = swe.Fluxes.HLLC[swe.Wavespeeds.Einfeldt]
    sd = swe.SpaceDisc(fl, dom, prim0)
    rng = np.random.default_rng(23)
    gauges = rng.choice(prim0.shape[0], 12, replace=False)
    fn = AD.gauge_misfit(gauges, prim0[gauges, 0] + 0.005)
    dt, nsteps = 2e-3, 6
    J, s0, bb, _ = AD.run_and_adjoint(sd, prim0, b0, nsteps, fn, scheme="SSPRK2", dt=dt,
                                      checkpoint_every=2, cell_bathymetry=cells)
    assert bb.shape == (prim0.shape[0],)

    def J_oracle(p, b):
        xyz = xyz0.copy()
        xyz[:, 2] = cells.to_nodes(b)
        o = O.OracleSolver(xyz, tables, "HLLC", "Einfeldt")
        o.set_state(p)
        o.run("SSPRK2", nsteps, dt=dt)
        return fn(o.state())[0]

    assert abs(J_oracle(prim0, b0) - J) <= 1e-13 * abs(J)
    checked = 0
    for trial in range(4):
        db = rng.standard_normal(b0.shape) * 1e-2
        dp = np.zeros_like(prim0)
        if trial % 2:
            dp[:, 0] = rng.standard_normal(prim0.shape[0]) * 1e-2  # initial eta
        cd = [(J_oracle(prim0 + e * dp, b0 + e * db) - J_oracle(prim0 - e * dp, b0 - e * db)) / (2 * e)
              for e in (1e-6, 2.5e-7)]
        if abs(cd[0] - cd[1]) > 1e-5 * abs(cd[0]) + 1e-12:
            print(f"cell bathymetry trial {trial}: the difference crossed a switch", flush=True)
            continue
        ad = float(np.sum(bb * db) + np.sum(s0 * dp))
        print(f"cell bathymetry trial {trial}: adjoint {ad:.12e} oracle fd {cd[0]:.12e} "
              f"rel {abs(ad - cd[0]) / abs(cd[0]):.2e}", flush=True)
        assert abs(ad - cd[0]) <= 1e-5 * abs(cd[0]) + 1e-12
        checked += 1
    assert checked >= 2


def test_autograd_function_cell_bathymetry_and_cfl_mode():
    """ShallowWaterRun with per-cell bathymetry equals run_and_adjoint with
    the same map, and a CFL-mode forward (dt <= 0) is differentiated with
    the step sizes it took."""
    torch = pytest.importorskip("torch")
    dom, prim0 = wet_case(n=12, seed=9)
    tables = dom.topology.tables()
    cells = AD.CellBathymetry(tables["tp"], dom.geometry.shape[0])
    b = dom.geometry[tables["tp"], 2].mean(1)
    sd = swe.SpaceDisc(swe.Fluxes.HLLC[swe.Wavespeeds.Einfeldt], dom, prim0)
    fn = lambda q: (float(np.sum(q[:, 0] ** 2)), np.stack([2 * q[:, 0], 0 * q[:, 0], 0 * q[:, 0]], 1))
    for dt in (2e-3, 0.0):
        p = torch.tensor(prim0, dtype=torch.float64, requires_grad=True)
        bt = torch.tensor(b, dtype=torch.float64, requires_grad=True)
        out = AD.ShallowWaterRun.apply(p, bt, sd, 10, "SSPRK2", dt, 4, cells)
        loss = (out[:, 0] ** 2).sum()
        loss.backward()
        dts = None
        if dt <= 0:
            sd.set_bathymetry(cells.to_nodes(b))
            sd.SetVolField(prim0)
            sd.run("SSPRK2", 10, dt=0.0)
            dts = sd.dts(10)
        J, s0, bb, _ = AD.run_and_adjoint(sd, prim0, b, 10, fn, dt=dt if dt > 0 else None, dts=dts,
                                          checkpoint_every=4, cell_bathymetry=cells)
        assert abs(loss.item() - J) <= 1e-12 * abs(J)
        assert np.allclose(p.grad.numpy(), s0, rtol=1e-12, atol=1e-15)
        assert np.allclose(bt.grad.numpy(), bb, rtol=1e-12, atol=1e-15)


def _inversion_batch(members, n, N, dt):
    """The stacked bathymetry-inversion batch: the domain, initial state,
    node bathymetry, gauge targets from one batched run over the members'
    true bathymetries, and the member offsets."""
    from swe_fvm_amd import workloads as W
    d = W.adjoint_inversion_batch(members=members, n=n)
    dom1, prim1 = d["member"]
    big, off = swe.stack_domains([dom1] * members)
    prim = np.tile(prim1, (members, 1))
    z0 = np.tile(d["z"], members)
    sd = swe.SpaceDisc(swe.Fluxes.HLLC[swe.Wavespeeds.Einfeldt], big, prim)
    sd.set_bathymetry(np.concatenate(d["z_true"]))
    sd.SetVolField(prim)
    sd.run("SSPRK2", N, dt=dt)
    fin = sd.GetVolField()
    targets = [(g + off["cells"][k], fin[g + off["cells"][k], 0].copy()) for k, g in enumerate(d["gauges"])]
    return d, big, off, sd, prim, z0, targets


def test_inversion_batch_full_size():
    """The bathymetry-inversion batch at full size: 250,632 cells per member,
    2 members stacked, 2,000 SSP-RK2 steps at dt = 2e-4, checkpoints every
    50.  Each member's slice of the batched final state and gradients equals
    that member run alone, bit for bit; the batched gradient satisfies the
    duality with the tangent-linear build over the full horizon to 1e-12 in
    directions that move both members."""
    from swe_fvm_amd import workloads as W
    B, N, dt = 2, 2000, 2e-4
    d, big, off, sd, prim, z0, targets = _inversion_batch(B, 354, N, dt)
    dom1, prim1 = d["member"]
    T1, V1 = prim1.shape[0], dom1.geometry.shape[0]
    assert T1 == 250_632
    fn = AD.batch_gauge_misfit(targets)
    J, s0, zb, ms = AD.run_and_adjoint(sd, prim, z0, N, fn, dt=dt, checkpoint_every=50)
    fin = sd.GetVolField()
    print(f"inversion batch {B} x {T1} cells, {N} steps: J {J:.6e}, sweeps {ms[0]:.0f} + {ms[1]:.0f} ms",
          flush=True)
    assert J > 0 and np.all(np.isfinite(s0)) and np.all(np.isfinite(zb))
    sd1 = swe.SpaceDisc(swe.Fluxes.HLLC[swe.Wavespeeds.Einfeldt], dom1, prim1)
    for k in range(B):
        cs, vs = slice(k * T1, (k + 1) * T1), slice(k * V1, (k + 1) * V1)
        f1 = AD.gauge_misfit(d["gauges"][k], targets[k][1])
        J1, a0, az, _ = AD.run_and_adjoint(sd1, prim1, d["z"], N, f1, dt=dt, checkpoint_every=50)
        same = (np.array_equal(sd1.GetVolField(), fin[cs]), np.array_equal(a0, s0[cs]),
                np.array_equal(az, zb[vs]))
        print(f"member {k}: J {J1:.6e} final/state-gradient/bathymetry-gradient bitwise {same}", flush=True)
        assert all(same)
        assert J1 == fn.parts[k]
    rng = np.random.default_rng(8)
    c = big.centroids()
    ybar = fn(fin)[1]
    taper = W.wall_taper(big.geometry[:, 0]) * W.wall_taper(big.geometry[:, 1])
    for trial in range(2):
        pdot = np.zeros_like(prim)
        pdot[:, 0] = W.smooth_field(c[:, 0], c[:, 1], rng, 1e-2)
        zdot = W.smooth_field(big.geometry[:, 0], big.geometry[:, 1], rng, 1e-2) * taper
        sd.set_bathymetry(z0)
        sd.SetVolField(prim)
        tdot, _ = sd.tangent_run(pdot, N, z_dot=zdot, dt=dt)
        lhs = float(np.sum(ybar * tdot))
        rhs = float(np.sum(s0 * pdot) + np.sum(zb * zdot))
        scale = float(np.sum(np.abs(ybar * tdot)))
        print(f"full-size duality {trial}: {lhs:.15e} {rhs:.15e} diff/scale {abs(lhs - rhs) / scale:.1e}",
              flush=True)
        assert abs(lhs - rhs) <= 1e-12 * scale


def test_inversion_batch_against_differences_of_the_oracle():
    """The inversion batch on a small mesh (2 members of 128 cells, 300
    steps, checkpoints every 50): the batched loss is the oracle's, and the
    GPU gradient matches central differences of the oracle's forward in
    smooth directions of the initial surface and the node bathymetry."""
    from oracle import oracle as O
    from swe_fvm_amd import workloads as W
    B, N, dt = 2, 300, 2e-3
    d, big, off, sd, prim, z0, targets = _inversion_batch(B, 8, N, dt)
    fn = AD.batch_gauge_misfit(targets)
    J, s0, zb, _ = AD.run_and_adjoint(sd, prim, z0, N, fn, dt=dt, checkpoint_every=50)
    xyz0 = big.geometry.copy()
    tables = big.topology.tables()

    def J_oracle(p, zz):
        xyz = xyz0.copy()
        xyz[:, 2] = zz
        o = O.OracleSolver(xyz, tables, "HLLC", "Einfeldt")
        o.set_state(p)
        o.run("SSPRK2", N, dt=dt)
        return fn(o.state())[0]

    Jo = J_oracle(prim, z0)
    print(f"inversion loss: GPU {J:.17e} oracle {Jo:.17e}", flush=True)
    assert abs(Jo - J) <= 1e-13 * abs(J)  # the forward is the oracle's
    rng = np.random.default_rng(21)
    c = big.centroids()
    taper = W.wall_taper(xyz0[:, 0]) * W.wall_taper(xyz0[:, 1])
    checked = 0
    for trial in range(6):
        dp = np.zeros_like(prim)
        dp[:, 0] = W.smooth_field(c[:, 0], c[:, 1], rng, 1e-2)
        dz = W.smooth_field(xyz0[:, 0], xyz0[:, 1], rng, 1e-2) * taper
        cd = [(J_oracle(prim + e * dp, z0 + e * dz) - J_oracle(prim - e * dp, z0 - e * dz)) / (2 * e)
              for e in (1e-6, 2.5e-7)]
        if abs(cd[0] - cd[1]) > 1e-5 * abs(cd[0]) + 1e-14:
            print(f"inversion trial {trial}: the difference crossed a switch", flush=True)
            continue
        ad = float(np.sum(s0 * dp) + np.sum(zb * dz))
        print(f"inversion trial {trial}: adjoint {ad:.12e} oracle fd {cd[0]:.12e} "
              f"rel {abs(ad - cd[0]) / abs(cd[0]):.2e}", flush=True)
        assert abs(ad - cd[0]) <= 1e-5 * abs(cd[0]) + 1e-14
        checked += 1
    assert checked >= 3


@pytest.mark.parametrize("scheme,n", [("Euler", 24), ("SSPRK2", 24), ("SSPRK3", 24), ("SSPRK2", 730)])
def test_adjoint_same_with_fused_recomputation(scheme, n):
    """The reverse sweep's recomputation runs the fused kernel storing each
    stage's fluxes and draining steps (one writer per edge and cell; below
    2^20 cells), the separate flux kernel and then the tile kernel storing
    the draining steps (above: n = 730 gives 1,065,800 cells), or the
    separate flux / draining / update kernels: the gradients are the same
    bit for bit, on a wet/dry mesh whose draining steps are active."""
    xyz, topo, prim = C.wet_dry_mesh(n, seed=5)
    dom = swe.Domain(xyz, topo)
    z = xyz[:, 2].copy()
    rng = np.random.default_rng(4)
    cvec = rng.standard_normal(prim.shape)
    fn = lambda p: (float(np.sum(cvec * p)), cvec)
    out = []
    for fused in (True, False):
        sd = swe.SpaceDisc(swe.Fluxes.HLLC[swe.Wavespeeds.Einfeldt], dom, prim, cor=0.2)
        assert sd.set_fused(fused) == fused
        # dt 2e-3 at n = 24, scaled with the cell size
        J, s0, zb, _ = AD.run_and_adjoint(sd, prim, z, 7, fn, scheme=scheme, dt=0.048 / n, checkpoint_every=3)
        out.append((J, s0, zb, sd.GetVolField()))
    (Ja, sa, za, fa), (Jb, sb, zb_, fb) = out
    assert np.all(np.isfinite(sa)) and np.all(np.isfinite(za))
    print(f"{scheme} T={prim.shape[0]}: fused recomputation bitwise={np.array_equal(sa, sb) and np.array_equal(za, zb_)}",
          flush=True)
    assert Ja == Jb and np.array_equal(fa, fb)
    assert np.array_equal(sa, sb) and np.array_equal(za, zb_)


def test_adjoint_timing_splits_the_reverse_sweep():
    """swe_adjoint_timing: the forward sweep as adjoint_run reports it, and
    the reverse sweep split into the segments' recomputation and the
    reverse stages, both positive, adding up to the reverse sweep."""
    xyz, topo, prim = C.wet_dry_mesh(20, seed=6)
    sd = swe.SpaceDisc(swe.Fluxes.HLLC[swe.Wavespeeds.Einfeldt], swe.Domain(xyz, topo), prim)
    assert sd.adjoint_timing() == {"forward_sweep_ms": 0.0, "recompute_ms": 0.0, "reverse_stages_ms": 0.0}
    bar = np.random.default_rng(2).standard_normal(prim.shape)
    _, _, ms = sd.adjoint_run(bar, 9, dt=1e-3, checkpoint_every=4)
    t = sd.adjoint_timing()
    print("adjoint timing:", ms, t, flush=True)
    assert t["forward_sweep_ms"] == ms[0]
    assert t["recompute_ms"] > 0 and t["reverse_stages_ms"] > 0
    assert abs(t["recompute_ms"] + t["reverse_stages_ms"] - ms[1]) <= 1e-9 * ms[1]
