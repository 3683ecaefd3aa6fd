"""The cell-state layout (swe_set_layout): SoA, AoSoA blocks of 64, 8, 2 and
AoS give bit-identical results on the single-mesh path (separate and fused
stages, every scheme, CFL and fixed steps), in ensembles (Euler's two-buffer
window copies included, members finishing at different times), across a
change of layout in the middle of a run and through the getters; the
autotuner's layout dimension keeps one of its candidates; the entry points
that need SoA refuse the others."""
import numpy as np
import pytest

import swe_fvm_amd as swe
from oracle import oracle as O
from swe_fvm_amd import workloads as W
from swe_fvm_amd.ensemble import Ensemble
from tests import _common as C

pytestmark = pytest.mark.gpu

FL = swe.Fluxes.HLLC[swe.Wavespeeds.Einfeldt]
BLOCKS = [64, 8, 2, 1]


def _case():
    xyz, topo, prim = C.wet_dry_mesh(30, seed=21)  # 1,800 cells: not a multiple of 64
    return swe.Domain(xyz, topo), prim


@pytest.mark.parametrize("scheme", ["Euler", "SSPRK2", "SSPRK3"])
def test_layouts_match_soa_and_the_oracle(scheme):
    dom, prim = _case()
    assert prim.shape[0] % 64 != 0
    ref = swe.SpaceDisc(FL, dom, prim, cor=0.1)
    ref.run(scheme, 5, dt=0.0, cfl=0.15)
    ref.run(scheme, 3, dt=1e-3)  # fixed steps: every stage fused
    want = ref.GetVolField()
    orc = O.OracleSolver(dom.geometry, dom.topology.tables(), "HLLC", "Einfeldt", cor=0.1)
    orc.set_state(prim)
    orc.run(scheme, 5, dt=0.0, cfl=0.15)
    orc.run(scheme, 3, dt=1e-3)
    assert np.array_equal(want, orc.state())
    for b in BLOCKS:
        sd = swe.SpaceDisc(FL, dom, prim, cor=0.1)
        sd.set_layout(b)
        assert sd.layout() == b
        assert np.array_equal(sd.GetVolField(), prim)  # the round trip through the layout
        sd.run(scheme, 5, dt=0.0, cfl=0.15)
        sd.run(scheme, 3, dt=1e-3)
        got = sd.GetVolField()
        print(f"{scheme} layout {b}: bitwise={np.array_equal(got, want)}", flush=True)
        assert np.array_equal(got, want)
        assert np.array_equal(sd.dts(3), ref.dts(3))


def test_layout_change_mid_run_and_getters():
    dom, prim = _case()
    ref = swe.SpaceDisc(FL, dom, prim)
    ref.run("SSPRK2", 6, dt=0.0, cfl=0.15)
    sd = swe.SpaceDisc(FL, dom, prim)
    sd.run("SSPRK2", 2, dt=0.0, cfl=0.15)
    sd.set_layout(8)
    sd.run("SSPRK2", 2, dt=0.0, cfl=0.15)
    sd.set_layout(1)
    sd.run("SSPRK2", 1, dt=0.0, cfl=0.15)
    sd.set_layout(0)
    sd.run("SSPRK2", 1, dt=0.0, cfl=0.15)
    assert np.array_equal(sd.GetVolField(), ref.GetVolField())
    # getters on a blocked layout equal those of the SoA solver
    sd.set_layout(64)
    for g in (sd, ref):
        g.ComputeInterfaceValues()
        g.ComputeFluxes()
    assert np.array_equal(sd.GetEdgField(), ref.GetEdgField())
    assert np.array_equal(sd.GetFluxes(), ref.GetFluxes())
    assert sd.GetMinLenToWavespeed() == ref.GetMinLenToWavespeed()
    assert np.array_equal(sd.Integrals(), ref.Integrals())
    td, tr = swe.TimeDisc(sd), swe.TimeDisc(ref)
    for i in (0, 17, prim.shape[0] - 1):
        assert np.array_equal(td.RHS(i, 1e-3), tr.RHS(i, 1e-3))


def test_soa_only_entry_points_refuse_other_layouts():
    dom, prim = _case()
    sd = swe.SpaceDisc(FL, dom, prim)
    sd.set_layout(8)
    with pytest.raises(ValueError):
        sd.adjoint_run(np.ones_like(prim), 2, dt=1e-3)
    with pytest.raises(ValueError):
        sd.tangent_run(np.ones_like(prim), 2, dt=1e-3)
    with pytest.raises(ValueError):
        sd.set_layout(3)
    sd.set_layout(0)
    s0, zb, _ = sd.adjoint_run(np.ones_like(prim), 2, dt=1e-3)
    assert np.all(np.isfinite(s0))


@pytest.mark.parametrize("scheme", ["Euler", "SSPRK2"])
def test_ensemble_layouts_match_soa(scheme):
    """Members of uneven size run to t_end (they finish at different times,
    so the active window shrinks and, for Euler, finished members are copied
    block by block into the other buffer)."""
    mem = [W.ensemble_member(n, 300 + k) for k, n in enumerate([6, 17, 9, 30, 12, 23, 4])]
    doms, prims = [d for d, _ in mem], [p for _, p in mem]
    ref = Ensemble(FL, doms, prims)
    ref.run(scheme, cfl=0.15, t_end=0.01, graph=True)
    want = ref.states()
    t_ref, st_ref = ref.times()
    assert len(set(st_ref.tolist())) > 1  # members took different numbers of steps
    for b in (64, 8, 1):
        ens = Ensemble(FL, doms, prims)
        ens.set_layout(b)
        ens.run(scheme, cfl=0.15, t_end=0.01, graph=True)
        got = ens.states()
        t, st = ens.times()
        assert np.array_equal(st, st_ref) and np.array_equal(t, t_ref)
        for k in range(len(mem)):
            assert np.array_equal(got[k], want[k]), (b, k)


def test_autotune_layout_keeps_a_candidate_and_the_results():
    dom, prim = _case()
    ref = swe.SpaceDisc(FL, dom, prim)
    ref.run("SSPRK2", 4, dt=0.0, cfl=0.15)
    sd = swe.SpaceDisc(FL, dom, prim)
    res = sd.autotune_layout(reps=2)
    print("autotune_layout:", res, flush=True)
    assert res["layout"] in swe.core.LAYOUTS and sd.layout() == res["layout"]
    assert set(res["stage_ms"]) == {"soa", "aosoa64", "aosoa8", "aosoa1"}
    assert all(v > 0 for v in res["stage_ms"].values())
    assert np.array_equal(sd.GetVolField(), prim)  # tuning leaves the state as it was
    sd.run("SSPRK2", 4, dt=0.0, cfl=0.15)
    assert np.array_equal(sd.GetVolField(), ref.GetVolField())
    mem = [W.ensemble_member(n, 500 + k) for k, n in enumerate([8, 14, 5])]
    ens = Ensemble(FL, [d for d, _ in mem], [p for _, p in mem])
    r2 = ens.autotune_layout(reps=2)
    assert r2["layout"] in swe.core.LAYOUTS and ens.layout() == r2["layout"]
