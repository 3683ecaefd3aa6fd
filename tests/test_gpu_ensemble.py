"""Ensembles: every member of a batched run is bit-identical to advancing it
alone (plain GPU solver, and the oracle for the clipped last step)."""
import numpy as np
import pytest

import swe_fvm_amd as swe
from swe_fvm_amd.ensemble import Ensemble
from swe_fvm_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FL = swe.Fluxes.HLLC[swe.Wavespeeds.Einfeldt]


def members():
    return [W.ensemble_member(n, 100 + k) for k, n in enumerate([6, 17, 9, 30, 12, 23, 4])]


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("scheme", ["SSPRK2", "Euler"])
def test_batched_members_equal_single_runs(scheme, graph):
    mem = members()
    ens = Ensemble(FL, [d for d, _ in mem], [p for _, p in mem])
    nsteps = 8
    done, act = ens.run(scheme, cfl=0.15, t_end=1e9, max_steps=nsteps, graph=graph)
    assert done == nsteps and act == len(mem)
    got = ens.states()
    t, steps = ens.times()
    assert np.all(steps == nsteps)
    for k, (dom, prim) in enumerate(mem):
        sd = swe.SpaceDisc(FL, dom, prim)
        sd.run(scheme, nsteps, dt=0.0, cfl=0.15)
        assert np.array_equal(got[k], sd.GetVolField()), k
        assert t[k] == np.sum(sd.dts(nsteps)) or abs(t[k] - np.sum(sd.dts(nsteps))) <= 1e-15 * t[k]


def test_members_stop_exactly_at_t_end():
    mem = members()
    ens = Ensemble(FL, [d for d, _ in mem], [p for _, p in mem])
    t_end = 0.02
    done, act = ens.run("SSPRK2", cfl=0.15, t_end=t_end, graph=True)
    assert act == 0
    t, steps = ens.times()
    assert np.all(t == t_end)
    got = ens.states()
    # replay member by member on the oracle: CFL steps, the last one clipped
    for k, (dom, prim) in enumerate(mem):
        orc = O.OracleSolver(dom.geometry, dom.topology.tables(), "HLLC", "Einfeldt")
        orc.set_state(prim)
        tk = 0.0
        for _ in range(int(steps[k])):
            orc.compute_interface_values()
            r = orc.compute_fluxes()
            d = 0.15 * r
            rest = t_end - tk
            if d < rest:
                orc.run("SSPRK2", 1, dt=0.0, cfl=0.15)
                tk = tk + d
            else:
                orc.run("SSPRK2", 1, dt=rest)
                tk = t_end
        assert tk == t_end
        assert np.array_equal(got[k], orc.state()), k
