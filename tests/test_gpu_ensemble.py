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


def _oracle_replay(dom, prim, scheme, t_end, nsteps):
    """A member advanced alone on the oracle: CFL steps, the last one
    clipped to land on t_end (k_batch_dt)."""
    orc = O.OracleSolver(dom.geometry, dom.topology.tables(), "HLLC", "Einfeldt")
    orc.set_state(prim)
    tk = 0.0
    for _ in range(nsteps):
        orc.compute_interface_values()
        d = 0.15 * orc.compute_fluxes()
        rest = t_end - tk
        if d < rest:
            orc.run(scheme, 1, dt=0.0, cfl=0.15)
            tk = tk + d
        else:
            orc.run(scheme, 1, dt=rest)
            tk = t_end
    return orc.state(), tk


@pytest.mark.parametrize("graph", [False, True])
def test_euler_members_finished_in_odd_chunks(graph):
    """Euler swaps the two state buffers every step; members that finish
    and leave the launch window must keep their final state whatever the
    parity of the steps taken after (odd max_steps per call)."""
    mem = members()
    ens = Ensemble(FL, [d for d, _ in mem], [p for _, p in mem])
    t_end = 0.015
    calls = 0
    while True:
        done, act = ens.run("Euler", cfl=0.15, t_end=t_end, max_steps=3, graph=graph)
        calls += 1
        if act == 0:
            break
        assert calls < 500
    t, steps = ens.times()
    assert np.all(t == t_end)
    assert len(set(steps.tolist())) > 1  # the members finish at different steps
    got = ens.states()
    for k, (dom, prim) in enumerate(mem):
        ref, tk = _oracle_replay(dom, prim, "Euler", t_end, int(steps[k]))
        assert tk == t_end
        assert np.array_equal(got[k], ref), k
    # further calls change nothing
    before = [g.copy() for g in got]
    ens.run("Euler", cfl=0.15, t_end=t_end, max_steps=1, graph=graph)
    assert all(np.array_equal(a, b) for a, b in zip(before, ens.states()))


def test_euler_graph_and_plain_runs_mixed():
    """An Euler graph captured on one state buffer is re-captured after an
    odd number of steps taken without it."""
    mem = members()
    ens = Ensemble(FL, [d for d, _ in mem], [p for _, p in mem])
    plan = [(True, 4), (False, 3), (True, 4), (True, 1), (False, 2), (True, 5)]
    for graph, n in plan:
        done, act = ens.run("Euler", cfl=0.15, t_end=1e9, max_steps=n, graph=graph)
        assert done == n and act == len(mem)
    nsteps = sum(n for _, n in plan)
    got = ens.states()
    for k, (dom, prim) in enumerate(mem):
        sd = swe.SpaceDisc(FL, dom, prim)
        sd.run("Euler", nsteps, dt=0.0, cfl=0.15)
        assert np.array_equal(got[k], sd.GetVolField()), k


def test_full_size_ensemble_members_equal_single_runs():
    """The benchmark's 256-member ensemble at full size (10k-340k cells per
    member, 19,999,564 cells in all), 8 graph-captured SSP-RK2 steps with
    per-member CFL step sizes: the smallest, a middle and the largest member
    equal their own single-mesh runs bit for bit, and so do their times."""
    import gc

    print("[full size] building the 256 members", flush=True)
    mem = W.ensemble(n_inst=256, total=20_000_000, seed=11)
    assert sum(p.shape[0] for _, p in mem) == 19_999_564
    ens = Ensemble(FL, [d for d, _ in mem], [p for _, p in mem])
    nsteps = 8
    done, act = ens.run("SSPRK2", cfl=0.15, t_end=1e9, max_steps=nsteps, graph=True)
    assert done == nsteps and act == len(mem)
    got = ens.states()
    t, steps = ens.times()
    assert np.all(steps == nsteps)
    del ens
    gc.collect()
    print("[full size] batched run done", flush=True)
    for k in (0, len(mem) // 2, len(mem) - 1):
        dom, prim = mem[k]
        sd = swe.SpaceDisc(FL, dom, prim)
        sd.run("SSPRK2", nsteps, dt=0.0, cfl=0.15)
        assert np.array_equal(got[k], sd.GetVolField()), k
        assert t[k] == np.sum(sd.dts(nsteps)) or abs(t[k] - np.sum(sd.dts(nsteps))) <= 1e-15 * t[k]
        del sd
