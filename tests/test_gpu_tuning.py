"""Work items per thread and block sizes (swe_set_tuning,
swe_set_tuning_blocks, swe_autotune): every choice gives bit-identical results, the autotuner leaves the state alone, and ensemble
handles are refused by the single-mesh entry points.  The cell layout:
morton_order of a shuffled mesh gives the same run, faster."""
import itertools

import numpy as np
import pytest

import swe_fvm_amd as swe
from swe_fvm_amd import workloads as W
from swe_fvm_amd.ensemble import Ensemble
from tests import _common as C

pytestmark = pytest.mark.gpu

FL = swe.Fluxes.HLLC[swe.Wavespeeds.Einfeldt]


def test_every_tuning_gives_identical_steps():
    xyz, topo, prim = C.wet_dry_mesh(40, seed=9)
    dom = swe.Domain(xyz, topo)
    ref = None
    for fi, ui, fb, ub in itertools.product((1, 2, 4), (1, 2), (256, 512), (256, 512)):
        sd = swe.SpaceDisc(FL, dom, prim)
        sd.set_tuning(fi, ui, fb, ub)
        assert sd.tuning() == (fi, ui, fb, ub)
        sd.run("SSPRK2", 6, dt=0.0, cfl=0.15)
        got = (sd.GetVolField(), sd.dts(6), sd.GetFluxes())
        print(f"tuning flux x{fi}/{fb} update x{ui}/{ub}", flush=True)
        if ref is None:
            ref = got
        else:
            for a, b in zip(got, ref):
                assert np.array_equal(a, b), (fi, ui, fb, ub)


def test_bad_tuning_rejected():
    xyz, topo, prim = C.wet_dry_mesh(8, seed=1)
    sd = swe.SpaceDisc(FL, swe.Domain(xyz, topo), prim)
    for fi, ui in ((3, 1), (0, 1), (2, 3), (8, 2)):
        with pytest.raises(ValueError):
            sd.set_tuning(fi, ui)
    for fb, ub in ((128, 256), (256, 1024), (0, 512)):
        with pytest.raises(ValueError):
            sd.set_tuning(2, 1, fb, ub)
    assert sd.tuning() == (2, 1, 256, 256)


def test_autotune_keeps_state_and_results():
    xyz, topo, prim = C.wet_dry_mesh(60, seed=4)
    dom = swe.Domain(xyz, topo)
    base = swe.SpaceDisc(FL, dom, prim)
    base.run("SSPRK2", 3, dt=0.0, cfl=0.15)
    sd = swe.SpaceDisc(FL, dom, prim)
    res = sd.autotune(reps=3)
    print("autotune:", res, flush=True)
    assert res["flux_items"] in (1, 2, 4) and res["update_items"] in (1, 2)
    assert res["flux_block"] in (256, 512) and res["update_block"] in (256, 512)
    assert sd.tuning() == (res["flux_items"], res["update_items"], res["flux_block"], res["update_block"])
    assert all(t > 0 for t in res["flux_ms"].values()) and all(t > 0 for t in res["update_ms"].values())
    assert np.array_equal(sd.GetVolField(), prim)
    sd.run("SSPRK2", 3, dt=0.0, cfl=0.15)
    assert np.array_equal(sd.GetVolField(), base.GetVolField())


def test_ensemble_retune_recaptures_graph():
    mem = [W.ensemble_member(n, 300 + k) for k, n in enumerate([7, 15, 11])]
    doms, prims = [d for d, _ in mem], [p for _, p in mem]
    ref = Ensemble(FL, doms, prims)
    ref.run("SSPRK2", cfl=0.15, t_end=1e9, max_steps=8, graph=True)
    ens = Ensemble(FL, doms, prims)
    ens.run("SSPRK2", cfl=0.15, t_end=1e9, max_steps=4, graph=True)
    ens.set_tuning(4, 2, 512, 512)
    ens.run("SSPRK2", cfl=0.15, t_end=1e9, max_steps=4, graph=True)
    for a, b in zip(ens.states(), ref.states()):
        assert np.array_equal(a, b)
    assert np.array_equal(ens.times()[0], ref.times()[0])


def test_single_mesh_calls_refuse_ensembles():
    mem = [W.ensemble_member(n, 400 + k) for k, n in enumerate([5, 6])]
    ens = Ensemble(FL, [d for d, _ in mem], [p for _, p in mem])
    L = swe._lib.load()
    assert L.swe_run(ens._h, 1, 0.0, 0.15, 1) == swe.core.SWE_ERR_ARG
    assert L.swe_compute_fluxes(ens._h, None) == swe.core.SWE_ERR_ARG
    assert L.swe_time_run(ens._h, 1, 0.0, 0.15, 1, None, None) == swe.core.SWE_ERR_ARG


def test_morton_order_speeds_up_a_shuffled_mesh():
    """A 320k-cell mesh with its cells in random order, and the same mesh
    after morton_order: the same run (HLL<Davis>, fixed dt) agrees cell by
    cell, and the Morton layout is faster (neighbour gathers stay in the
    workgroup's L2 lines and LDS tile)."""
    xyz, topo = swe.jittered_mesh(400, 0.25, 2)
    dom = swe.Domain(xyz, topo)
    T = topo.NumTriangles()
    rng = np.random.default_rng(1)
    shuf, _, _ = swe.reorder_cells(dom, rng.permutation(T))
    mort, p2, _ = swe.morton_order(shuf)
    c = shuf.centroids()
    prim = np.stack([1.0 + 0.2 * np.exp(-((c[:, 0] - 0.5) ** 2 + (c[:, 1] - 0.5) ** 2) / 0.01),
                     np.zeros(T), np.zeros(T)], 1)
    fl = swe.Fluxes.HLL[swe.Wavespeeds.Davis]
    sds = swe.SpaceDisc(fl, shuf, prim)
    sdm = swe.SpaceDisc(fl, mort, prim[p2])
    ms = []
    for sd, p0 in ((sds, prim), (sdm, prim[p2])):
        sd.run("SSPRK2", 3, dt=2e-4)
        sd.SetVolField(p0)
        ms.append(sd.time_run("SSPRK2", 50, dt=2e-4))
    d = np.max(np.abs(sdm.GetVolField() - sds.GetVolField()[p2]))
    print(f"{T} cells, 50 steps: shuffled {ms[0]:.1f} ms, Morton {ms[1]:.1f} ms ({ms[0] / ms[1]:.2f}x), "
          f"max diff {d:.1e}", flush=True)
    assert d <= 1e-13
    assert ms[1] < ms[0] / 1.2
